// rdm.cpp -- reduced density matrices from a primal vector (see rdm.hpp).
#include "rdm.hpp"

#include <stdexcept>

namespace v2h {

namespace {

// D2ab(pq, rs), p q r s Pitzer indices with equal pair irreps
double d2ab(const Layout& L, const double* x, int h, int pq, int rs) {
    return x[L.off[D2AB][h] + (int64_t)pq * L.gab[h] + rs];
}

}  // namespace

void opdm_dense(const Layout& L, const double* x, const std::string& type, double* out) {
    if (type != "A" && type != "B" && type != "SUM") throw std::invalid_argument("invalid opdm type: " + type);
    const int n = L.n;
    for (int e = 0; e < n * n; e++) out[e] = 0.0;
    for (int p = 0; p < n; p++)
        for (int q = 0; q < n; q++) {
            if (L.sym[p] != L.sym[q]) continue;
            const double a = x[L.one_body(D1A, p, q)];
            const double b = x[L.one_body(D1B, p, q)];
            out[L.to_user[p] * n + L.to_user[q]] = type == "A" ? a : (type == "B" ? b : a + b);
        }
}

std::vector<OpdmEntry> opdm_sparse(const Layout& L, const double* x, const std::string& type) {
    if (type != "A" && type != "B" && type != "SUM") throw std::invalid_argument("invalid opdm type: " + type);
    std::vector<OpdmEntry> out;
    for (int h = 0; h < L.H; h++)
        for (int i = 0; i < L.amopi[h]; i++)
            for (int j = 0; j < L.amopi[h]; j++) {
                const int p = L.poff[h] + i, q = L.poff[h] + j;
                const double a = x[L.one_body(D1A, p, q)], b = x[L.one_body(D1B, p, q)];
                out.push_back({L.to_user[p], L.to_user[q], type == "A" ? a : (type == "B" ? b : a + b)});
            }
    return out;
}

std::vector<TpdmEntry> tpdm_sparse(const Layout& L, const double* x, const std::string& type) {
    if (type != "AA" && type != "BB" && type != "AB" && type != "BA" && type != "SUM")
        throw std::invalid_argument("invalid tpdm type: " + type);
    std::vector<TpdmEntry> out;
    for (int h = 0; h < L.H; h++) {
        const int g = L.gab[h];
        for (int ij = 0; ij < g; ij++) {
            const int i = L.ab_pairs[h][ij][0], j = L.ab_pairs[h][ij][1];
            const int ji = L.ab(h, j, i);
            for (int kl = 0; kl < g; kl++) {
                const int k = L.ab_pairs[h][kl][0], l = L.ab_pairs[h][kl][1];
                const int lk = L.ab(h, l, k);
                const double vab = d2ab(L, x, h, ij, kl);
                const double vba = d2ab(L, x, h, ji, lk);
                double vaa = 0.0, vbb = 0.0;
                if (i != j && k != l) {
                    const int64_t e = (int64_t)L.aa(h, i, j) * L.gaa[h] + L.aa(h, k, l);
                    const double s = ((i < j) ? 1.0 : -1.0) * ((k < l) ? 1.0 : -1.0);
                    vaa = s * x[L.off[D2AA][h] + e];
                    vbb = s * x[L.off[D2BB][h] + e];
                }
                double v;
                if (type == "AA") v = vaa;
                else if (type == "BB") v = vbb;
                else if (type == "AB") v = vab;
                else if (type == "BA") v = vba;
                else v = vaa + vbb + vab + vba;
                out.push_back({L.to_user[i], L.to_user[j], L.to_user[k], L.to_user[l], v});
            }
        }
    }
    return out;
}

void tpdm_dense(const Layout& L, const double* x, double* out) {
    const int n = L.n;
    const int64_t n2 = (int64_t)n * n;
    for (int64_t e = 0; e < n2 * n2; e++) out[e] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const int h = L.pair_irrep(i, j);
            const int ij = L.ab(h, i, j), ji = L.ab(h, j, i);
            for (int k = 0; k < n; k++)
                for (int l = 0; l < n; l++) {
                    if (L.pair_irrep(k, l) != h) continue;
                    const int kl = L.ab(h, k, l), lk = L.ab(h, l, k);
                    double v = d2ab(L, x, h, ij, kl) + d2ab(L, x, h, ji, lk);
                    if (i != j && k != l) {
                        const int64_t e = (int64_t)L.aa(h, i, j) * L.gaa[h] + L.aa(h, k, l);
                        const double s = ((i > j) ? -1.0 : 1.0) * ((k > l) ? -1.0 : 1.0);
                        v += s * (x[L.off[D2AA][h] + e] + x[L.off[D2BB][h] + e]);
                    }
                    const int64_t ui = L.to_user[i], uj = L.to_user[j], uk = L.to_user[k], ul = L.to_user[l];
                    out[(ui * n + uk) * n2 + uj * n + ul] = v;
                }
        }
}

}  // namespace v2h
