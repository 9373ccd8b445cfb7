// constraints.cpp -- DQG + spin rows of A as (column, coefficient) lists.
//
// Row families and their order follow hilbert's bpsdp_Au
// (v2rdm_solver.cc:2995-3050): D2 (traces, hermiticity, D1/Q1, contractions;
// d2.cc:314-594), spin (spin.cc:276-508), Q2 (q2.cc:170-317), G2
// (g2.cc:245-441), then T1, T2 and D3 when asked for (t1.cc:419-793, the
// slow path t2.cc:1236-1843, d3.cc:45-406; dispatched at
// v2rdm_solver.cc:3015-3044).  Each row is written here as the linear
// functional it imposes on the primal vector, with its b value.
#include "constraints.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace v2h {

namespace {

class RowWriter {
  public:
    explicit RowWriter(Csr& out) : out_(out) { out_.ptr.assign(1, 0); }
    void add(int64_t col, double v) { cur_.push_back({col, v}); }
    void end(double b) {
        std::sort(cur_.begin(), cur_.end(),
                  [](const Term& a, const Term& c) { return a.col < c.col; });
        size_t k = 0;
        while (k < cur_.size()) {
            int64_t c = cur_[k].col;
            double v = 0.0;
            for (; k < cur_.size() && cur_[k].col == c; k++) v += cur_[k].v;
            if (v != 0.0) {
                out_.col.push_back((int32_t)c);
                out_.val.push_back(v);
            }
        }
        cur_.clear();
        out_.rhs.push_back(b);
        out_.ptr.push_back((int64_t)out_.col.size());
    }

  private:
    struct Term { int64_t col; double v; };
    Csr& out_;
    std::vector<Term> cur_;
};

inline double sgn(int a, int b) { return a < b ? 1.0 : -1.0; }

}  // namespace

Csr build_constraints(const Layout& L) {
    Csr A;
    A.ncols = L.n_primal;
    RowWriter R(A);
    const int n = L.n, H = L.H;
    const double na = L.na, nb = L.nb;
    const auto& d = L.desc;

    // element helpers (orbital indices in Pitzer order)
    auto d2ab = [&](int i, int j, int k, int l) -> int64_t {
        int h = L.pair_irrep(i, j);
        int a = L.ab(h, i, j), b = L.ab(h, k, l);
        if (a < 0 || b < 0) throw std::logic_error("D2ab element outside its irrep block");
        return L.off[D2AB][h] + (int64_t)a * L.gab[h] + b;
    };
    auto gem = [&](int f, int h, int a, int b, int dim) -> int64_t {
        return L.off[f][h] + (int64_t)a * dim + b;
    };
    // D2aa / D2bb element for ordered pairs (i,j),(k,l): index and sign
    auto d2same = [&](int f, int i, int j, int k, int l, double& s) -> int64_t {
        int h = L.pair_irrep(i, j);
        s = sgn(i, j) * sgn(k, l);
        return L.off[f][h] + (int64_t)L.aa(h, i, j) * L.gaa[h] + L.aa(h, k, l);
    };

    // ---- traces
    auto trace_ab = [&](double w) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) R.add(d2ab(i, j, i, j), w);
    };
    auto trace_same = [&](int f) {
        double s;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                if (i != j) R.add(d2same(f, i, j, i, j, s), 1.0);
    };
    if (d.sz) {
        trace_ab(1.0); R.end(na * nb);
        trace_same(D2AA); R.end(na * (na - 1.0));
        trace_same(D2BB); R.end(nb * (nb - 1.0));
    } else {
        trace_ab(2.0); trace_same(D2AA); trace_same(D2BB);
        R.end((na + nb) * (na + nb - 1.0));
    }

    // ---- hermiticity: X(ij,kl) - X(kl,ij) = 0 for D2aa, D2bb, D2ab
    for (int f : {D2AA, D2BB, D2AB})
        for (int h = 0; h < H; h++) {
            int g = f == D2AB ? L.gab[h] : L.gaa[h];
            for (int a = 0; a < g; a++)
                for (int b = 0; b < g; b++) {
                    R.add(gem(f, h, a, b, g), 1.0);
                    R.add(gem(f, h, b, a, g), -1.0);
                    R.end(0.0);
                }
        }

    // ---- D1(j,i) + Q1(i,j) = delta_ij
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < H; h++) {
            int m = L.amopi[h];
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    R.add(L.off[s ? D1B : D1A][h] + (int64_t)j * m + i, 1.0);
                    R.add(L.off[s ? Q1B : Q1A][h] + (int64_t)i * m + j, 1.0);
                    R.end(i == j ? 1.0 : 0.0);
                }
        }

    // ---- contractions to D1
    // ab part: alpha -> sum_k D2ab(ik,jk); beta -> sum_k D2ab(ki,kj)
    auto contract_ab = [&](int s, int ii, int jj) {
        for (int k = 0; k < n; k++)
            R.add(s == 0 ? d2ab(ii, k, jj, k) : d2ab(k, ii, k, jj), -1.0);
    };
    auto contract_same = [&](int s, int ii, int jj) {
        double sg;
        for (int k = 0; k < n; k++) {
            if (k == ii || k == jj) continue;
            int64_t e = d2same(s ? D2BB : D2AA, ii, k, jj, k, sg);
            R.add(e, -sg);
        }
    };
    auto contraction_block = [&](int s, double pre, bool with_ab, bool with_same) {
        for (int h = 0; h < H; h++) {
            int m = L.amopi[h], o = L.poff[h];
            for (int i = 0; i < m; i++)
                for (int j = 0; j < m; j++) {
                    R.add(L.off[s ? D1B : D1A][h] + (int64_t)i * m + j, pre);
                    if (with_ab) contract_ab(s, i + o, j + o);
                    if (with_same) contract_same(s, i + o, j + o);
                    R.end(0.0);
                }
        }
    };
    if (d.sz) {
        contraction_block(0, nb, true, false);
        contraction_block(1, na, true, false);
        contraction_block(0, na - 1.0, false, true);
        contraction_block(1, nb - 1.0, false, true);
    } else {
        contraction_block(0, na + nb - 1.0, true, true);
        contraction_block(1, na + nb - 1.0, true, true);
    }

    // ---- spin
    if (d.spin) {
        double ms = (L.mult - 1.0) / 2.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) R.add(d2ab(i, j, j, i), 1.0);
        R.end(0.5 * (na + nb) + ms * ms - ms * (ms + 1.0));

        if (L.singlet()) {
            for (int h = 0; h < H; h++) {  // D1a = D1b
                int64_t mm = (int64_t)L.amopi[h] * L.amopi[h];
                for (int64_t t = 0; t < mm; t++) {
                    R.add(L.off[D1A][h] + t, 1.0);
                    R.add(L.off[D1B][h] + t, -1.0);
                    R.end(0.0);
                }
            }
            for (int h = 0; h < H; h++) {  // D2aa = D2bb
                int64_t gg = (int64_t)L.gaa[h] * L.gaa[h];
                for (int64_t t = 0; t < gg; t++) {
                    R.add(L.off[D2AA][h] + t, 1.0);
                    R.add(L.off[D2BB][h] + t, -1.0);
                    R.end(0.0);
                }
            }
            for (int f : {D2AA, D2BB})  // D2ss(ij,kl) = 1/2 [ab(ij,kl) - ab(ji,kl) - ab(ij,lk) + ab(ji,lk)]
                for (int h = 0; h < H; h++) {
                    int g = L.gaa[h];
                    for (int a = 0; a < g; a++) {
                        auto [i, j] = L.aa_pairs[h][a];
                        for (int b = 0; b < g; b++) {
                            auto [k, l] = L.aa_pairs[h][b];
                            R.add(gem(f, h, a, b, g), 1.0);
                            R.add(d2ab(i, j, k, l), -0.5);
                            R.add(d2ab(j, i, k, l), 0.5);
                            R.add(d2ab(i, j, l, k), 0.5);
                            R.add(d2ab(j, i, l, k), -0.5);
                            R.end(0.0);
                        }
                    }
                }
            for (int h = 0; h < H; h++) {  // D200 from D2ab
                int g = L.gab[h];
                for (int a = 0; a < g; a++) {
                    auto [i, j] = L.ab_pairs[h][a];
                    double dij = i == j ? std::sqrt(2.0) : 1.0;
                    for (int b = 0; b < g; b++) {
                        auto [k, l] = L.ab_pairs[h][b];
                        double dkl = k == l ? std::sqrt(2.0) : 1.0;
                        double f = -0.5 / (dij * dkl);
                        R.add(gem(D200, h, a, b, g), 1.0);
                        R.add(d2ab(i, j, k, l), f);
                        R.add(d2ab(j, i, k, l), f);
                        R.add(d2ab(i, j, l, k), f);
                        R.add(d2ab(j, i, l, k), f);
                        R.end(0.0);
                    }
                }
            }
            for (int h = 0; h < H; h++) {  // D2ab(ij,kl) = D2ab(ji,lk)
                int g = L.gab[h];
                for (int a = 0; a < g; a++) {
                    auto [i, j] = L.ab_pairs[h][a];
                    for (int b = 0; b < g; b++) {
                        auto [k, l] = L.ab_pairs[h][b];
                        R.add(d2ab(i, j, k, l), 1.0);
                        R.add(d2ab(j, i, l, k), -1.0);
                        R.end(0.0);
                    }
                }
            }
        } else {
            // (2g x 2g) D200 block, quadrants 00, 01, 10, 11 (row-major over the block)
            for (int h = 0; h < H; h++) {
                int g = L.gab[h], G = 2 * g;
                for (int I = 0; I < G; I++)
                    for (int K = 0; K < G; K++) {
                        int a = I % g, b = K % g;
                        auto [i, j] = L.ab_pairs[h][a];
                        auto [k, l] = L.ab_pairs[h][b];
                        double dij = i == j ? std::sqrt(2.0) : 1.0;
                        double dkl = k == l ? std::sqrt(2.0) : 1.0;
                        // signs of ab(ij,kl), ab(ji,kl), ab(ij,lk), ab(ji,lk) and the prefactor
                        double f, s1, s2, s3, s4;
                        if (I < g && K < g)       { f = 0.5 / (dij * dkl); s1 = s2 = s3 = s4 = 1.0; }
                        else if (I < g)           { f = 0.5 / dij; s1 = 1; s2 = 1; s3 = -1; s4 = -1; }
                        else if (K < g)           { f = 0.5 / dkl; s1 = 1; s2 = -1; s3 = 1; s4 = -1; }
                        else                      { f = 0.5; s1 = 1; s2 = -1; s3 = -1; s4 = 1; }
                        R.add(gem(D200, h, I, K, G), 1.0);
                        R.add(d2ab(i, j, k, l), -f * s1);
                        R.add(d2ab(j, i, k, l), -f * s2);
                        R.add(d2ab(i, j, l, k), -f * s3);
                        R.add(d2ab(j, i, l, k), -f * s4);
                        R.end(0.0);
                    }
            }
        }
        if (d.g2) {  // maximal spin: sum_i G2ba(kl,ii) = 0 and sum_i G2ba(ii,kl) = 0, irrep 0 only
            int g = L.gab[0];
            for (int t = 0; t < 2; t++)
                for (int a = 0; a < g; a++) {
                    for (int i = 0; i < n; i++) {
                        int ii = L.ab(0, i, i);
                        R.add(t == 0 ? gem(G2BA, 0, a, ii, g) : gem(G2BA, 0, ii, a, g), 1.0);
                    }
                    R.end(0.0);
                }
        }
    }

    // ---- Q2
    if (d.q2) {
        for (int h = 0; h < H; h++) {
            int g = L.gab[h];
            for (int a = 0; a < g; a++) {
                auto [i, j] = L.ab_pairs[h][a];
                for (int b = 0; b < g; b++) {
                    auto [k, l] = L.ab_pairs[h][b];
                    R.add(gem(D2AB, h, a, b, g), 1.0);
                    R.add(gem(Q2AB, h, a, b, g), -1.0);
                    if (j == l) R.add(L.one_body(D1A, k, i), -1.0);
                    if (i == k) R.add(L.one_body(D1B, l, j), -1.0);
                    R.end(-(double)(i == k && j == l));
                }
            }
        }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < H; h++) {
                int g = L.gaa[h];
                int fd = s ? D2BB : D2AA, fq = s ? Q2BB : Q2AA, f1 = s ? D1B : D1A;
                for (int a = 0; a < g; a++) {
                    auto [i, j] = L.aa_pairs[h][a];
                    for (int b = 0; b < g; b++) {
                        auto [k, l] = L.aa_pairs[h][b];
                        R.add(gem(fd, h, a, b, g), 1.0);
                        R.add(gem(fq, h, a, b, g), -1.0);
                        if (j == l) R.add(L.one_body(f1, k, i), -1.0);
                        if (j == k) R.add(L.one_body(f1, l, i), 1.0);
                        if (i == l) R.add(L.one_body(f1, k, j), 1.0);
                        if (i == k) R.add(L.one_body(f1, l, j), -1.0);
                        R.end(-(double)(i == k && j == l) + (double)(i == l && j == k));
                    }
                }
            }
    }

    // ---- G2
    if (d.g2) {
        for (int s = 0; s < 2; s++)  // G2ab, G2ba
            for (int h = 0; h < H; h++) {
                int g = L.gab[h];
                for (int a = 0; a < g; a++) {
                    auto [i, j] = L.ab_pairs[h][a];
                    for (int b = 0; b < g; b++) {
                        auto [k, l] = L.ab_pairs[h][b];
                        R.add(gem(s ? G2BA : G2AB, h, a, b, g), -1.0);
                        if (j == l) R.add(L.one_body(s ? D1B : D1A, i, k), 1.0);
                        R.add(s ? d2ab(l, i, j, k) : d2ab(i, l, k, j), -1.0);
                        R.end(0.0);
                    }
                }
            }
        for (int h = 0; h < H; h++) {
            int g = L.gab[h], G = 2 * g;
            for (int I = 0; I < G; I++)
                for (int K = 0; K < G; K++) {
                    auto [i, j] = L.ab_pairs[h][I % g];
                    auto [k, l] = L.ab_pairs[h][K % g];
                    R.add(gem(G2AA, h, I, K, G), -1.0);
                    if ((I < g) == (K < g)) {  // aaaa or bbbb
                        int beta = I >= g;
                        if (j == l) R.add(L.one_body(beta ? D1B : D1A, i, k), 1.0);
                        if (i != l && k != j) {
                            double sg;
                            int64_t e = d2same(beta ? D2BB : D2AA, i, l, k, j, sg);
                            R.add(e, -sg);
                        }
                    } else if (I < g) {      // aabb
                        R.add(d2ab(i, l, j, k), 1.0);
                    } else {                 // bbaa
                        R.add(d2ab(l, i, k, j), 1.0);
                    }
                    R.end(0.0);
                }
        }
    }

    // ---- T1: rows T1aab, T1bba, T1aaa, T1bbb (the order T1_constraints_Au
    // writes them), each -T1(ijk,lmn) plus the Q2, D2, G2 and D1 elements
    // t1.cc names under the same Kronecker conditions; b = 0
    // (v2rdm_solver.cc:2076-2112)
    if (d.t1) {
        auto sy = [&](int p) { return L.sym[p]; };
        for (int s = 0; s < 2; s++)  // T1aab, T1bba
            for (int h = 0; h < H; h++) {
                const auto& T = L.aab_trips[h];
                const int t = (int)T.size();
                for (int a = 0; a < t; a++) {
                    const auto [i, j, k] = T[a];
                    for (int b = 0; b < t; b++) {
                        const auto [l, m, nn] = T[b];
                        R.add(gem(s ? T1BBA : T1AAB, h, a, b, t), -1.0);
                        if (k == nn) {
                            const int hij = sy(i) ^ sy(j);
                            R.add(gem(s ? Q2BB : Q2AA, hij, L.aa(hij, i, j), L.aa(hij, l, m), L.gaa[hij]), 1.0);
                        }
                        if (j == l) {
                            const int hki = sy(k) ^ sy(i);
                            const int nm = s ? L.ab(hki, nn, m) : L.ab(hki, m, nn);
                            const int ki = s ? L.ab(hki, k, i) : L.ab(hki, i, k);
                            R.add(gem(D2AB, hki, nm, ki, L.gab[hki]), -1.0);
                        }
                        if (l == i) {
                            const int hkj = sy(k) ^ sy(j);
                            const int nm = s ? L.ab(hkj, nn, m) : L.ab(hkj, m, nn);
                            const int kj = s ? L.ab(hkj, k, j) : L.ab(hkj, j, k);
                            R.add(gem(D2AB, hkj, nm, kj, L.gab[hkj]), 1.0);
                        }
                        if (j == m) {
                            const int hni = sy(nn) ^ sy(i);
                            R.add(gem(s ? G2AB : G2BA, hni, L.ab(hni, nn, i), L.ab(hni, k, l), L.gab[hni]), -1.0);
                        }
                        if (i == m) {
                            const int hkl = sy(k) ^ sy(l);
                            R.add(gem(s ? G2AB : G2BA, hkl, L.ab(hkl, nn, j), L.ab(hkl, k, l), L.gab[hkl]), 1.0);
                        }
                        R.end(0.0);
                    }
                }
            }
        for (int s = 0; s < 2; s++)  // T1aaa, T1bbb
            for (int h = 0; h < H; h++) {
                const auto& T = L.aaa_trips[h];
                const int t = (int)T.size();
                const int q2 = s ? Q2BB : Q2AA, d2 = s ? D2BB : D2AA, d1 = s ? D1B : D1A;
                // G2aa element (a, b) of the aa (s = 0) or bb (s = 1) quadrant
                auto g2q = [&](int hh, int a, int b) {
                    const int g = L.gab[hh];
                    return gem(G2AA, hh, a + s * g, b + s * g, 2 * g);
                };
                auto same = [&](int f, int hh, int p, int q, int r, int u) {  // pair indices, no sign
                    return gem(f, hh, L.aa(hh, p, q), L.aa(hh, r, u), L.gaa[hh]);
                };
                for (int a = 0; a < t; a++) {
                    const auto [i, j, k] = T[a];
                    for (int b = 0; b < t; b++) {
                        const auto [l, m, nn] = T[b];
                        R.add(gem(s ? T1BBB : T1AAA, h, a, b, t), -1.0);
                        if (k == nn) R.add(same(q2, sy(i) ^ sy(j), i, j, l, m), 1.0);
                        if (j == nn && (sy(i) ^ sy(k)) == (sy(l) ^ sy(m))) R.add(same(q2, sy(i) ^ sy(k), i, k, l, m), -1.0);
                        if (i == nn && (sy(j) ^ sy(k)) == (sy(l) ^ sy(m))) R.add(same(q2, sy(j) ^ sy(k), j, k, l, m), 1.0);
                        if (l == k && (sy(j) ^ sy(i)) == (sy(nn) ^ sy(m))) R.add(same(d2, sy(j) ^ sy(i), nn, m, j, i), 1.0);
                        if (j == l) R.add(same(d2, sy(k) ^ sy(i), nn, m, k, i), -1.0);
                        if (l == i) R.add(same(d2, sy(k) ^ sy(j), nn, m, k, j), 1.0);
                        if (k == m) {
                            if (j == l) R.add(L.one_body(d1, nn, i), -1.0);
                            const int hni = sy(nn) ^ sy(i);
                            R.add(g2q(hni, L.ab(hni, nn, i), L.ab(hni, j, l)), 1.0);
                        }
                        if (j == m) {
                            if (k == l) R.add(L.one_body(d1, nn, i), 1.0);
                            const int hni = sy(nn) ^ sy(i);
                            R.add(g2q(hni, L.ab(hni, nn, i), L.ab(hni, k, l)), -1.0);
                        }
                        if (i == m) {
                            if (k == l) R.add(L.one_body(d1, nn, j), -1.0);
                            const int hkl = sy(k) ^ sy(l);
                            R.add(g2q(hkl, L.ab(hkl, nn, j), L.ab(hkl, k, l)), 1.0);
                        }
                        R.end(0.0);
                    }
                }
            }
    }

    // ---- T2: rows T2aab, T2bba, T2aaa, T2bbb (T2_constraints_Au_slow's
    // order); T2aaa / T2bbb are one block each over the aab triplets then all
    // triplets.  Each row is -T2(ijk,lmn) plus the D2 and D1 elements t2.cc
    // names; b = 0 (v2rdm_solver.cc:2113-2150)
    if (d.t2) {
        auto sy = [&](int p) { return L.sym[p]; };
        auto dab = [&](int hh, int a, int b) { return gem(D2AB, hh, a, b, L.gab[hh]); };
        auto dss = [&](int f, int hh, int a, int b) { return gem(f, hh, a, b, L.gaa[hh]); };
        // signed same-spin D2 element of pairs (x, y), (u, v): skipped when a
        // pair repeats an orbital, sign flipped for x > y and for u > v
        auto signed_d2 = [&](int f, int x, int y, int u, int v, double c0) {
            if (x == y || u == v) return;
            const int hh = sy(x) ^ sy(y);
            double c = c0;
            if (x > y) c = -c;
            if (u > v) c = -c;
            R.add(dss(f, hh, L.aa(hh, x, y), L.aa(hh, u, v)), c);
        };
        for (int s = 0; s < 2; s++)  // T2aab, T2bba
            for (int h = 0; h < H; h++) {
                const auto& T = L.aab_trips[h];
                const int t = (int)T.size();
                const int d2s = s ? D2BB : D2AA, d1o = s ? D1A : D1B;
                for (int a = 0; a < t; a++) {
                    const auto [i, j, k] = T[a];
                    for (int b = 0; b < t; b++) {
                        const auto [l, m, nn] = T[b];
                        R.add(gem(s ? T2BBA : T2AAB, h, a, b, t), -1.0);
                        if (k == nn) {
                            const int hij = sy(i) ^ sy(j);
                            R.add(dss(d2s, hij, L.aa(hij, i, j), L.aa(hij, l, m)), 1.0);
                        }
                        if (j == m && i == l) R.add(L.one_body(d1o, nn, k), 1.0);
                        if (j == l && i == m) R.add(L.one_body(d1o, nn, k), -1.0);
                        if (i == l) {
                            const int hh = sy(nn) ^ sy(j);
                            R.add(s ? dab(hh, L.ab(hh, nn, j), L.ab(hh, k, m)) : dab(hh, L.ab(hh, j, nn), L.ab(hh, m, k)), -1.0);
                        }
                        if (j == l) {
                            const int hh = sy(nn) ^ sy(i);
                            R.add(s ? dab(hh, L.ab(hh, nn, i), L.ab(hh, k, m)) : dab(hh, L.ab(hh, i, nn), L.ab(hh, m, k)), 1.0);
                        }
                        if (i == m) {
                            const int hh = sy(nn) ^ sy(j);
                            R.add(s ? dab(hh, L.ab(hh, nn, j), L.ab(hh, k, l)) : dab(hh, L.ab(hh, j, nn), L.ab(hh, l, k)), 1.0);
                        }
                        if (j == m) {
                            const int hh = sy(nn) ^ sy(i);
                            R.add(s ? dab(hh, L.ab(hh, nn, i), L.ab(hh, k, l)) : dab(hh, L.ab(hh, i, nn), L.ab(hh, l, k)), -1.0);
                        }
                        R.end(0.0);
                    }
                }
            }
        for (int s = 0; s < 2; s++)  // T2aaa (aaa + abb parts), T2bbb (bbb + baa parts)
            for (int h = 0; h < H; h++) {
                const auto& A3 = L.aab_trips[h];
                const auto& B3 = L.aba_trips[h];
                const int ta = (int)A3.size(), tb = (int)B3.size(), D = ta + tb;
                const int f = s ? T2BBB : T2AAA;
                const int d2s = s ? D2BB : D2AA, d2o = s ? D2AA : D2BB, d1s = s ? D1B : D1A, d1o = s ? D1A : D1B;
                for (int a = 0; a < D; a++) {
                    const auto [i, j, k] = a < ta ? A3[a] : B3[a - ta];
                    for (int b = 0; b < D; b++) {
                        const auto [l, m, nn] = b < ta ? A3[b] : B3[b - ta];
                        R.add(gem(f, h, a, b, D), -1.0);
                        if (a < ta && b < ta) {         // aab x aab
                            if (k == nn) {
                                const int hij = sy(i) ^ sy(j);
                                R.add(dss(d2s, hij, L.aa(hij, i, j), L.aa(hij, l, m)), 1.0);
                            }
                            if (j == m && i == l) R.add(L.one_body(d1s, nn, k), 1.0);
                            if (j == l && i == m) R.add(L.one_body(d1s, nn, k), -1.0);
                            if (i == l) signed_d2(d2s, nn, j, k, m, -1.0);
                            if (j == l) signed_d2(d2s, nn, i, k, m, 1.0);
                            if (i == m) signed_d2(d2s, nn, j, k, l, 1.0);
                            if (j == m) signed_d2(d2s, nn, i, k, l, -1.0);
                        } else if (a < ta) {            // aab x aba
                            if (i == l) {
                                const int hh = sy(j) ^ sy(nn);
                                R.add(s ? dab(hh, L.ab(hh, nn, j), L.ab(hh, m, k)) : dab(hh, L.ab(hh, j, nn), L.ab(hh, k, m)), 1.0);
                            }
                            if (j == l) {
                                const int hh = sy(i) ^ sy(nn);
                                R.add(s ? dab(hh, L.ab(hh, nn, i), L.ab(hh, m, k)) : dab(hh, L.ab(hh, i, nn), L.ab(hh, k, m)), -1.0);
                            }
                        } else if (b < ta) {            // aba x aab
                            if (i == l) {
                                const int hh = sy(j) ^ sy(nn);
                                R.add(s ? dab(hh, L.ab(hh, j, nn), L.ab(hh, k, m)) : dab(hh, L.ab(hh, nn, j), L.ab(hh, m, k)), 1.0);
                            }
                            if (i == m) {
                                const int hh = sy(j) ^ sy(nn);
                                R.add(s ? dab(hh, L.ab(hh, j, nn), L.ab(hh, k, l)) : dab(hh, L.ab(hh, nn, j), L.ab(hh, l, k)), -1.0);
                            }
                        } else {                        // aba x aba
                            if (k == nn) {
                                const int hij = sy(i) ^ sy(j);
                                R.add(s ? dab(hij, L.ab(hij, j, i), L.ab(hij, m, l)) : dab(hij, L.ab(hij, i, j), L.ab(hij, l, m)), 1.0);
                            }
                            if (j == m && i == l) R.add(L.one_body(d1o, nn, k), 1.0);
                            if (i == l) signed_d2(d2o, nn, j, k, m, -1.0);
                            if (j == m) {
                                const int hh = sy(nn) ^ sy(i);
                                R.add(s ? dab(hh, L.ab(hh, nn, i), L.ab(hh, k, l)) : dab(hh, L.ab(hh, i, nn), L.ab(hh, l, k)), -1.0);
                            }
                        }
                        R.end(0.0);
                    }
                }
            }
    }

    // ---- D3 (d3.cc:45-406): contractions to D2 (with CONSTRAIN_SZ: aaa->aa,
    // bbb->bb, aab->aa, bba->bb, aab->ab, bba->ab; without: (aaa+aab)->aa,
    // (bbb+bba)->bb, (aab+bba)->ab), then for singlets with CONSTRAIN_SPIN
    // D3aab = D3bba and D3aaa / D3bbb from D3aab / D3bba; b = 0
    if (d.d3) {
        auto sy = [&](int p) { return L.sym[p]; };
        const int nsame = d.sz ? 2 : 1;
        for (int pass = 0; pass < 2 * nsame; pass++) {
            const int spin = pass % 2;
            const int part = d.sz ? pass / 2 : -1;  // 0: from D3aaa / bbb, 1: from D3aab / bba, -1: both
            const int d2 = spin ? D2BB : D2AA, f3s = spin ? D3BBB : D3AAA, f3m = spin ? D3BBA : D3AAB;
            const double coef = part == 0 ? (spin ? nb : na) - 2.0 : part == 1 ? (spin ? na : nb) : na + nb - 2.0;
            for (int h = 0; h < H; h++) {
                const int g = L.gaa[h];
                for (int a = 0; a < g; a++) {
                    const auto [i, j] = L.aa_pairs[h][a];
                    for (int b = 0; b < g; b++) {
                        const auto [k, l] = L.aa_pairs[h][b];
                        R.add(gem(d2, h, a, b, g), coef);
                        for (int q = 0; q < n; q++) {
                            const int h2 = h ^ sy(q);
                            if (part != 0) R.add(gem(f3m, h2, L.aab(h2, i, j, q), L.aab(h2, k, l, q), L.taab(h2)), -1.0);
                            if (part != 1 && i != q && j != q && k != q && l != q) {
                                double c = -1.0;
                                if (i < q && q < j) c = -c;
                                if (k < q && q < l) c = -c;
                                R.add(gem(f3s, h2, L.aaa(h2, i, j, q), L.aaa(h2, k, l, q), L.taaa(h2)), c);
                            }
                        }
                        R.end(0.0);
                    }
                }
            }
        }
        for (int pass = 0; pass < nsame; pass++) {
            const double coef = d.sz ? (pass == 0 ? na - 1.0 : nb - 1.0) : na + nb - 2.0;
            for (int h = 0; h < H; h++) {
                const int g = L.gab[h];
                for (int a = 0; a < g; a++) {
                    const auto [i, j] = L.ab_pairs[h][a];
                    for (int b = 0; b < g; b++) {
                        const auto [k, l] = L.ab_pairs[h][b];
                        R.add(gem(D2AB, h, a, b, g), coef);
                        for (int q = 0; q < n; q++) {
                            const int h2 = h ^ sy(q);
                            if ((!d.sz || pass == 0) && i != q && k != q) {
                                double c = -1.0;
                                if (q < i) c = -c;
                                if (q < k) c = -c;
                                R.add(gem(D3AAB, h2, L.aab(h2, i, q, j), L.aab(h2, k, q, l), L.taab(h2)), c);
                            }
                            if ((!d.sz || pass == 1) && j != q && l != q) {
                                double c = -1.0;
                                if (q < j) c = -c;
                                if (q < l) c = -c;
                                R.add(gem(D3BBA, h2, L.aab(h2, j, q, i), L.aab(h2, l, q, k), L.taab(h2)), c);
                            }
                        }
                        R.end(0.0);
                    }
                }
            }
        }
        if (d.spin && L.singlet()) {
            for (int h = 0; h < H; h++) {
                const int t = L.taab(h);
                for (int a = 0; a < t; a++)
                    for (int b = 0; b < t; b++) {
                        R.add(gem(D3AAB, h, a, b, t), 1.0);
                        R.add(gem(D3BBA, h, a, b, t), -1.0);
                        R.end(0.0);
                    }
            }
            for (int s = 0; s < 2; s++)
                for (int h = 0; h < H; h++) {
                    const auto& T = L.aaa_trips[h];
                    const int t = (int)T.size(), ta = L.taab(h);
                    for (int a = 0; a < t; a++) {
                        const auto [p, q, r] = T[a];
                        const int rows3[3] = {L.aab(h, p, q, r), L.aab(h, p, r, q), L.aab(h, q, r, p)};
                        for (int b = 0; b < t; b++) {
                            const auto [s1, t1, v1] = T[b];
                            const int cols3[3] = {L.aab(h, s1, t1, v1), L.aab(h, s1, v1, t1), L.aab(h, t1, v1, s1)};
                            R.add(gem(s ? D3BBB : D3AAA, h, a, b, t), 1.0);
                            for (int x = 0; x < 3; x++)
                                for (int y = 0; y < 3; y++)
                                    R.add(gem(s ? D3BBA : D3AAB, h, rows3[x], cols3[y], ta),
                                          ((x + y) % 2 == 0 ? -1.0 : 1.0) / 3.0);
                            R.end(0.0);
                        }
                    }
                }
        }
    }

    A.nrows = (int64_t)A.rhs.size();
    return A;
}

Csr Csr::transpose() const {
    Csr T;
    T.nrows = ncols;
    T.ncols = nrows;
    T.ptr.assign(ncols + 1, 0);
    for (int32_t c : col) T.ptr[c + 1]++;
    for (int64_t c = 0; c < ncols; c++) T.ptr[c + 1] += T.ptr[c];
    T.col.resize(col.size());
    T.val.resize(val.size());
    std::vector<int64_t> cur(T.ptr.begin(), T.ptr.end() - 1);
    for (int64_t r = 0; r < nrows; r++)
        for (int64_t e = ptr[r]; e < ptr[r + 1]; e++) {
            int64_t dst = cur[col[e]]++;
            T.col[dst] = (int32_t)r;
            T.val[dst] = val[e];
        }
    return T;
}

std::vector<double> hf_guess(const Layout& L, const Csr& A, int nocc_alpha, int nocc_beta) {
    std::vector<double> x(L.n_primal, 0.0);
    const int n = L.n;
    std::vector<double> oa(n), ob(n);
    for (int p = 0; p < n; p++) {
        oa[p] = L.to_user[p] < nocc_alpha ? 1.0 : 0.0;
        ob[p] = L.to_user[p] < nocc_beta ? 1.0 : 0.0;
    }
    for (int p = 0; p < n; p++) {
        x[L.one_body(D1A, p, p)] = oa[p];
        x[L.one_body(D1B, p, p)] = ob[p];
        x[L.one_body(Q1A, p, p)] = 1.0 - oa[p];
        x[L.one_body(Q1B, p, p)] = 1.0 - ob[p];
    }
    for (int h = 0; h < L.H; h++) {
        for (int ij = 0; ij < L.gab[h]; ij++) {
            const int i = L.ab_pairs[h][ij][0], j = L.ab_pairs[h][ij][1];
            x[L.off[D2AB][h] + (int64_t)ij * L.gab[h] + ij] = oa[i] * ob[j];
        }
        for (int ij = 0; ij < L.gaa[h]; ij++) {
            const int i = L.aa_pairs[h][ij][0], j = L.aa_pairs[h][ij][1];
            x[L.off[D2AA][h] + (int64_t)ij * L.gaa[h] + ij] = oa[i] * oa[j];
            x[L.off[D2BB][h] + (int64_t)ij * L.gaa[h] + ij] = ob[i] * ob[j];
        }
    }
    // family of every primal index
    std::vector<int8_t> fam(L.n_primal, -1);
    for (const Block& b : L.blocks)
        for (int64_t e = 0; e < (int64_t)b.dim * b.dim; e++) fam[b.off + e] = (int8_t)b.fam;
    // Q2, G2 and T1 elements are derived: every Q2 / G2 row holds exactly one
    // Q2 / G2 element, and every T1 row one T1 element besides Q2 / G2
    // elements that earlier rows have fixed; solve each row for its one
    // unknown, in row order
    std::vector<char> known(L.n_primal, 1);
    // the D3 blocks stay zero: Guess() has no D3 part (v2rdm_solver.cc:1808-1823)
    for (int64_t c = 0; c < L.n_primal; c++) known[c] = fam[c] < Q2AB || fam[c] >= D3AAA;
    for (int64_t r = 0; r < A.nrows; r++) {
        int64_t e = -1;
        int nd = 0;
        for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (!known[A.col[k]]) {
                e = k;
                nd++;
            }
        if (nd != 1) continue;
        double s = A.rhs[r];
        for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; k++)
            if (k != e) s -= A.val[k] * x[A.col[k]];
        x[A.col[e]] = s / A.val[e];
        known[A.col[e]] = 1;
    }
    return x;
}

CostMap build_cost_map(const Layout& L) {
    CostMap M;
    M.i1.assign(L.n_primal, -1);
    M.i2.assign(L.n_primal, -1);
    const int n = L.n;
    const int64_t nn = (int64_t)n * n;
    auto u = [&](int p) { return L.to_user[p]; };
    auto hidx = [&](int p, int q) { return (int64_t)u(p) * n + u(q); };
    auto gidx = [&](int p, int q, int r, int s) {
        return nn + (((int64_t)u(p) * n + u(q)) * n + u(r)) * n + u(s);
    };
    if (nn + nn * nn >= ((int64_t)1 << 31)) throw std::invalid_argument("integral arrays too large");
    for (int h = 0; h < L.H; h++) {
        int m = L.amopi[h], o = L.poff[h];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                M.i1[L.off[D1A][h] + (int64_t)i * m + j] = (int32_t)hidx(i + o, j + o);
                M.i1[L.off[D1B][h] + (int64_t)i * m + j] = (int32_t)hidx(i + o, j + o);
            }
        int g = L.gab[h];
        for (int a = 0; a < g; a++) {
            auto [i, j] = L.ab_pairs[h][a];
            for (int b = 0; b < g; b++) {
                auto [k, l] = L.ab_pairs[h][b];
                M.i1[L.off[D2AB][h] + (int64_t)a * g + b] = (int32_t)gidx(i, k, j, l);
            }
        }
        int ga = L.gaa[h];
        for (int a = 0; a < ga; a++) {
            auto [i, j] = L.aa_pairs[h][a];
            for (int b = 0; b < ga; b++) {
                auto [k, l] = L.aa_pairs[h][b];
                for (int f : {D2AA, D2BB}) {
                    M.i1[L.off[f][h] + (int64_t)a * ga + b] = (int32_t)gidx(i, k, j, l);
                    M.i2[L.off[f][h] + (int64_t)a * ga + b] = (int32_t)gidx(i, l, j, k);
                }
            }
        }
    }
    return M;
}

}  // namespace v2h
