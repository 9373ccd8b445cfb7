// layout.cpp -- tables and primal layout of a DQG v2RDM problem.
//
// Follows edeprince3/hilbert:
//   basis.cc:47-76      irrep product = XOR of Cotton indices
//   basis.cc:104-111    orbitals of an irrep are contiguous (Pitzer order)
//   basis.cc:427-533    geminal enumeration: outer loop over the second
//                       orbital, inner over the first, pair = (inner, outer);
//                       aa pairs keep first < second in the same order
//   v2rdm_solver.cc:3729-3846  primal blocks D2ab, D2aa, D2bb, D200, D1a, D1b,
//                       Q1a, Q1b, Q2ab, Q2aa, Q2bb, G2ab, G2ba, G2aa(2g)
//   basis.cc:642-744    triplets (T1): loop order i, j, k within an irrep;
//                       aab keeps i < j, aaa keeps i < j < k
//   v2rdm_solver.cc:3848-3873  T1aaa, T1bbb, T1aab, T1bba after G2
//   v2rdm_solver.cc:3879-3908  T2aaa, T2bbb (trip_aab + trip_aba), T2aab,
//                       T2bba after T1
//   v2rdm_solver.cc:4000-4026  D3aaa, D3bbb, D3aab, D3bba after T2
#include "layout.hpp"

#include <algorithm>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace v2h {

std::string ProblemDesc::key() const {
    std::ostringstream k;
    k << nmo << ':' << nalpha << ':' << nbeta << ':' << (multiplicity ? multiplicity : nalpha - nbeta + 1)
      << ':' << nirrep << ':' << q2 << g2 << spin << sz << t1 << t2 << d3 << ':';
    std::vector<int> s = irreps.empty() ? std::vector<int>(nmo, 0) : irreps;
    std::sort(s.begin(), s.end());  // the tables depend only on the irrep counts
    for (int v : s) k << v << ',';
    return k.str();
}

Layout::Layout(const ProblemDesc& d) : desc(d) {
    n = d.nmo;
    H = d.nirrep;
    na = d.nalpha;
    nb = d.nbeta;
    mult = d.multiplicity ? d.multiplicity : na - nb + 1;
    if (n <= 0) throw std::invalid_argument("nmo must be positive");
    if (H != 1 && H != 2 && H != 4 && H != 8) throw std::invalid_argument("nirrep must be 1, 2, 4 or 8");
    if (na < 0 || nb < 0 || na > n || nb > n) throw std::invalid_argument("electron counts out of range");
    if ((d.t1 || d.t2) && !(d.q2 && d.g2))
        throw std::invalid_argument("the T1 and T2 conditions need Q2 and G2 (POSITIVITY DQGT1, DQGT2, DQGT1T2)");
    std::vector<int> irr = d.irreps.empty() ? std::vector<int>(n, 0) : d.irreps;
    if ((int)irr.size() != n) throw std::invalid_argument("orbital_irreps has the wrong length");
    for (int v : irr)
        if (v < 0 || v >= H) throw std::invalid_argument("orbital irrep out of range");

    // stable sort of the user orbitals by irrep gives the Pitzer order
    to_user.resize(n);
    std::iota(to_user.begin(), to_user.end(), 0);
    std::stable_sort(to_user.begin(), to_user.end(), [&](int a, int b) { return irr[a] < irr[b]; });
    sym.resize(n);
    for (int p = 0; p < n; p++) sym[p] = irr[to_user[p]];
    amopi.fill(0);
    for (int p = 0; p < n; p++) amopi[sym[p]]++;
    for (int h = 0, acc = 0; h < H; h++) { poff[h] = acc; acc += amopi[h]; }

    for (int h = 0; h < H; h++) {
        ab_index[h].assign((size_t)n * n, -1);
        aa_index[h].assign((size_t)n * n, -1);
        ab_pairs[h].clear();
        aa_pairs[h].clear();
        for (int second = 0; second < n; second++)
            for (int first = 0; first < n; first++) {
                if ((sym[first] ^ sym[second]) != h) continue;
                ab_index[h][first * n + second] = (int)ab_pairs[h].size();
                ab_pairs[h].push_back({first, second});
                if (first < second) {
                    int k = (int)aa_pairs[h].size();
                    aa_index[h][first * n + second] = k;
                    aa_index[h][second * n + first] = k;
                    aa_pairs[h].push_back({first, second});
                }
            }
        gab[h] = (int)ab_pairs[h].size();
        gaa[h] = (int)aa_pairs[h].size();
    }

    if (d.t1 || d.t2 || d.d3)
        for (int h = 0; h < H; h++) {
            aab_index[h].assign((size_t)n * n * n, -1);
            aaa_index[h].assign((size_t)n * n * n, -1);
            auto at = [&](std::vector<int>& t, int i, int j, int k) -> int& { return t[((size_t)i * n + j) * n + k]; };
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++)
                    for (int k = 0; k < n; k++) {
                        if ((sym[i] ^ sym[j] ^ sym[k]) != h) continue;
                        aba_trips[h].push_back({i, j, k});
                        if (i >= j) continue;
                        const int c = (int)aab_trips[h].size();
                        at(aab_index[h], i, j, k) = at(aab_index[h], j, i, k) = c;
                        aab_trips[h].push_back({i, j, k});
                        if (j >= k) continue;
                        const int e = (int)aaa_trips[h].size();
                        for (auto [a, b, c2] : {std::array<int, 3>{i, j, k}, {i, k, j}, {j, i, k}, {j, k, i}, {k, i, j},
                                                {k, j, i}})
                            at(aaa_index[h], a, b, c2) = e;
                        aaa_trips[h].push_back({i, j, k});
                    }
        }

    int64_t o = 0;
    auto push = [&](int f, int h, int dim) {
        off[f][h] = o;
        blocks.push_back({f, h, dim, o});
        o += (int64_t)dim * dim;
    };
    for (int h = 0; h < H; h++) push(D2AB, h, gab[h]);
    for (int h = 0; h < H; h++) push(D2AA, h, gaa[h]);
    for (int h = 0; h < H; h++) push(D2BB, h, gaa[h]);
    if (d.spin)
        for (int h = 0; h < H; h++) push(D200, h, singlet() ? gab[h] : 2 * gab[h]);
    for (int f : {D1A, D1B, Q1A, Q1B})
        for (int h = 0; h < H; h++) push(f, h, amopi[h]);
    if (d.q2) {
        for (int h = 0; h < H; h++) push(Q2AB, h, gab[h]);
        for (int h = 0; h < H; h++) push(Q2AA, h, gaa[h]);
        for (int h = 0; h < H; h++) push(Q2BB, h, gaa[h]);
    }
    if (d.g2) {
        for (int h = 0; h < H; h++) push(G2AB, h, gab[h]);
        for (int h = 0; h < H; h++) push(G2BA, h, gab[h]);
        for (int h = 0; h < H; h++) push(G2AA, h, 2 * gab[h]);
    }
    if (d.t1) {
        for (int h = 0; h < H; h++) push(T1AAA, h, taaa(h));
        for (int h = 0; h < H; h++) push(T1BBB, h, taaa(h));
        for (int h = 0; h < H; h++) push(T1AAB, h, taab(h));
        for (int h = 0; h < H; h++) push(T1BBA, h, taab(h));
    }
    if (d.t2) {
        for (int h = 0; h < H; h++) push(T2AAA, h, taab(h) + taba(h));
        for (int h = 0; h < H; h++) push(T2BBB, h, taab(h) + taba(h));
        for (int h = 0; h < H; h++) push(T2AAB, h, taab(h));
        for (int h = 0; h < H; h++) push(T2BBA, h, taab(h));
    }
    if (d.d3) {
        for (int h = 0; h < H; h++) push(D3AAA, h, taaa(h));
        for (int h = 0; h < H; h++) push(D3BBB, h, taaa(h));
        for (int h = 0; h < H; h++) push(D3AAB, h, taab(h));
        for (int h = 0; h < H; h++) push(D3BBA, h, taab(h));
    }
    n_primal = o;
    if (n_primal >= (int64_t)1 << 24)
        throw std::invalid_argument("problem too large: more than 2^24 primal variables per instance");
}

}  // namespace v2h
