// layout.hpp -- geminal tables, primal block layout and dual row layout of one
// DQG v2RDM problem (host side, built once per problem structure).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace v2h {

constexpr int kMaxIrrep = 8;

// Problem description in the user's orbital order.
struct ProblemDesc {
    int nmo = 0;            // active orbitals
    int nalpha = 0;         // active alpha electrons
    int nbeta = 0;          // active beta electrons
    int multiplicity = 0;   // 2S+1; 0 selects nalpha - nbeta + 1
    int nirrep = 1;         // 1, 2, 4 or 8
    std::vector<int> irreps;  // per orbital (user order); empty = all 0
    bool q2 = true, g2 = true, spin = true, sz = true;
    bool t1 = false;          // POSITIVITY DQGT1 (needs q2 and g2)
    bool t2 = false;          // POSITIVITY DQGT2 (needs q2 and g2); both: DQGT1T2
    bool d3 = false;          // CONSTRAIN_D3

    // Everything that determines the tables (structures with equal keys share them).
    std::string key() const;
};

// Primal block families in layout order (v2rdm_solver.cc:3729-3846).
enum Fam : int {
    D2AB = 0, D2AA, D2BB, D200, D1A, D1B, Q1A, Q1B, Q2AB, Q2AA, Q2BB, G2AB, G2BA, G2AA,
    T1AAA, T1BBB, T1AAB, T1BBA,  // after G2 (v2rdm_solver.cc:3848-3873)
    T2AAA, T2BBB, T2AAB, T2BBA,  // after T1 (v2rdm_solver.cc:3879-3908)
    D3AAA, D3BBB, D3AAB, D3BBA,  // after T2 (v2rdm_solver.cc:4000-4026)
    NFAM
};

struct Block {
    int fam, irrep, dim;
    int64_t off;
};

struct Layout {
    ProblemDesc desc;
    int n = 0, H = 1, na = 0, nb = 0, mult = 1;
    std::vector<int> to_user;   // Pitzer position -> user orbital index
    std::vector<int> sym;       // irrep of each Pitzer orbital
    std::array<int, kMaxIrrep> amopi{}, poff{}, gab{}, gaa{};
    // pairs[h][k] = (first, second); index tables [h][i*n + j] (-1 when absent)
    std::array<std::vector<std::array<int, 2>>, kMaxIrrep> ab_pairs, aa_pairs;
    std::array<std::vector<int>, kMaxIrrep> ab_index, aa_index;
    // T1 / T2 only: triplets of irrep h (basis.cc:642-744), aba all of them,
    // aab with i < j, aaa with i < j < k; index tables [h][(i*n + j)*n + k]
    // (-1 when absent)
    std::array<std::vector<std::array<int, 3>>, kMaxIrrep> aab_trips, aaa_trips, aba_trips;
    std::array<std::vector<int>, kMaxIrrep> aab_index, aaa_index;
    std::array<std::array<int64_t, kMaxIrrep>, NFAM> off{};
    std::vector<Block> blocks;  // in layout order, zero-size blocks included
    int64_t n_primal = 0, n_dual = 0;

    explicit Layout(const ProblemDesc& d);

    int ab(int h, int i, int j) const { return ab_index[h][i * n + j]; }
    int aa(int h, int i, int j) const { return aa_index[h][i * n + j]; }
    int taab(int h) const { return (int)aab_trips[h].size(); }
    int taaa(int h) const { return (int)aaa_trips[h].size(); }
    int taba(int h) const { return (int)aba_trips[h].size(); }
    int aab(int h, int i, int j, int k) const { return aab_index[h][((size_t)i * n + j) * n + k]; }
    int aaa(int h, int i, int j, int k) const { return aaa_index[h][((size_t)i * n + j) * n + k]; }
    int pair_irrep(int i, int j) const { return sym[i] ^ sym[j]; }
    // element (i, j) of a one-body block of family f; i, j in the same irrep
    int64_t one_body(int f, int i, int j) const {
        int h = sym[i];
        return off[f][h] + (int64_t)(i - poff[h]) * amopi[h] + (j - poff[h]);
    }
    bool singlet() const { return na == nb; }
};

}  // namespace v2h
