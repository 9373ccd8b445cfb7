// fcidump.cpp -- the binary integral and RDM files of v2RDMSolver::FCIDUMP
// (v2rdm_casscf/fcidump.cc:60-243) for one instance of a batch.
//
// Both files are sequences of 24-byte records, a double and four int32
// orbital labels (1-based, 0 = unused), in the order the reference writes
// them.  The integral file holds every symmetry-allowed (pq|rs) with p <= q,
// r <= s and INDEX(p,q) <= INDEX(r,s), then h_pq over each irrep block, then
// the nuclear repulsion; its two-electron values come from one device pass,
// the pair Gram matrix G = Qmo^T Qmo of the density-fitted integrals
// (k_pair_gram), copied back once.  The RDM file holds the spin-free 2-RDM
// over the ab geminals and the spin-summed 1-RDM of the instance's current
// primal vector, with labels relative to the core.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "df.hpp"
#include "solver.hpp"

namespace v2h {

namespace {

struct RecordFile {
    std::vector<char> buf;
    void put(double v, int a, int b, int c, int d) {
        const size_t o = buf.size();
        buf.resize(o + 24);
        std::memcpy(&buf[o], &v, 8);
        const int l[4] = {a, b, c, d};
        std::memcpy(&buf[o + 8], l, 16);
    }
    int64_t count() const { return (int64_t)(buf.size() / 24); }
    void write(const char* path) const {
        FILE* f = std::fopen(path, "wb");
        if (!f) throw std::invalid_argument(std::string("write_fcidump: cannot open ") + path);
        const size_t n = buf.empty() ? 0 : std::fwrite(buf.data(), 1, buf.size(), f);
        const bool bad = std::fclose(f) != 0 || n != buf.size();
        if (bad) throw std::runtime_error(std::string("write_fcidump: short write to ") + path);
    }
};

int64_t tri(int64_t a, int64_t b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

}  // namespace

void Batch::write_fcidump(int i, const FcidumpArgs& a, int64_t* n_int, int64_t* n_rdm) {
    const Layout& L = layout(i);
    const int n = L.n, nfull = a.nfull;
    if (nfull < 1 || a.nq < 1 || a.ncore < 0 || a.ncore + n > nfull)
        throw std::invalid_argument("write_fcidump: need nfull >= ncore + nmo, nq >= 1");
    if (!a.int_path || !a.rdm_path) throw std::invalid_argument("write_fcidump: null file name");
    // space of every orbital: 0 core, 1 active, 2 virtual
    std::vector<int> space(nfull, 2);
    auto take = [&](int k, int sp) {
        if (k < 0 || k >= nfull) throw std::invalid_argument("write_fcidump: orbital index out of range");
        if (space[k] != 2) throw std::invalid_argument("write_fcidump: orbital listed twice (core and active are disjoint)");
        space[k] = sp;
    };
    for (int c = 0; c < a.ncore; c++) take(a.core[c], 0);
    for (int p = 0; p < n; p++) take(a.active[p], 1);
    auto irr = [&](int p) { return a.full_irreps ? a.full_irreps[p] : 0; };
    int nirrep = 1;
    for (int p = 0; p < nfull; p++) {
        if (irr(p) < 0 || irr(p) >= kMaxIrrep) throw std::invalid_argument("write_fcidump: irrep label out of range");
        while (irr(p) >= nirrep) nirrep *= 2;
    }
    for (int u = 0; u < n; u++) {
        const int pz = (int)(std::find(L.to_user.begin(), L.to_user.end(), u) - L.to_user.begin());
        if (irr(a.active[u]) != L.sym[pz])
            throw std::invalid_argument("write_fcidump: active orbital irreps differ from the instance's");
    }
    // labels: core, active, virtual; by irrep, then index (fcidump.cc:83-107)
    std::vector<int> map(nfull, 0);
    int count = 0;
    for (int sp = 0; sp < 3; sp++)
        for (int h = 0; h < nirrep; h++)
            for (int p = 0; p < nfull; p++)
                if (space[p] == sp && irr(p) == h) map[p] = ++count;

    // two-electron integrals: G = Qmo^T Qmo on the device, one copy back
    const int64_t ntri = (int64_t)nfull * (nfull + 1) / 2;
    if ((double)ntri * (double)ntri * 8.0 > 64.0 * (1 << 30))
        throw std::invalid_argument("write_fcidump: the pair Gram matrix of " + std::to_string(nfull) +
                                    " orbitals exceeds 64 GiB (ntri^2 doubles, on the device and the host)");
    std::vector<double> G((size_t)(ntri * ntri));
    {
        DBuf<double> dq, dg;
        dq.alloc((size_t)a.nq * ntri);
        dg.alloc((size_t)(ntri * ntri));
        V2H_HIP(hipMemcpyAsync(dq.p, a.qmo, sizeof(double) * a.nq * ntri, hipMemcpyHostToDevice, stream_));
        tstart(kKindGram);
        launch_pair_gram(dq.p, ntri, a.nq, dg.p, stream_);
        tstop(kKindGram);
        V2H_HIP(hipGetLastError());
        V2H_HIP(hipMemcpyAsync(G.data(), dg.p, sizeof(double) * ntri * ntri, hipMemcpyDeviceToHost, stream_));
        V2H_HIP(hipStreamSynchronize(stream_));
        tflush();
    }
    RecordFile fi;
    fi.buf.reserve((size_t)(ntri * (ntri + 1) / 2 / nirrep + nfull * nfull + 1) * 24);
    for (int p = 0; p < nfull; p++)
        for (int q = p; q < nfull; q++) {
            const int hpq = irr(p) ^ irr(q);
            const int64_t pq = tri(p, q);
            const double* row = &G[(size_t)(pq * ntri)];
            for (int r = 0; r < nfull; r++)
                for (int s = r; s < nfull; s++) {
                    if ((irr(r) ^ irr(s)) != hpq) continue;
                    const int64_t rs = tri(r, s);
                    if (pq > rs) continue;
                    fi.put(row[rs], map[p], map[q], map[r], map[s]);
                }
        }
    for (int h = 0; h < nirrep; h++)
        for (int p = 0; p < nfull; p++) {
            if (irr(p) != h) continue;
            for (int q = 0; q < nfull; q++)
                if (irr(q) == h) fi.put(a.h_full[(size_t)p * nfull + q], map[p], map[q], 0, 0);
        }
    fi.put(a.enuc, 0, 0, 0, 0);

    // RDMs of the current primal vector (fcidump.cc:170-243)
    std::vector<double> x(n_primal(i));
    get_primal(i, x.data());
    const int ninact = a.ncore;
    auto lab = [&](int pz) { return map[a.active[L.to_user[pz]]] - ninact; };
    RecordFile fr;
    for (int h = 0; h < L.H; h++) {
        const int g = L.gab[h], ga = L.gaa[h];
        for (int ij = 0; ij < g; ij++) {
            const int pi = L.ab_pairs[h][ij][0], pj = L.ab_pairs[h][ij][1], ji = L.ab(h, pj, pi);
            for (int kl = 0; kl < g; kl++) {
                const int pk = L.ab_pairs[h][kl][0], pl = L.ab_pairs[h][kl][1], lk = L.ab(h, pl, pk);
                double dum = x[L.off[D2AB][h] + (int64_t)ij * g + kl] + x[L.off[D2AB][h] + (int64_t)ji * g + lk];
                if (pi != pj && pk != pl) {
                    const int ija = L.aa(h, pi, pj), kla = L.aa(h, pk, pl);
                    int sg = 1;
                    if (pi > pj) sg = -sg;
                    if (pk > pl) sg = -sg;
                    dum += sg * x[L.off[D2AA][h] + (int64_t)ija * ga + kla];
                    dum += sg * x[L.off[D2BB][h] + (int64_t)ija * ga + kla];
                }
                fr.put(dum, lab(pi), lab(pj), lab(pk), lab(pl));
            }
        }
    }
    for (int h = 0; h < L.H; h++)
        for (int ij = 0; ij < L.gab[h]; ij++) {
            const int pi = L.ab_pairs[h][ij][0], pj = L.ab_pairs[h][ij][1];
            if (L.sym[pi] != L.sym[pj]) continue;
            const double dum = x[L.one_body(D1A, pi, pj)] + x[L.one_body(D1B, pi, pj)];
            fr.put(dum, lab(pi), lab(pj), 0, 0);
        }
    fi.write(a.int_path);
    fr.write(a.rdm_path);
    if (n_int) *n_int = fi.count();
    if (n_rdm) *n_rdm = fr.count();
}

}  // namespace v2h
