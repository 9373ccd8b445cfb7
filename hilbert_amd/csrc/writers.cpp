// writers.cpp -- PSIO RDM records and the IWL TPDM (see writers.hpp).
#include "writers.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace v2h {

static_assert(sizeof(OpdmEntry) == 16, "opdm record: int i, j; double value (v2rdm_solver.h:53-57)");
static_assert(sizeof(TpdmEntry) == 24, "tpdm record: int i, j, k, l; double value (v2rdm_solver.h:59-65)");

namespace {

struct Ctx {
    const Layout& L;
    const double* x;
    const FullSpace& f;
    int full(int p) const { return f.active[L.to_user[p]]; }  // Pitzer position -> full-space label
    double d1(bool beta, int p, int q) const { return x[L.one_body(beta ? D1B : D1A, p, q)]; }
    double d2ab(int h, int ij, int kl) const { return x[L.off[D2AB][h] + (int64_t)ij * L.gab[h] + kl]; }
    double d2same(bool beta, int h, int ij, int kl) const {
        return x[L.off[beta ? D2BB : D2AA][h] + (int64_t)ij * L.gaa[h] + kl];
    }
};

void check(const Layout& L, const FullSpace& f) {
    if ((int)f.active.size() != L.n) throw std::invalid_argument("one full-space label per active orbital needed");
}

// The active-active part: geminal pairs of every irrep, as basis.cc orders them.
template <class Emit>
void active_pairs(const Ctx& c, Emit&& emit) {
    const Layout& L = c.L;
    for (int h = 0; h < L.H; h++)
        for (int ij = 0; ij < L.gab[h]; ij++) {
            const int i = L.ab_pairs[h][ij][0], j = L.ab_pairs[h][ij][1];
            for (int kl = 0; kl < L.gab[h]; kl++) {
                const int k = L.ab_pairs[h][kl][0], l = L.ab_pairs[h][kl][1];
                const double vab = c.d2ab(h, ij, kl);
                double vaa = 0.0, vbb = 0.0;
                const bool same = i != j && k != l;
                if (same) {
                    const int ija = L.aa(h, i, j), kla = L.aa(h, k, l);
                    const double s = ((i < j) ? 1.0 : -1.0) * ((k < l) ? 1.0 : -1.0);
                    vaa = s * c.d2same(false, h, ija, kla);
                    vbb = s * c.d2same(true, h, ija, kla);
                }
                emit(c.full(i), c.full(j), c.full(k), c.full(l), vab, same, vaa, vbb);
            }
        }
}

// The core-active part: every core orbital with every active pair (j, l) of
// one irrep, D1 values.
template <class Emit>
void core_active(const Ctx& c, Emit&& emit) {
    const Layout& L = c.L;
    for (const int ci : c.f.core)
        for (int h = 0; h < L.H; h++)
            for (int j = 0; j < L.amopi[h]; j++)
                for (int l = 0; l < L.amopi[h]; l++) {
                    const int p = L.poff[h] + j, q = L.poff[h] + l;
                    emit(ci, c.full(p), c.full(q), c.d1(false, p, q), c.d1(true, p, q));
                }
}

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

File open_out(const std::string& path) {
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) throw std::invalid_argument("cannot open " + path + " for writing");
    return f;
}

void put(FILE* f, const void* p, size_t n, const std::string& what) {
    if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("write failed: " + what);
}

// One keyed entry of a PSIO unit: an 80-byte key (PSIO_KEYLEN), the byte
// count, the bytes.
void entry(FILE* f, const char* key, const void* data, int64_t nbytes) {
    char k[80] = {};
    std::strncpy(k, key, sizeof(k) - 1);
    put(f, k, sizeof(k), key);
    put(f, &nbytes, sizeof(nbytes), key);
    put(f, data, (size_t)nbytes, key);
}

// The "IWL Buffers" entry: buffers of kIwlIntsPerBuf values, each {int
// lastbuf, int inbuf, int16 labels[4 n], double values[n]}, unused slots zero,
// the last one (possibly empty) with lastbuf = 1; values at or below the
// cutoff in magnitude are dropped, as IWL::write_value drops them.
int64_t iwl_entry(FILE* f, const std::vector<TpdmEntry>& recs) {
    constexpr int n = kIwlIntsPerBuf;
    std::vector<int16_t> lab(4 * (size_t)n);
    std::vector<double> val(n);
    std::vector<char> buf;
    int inbuf = 0;
    int64_t kept = 0;
    auto flush = [&](int lastbuf) {
        for (int i = inbuf; i < n; i++) {
            lab[4 * i] = lab[4 * i + 1] = lab[4 * i + 2] = lab[4 * i + 3] = 0;
            val[i] = 0.0;
        }
        const size_t off = buf.size();
        buf.resize(off + 2 * sizeof(int) + lab.size() * sizeof(int16_t) + val.size() * sizeof(double));
        char* q = buf.data() + off;
        std::memcpy(q, &lastbuf, sizeof(int));
        std::memcpy(q + sizeof(int), &inbuf, sizeof(int));
        std::memcpy(q + 2 * sizeof(int), lab.data(), lab.size() * sizeof(int16_t));
        std::memcpy(q + 2 * sizeof(int) + lab.size() * sizeof(int16_t), val.data(), val.size() * sizeof(double));
        inbuf = 0;
    };
    for (const auto& r : recs) {
        if (!(std::fabs(r.value) > kIwlCutoff)) continue;
        if (r.i > 32767 || r.j > 32767 || r.k > 32767 || r.l > 32767)
            throw std::invalid_argument("IWL labels are 16-bit");
        lab[4 * inbuf] = (int16_t)r.i;
        lab[4 * inbuf + 1] = (int16_t)r.j;
        lab[4 * inbuf + 2] = (int16_t)r.k;
        lab[4 * inbuf + 3] = (int16_t)r.l;
        val[inbuf] = r.value;
        kept++;
        if (++inbuf == n) flush(0);
    }
    flush(1);
    entry(f, "IWL Buffers", buf.data(), (int64_t)buf.size());
    return kept;
}

}  // namespace

std::vector<OpdmEntry> psio_opdm(const Layout& L, const double* x, const FullSpace& f, bool beta) {
    check(L, f);
    const Ctx c{L, x, f};
    std::vector<OpdmEntry> out;
    // active-active (write_opdm.cc:57-84)
    for (int h = 0; h < L.H; h++)
        for (int i = 0; i < L.amopi[h]; i++)
            for (int j = 0; j < L.amopi[h]; j++) {
                const int p = L.poff[h] + i, q = L.poff[h] + j;
                out.push_back({c.full(p), c.full(q), c.d1(beta, p, q)});
            }
    // core-core (write_opdm.cc:87-106)
    for (const int ci : f.core) out.push_back({ci, ci, 1.0});
    return out;
}

std::vector<TpdmEntry> psio_tpdm(const Layout& L, const double* x, const FullSpace& f, int part,
                                 bool active_only) {
    check(L, f);
    if (part < 0 || part > 2) throw std::invalid_argument("part: 0 aa, 1 bb, 2 ab");
    const Ctx c{L, x, f};
    const bool beta = part == 1;
    std::vector<TpdmEntry> out;
    // active-active (write_tpdm.cc:60-103, 536-584)
    active_pairs(c, [&](int i, int j, int k, int l, double vab, bool same, double vaa, double vbb) {
        if (part == 2)
            out.push_back({i, j, k, l, vab});
        else if (same)
            out.push_back({i, j, k, l, beta ? vbb : vaa});
    });
    if (active_only) return out;
    // core-core (write_tpdm.cc:106-148)
    for (const int ci : f.core)
        for (const int cj : f.core) {
            if (part == 2) {
                out.push_back({ci, cj, ci, cj, 1.0});
            } else if (ci != cj) {
                out.push_back({ci, cj, ci, cj, 1.0});
                out.push_back({ci, cj, cj, ci, -1.0});
            }
        }
    // core-active (write_tpdm.cc:151-256): D2(ij;il) = D1(j,l) and its images
    core_active(c, [&](int i, int j, int l, double va, double vb) {
        if (part == 2) {
            out.push_back({i, j, i, l, vb});  // ab (ij;il)
            out.push_back({j, i, l, i, va});  // ba (ji;li)
        } else {
            const double v = beta ? vb : va;
            out.push_back({i, j, i, l, v});    // ij;il
            out.push_back({i, j, l, i, -v});   // ij;li
            out.push_back({j, i, l, i, v});    // ji;li
            out.push_back({j, i, i, l, -v});   // ji;il
        }
    });
    return out;
}

std::vector<TpdmEntry> psio_tpdm_spin_free(const Layout& L, const double* x, const FullSpace& f) {
    check(L, f);
    const Ctx c{L, x, f};
    std::vector<TpdmEntry> out;
    // active-active (write_tpdm.cc:326-374): D2ab(ij;kl) + D2ab(ji;lk) plus
    // sg (D2aa + D2bb)(ij;kl) over the stored pair, where hilbert takes sg = -1
    // as soon as i > j or k > l (so also when both hold)
    for (int h = 0; h < L.H; h++)
        for (int ij = 0; ij < L.gab[h]; ij++) {
            const int i = L.ab_pairs[h][ij][0], j = L.ab_pairs[h][ij][1];
            const int ji = L.ab(h, j, i);
            for (int kl = 0; kl < L.gab[h]; kl++) {
                const int k = L.ab_pairs[h][kl][0], l = L.ab_pairs[h][kl][1];
                const int lk = L.ab(h, l, k);
                double v = c.d2ab(h, ij, kl) + c.d2ab(h, ji, lk);
                if (i != j && k != l) {
                    const int ija = L.aa(h, i, j), kla = L.aa(h, k, l);
                    const double sg = (i > j || k > l) ? -1.0 : 1.0;
                    v += sg * c.d2same(false, h, ija, kla);
                    v += sg * c.d2same(true, h, ija, kla);
                }
                out.push_back({c.full(i), c.full(j), c.full(k), c.full(l), v});
            }
        }
    // core-core (write_tpdm.cc:377-420)
    for (const int ci : f.core)
        for (const int cj : f.core) {
            out.push_back({ci, cj, ci, cj, ci != cj ? 4.0 : 2.0});
            if (ci != cj) out.push_back({ci, cj, cj, ci, -2.0});
        }
    // core-active (write_tpdm.cc:423-486)
    core_active(c, [&](int i, int j, int l, double va, double vb) {
        const double d = va + vb + vb + va, e = -va - vb;
        out.push_back({i, j, i, l, d});  // ij;il
        out.push_back({i, j, l, i, e});  // ij;li
        out.push_back({j, i, l, i, d});  // ji;li
        out.push_back({j, i, i, l, e});  // ji;il
    });
    return out;
}

std::vector<TpdmEntry> iwl_tpdm(const Layout& L, const double* x, const FullSpace& f, int part) {
    check(L, f);
    if (part < 0 || part > 2) throw std::invalid_argument("part: 0 aa, 1 bb, 2 ab");
    const Ctx c{L, x, f};
    const bool beta = part == 1;
    std::vector<TpdmEntry> out;
    // active-active (write_tpdm_iwl.cc:56-95): labels (i, k, j, l), aa and bb x 1/4
    active_pairs(c, [&](int i, int j, int k, int l, double vab, bool same, double vaa, double vbb) {
        if (part == 2)
            out.push_back({i, k, j, l, vab});
        else if (same)
            out.push_back({i, k, j, l, 0.25 * (beta ? vbb : vaa)});
    });
    // core-core (write_tpdm_iwl.cc:98-126)
    for (const int ci : f.core)
        for (const int cj : f.core) {
            if (part == 2) {
                out.push_back({ci, ci, cj, cj, 1.0});
            } else if (ci != cj) {
                out.push_back({ci, ci, cj, cj, 0.25});
                out.push_back({ci, cj, cj, ci, -0.25});
            }
        }
    // core-active (write_tpdm_iwl.cc:129-183)
    core_active(c, [&](int i, int j, int l, double va, double vb) {
        if (part == 2) {
            out.push_back({i, i, j, l, vb});  // ab (ij;il)
            out.push_back({j, l, i, i, va});  // ba (ji;li)
        } else {
            const double v = 0.25 * (beta ? vb : va);
            out.push_back({i, i, j, l, v});    // ij;il
            out.push_back({i, l, j, i, -v});   // ij;li
            out.push_back({j, l, i, i, v});    // ji;li
            out.push_back({j, i, i, l, -v});   // ji;il
        }
    });
    return out;
}

namespace {

// The orbital lists a TPDM file ends with (write_tpdm.cc:262-300).
void orbital_lists(FILE* f, const Layout& L, const FullSpace& fs) {
    std::vector<int> act;
    for (int p = 0; p < L.n; p++) act.push_back(fs.active[L.to_user[p]]);
    const int namo = L.n, ninact = (int)fs.core.size();
    entry(f, "NUMBER ACTIVE ORBITALS", &namo, sizeof(int));
    entry(f, "ACTIVE ORBITALS", act.data(), (int64_t)act.size() * (int64_t)sizeof(int));
    entry(f, "NUMBER INACTIVE ORBITALS", &ninact, sizeof(int));
    entry(f, "INACTIVE ORBITALS", fs.core.data(), (int64_t)fs.core.size() * (int64_t)sizeof(int));
}

// A record entry and its "length" entry (write_opdm.cc:109-111, write_tpdm.cc:259-261).
template <class Rec>
int64_t records(FILE* f, const char* key, const std::vector<Rec>& recs) {
    const int64_t n = (int64_t)recs.size();
    entry(f, key, recs.data(), n * (int64_t)sizeof(Rec));
    entry(f, "length", &n, sizeof(n));
    return n;
}

void close_out(File& f, const std::string& path) {
    FILE* raw = f.release();
    if (std::fclose(raw) != 0) throw std::runtime_error("write failed: " + path);
}

}  // namespace

void write_rdm_files(const Layout& L, const double* x, const FullSpace& f, const std::string& prefix, int which,
                     int64_t* counts) {
    check(L, f);
    if (which & ~kRdmAll) throw std::invalid_argument("unknown RDM file flags");
    for (int k = 0; k < kRdmCounts; k++) counts[k] = -1;
    const char* tpdm_keys[3] = {"D2aa", "D2bb", "D2ab"};
    const char* tpdm_ext[3] = {".d2aa", ".d2bb", ".d2ab"};
    // TPDM_WRITE, then TPDM_WRITE_FULL over the same files (v2rdm_solver.cc:1474-1480)
    for (const bool full : {false, true}) {
        if (!(which & (full ? kRdmTpdmFull : kRdmTpdmActive))) continue;
        for (int part = 0; part < 3; part++) {
            const std::string path = prefix + tpdm_ext[part];
            File fo = open_out(path);
            counts[2 + part] = records(fo.get(), tpdm_keys[part], psio_tpdm(L, x, f, part, !full));
            if (full) orbital_lists(fo.get(), L, f);
            close_out(fo, path);
        }
    }
    if (which & kRdmTpdmSpinFree) {  // v2rdm_solver.cc:1481-1484
        const std::string path = prefix + ".d2";
        File fo = open_out(path);
        counts[5] = records(fo.get(), "D2", psio_tpdm_spin_free(L, x, f));
        orbital_lists(fo.get(), L, f);
        close_out(fo, path);
    }
    if (which & kRdmOpdmFull) {  // v2rdm_solver.cc:1485-1487
        const char* keys[2] = {"D1a", "D1b"};
        const char* ext[2] = {".d1a", ".d1b"};
        for (int s = 0; s < 2; s++) {
            const std::string path = prefix + ext[s];
            File fo = open_out(path);
            counts[s] = records(fo.get(), keys[s], psio_opdm(L, x, f, s == 1));
            close_out(fo, path);
        }
    }
    if (which & kRdmTpdmIwl) {  // v2rdm_solver.cc:1508-1509
        const char* ext[3] = {".aa.iwl", ".bb.iwl", ".ab.iwl"};
        for (int part = 0; part < 3; part++) {
            const std::string path = prefix + ext[part];
            File fo = open_out(path);
            counts[6 + part] = iwl_entry(fo.get(), iwl_tpdm(L, x, f, part));
            close_out(fo, path);
        }
    }
}

}  // namespace v2h
