// kernels.hip -- CDNA4 (gfx950) kernels of the batched BPSDP solve.
//
// * Gather kernels over SELL-64 tables (one wavefront per 64 rows, or per
//   long row).  Every CG step is three launches:
//     k_cg_at     : q = A^T p, with p = (r - alpha w) + beta p formed on the fly
//     k_cg_a      : w = A q, own-row updates of y, r, p, per-workgroup dot partials
//     k_cg_reduce : one workgroup per instance sums the partials in a fixed
//                   order (deterministic) and advances the CG state
// * k_pre (+ k_pre_reduce): A [v1, x, A^T y] in one pass (rhs pieces, A x),
//   then the end-of-iteration bookkeeping (errors, energies, mu, convergence,
//   CG threshold).
// * k_post : A^T y and U = mu x + A^T y - c.
// * k_project<NT, GLOBAL> : PSD projection of one block per workgroup by
//   parallel cyclic Jacobi on the packed upper triangle (in LDS, or in global
//   memory for blocks too large for LDS); the rotations are recorded and
//   replayed in reverse on diag(max(L,0)) to form U+ without an explicit
//   eigenvector matrix.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "kernels.hpp"

namespace v2h {

namespace {

constexpr int kWG = 256;         // gather workgroup: 4 wavefronts
constexpr uint32_t kMarker = 255;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup id remapped so that each XCD (blocks are dealt round-robin over
// the 8 XCDs) runs a contiguous range of logical workgroups, keeping one
// instance's vectors in one XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int n) {
    int per = n >> 3, rem = n & 7;
    int x = bid & 7, loc = bid >> 3;
    return x * per + (x < rem ? x : rem) + loc;
}

// Logical workgroup of this block under the launch plan (runs of the
// workgroups of selected instances; identity without a plan).
__device__ __forceinline__ int unit_wg(const UnitsDev& U) {
    const int c = U.remap ? xcd_remap(blockIdx.x, U.grid) : (int)blockIdx.x;
    if (U.plan == nullptr) return c;
    int lo = 0, hi = U.plan_m;  // plan[lo] <= c < plan[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (U.plan[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return U.plan[U.plan_m + 1 + lo] + (c - U.plan[lo]);
}

// The CG vector recurrences, shared by the gather (on the fly) and the
// owning-row update so both produce identical bits.
__device__ __forceinline__ double r0_of(double av1, double b, double ax, double aaty, double mu) {
    return __fma_rn(mu, b - ax, av1) - aaty;
}
__device__ __forceinline__ double r_next(double r, double w, double al) { return __fma_rn(-al, w, r); }
__device__ __forceinline__ double p_next(double rn, double p, double be) { return __fma_rn(be, p, rn); }

struct CgScalars {
    double al, be;
};
__device__ __forceinline__ CgScalars cg_scalars(const InstState& st) {
    CgScalars s{0.0, 0.0};
    if (st.cg_iter > 0) {
        s.al = st.rr / st.pw;
        double rr1 = st.rr - 2.0 * s.al * st.rw + s.al * s.al * st.ww;
        s.be = rr1 / st.rr;
    }
    return s;
}

// Gather over one SELL slice (lane = row) or one long row (whole wave).
// Load(col, coef) accumulates into the caller's accumulators.
template <class Load>
__device__ __forceinline__ int64_t gather_unit(const SellDev& S, const double* coef, int unit, int lane,
                                               Load&& load, bool& is_long) {
    is_long = (unit & kLongFlag) != 0;
    if (!is_long) {
        uint32_t beg = S.slice_start[unit], end = S.slice_start[unit + 1];
        const uint32_t* e = S.ent + (size_t)beg * 64 + lane;
        int w = (int)(end - beg);
        bool skip = false;
        int k = 0;
        for (; k + 4 <= w; k += 4) {
            uint32_t v0 = e[(size_t)(k + 0) * 64], v1 = e[(size_t)(k + 1) * 64];
            uint32_t v2 = e[(size_t)(k + 2) * 64], v3 = e[(size_t)(k + 3) * 64];
            skip |= (v0 >> 24) == kMarker;
            load(v0 & 0xFFFFFFu, coef[v0 >> 24]);
            load(v1 & 0xFFFFFFu, coef[v1 >> 24]);
            load(v2 & 0xFFFFFFu, coef[v2 >> 24]);
            load(v3 & 0xFFFFFFu, coef[v3 >> 24]);
        }
        for (; k < w; k++) {
            uint32_t v = e[(size_t)k * 64];
            skip |= (v >> 24) == kMarker;
            load(v & 0xFFFFFFu, coef[v >> 24]);
        }
        int64_t row = (int64_t)unit * 64 + lane;
        return (skip || row >= S.nrows) ? -1 : row;
    }
    int li = unit & ~kLongFlag;
    uint32_t beg = S.long_start[li], end = S.long_start[li + 1];
    for (uint32_t e = beg + lane; e < end; e += 64) {
        uint32_t v = S.long_ent[e];
        load(v & 0xFFFFFFu, coef[v >> 24]);
    }
    return S.long_row[li];  // caller reduces the accumulators over the wave
}

template <int N>
__device__ __forceinline__ void block_partials(const double (&v)[N], double* red, double* out) {
    // red: [4][N] in LDS; out: N partials written by thread 0
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double s[N];
#pragma unroll
    for (int i = 0; i < N; i++) s[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; i++) red[wv * N + i] = s[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; i++) out[i] = ((red[i] + red[N + i]) + red[2 * N + i]) + red[3 * N + i];
    }
}

// Sum the nwg partial records (N doubles each) of one instance in a fixed
// order; every thread of the workgroup receives tot[N].  Runs in its own
// launch after the producing kernel: the kernel boundary makes every XCD's
// writes visible without device-scope fences inside the producer.
template <int N>
__device__ void reduce_partials(const double* parts, int nwg, double* red, double (&tot)[N]) {
    double s[N];
#pragma unroll
    for (int i = 0; i < N; i++) s[i] = 0.0;
    for (int k = threadIdx.x; k < nwg; k += blockDim.x) {
#pragma unroll
        for (int i = 0; i < N; i++) s[i] += parts[(size_t)k * N + i];
    }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; i++) s[i] = wave_sum(s[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; i++) red[wv * N + i] = s[i];
    }
    __syncthreads();
    int nw = blockDim.x >> 6;
#pragma unroll
    for (int i = 0; i < N; i++) {
        double t = 0.0;
        for (int q = 0; q < nw; q++) t += red[q * N + i];
        tot[i] = t;
    }
}

__device__ __forceinline__ void load_coef(double* dst, const double* src) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

// ------------------------------------------------------------------------
// Gather launches.  Every wavefront starts from its UnitRec (one scalar load:
// table entries, coefficient table, first row), so its table loads, the
// workgroup's coefficient table and the instance state are in flight together;
// the first kPre entries of each lane are loaded before the set-up barrier.

constexpr int kPre = 4;

struct WaveGather {
    UnitRec R;
    int lane;
    bool is_long;
    uint32_t pre[kPre];
    __device__ __forceinline__ WaveGather(const UnitRec& r, int l) : R(r), lane(l) {
        is_long = (R.meta & kRecLong) != 0u;
        const int w = is_long ? 0 : (int)(R.meta & 0xFFFFu);
#pragma unroll
        for (int k = 0; k < kPre; k++) pre[k] = (k < w) ? R.ent[(size_t)k * 64 + lane] : 0u;
    }
    // first row of this lane (slices), before the entries are known
    __device__ __forceinline__ int64_t lane_row() const {
        const int nvalid = (int)((R.meta >> 16) & 0x7Fu);
        return (!is_long && lane < nvalid) ? (int64_t)R.row + lane : -1;
    }
    // Load(col, coef) over the lane's entries; returns the row the lane
    // writes (slices: -1 for padding lanes and long-row markers; long rows:
    // the row, after the caller's wave reduction, on lane 0).
    template <class Load>
    __device__ __forceinline__ int64_t run(const double* coef, Load&& load) {
        if (!is_long) {
            const int w = (int)(R.meta & 0xFFFFu);
            const bool skip = w > 0 && (pre[0] >> 24) == kMarker;
#pragma unroll
            for (int k = 0; k < kPre; k++)
                if (k < w) load(pre[k] & 0xFFFFFFu, coef[pre[k] >> 24]);
            const uint32_t* e = R.ent + lane;
            int k = kPre;
            for (; k + 4 <= w; k += 4) {
                const uint32_t v0 = e[(size_t)(k + 0) * 64], v1 = e[(size_t)(k + 1) * 64];
                const uint32_t v2 = e[(size_t)(k + 2) * 64], v3 = e[(size_t)(k + 3) * 64];
                load(v0 & 0xFFFFFFu, coef[v0 >> 24]);
                load(v1 & 0xFFFFFFu, coef[v1 >> 24]);
                load(v2 & 0xFFFFFFu, coef[v2 >> 24]);
                load(v3 & 0xFFFFFFu, coef[v3 >> 24]);
            }
            for (; k < w; k++) {
                const uint32_t v = e[(size_t)k * 64];
                load(v & 0xFFFFFFu, coef[v >> 24]);
            }
            return skip ? -1 : lane_row();
        }
        const int n = (int)(R.meta & ~kRecLong);
        for (int k = lane; k < n; k += 64) {
            const uint32_t v = R.ent[k];
            load(v & 0xFFFFFFu, coef[v >> 24]);
        }
        return R.row;
    }
};

// Coefficient table of the workgroup's structure into LDS (every wavefront of
// a workgroup belongs to one instance).
__device__ __forceinline__ void put_coef(double* coef, double v) {
    coef[threadIdx.x] = v;
    __syncthreads();
}

// CG step, part 1: q = A^T p
__global__ __launch_bounds__(kWG) void k_cg_at(BatchDev B, UnitsDev U) {
    static_assert(kWG == 256, "one coefficient per thread");
    const int wg = unit_wg(U);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int inst = U.wg_inst[wg];
    const UnitRec R = U.rec[(size_t)wg * 4 + wv];
    const double cv = R.coef[threadIdx.x];
    WaveGather G(R, lane);
    const InstState& st = B.st[inst];
    const InstDev I = B.inst[inst];
    if (!st.active || st.cg_done) return;
    __shared__ double coef[256];
    put_coef(coef, cv);
    const CgScalars cs = cg_scalars(st);
    const double mu = st.mu;
    double acc = 0.0;
    int64_t row;
    if (st.cg_iter == 0) {
        const double *av1 = B.av1 + I.doff, *ax = B.ax + I.doff, *aaty = B.aaty + I.doff;
        const double* b = B.structs[I.sid].b;
        row = G.run(coef, [&](uint32_t col, double cf) {
            acc = __fma_rn(cf, r0_of(av1[col], b[col], ax[col], aaty[col], mu), acc);
        });
    } else {
        const CgRow* cg = B.cg + I.doff;
        row = G.run(coef, [&](uint32_t col, double cf) {
            const double2 rw = *reinterpret_cast<const double2*>(&cg[col].r);
            acc = __fma_rn(cf, p_next(r_next(rw.x, rw.y, cs.al), cg[col].p, cs.be), acc);
        });
    }
    if (G.is_long) {
        acc = wave_sum(acc);
        if (lane != 0) row = -1;
    }
    if (row >= 0) B.q[I.poff + row] = acc;
}

// CG step, part 2: w = A q, own-row updates of y, r, p and the dot partials.
__global__ __launch_bounds__(kWG) void k_cg_a(BatchDev B, UnitsDev U) {
    const int wg = unit_wg(U);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int inst = U.wg_inst[wg];
    const UnitRec R = U.rec[(size_t)wg * 4 + wv];
    const double cv = R.coef[threadIdx.x];
    WaveGather G(R, lane);
    const InstState& st = B.st[inst];
    const InstDev I = B.inst[inst];
    if (!st.active || st.cg_done) return;
    __shared__ double coef[256];
    __shared__ double red[4 * 4];
    put_coef(coef, cv);
    const CgScalars cs = cg_scalars(st);
    const double mu = st.mu;
    const int s = st.cg_iter;
    // the own row's state is loaded with the gather
    const int64_t r0 = G.lane_row();
    CgRow c{0.0, 0.0, 0.0, 0.0};
    if (r0 >= 0) c = B.cg[I.doff + r0];
    const double* q = B.q + I.poff;
    double acc = 0.0;
    int64_t row = G.run(coef, [&](uint32_t col, double cf) { acc = __fma_rn(cf, q[col], acc); });
    if (G.is_long) {
        acc = wave_sum(acc);
        if (lane != 0) row = -1;
        if (row >= 0) c = B.cg[I.doff + row];
    }
    double d4[4] = {0.0, 0.0, 0.0, 0.0};
    if (row >= 0) {
        const int64_t g = I.doff + row;
        double rn, pn;
        if (s == 0) {
            rn = r0_of(B.av1[g], B.structs[I.sid].b[row], B.ax[g], B.aaty[g], mu);
            pn = rn;
        } else {
            c.y = __fma_rn(cs.al, c.p, c.y);
            rn = r_next(c.r, c.w, cs.al);
            pn = p_next(rn, c.p, cs.be);
        }
        c.r = rn;
        c.w = acc;
        c.p = pn;
        B.cg[g] = c;
        d4[0] = rn * rn;
        d4[1] = pn * acc;
        d4[2] = rn * acc;
        d4[3] = acc * acc;
    }
    block_partials<4>(d4, red, B.partials + (size_t)wg * 4);
}

// CG step, part 3 (one workgroup per instance): reduce the dots of k_cg_a and
// advance the CG state.
__global__ __launch_bounds__(kWG) void k_cg_reduce(BatchDev B, UnitsDev U) {
    const int inst = blockIdx.x;
    InstState& st = B.st[inst];
    if (!st.active || st.cg_done || !in_pgroup(B, inst)) return;
    __shared__ double red[4 * 4];
    const int s = st.cg_iter;
    double tot[4];
    reduce_partials<4>(B.partials + (size_t)U.wg_begin[inst] * 4, U.wg_begin[inst + 1] - U.wg_begin[inst], red, tot);
    if (threadIdx.x == 0) {
        st.rr = tot[0];
        st.pw = tot[1];
        st.rw = tot[2];
        st.ww = tot[3];
        bool go_on = (s < B.opt.cg_maxiter) && (tot[0] > st.cg_thr) && (tot[1] > 0.0);
        if (!isfinite(tot[0]) || !isfinite(tot[1])) {
            st.status |= kStatusNonFinite;
            go_on = false;
        }
        if (go_on) {
            st.cg_iter = s + 1;
        } else {
            st.cg_done = 1;
            st.cg_last = s;
            st.iiter += s;
        }
    }
}

// Start of an outer iteration: A [v1, x, aty] -> av1, ax, aaty; finishes the
// previous iteration (errors, energies, mu, convergence) and sets up the CG.
__global__ __launch_bounds__(kWG) void k_pre(BatchDev B, UnitsDev U) {
    const int wg = unit_wg(U);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int inst = U.wg_inst[wg];
    const UnitRec R = U.rec[(size_t)wg * 4 + wv];
    const double cv = R.coef[threadIdx.x];
    WaveGather G(R, lane);
    const InstState& st = B.st[inst];
    const InstDev I = B.inst[inst];
    if (!st.active) return;
    __shared__ double coef[256];
    __shared__ double red[4 * 4];
    put_coef(coef, cv);
    const double *v1 = B.v1 + I.poff, *x = B.x + I.poff, *aty = B.aty + I.poff;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    int64_t row = G.run(coef, [&](uint32_t col, double cf) {
        a0 = __fma_rn(cf, v1[col], a0);
        a1 = __fma_rn(cf, x[col], a1);
        a2 = __fma_rn(cf, aty[col], a2);
    });
    if (G.is_long) {
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        a2 = wave_sum(a2);
        if (lane != 0) row = -1;
    }
    double d4[4] = {0.0, 0.0, 0.0, 0.0};
    if (row >= 0) {
        const int64_t g = I.doff + row;
        B.av1[g] = a0;
        B.ax[g] = a1;
        B.aaty[g] = a2;
        const double bv = B.structs[I.sid].b[row];
        const double e = bv - a1;
        d4[0] = a0 * a0;
        d4[1] = a0 * e;
        d4[2] = e * e;
        d4[3] = bv * B.cg[g].y;
    }
    block_partials<4>(d4, red, B.partials + (size_t)wg * 4);
}

// Iteration start, part 2 (one workgroup per instance): book-keeping of the
// finished iteration (errors, energies, mu, convergence) and CG set-up.
__global__ __launch_bounds__(kWG) void k_pre_reduce(BatchDev B, UnitsDev U) {
    const int inst = blockIdx.x;
    InstState& st = B.st[inst];
    if (!st.active || !in_pgroup(B, inst)) return;
    const InstDev I = B.inst[inst];
    __shared__ double red[4 * 4];
    double tot[4];
    reduce_partials<4>(B.partials + (size_t)U.wg_begin[inst] * 4, U.wg_begin[inst + 1] - U.wg_begin[inst], red, tot);
    if (threadIdx.x == 0) {
        const SolverOptions& o = B.opt;
        if (st.pending) {
            double ep = sqrt(tot[2]) / st.bnorm;
            double Ep = 0.0, dres = 0.0;
            for (int k = 0; k < I.nblocks; k++) {
                Ep += B.proj_part[2 * (size_t)(I.block_first + k)];
                dres += B.proj_part[2 * (size_t)(I.block_first + k) + 1];
            }
            double ed = sqrt(dres) / st.cnorm;
            st.eps_p = ep;
            st.eps_d = ed;
            st.e_primal = Ep;
            st.e_dual = tot[3];
            st.oiter += 1;
            st.maxiter_left -= 1;
            if (o.mu_update_frequency > 0 && st.oiter % o.mu_update_frequency == 0 && ed > 0.0 && ep > 0.0)
                st.mu *= ep / ed;
            if (fabs(Ep - tot[3]) < o.e_convergence && ep < o.r_convergence && ed < o.r_convergence) {
                st.converged = 1;
                st.active = 0;
            }
            if (!isfinite(Ep) || !isfinite(ep) || !isfinite(ed)) {
                st.status |= kStatusNonFinite;
                st.active = 0;
            }
            if (st.oiter >= o.maxiter || st.maxiter_left <= 0) st.active = 0;
        }
        st.pending = 0;
        double e = st.oiter == 0 ? 1.0 : fmax(st.eps_p, st.eps_d);
        double tol = o.cg_tol_factor * e;
        if (tol > kCgTolCap) tol = kCgTolCap;
        if (tol < o.cg_convergence) tol = o.cg_convergence;
        double mu = st.mu;
        double rhs2 = tot[0] + 2.0 * mu * tot[1] + mu * mu * tot[2];
        st.cg_thr = tol * tol * rhs2;
        st.cg_iter = 0;
        st.cg_done = 0;
    }
}

// After the CG: aty = A^T y, u = mu x + aty - c.
__global__ __launch_bounds__(kWG) void k_post(BatchDev B, UnitsDev U) {
    const int wg = unit_wg(U);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int inst = U.wg_inst[wg];
    const UnitRec R = U.rec[(size_t)wg * 4 + wv];
    const double cv = R.coef[threadIdx.x];
    WaveGather G(R, lane);
    const InstState& st = B.st[inst];
    const InstDev I = B.inst[inst];
    if (!st.active) return;
    __shared__ double coef[256];
    put_coef(coef, cv);
    const CgRow* cg = B.cg + I.doff;
    double acc = 0.0;
    int64_t row = G.run(coef, [&](uint32_t col, double cf) { acc = __fma_rn(cf, cg[col].y, acc); });
    if (G.is_long) {
        acc = wave_sum(acc);
        if (lane != 0) row = -1;
    }
    if (row >= 0) {
        const int64_t g = I.poff + row;
        B.aty[g] = acc;
        B.u[g] = __fma_rn(st.mu, B.x[g], acc) - B.c[g];
    }
}

// Plain out = A in  or  out = A^T in, for one instance (API and tests).
__global__ __launch_bounds__(kWG) void k_spmv_plain(SellDev S, const int4* units, int wg0, const double* in,
                                                    double* out) {
    __shared__ double coef[256];
    load_coef(coef, S.coef);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int unit = ((const int*)&units[wg0 + blockIdx.x])[wv];
    if (unit < 0) return;
    double acc = 0.0;
    bool lng;
    int64_t row = gather_unit(S, coef, unit, lane, [&](uint32_t col, double cf) { acc = __fma_rn(cf, in[col], acc); }, lng);
    if (lng) {
        acc = wave_sum(acc);
        if (lane != 0) row = -1;
    }
    if (row >= 0) out[row] = acc;
}

// ------------------------------------------------------------------------
// PSD projection by parallel cyclic Jacobi (round-robin ordering).
//
// Packed upper triangle: element (i, j), i <= j, at j (j + 1) / 2 + i.
__device__ __forceinline__ int pk(int i, int j) { return i <= j ? (j * (j + 1) >> 1) + i : (i * (i + 1) >> 1) + j; }

// Pair k of round t for dp (even) indices: the circle method with index
// dp - 1 fixed.  Returns p < q.
__device__ __forceinline__ void round_pair(int t, int k, int dp, int& p, int& q) {
    int M = dp - 1;
    int a, b;
    if (k == 0) {
        a = t;
        b = M;
    } else {
        a = t + k;
        if (a >= M) a -= M;
        b = t - k;
        if (b < 0) b += M;
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// Triangular tile index e -> (P, Q), P <= Q.
__device__ __forceinline__ void tri_pair(int e, int& P, int& Q) {
    int qq = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
    while ((qq + 1) * (qq + 2) / 2 <= e) qq++;
    while (qq * (qq + 1) / 2 > e) qq--;
    Q = qq;
    P = e - qq * (qq + 1) / 2;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < NT / 64; q++) t += red[q];
    return t;
}

// One round of two-sided rotations on the packed matrix A.  sgn = +1 applies
// G_P T G_Q^T (forward), sgn = -1 applies G_P^T T G_Q (replay).  In the
// forward direction the 2x2 diagonal tiles are set to their closed form.
template <int NT>
__device__ void apply_round(double* A, const int* pp, const int* qq, const double* rc, const double* rs,
                            int m, double sgn, bool forward) {
    const int ntile = m * (m + 1) / 2;
    for (int e = threadIdx.x; e < ntile; e += NT) {
        int P, Q;
        tri_pair(e, P, Q);
        const double cP = rc[P], sP = sgn * rs[P];
        const int p = pp[P], q = qq[P];
        if (P == Q) {
            const int ipp = pk(p, p), ipq = pk(p, q), iqq = pk(q, q);
            double app = A[ipp], apq = A[ipq], aqq = A[iqq];
            if (forward) {
                if (rs[P] != 0.0) {
                    double t = rs[P] / rc[P];
                    A[ipp] = app - t * apq;
                    A[iqq] = aqq + t * apq;
                    A[ipq] = 0.0;
                }
            } else {
                // T = G^T T' G with G = [[c, -s], [s, c]] (s already negated by sgn)
                double r00 = cP * app - sP * apq, r01 = cP * apq - sP * aqq;
                double r10 = sP * app + cP * apq, r11 = sP * apq + cP * aqq;
                A[ipp] = cP * r00 - sP * r01;
                A[ipq] = sP * r00 + cP * r01;
                A[iqq] = sP * r10 + cP * r11;
            }
            continue;
        }
        const double cQ = rc[Q], sQ = sgn * rs[Q];
        const int r = pp[Q], s = qq[Q];
        const int i0 = pk(p, r), i1 = pk(p, s), i2 = pk(q, r), i3 = pk(q, s);
        double a0 = A[i0], a1 = A[i1], a2 = A[i2], a3 = A[i3];
        // rows (p, q) by P
        double b0 = cP * a0 - sP * a2, b1 = cP * a1 - sP * a3;
        double b2 = sP * a0 + cP * a2, b3 = sP * a1 + cP * a3;
        // columns (r, s) by Q
        A[i0] = cQ * b0 - sQ * b1;
        A[i1] = sQ * b0 + cQ * b1;
        A[i2] = cQ * b2 - sQ * b3;
        A[i3] = sQ * b2 + cQ * b3;
    }
}

// Workgroup barrier that orders LDS only: outstanding global stores (the
// rotation records) and loads (the prefetched rotations) stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
template <bool GLOBAL>
__device__ __forceinline__ void round_sync() {
    if constexpr (GLOBAL) __syncthreads();
    else lds_barrier();
}

// GLOBAL: the packed matrix lives in global memory (blocks too large for LDS).
template <int NT, bool GLOBAL>
__global__ __launch_bounds__(NT) void k_project(BatchDev B, const ProjItem* items, double* rec_base, int max_sweeps,
                                                double* gmat) {
    extern __shared__ double smem[];
    const ProjItem it = items[blockIdx.x];
    InstState& st = B.st[it.inst];
    if (!st.active || !in_pgroup(B, it.inst)) return;
    const InstDev I = B.inst[it.inst];
    const int d = it.d, dp = d + (d & 1), m = dp >> 1;
    const int np = dp * (dp + 1) / 2;
    double* A = GLOBAL ? gmat + it.mat_off : smem;
    double* rc = GLOBAL ? smem : A + np;
    double* rs = rc + m;
    double* red = rs + m;
    int* pp = (int*)(red + NT / 64);
    int* qq = pp + m;
    const double mu = st.mu;
    const double* U = B.u + I.poff + it.off;
    double* rec = rec_base + it.rec_off;

    // symmetrised upper triangle, padded to even size with a zero row
    double nrm = 0.0;
    for (int e = threadIdx.x; e < dp * dp; e += NT) {
        int i = e / dp, j = e - i * dp;
        if (i > j) continue;
        double v = (j < d) ? 0.5 * (U[(size_t)i * d + j] + U[(size_t)j * d + i]) : 0.0;
        A[pk(i, j)] = v;
        nrm += (i == j ? 1.0 : 2.0) * v * v;
    }
    nrm = block_sum<NT>(nrm, red);
    const double thr = 1e-30 * nrm;  // (1e-15 ||A||_F)^2

    int nsweep = 0;
    for (; nsweep < max_sweeps; nsweep++) {
        double off = 0.0;
        for (int e = threadIdx.x; e < np; e += NT) {
            // e -> (i, j) column-packed, i < j contributes twice
            int j = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
            while ((j + 1) * (j + 2) / 2 <= e) j++;
            while (j * (j + 1) / 2 > e) j--;
            int i = e - j * (j + 1) / 2;
            if (i != j) off += 2.0 * A[e] * A[e];
        }
        off = block_sum<NT>(off, red);
        if (off <= thr) break;
        // one pair per thread (m <= NT is guaranteed by the host)
        const int k = threadIdx.x;
        for (int t = 0; t < dp - 1; t++) {
            if (k < m) {
                int p, q;
                round_pair(t, k, dp, p, q);
                pp[k] = p;
                qq[k] = q;
                double apq = A[pk(p, q)];
                double c = 1.0, s = 0.0;
                if (apq != 0.0) {
                    double app = A[pk(p, p)], aqq = A[pk(q, q)];
                    double th = (aqq - app) / (2.0 * apq);
                    double tn = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    c = 1.0 / sqrt(tn * tn + 1.0);
                    s = tn * c;
                }
                rc[k] = c;
                rs[k] = s;
                // recorded for the replay; the LDS-only barrier below does not
                // wait for these stores
                size_t ro = ((size_t)nsweep * (dp - 1) + t) * m + k;
                rec[2 * ro] = c;
                rec[2 * ro + 1] = s;
            }
            round_sync<GLOBAL>();
            apply_round<NT>(A, pp, qq, rc, rs, m, 1.0, true);
            round_sync<GLOBAL>();
        }
    }
    if (nsweep == max_sweeps && threadIdx.x == 0) atomicOr(&st.status, (int)kStatusJacobiNotConverged);

    // U+ = J diag(max(L, 0)) J^T by replaying the rounds backwards
    __syncthreads();  // also completes the rotation-record stores
    for (int e = threadIdx.x; e < np; e += NT) {
        int j = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
        while ((j + 1) * (j + 2) / 2 <= e) j++;
        while (j * (j + 1) / 2 > e) j--;
        int i = e - j * (j + 1) / 2;
        double v = A[e];
        A[e] = (i == j && v > 0.0) ? v : 0.0;
    }
    __syncthreads();
    {
        // round idx = sweep * (dp - 1) + t, replayed from last to first; the
        // next round's rotation is fetched while the current one is applied
        const int k = threadIdx.x;
        const int nround = nsweep * (dp - 1);
        double nc = 1.0, ns = 0.0;
        if (nround > 0 && k < m) {
            size_t ro = (size_t)(nround - 1) * m + k;
            nc = rec[2 * ro];
            ns = rec[2 * ro + 1];
        }
        for (int idx = nround - 1; idx >= 0; idx--) {
            if (k < m) {
                int p, q;
                round_pair(idx % (dp - 1), k, dp, p, q);
                pp[k] = p;
                qq[k] = q;
                rc[k] = nc;
                rs[k] = ns;
                if (idx > 0) {
                    size_t ro = (size_t)(idx - 1) * m + k;
                    nc = rec[2 * ro];
                    ns = rec[2 * ro + 1];
                }
            }
            round_sync<GLOBAL>();
            apply_round<NT>(A, pp, qq, rc, rs, m, -1.0, false);
            round_sync<GLOBAL>();
        }
    }

    // x = U+ / mu, z = U+ - sym(U), v1 = c - z; partial c.x and dual residual
    double ex = 0.0, dr = 0.0;
    double* X = B.x + I.poff + it.off;
    double* Z = B.z + I.poff + it.off;
    double* V1 = B.v1 + I.poff + it.off;
    const double* Cc = B.c + I.poff + it.off;
    const double inv_mu = 1.0 / mu;
    for (int e = threadIdx.x; e < d * d; e += NT) {
        int i = e / d, j = e - i * d;
        double up = A[pk(i, j)];
        double uij = U[e];
        double sym = 0.5 * (uij + U[(size_t)j * d + i]);
        double xn = up * inv_mu;
        double zn = up - sym;
        double xo = X[e];
        double res = (uij - mu * xo) + zn;
        X[e] = xn;
        Z[e] = zn;
        double ce = Cc[e];
        V1[e] = ce - zn;
        ex += ce * xn;
        dr += res * res;
    }
    ex = block_sum<NT>(ex, red);
    dr = block_sum<NT>(dr, red);
    if (threadIdx.x == 0) {
        st.pending = 1;
        B.proj_part[2 * (size_t)it.gblock] = ex;
        B.proj_part[2 * (size_t)it.gblock + 1] = dr;
    }
}

// ------------------------------------------------------------------------
// small kernels

__global__ void k_fill(double* p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

__global__ void k_copy_strided(double* dst, int64_t dst_stride, const double* src, int64_t src_stride, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i * dst_stride] = src[i * src_stride];
}

__global__ void k_copy(double* dst, const double* src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) dst[i] = src[i];
}

// c[e] = src[i1[e]] - src[i2[e]] (index < 0 reads 0)
__global__ void k_pack_cost(const int32_t* i1, const int32_t* i2, const double* src, double* c, int64_t n) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; e < n; e += st) {
        int32_t a = i1[e], b = i2[e];
        double v = a >= 0 ? src[a] : 0.0;
        if (b >= 0) v -= src[b];
        c[e] = v;
    }
}

// dsrc[t] = sum over the primal elements e that read src[t] of +-xbar[e]
__global__ void k_cost_adjoint(const int32_t* ptr, const int32_t* ent, const double* xbar, double* dsrc, int64_t nsrc) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; t < nsrc; t += st) {
        double v = 0.0;
        for (int32_t k = ptr[t]; k < ptr[t + 1]; k++) {
            uint32_t e = (uint32_t)ent[k];
            double xv = xbar[e & 0x7FFFFFFFu];
            v += (e & 0x80000000u) ? -xv : xv;
        }
        dsrc[t] = v;
    }
}

// out[0] = sum_i a[i] b[i] in one workgroup (fixed order)
// a . b in two launches with a fixed summation order: kDotParts workgroups
// each reduce one contiguous chunk into out[1 + part], then one workgroup
// sums the parts in order into out[0].
constexpr int kDotParts = 256;
__global__ __launch_bounds__(256) void k_dot_part(const double* a, const double* b, int64_t n, double* out) {
    __shared__ double red[4];
    const int64_t chunk = (n + kDotParts - 1) / kDotParts;
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s = fma(a[i], b[i], s);
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) out[1 + blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_dot_final(double* out) {
    __shared__ double red[4];
    const double s = block_sum<256>(out[1 + threadIdx.x], red);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

// ------------------------------------------------------------------------
// launchers

void launch_cg_at(const BatchDev& B, const UnitsDev& U, hipStream_t s) {
    if (U.grid > 0) hipLaunchKernelGGL(k_cg_at, dim3(U.grid), dim3(kWG), 0, s, B, U);
}
void launch_cg_a(const BatchDev& B, const UnitsDev& U, hipStream_t s) {
    if (U.grid <= 0) return;
    hipLaunchKernelGGL(k_cg_a, dim3(U.grid), dim3(kWG), 0, s, B, U);
    hipLaunchKernelGGL(k_cg_reduce, dim3(U.ninst), dim3(kWG), 0, s, B, U);
}
void launch_pre(const BatchDev& B, const UnitsDev& U, hipStream_t s) {
    // every instance: k_pre_reduce also finishes the bookkeeping of instances
    // outside the plan
    if (U.nwg <= 0) return;
    if (U.grid > 0) hipLaunchKernelGGL(k_pre, dim3(U.grid), dim3(kWG), 0, s, B, U);
    hipLaunchKernelGGL(k_pre_reduce, dim3(U.ninst), dim3(kWG), 0, s, B, U);
}
void launch_post(const BatchDev& B, const UnitsDev& U, hipStream_t s) {
    if (U.grid > 0) hipLaunchKernelGGL(k_post, dim3(U.grid), dim3(kWG), 0, s, B, U);
}
void launch_spmv_plain(const SellDev& S, const int4* units, int wg0, int nwg, const double* in, double* out,
                       hipStream_t s) {
    if (nwg > 0) hipLaunchKernelGGL(k_spmv_plain, dim3(nwg), dim3(kWG), 0, s, S, units, wg0, in, out);
}

size_t project_lds_bytes(int d, int nt, bool global_matrix) {
    int dp = d + (d & 1), m = dp / 2;
    size_t mat = global_matrix ? 0 : (size_t)dp * (dp + 1) / 2;
    return sizeof(double) * (mat + 2 * (size_t)m + nt / 64) + sizeof(int) * 2 * (size_t)m;
}

void launch_project(int nt, bool global_matrix, const BatchDev& B, const ProjItem* items, int nitems, int max_d,
                    double* rec, int max_sweeps, double* gmat, hipStream_t s) {
    if (nitems <= 0) return;
    size_t lds = project_lds_bytes(max_d, nt, global_matrix);
    if (global_matrix) {
        hipLaunchKernelGGL((k_project<1024, true>), dim3(nitems), dim3(1024), lds, s, B, items, rec, max_sweeps, gmat);
        return;
    }
    switch (nt) {
        case 64:
            hipLaunchKernelGGL((k_project<64, false>), dim3(nitems), dim3(64), lds, s, B, items, rec, max_sweeps, gmat);
            break;
        case 256:
            hipLaunchKernelGGL((k_project<256, false>), dim3(nitems), dim3(256), lds, s, B, items, rec, max_sweeps, gmat);
            break;
        case 512:
            hipLaunchKernelGGL((k_project<512, false>), dim3(nitems), dim3(512), lds, s, B, items, rec, max_sweeps, gmat);
            break;
        default:
            hipLaunchKernelGGL((k_project<1024, false>), dim3(nitems), dim3(1024), lds, s, B, items, rec, max_sweeps,
                               gmat);
            break;
    }
}

static int grid_for(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

void launch_fill(double* p, int64_t n, double v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
}
void launch_copy_strided(double* dst, int64_t dst_stride, const double* src, int64_t src_stride, int64_t n,
                         hipStream_t s) {
    if (n > 0)
        hipLaunchKernelGGL(k_copy_strided, dim3(grid_for(n)), dim3(256), 0, s, dst, dst_stride, src, src_stride, n);
}
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(256), 0, s, dst, src, n);
}
void launch_pack_cost(const int32_t* i1, const int32_t* i2, const double* src, double* c, int64_t n, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_pack_cost, dim3(grid_for(n)), dim3(256), 0, s, i1, i2, src, c, n);
}
void launch_cost_adjoint(const int32_t* ptr, const int32_t* ent, const double* xbar, double* dsrc, int64_t nsrc,
                         hipStream_t s) {
    if (nsrc > 0) hipLaunchKernelGGL(k_cost_adjoint, dim3(grid_for(nsrc)), dim3(256), 0, s, ptr, ent, xbar, dsrc, nsrc);
}
void launch_dot1(const double* a, const double* b, int64_t n, double* out, hipStream_t s) {
    static_assert(kDotParts == 256, "k_dot_final sums one part per thread");
    hipLaunchKernelGGL(k_dot_part, dim3(kDotParts), dim3(256), 0, s, a, b, n, out);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, out);
}

}  // namespace v2h

namespace v2h {
void set_project_lds_limit(size_t bytes) {
    V2H_HIP(hipFuncSetAttribute((const void*)k_project<64, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
    V2H_HIP(hipFuncSetAttribute((const void*)k_project<256, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
    V2H_HIP(hipFuncSetAttribute((const void*)k_project<512, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
    V2H_HIP(hipFuncSetAttribute((const void*)k_project<1024, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes));
}
}  // namespace v2h
