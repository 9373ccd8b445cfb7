// refine.hip -- the warm-started PSD projection (see refine.hpp).
//
// Every product of the refinement is a d x d x d FP64 GEMM of one block.  A
// launch covers every block of the refined classes: one workgroup per 32 x 32
// output tile, four wavefronts of one 16 x 16 v_mfma_f64_16x16x4f64 tile
// each, operands staged through LDS in chunks of 32 along K.  The per-block
// scalar work (eigenvalue estimates, corrections, acceptance) and the
// eigenvectors of the fallback are one workgroup per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "refine.hpp"

namespace v2h {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kT = 32;    // output tile (rows and columns)
constexpr int kKc = 32;   // K chunk staged in LDS
constexpr int kNT = 256;  // threads of every kernel here
constexpr int kNB = 16;   // reflectors per compact-WY product in the eigenvector formation

enum AMode : int { kAN = 0, kAT = 1, kAScaled = 2 };  // op(A)(m, k): A[m][k], A[k][m], A[m][k] s[k]
enum BMode : int { kBN = 0, kBT = 1 };                // op(B)(k, n): B[k][n], B[n][k]
enum Epi : int { kEpiStore = 0, kEpiAdd = 1, kEpiFinal = 2 };
enum Gate : int { kGateRefining = 0, kGateUpdate = 1, kGateFinal = 2 };

// One product C = [D +] op(A) op(B) of every block, operands named by slot.
struct Prod {
    int a, b, c, dadd;  // slots; a < 0: the accepted X of the block
    int am, bm, epi, pad;
};

// NaN-propagating maximum (a NaN correction must reject the block).
__device__ __forceinline__ double nmax(double a, double b) { return (a != a) ? a : ((b != b) ? b : fmax(a, b)); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_nmax(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Workgroup reductions in a fixed order; every thread receives the result.
__device__ double wg_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return t;
}
__device__ double wg_nmax(double v, double* red) {
    v = wave_nmax(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double t = nmax(nmax(red[0], red[1]), nmax(red[2], red[3]));
    __syncthreads();
    return t;
}

__device__ __forceinline__ double* slot(const RefDev& R, const RefBlock& b, int s) {
    return R.work + b.w + (int64_t)s * b.d * b.d;
}
__device__ __forceinline__ double* lam_of(const RefDev& R, const RefBlock& b) {
    return R.work + b.w + (int64_t)kRefSlots * b.d * b.d;
}

__device__ __forceinline__ bool gate_ok(const RefState& rs, int gate, int k) {
    if (gate == kGateRefining) return rs.phase == 1;
    if (gate == kGateUpdate) return rs.phase == 1 || (rs.phase == 2 && rs.cstep == k);
    return rs.phase == 2;
}

// ---------------------------------------------------------------- products
// One 32 x 32 tile of one block's product.
__device__ void ref_tile(const BatchDev& B, const RefDev& R, const Prod& p, const RefJob& job, const RefBlock& b,
                         const RefState& rs, double (*As)[kT + 1], double (*Bs)[kT + 1], double* red) {
    const int d = b.d;
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int aslot = p.a >= 0 ? p.a : rs.xfinal;
    const double* A = slot(R, b, aslot);
    const double* Bm = slot(R, b, p.b >= 0 ? p.b : rs.xfinal);
    const double* sc = lam_of(R, b);
    const double mu = B.st[b.inst].mu;
    const double inv_mu = 1.0 / mu;
    const int m0 = job.ti * kT, n0 = job.tj * kT;
    const int wm = (w >> 1) * 16, wn = (w & 1) * 16;
    const int li = l & 15, lk = l >> 4;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    // the loads of a chunk go to registers first, so that the next chunk's
    // loads are in flight during this chunk's products
    constexpr int kPer = (kT * kKc) / kNT;
    double ra[kPer], rb[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int idx = t + kNT * e;
            // consecutive threads read consecutive addresses of the stored matrix
            int m, kk;
            if (p.am == kAT) {
                m = idx & 31;
                kk = idx >> 5;
            } else {
                kk = idx & 31;
                m = idx >> 5;
            }
            const int gm = m0 + m, gk = k0 + kk;
            double av = 0.0;
            if (gm < d && gk < d) {
                av = p.am == kAT ? A[(int64_t)gk * d + gm] : A[(int64_t)gm * d + gk];
                if (p.am == kAScaled) av *= fmax(sc[gk], 0.0) * inv_mu;
            }
            ra[e] = av;
            int n, kb;
            if (p.bm == kBT) {
                kb = idx & 31;
                n = idx >> 5;
            } else {
                n = idx & 31;
                kb = idx >> 5;
            }
            const int gn = n0 + n, gkb = k0 + kb;
            rb[e] = (gn < d && gkb < d) ? (p.bm == kBT ? Bm[(int64_t)gn * d + gkb] : Bm[(int64_t)gkb * d + gn]) : 0.0;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int e = 0; e < kPer; e++) {
            const int idx = t + kNT * e;
            if (p.am == kAT)
                As[idx >> 5][idx & 31] = ra[e];
            else
                As[idx & 31][idx >> 5] = ra[e];
            if (p.bm == kBT)
                Bs[idx & 31][idx >> 5] = rb[e];
            else
                Bs[idx >> 5][idx & 31] = rb[e];
        }
    };
    // a wavefront whose 16 x 16 part lies outside the block only helps load
    const bool live = (m0 + wm < d) && (n0 + wn < d);
    fetch(0);
    for (int k0 = 0; k0 < d; k0 += kKc) {
        stage();
        __syncthreads();
        if (k0 + kKc < d) fetch(k0 + kKc);
        // k steps of 4 up to the end of the matrix (zero-filled beyond d)
        const int kend = min(kKc, (d - k0 + 3) & ~3);
        if (live)
            for (int kk = 0; kk < kend; kk += 4) {
                const double a = As[kk + lk][wm + li];
                const double bb = Bs[kk + lk][wn + li];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
            }
        __syncthreads();
    }
    if (p.epi != kEpiFinal) {
        double* C = slot(R, b, p.c);
        const double* D = p.epi == kEpiAdd ? slot(R, b, p.dadd >= 0 ? p.dadd : rs.xfinal) : nullptr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = m0 + wm + 4 * r + lk, n = n0 + wn + li;
            if (m < d && n < d) {
                const int64_t e = (int64_t)m * d + n;
                C[e] = D ? D[e] + acc[r] : acc[r];
            }
        }
        return;
    }
    // x = X diag(l+) X^T / mu (acc), z = U+ - sym(U), v1 = c - z; the tile's
    // c . x and dual residual; the accepted X becomes the block's V
    const InstDev I = B.inst[b.inst];
    double* X = B.x + I.poff + b.off;
    double* Z = B.z + I.poff + b.off;
    double* V1 = B.v1 + I.poff + b.off;
    const double* Cc = B.c + I.poff + b.off;
    const double* U = B.u + I.poff + b.off;
    const double* Us = slot(R, b, kSlotU);
    double* V = slot(R, b, kSlotV);
    double ex = 0.0, dr = 0.0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + 4 * r + lk, n = n0 + wn + li;
        if (m < d && n < d) {
            const int64_t e = (int64_t)m * d + n;
            const double xn = acc[r];
            const double zn = mu * xn - Us[e];
            const double res = (U[e] - mu * X[e]) + zn;
            X[e] = xn;
            Z[e] = zn;
            const double ce = Cc[e];
            V1[e] = ce - zn;
            ex += ce * xn;
            dr += res * res;
            V[e] = A[e];
        }
    }
    ex = wg_sum(ex, red);
    dr = wg_sum(dr, red);
    if (t == 0) {
        R.part[b.part + 2 * job.tile] = ex;
        R.part[b.part + 2 * job.tile + 1] = dr;
    }
}

// One workgroup per tile job; blockIdx.y selects the product.  The jobs of a
// block are contiguous and workgroups are dispatched round-robin over the 8
// XCDs, so job j = the XCD-major remap of the workgroup index: the tiles of one
// block run on one XCD and share its L2.
__device__ __forceinline__ int xcd_job(int bid, int n) {
    const int per = n >> 3, rem = n & 7;
    const int x = bid & 7, loc = bid >> 3;
    return x * per + (x < rem ? x : rem) + loc;
}

__global__ __launch_bounds__(kNT) void k_ref_gemm(BatchDev B, RefDev R, Prod p0, Prod p1, int gate, int k) {
    __shared__ double As[kKc][kT + 1];  // As[k][m]
    __shared__ double Bs[kKc][kT + 1];  // Bs[k][n]
    __shared__ double red[4];
    const RefJob job = R.jobs[xcd_job(blockIdx.x, gridDim.x)];
    const RefState& rs = R.rs[job.blk];
    if (!gate_ok(rs, gate, k)) return;
    const RefBlock b = R.blk[job.blk];
    if (!in_pgroup(B, b.inst)) return;  // accepted in the other group's projection
    ref_tile(B, R, blockIdx.y == 0 ? p0 : p1, job, b, rs, As, Bs, red);
}

// ---------------------------------------------------------------- per block
// Start of an outer iteration, one thread per block: phase, backoff, route.
__global__ void k_ref_state(BatchDev B, RefDev R) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R.nblk) return;
    const RefBlock b = R.blk[i];
    if (!in_pgroup(B, b.inst)) return;
    RefState& rs = R.rs[i];
    const bool active = B.st[b.inst].active != 0;
    const int wait = rs.wait;
    const bool run = active && rs.vvalid && wait == 0;
    rs.phase = run ? 1 : 0;
    rs.live = active ? 1 : 0;
    if (active && wait > 0) rs.wait = wait - 1;
    R.route[b.gblock] = run ? kRouteRefining : kRouteFallback;
}

// sym(U) of one 32 x 32 tile of a refining block; the transposed tile goes
// through LDS, so that both reads are coalesced.
__global__ __launch_bounds__(kNT) void k_ref_sym(BatchDev B, RefDev R) {
    __shared__ double tt[32][33];
    const RefJob job = R.jobs[xcd_job(blockIdx.x, gridDim.x)];
    if (R.rs[job.blk].phase != 1) return;
    const RefBlock b = R.blk[job.blk];
    const int d = b.d, tc = threadIdx.x & 31, tr = threadIdx.x >> 5;
    const double* U = B.u + B.inst[b.inst].poff + b.off;
    double* Us = slot(R, b, kSlotU);
    for (int r = tr; r < 32; r += kNT / 32) {  // tile (tj, ti) of U
        const int gi = job.tj * 32 + r, gj = job.ti * 32 + tc;
        tt[r][tc] = (gi < d && gj < d) ? U[(int64_t)gi * d + gj] : 0.0;
    }
    __syncthreads();
    for (int r = tr; r < 32; r += kNT / 32) {
        const int gi = job.ti * 32 + r, gj = job.tj * 32 + tc;
        if (gi < d && gj < d) Us[(int64_t)gi * d + gj] = 0.5 * (U[(int64_t)gi * d + gj] + tt[tc][r]);
    }
}

// Corrections E of one 32 x 32 tile from S = X^T U X and G = X^T X, with the
// eigenvalue estimates l_i = s_ii / g_ii of the block; E goes to slot K (free
// once S is formed; G and S stay intact while other tiles read their
// diagonals); the tile's largest correction (separated pairs) and largest
// coupling left inside clusters go to its partial slots; diagonal tiles store
// their l.
__global__ __launch_bounds__(kNT) void k_ref_etile(RefDev R, RefParams P) {
    __shared__ double li_s[32], lj_s[32];
    __shared__ double red[4];
    const RefJob job = R.jobs[xcd_job(blockIdx.x, gridDim.x)];
    if (R.rs[job.blk].phase != 1) return;
    const RefBlock b = R.blk[job.blk];
    const int d = b.d, t = threadIdx.x;
    const double* S = slot(R, b, kSlotS);
    const double* G = slot(R, b, kSlotG);
    double* E = slot(R, b, kSlotK);
    // block scale max |l|: every tile reads the d diagonal pairs itself
    double scale = 0.0;
    for (int i = t; i < d; i += kNT) scale = nmax(scale, fabs(S[(int64_t)i * d + i] / G[(int64_t)i * d + i]));
    scale = wg_nmax(scale, red);
    const double om = P.cluster_rel * scale;
    const int i0 = job.ti * 32, j0 = job.tj * 32;
    if (t < 32) {
        const int i = i0 + t, j = j0 + t;
        li_s[t] = i < d ? S[(int64_t)i * d + i] / G[(int64_t)i * d + i] : 0.0;
        lj_s[t] = j < d ? S[(int64_t)j * d + j] / G[(int64_t)j * d + j] : 0.0;
        if (job.ti == job.tj && i < d) lam_of(R, b)[i] = li_s[t];
    }
    __syncthreads();
    double me = 0.0, mc = 0.0;
    const int c = t & 31;
    const int j = j0 + c;
    for (int r = t >> 5; r < 32; r += kNT / 32) {
        const int i = i0 + r;
        if (i >= d || j >= d) continue;
        const int64_t e = (int64_t)i * d + j;
        const double rr = (i == j ? 1.0 : 0.0) - G[e];
        double ev;
        if (i == j) {
            ev = 0.5 * rr;
        } else {
            const double dl = lj_s[c] - li_s[r];
            const double sv = S[e];
            if (fabs(dl) > om) {
                ev = (sv + lj_s[c] * rr) / dl;
                me = nmax(me, fabs(ev));
            } else {
                ev = 0.5 * rr;
                mc = nmax(mc, fabs(sv));
            }
        }
        E[e] = ev;
    }
    me = wg_nmax(me, red);
    mc = wg_nmax(mc, red);
    if (t == 0) {
        R.part[b.part + 2 * job.tile] = me;
        R.part[b.part + 2 * job.tile + 1] = scale > 0.0 ? mc / scale : mc;
    }
}

// Acceptance of step k from the tile maxima, one thread per block.
__global__ void k_ref_decide(RefDev R, RefParams P, int k, int xout) {
    const int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= R.nblk) return;
    RefState& rs = R.rs[bi];
    if (rs.phase != 1) return;
    const RefBlock b = R.blk[bi];
    double me = 0.0, cpl = 0.0;
    for (int q = 0; q < b.nt * b.nt; q++) {
        me = nmax(me, R.part[b.part + 2 * q]);
        cpl = nmax(cpl, R.part[b.part + 2 * q + 1]);
    }
    rs.max_e = me;
    rs.cpl = cpl;
    const int64_t d3 = 2 * (int64_t)b.d * b.d * b.d;
    bool reject = false;
    if (!(me <= 0.25) || !(cpl == cpl)) {
        reject = true;
        rs.rej_div += 1;
    } else if (me <= P.accept_e && cpl <= P.accept_cpl) {
        rs.phase = 2;
        rs.cstep = k;
        rs.xfinal = xout;
        rs.backoff = 0;
    } else if (k >= P.max_steps) {
        reject = true;
        if (me > P.accept_e)
            rs.rej_e += 1;
        else
            rs.rej_c += 1;
    }
    rs.flops += (reject ? 3 : 4) * d3;  // a rejected block skips X + XE
    if (reject) {
        rs.phase = 3;
        const int bo = min(2 * rs.backoff + 2, kMaxBackoff);
        rs.backoff = bo;
        rs.wait = bo;
        R.route[b.gblock] = kRouteRejected;
    }
}

// Block sums of the accepted blocks (tile partials in a fixed order).
__global__ __launch_bounds__(kNT) void k_ref_close(BatchDev B, RefDev R) {
    __shared__ double red[4];
    const RefBlock b = R.blk[blockIdx.x];
    if (!in_pgroup(B, b.inst)) return;
    RefState& rs = R.rs[blockIdx.x];
    if (rs.phase != 2) {  // projected by the tridiagonal fallback this iteration
        if (threadIdx.x == 0 && B.st[b.inst].active) rs.fallback += 1;
        return;
    }
    const int nt2 = b.nt * b.nt;
    double ex = 0.0, dr = 0.0;
    for (int q = threadIdx.x; q < nt2; q += kNT) {
        ex += R.part[b.part + 2 * q];
        dr += R.part[b.part + 2 * q + 1];
    }
    ex = wg_sum(ex, red);
    dr = wg_sum(dr, red);
    if (threadIdx.x == 0) {
        B.proj_part[2 * (size_t)b.gblock] = ex;
        B.proj_part[2 * (size_t)b.gblock + 1] = dr;
        B.st[b.inst].pending = 1;
        R.route[b.gblock] = kRouteAccepted;
        rs.refined += 1;
        rs.flops += 2 * (int64_t)b.d * b.d * b.d;
    }
}

// Eigenvectors of a block that took the tridiagonal path this iteration:
// V = Q S with S the tridiagonal eigenvectors (normalised) and Q = H_0 H_1 ..
// H_{d-3} the reflectors of the reduction, applied as compact WY products
// (I - Y T Y^T) of kNB reflectors each, the last ones first.  Thread l keeps
// column l.
__global__ __launch_bounds__(kNT) void k_ref_form_v(BatchDev B, RefDev R) {
    __shared__ double Y[kNB][256];
    __shared__ double gram[kNB][kNB];
    __shared__ double Tm[kNB][kNB];
    __shared__ double tau[kNB];
    const RefBlock b = R.blk[blockIdx.x];
    if (!in_pgroup(B, b.inst)) return;
    RefState& rs = R.rs[blockIdx.x];
    // blocks that took the fallback and try the refinement next iteration
    const int route = R.route[b.gblock];
    if (!rs.live || (route != kRouteFallback && route != kRouteRejected) || rs.wait > 0) return;
    const int d = b.d, t = threadIdx.x;
    const int64_t dd = (int64_t)d * d;
    const double* hv = R.tri_work + b.tri;  // row k: reflector k (v[k + 1] = 1), tau_k at [k][k]
    const double* Sv = hv + dd;             // column l: eigenvector l of T (unnormalised)
    double* V = slot(R, b, kSlotV);
    if (t < d) {
        double s2 = 0.0;
        for (int i = 0; i < d; i++) {
            const double s = Sv[(int64_t)i * d + t];
            s2 = fma(s, s, s2);
        }
        const double inv = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
        for (int i = 0; i < d; i++) V[(int64_t)i * d + t] = Sv[(int64_t)i * d + t] * inv;
    }
    const int nref = d - 2;  // reflectors 0 .. d - 3
    for (int kc = nref > 0 ? ((nref - 1) / kNB) * kNB : -1; kc >= 0; kc -= kNB) {
        const int nb = min(kNB, nref - kc);
        __syncthreads();
        for (int e = t; e < kNB * 256; e += kNT) {
            const int j = e >> 8, i = e & 255;
            const int kr = kc + j;
            double v = 0.0;
            if (j < nb && i < d && i > kr && hv[(int64_t)kr * d + kr] != 0.0) v = hv[(int64_t)kr * d + i];
            Y[j][i] = v;
        }
        if (t < kNB) tau[t] = t < nb ? hv[(int64_t)(kc + t) * d + kc + t] : 0.0;
        __syncthreads();
        for (int q = t; q < kNB * kNB; q += kNT) {
            const int a = q / kNB, c = q - a * kNB;
            double s = 0.0;
            if (a < c)
                for (int i = kc + 1; i < d; i++) s = fma(Y[a][i], Y[c][i], s);
            gram[a][c] = s;
            Tm[a][c] = 0.0;
        }
        __syncthreads();
        // T of H_kc .. H_kc+nb-1 = I - Y T Y^T, column by column: T_ii = tau_i,
        // T(0:i, i) = -tau_i T(0:i, 0:i) (Y^T y_i)(0:i)
        for (int i = 0; i < kNB; i++) {
            if (t < i) {
                double s = 0.0;
                for (int c = t; c < i; c++) s = fma(Tm[t][c], gram[c][i], s);
                Tm[t][i] = -tau[i] * s;
            }
            if (t == i) Tm[i][i] = tau[i];
            __syncthreads();
        }
        if (t < d) {
            double wv[kNB];
#pragma unroll
            for (int j = 0; j < kNB; j++) wv[j] = 0.0;
            for (int i = kc + 1; i < d; i++) {
                const double m = V[(int64_t)i * d + t];
#pragma unroll
                for (int j = 0; j < kNB; j++) wv[j] = fma(Y[j][i], m, wv[j]);
            }
            double w2[kNB];
#pragma unroll
            for (int a = 0; a < kNB; a++) {
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < kNB; c++) s = fma(Tm[a][c], wv[c], s);
                w2[a] = s;
            }
            for (int i = kc + 1; i < d; i++) {
                double m = V[(int64_t)i * d + t];
#pragma unroll
                for (int j = 0; j < kNB; j++) m = fma(-Y[j][i], w2[j], m);
                V[(int64_t)i * d + t] = m;
            }
        }
    }
    if (t == 0) rs.vvalid = 1;
}

__global__ void k_ref_forget(RefDev R, int b0, int b1) {
    const int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b1) R.rs[i].vvalid = 0;
}

}  // namespace

int gemm_grid(const RefDev& R) { return R.njobs; }

void launch_ref_state(const BatchDev& B, const RefDev& R, hipStream_t s) {
    if (R.nblk > 0) hipLaunchKernelGGL(k_ref_state, dim3((R.nblk + 255) / 256), dim3(256), 0, s, B, R);
}

void launch_ref_sym(const BatchDev& B, const RefDev& R, hipStream_t s) {
    if (R.nblk > 0) hipLaunchKernelGGL(k_ref_sym, dim3(R.njobs), dim3(kNT), 0, s, B, R);
}

void launch_ref_step(const BatchDev& B, const RefDev& R, const RefParams& P, int k, hipStream_t s) {
    if (R.nblk == 0) return;
    // X alternates between two slots: V -> Xa -> Xb -> Xa -> ...
    const int xin = k == 1 ? kSlotV : (k % 2 == 0 ? kSlotXa : kSlotXb);
    const int xout = k % 2 == 1 ? kSlotXa : kSlotXb;
    // K = U X, G = X^T X
    const Prod pk{kSlotU, xin, kSlotK, -1, kAN, kBN, kEpiStore, 0};
    const Prod pg{xin, xin, kSlotG, -1, kAT, kBN, kEpiStore, 0};
    const int g = gemm_grid(R);
    hipLaunchKernelGGL(k_ref_gemm, dim3(g, 2), dim3(kNT), 0, s, B, R, pk, pg, (int)kGateRefining, k);
    // S = X^T K
    const Prod ps{xin, kSlotK, kSlotS, -1, kAT, kBN, kEpiStore, 0};
    hipLaunchKernelGGL(k_ref_gemm, dim3(g, 1), dim3(kNT), 0, s, B, R, ps, ps, (int)kGateRefining, k);
    hipLaunchKernelGGL(k_ref_etile, dim3(R.njobs), dim3(kNT), 0, s, R, P);
    hipLaunchKernelGGL(k_ref_decide, dim3((R.nblk + 255) / 256), dim3(256), 0, s, R, P, k, xout);
    // X_out = X + X E (E in slot K)
    const Prod pu{xin, kSlotK, xout, xin, kAN, kBN, kEpiAdd, 0};
    hipLaunchKernelGGL(k_ref_gemm, dim3(g, 1), dim3(kNT), 0, s, B, R, pu, pu, (int)kGateUpdate, k);
}

void launch_ref_final(const BatchDev& B, const RefDev& R, hipStream_t s) {
    if (R.nblk == 0) return;
    const Prod pf{-1, -1, -1, -1, kAScaled, kBT, kEpiFinal, 0};
    hipLaunchKernelGGL(k_ref_gemm, dim3(gemm_grid(R), 1), dim3(kNT), 0, s, B, R, pf, pf, (int)kGateFinal, 0);
    hipLaunchKernelGGL(k_ref_close, dim3(R.nblk), dim3(kNT), 0, s, B, R);
}

void launch_ref_form_v(const BatchDev& B, const RefDev& R, hipStream_t s) {
    if (R.nblk > 0) hipLaunchKernelGGL(k_ref_form_v, dim3(R.nblk), dim3(kNT), 0, s, B, R);
}

void launch_ref_forget(const RefDev& R, int b0, int b1, hipStream_t s) {
    if (b1 > b0) hipLaunchKernelGGL(k_ref_forget, dim3((b1 - b0 + 255) / 256), dim3(256), 0, s, R, b0, b1);
}

// ---------------------------------------------------------------- host side
void Refiner::build(const std::vector<Item>& items, int count, int nblocks_total, hipStream_t s) {
    blk_.clear();
    inst_begin_.assign(count + 1, 0);
    std::vector<RefJob> jobs;
    int64_t w = 0, part = 0;
    int cur = 0;
    for (const Item& it : items) {
        if (it.d > 256) throw std::logic_error("refined projection: block larger than 256");
        while (cur < it.inst) inst_begin_[++cur] = (int)blk_.size();
        RefBlock b{};
        b.inst = it.inst;
        b.d = it.d;
        b.gblock = it.gblock;
        b.nt = (it.d + kT - 1) / kT;
        b.off = it.off;
        b.tri = it.tri;
        b.w = w;
        w += (int64_t)kRefSlots * it.d * it.d + ((it.d + 1) & ~1);
        b.part = part;
        part += 2 * (int64_t)b.nt * b.nt;
        for (int i = 0; i < b.nt; i++)
            for (int j = 0; j < b.nt; j++) jobs.push_back(RefJob{(int)blk_.size(), i, j, i * b.nt + j});
        blk_.push_back(b);
    }
    while (cur < count) inst_begin_[++cur] = (int)blk_.size();
    njobs_ = (int)jobs.size();
    d_blk_.upload(blk_);
    d_jobs_.upload(jobs);
    d_rs_.alloc(std::max<size_t>(blk_.size(), 1));
    V2H_HIP(hipMemsetAsync(d_rs_.p, 0, sizeof(RefState) * d_rs_.n, s));
    route_.alloc(std::max(nblocks_total, 1));
    V2H_HIP(hipMemsetAsync(route_.p, 0, sizeof(int) * route_.n, s));
    work_.alloc(std::max<int64_t>(w, 1));
    part_.alloc(std::max<int64_t>(part, 1));
    V2H_HIP(hipStreamSynchronize(s));
}

RefDev Refiner::dev() const {
    RefDev R{};
    R.blk = d_blk_.p;
    R.jobs = d_jobs_.p;
    R.rs = d_rs_.p;
    R.route = route_.p;
    R.work = work_.p;
    R.part = part_.p;
    R.tri_work = tri_work_;
    R.nblk = (int)blk_.size();
    R.njobs = njobs_;
    return R;
}

void Refiner::begin(const BatchDev& B, hipStream_t s) const {
    if (!blk_.empty()) launch_ref_state(B, dev(), s);
}

void Refiner::refine(const BatchDev& B, hipStream_t s) const {
    if (blk_.empty()) return;
    const RefDev R = dev();
    launch_ref_sym(B, R, s);
    for (int k = 1; k <= params_.max_steps; k++) launch_ref_step(B, R, params_, k, s);
    launch_ref_final(B, R, s);
}

void Refiner::seed(const BatchDev& B, hipStream_t s) const {
    if (!blk_.empty()) launch_ref_form_v(B, dev(), s);
}

void Refiner::forget(int inst, hipStream_t s) const {
    if (!blk_.empty() && inst >= 0 && inst + 1 < (int)inst_begin_.size())
        launch_ref_forget(dev(), inst_begin_[inst], inst_begin_[inst + 1], s);
}

void Refiner::stats(int64_t* out) const {
    for (int k = 0; k < kStats; k++) out[k] = 0;
    if (blk_.empty()) return;
    std::vector<RefState> rs(blk_.size());
    V2H_HIP(hipMemcpy(rs.data(), d_rs_.p, sizeof(RefState) * rs.size(), hipMemcpyDeviceToHost));
    for (const auto& r : rs) {
        out[0] += r.refined;
        out[1] += r.fallback;
        out[2] += r.rej_div;
        out[3] += r.rej_e;
        out[4] += r.rej_c;
        out[5] += r.flops;
    }
}

}  // namespace v2h
