// eig_tri.hip -- PSD projection of one symmetric block per workgroup through
// a tridiagonal eigendecomposition (gfx950).
//
//   U  = Q T Q^T        Householder reduction; the matrix is held in registers,
//                       2-D cyclic over the workgroup (thread (tr, tc) owns rows
//                       tr + PR r and columns tc + 32 c), one reflector per step
//   T  = S L S^T        eigenvalues by Sturm counts with bracketed Newton steps
//                       (warm-started from the previous call's eigenvalues of the
//                       same block), eigenvectors by two steps of inverse
//                       iteration, orthogonalised within clusters of close
//                       eigenvalues
//   T+ = S L+ S^T       staged through LDS into the register matrix (or
//                       T - S L- S^T when there are fewer non-positive eigenvalues)
//   U+ = Q T+ Q^T       the reflectors applied two-sidedly in reverse order
//
// followed by the BPSDP update of the block: x = U+ / mu, z = U+ - sym(U),
// v1 = c - z and the block's partial c.x and dual-residual sums.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "kernels.hpp"

namespace v2h {

namespace {

constexpr double kEps = DBL_EPSILON;
// Consecutive eigenvalues closer than kClusterRel |lambda| + kClusterAbs ||T||
// share a cluster; its eigenvectors are orthogonalised against each other.
constexpr double kClusterRel = 1e-5;
constexpr double kClusterAbs = 1e-11;
constexpr int kStageMax = 16;    // eigenvectors per LDS chunk when forming T+
constexpr int kMaxPasses = 160;  // Sturm passes per eigenvalue (bisection needs about 60)
constexpr int kChunk = 8;        // inverse-iteration loads issued ahead of their use

__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Lane index recomputed where it is used (the asm keeps the compiler from
// holding it, and values derived from it, in registers across the loops of
// the projection, where they would otherwise be spilled and reloaded).
__device__ __forceinline__ int lane_now() {
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// Compiler-only fence between the rows of a register tile: keeps the
// scheduler from hoisting every row's memory operations at once, which would
// exceed the register budget.
__device__ __forceinline__ void row_fence() { asm volatile("" ::: "memory"); }

// ---- cross-lane exchange of doubles (DPP within rows of 16 lanes, ds_swizzle
// within 32).  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int PATTERN>
__device__ __forceinline__ double swz(double v) {
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), PATTERN);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), PATTERN);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
constexpr int kQuadXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowRor8 = 0x128;      // rotate by 8 in a row of 16 = lane xor 8
constexpr int kSwzXor4 = (4 << 10) | 0x1F;
constexpr int kSwzXor16 = (16 << 10) | 0x1F;

// Sum over the 64 lanes; every lane receives the same total.
__device__ __forceinline__ double wsum(double v) {
    v += dpp<kQuadXor1>(v);
    v += dpp<kQuadXor2>(v);
    v += dpp<kRowHalfMirror>(v);
    v += dpp<kRowMirror>(v);
    v += swz<kSwzXor16>(v);
    return lane_value(v, 0) + lane_value(v, 32);
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wsum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < NT / 64; q++) t += red[q];
    return t;
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    v = wmax(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; q++) t = fmax(t, red[q]);
    return t;
}
template <int NT>
__device__ __forceinline__ double block_min(double v, double* red) {
    v = wmin(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; q++) t = fmin(t, red[q]);
    return t;
}

// Sums v[0..7] over the 32 lanes of a half-wavefront with 9 exchanges
// (reduce-scatter).  Lane L ends up with the total of entry
// 4 b4 + 2 b3 + b2 (b = bits of L); the four lanes differing in bits 0-1 agree.
__device__ __forceinline__ double reduce_scatter8(double (&v)[8]) {
    const int lane = threadIdx.x & 63;
    {
        const bool hi = lane & 16;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double send = hi ? v[k] : v[k + 4];
            double keep = hi ? v[k + 4] : v[k];
            v[k] = keep + swz<kSwzXor16>(send);
        }
    }
    {
        const bool hi = lane & 8;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            double send = hi ? v[k] : v[k + 2];
            double keep = hi ? v[k + 2] : v[k];
            v[k] = keep + dpp<kRowRor8>(send);
        }
    }
    {
        const bool hi = lane & 4;
        double send = hi ? v[0] : v[1];
        double keep = hi ? v[1] : v[0];
        v[0] = keep + swz<kSwzXor4>(send);
    }
    v[0] += dpp<kQuadXor2>(v[0]);
    v[0] += dpp<kQuadXor1>(v[0]);
    return v[0];
}

// 1/q to about 1 ulp: hardware reciprocal and one Newton step.
__device__ __forceinline__ double recip(double q) {
    const double r = __builtin_amdgcn_rcp(q);
    const double e = fma(-q, r, 1.0);
    return fma(r, e, r);
}

// Deterministic start vectors for inverse iteration, uniform in [-1, 1).
__device__ __forceinline__ double start_value(uint32_t l, uint32_t i) {
    uint32_t h = (l * 0x9E3779B1u) ^ ((i + 0x7F4A7C15u) * 0x85EBCA77u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)(h & 0xFFFFFFu) * (1.0 / 8388608.0) - 1.0;
}

// Sturm count of T - s I (eigenvalues below s) and S = d/ds log|det(T - s I)|
// = sum_j 1 / (s - lambda_j), from the pivots q_i of the LDL^T recurrence.
__device__ int sturm(double s, int d, const double* dd, const double* e2, double pivmin, double& S) {
    double q = dd[0] - s;
    if (fabs(q) < pivmin) q = -pivmin;
    double r = recip(q);
    int c = q < 0.0;
    double dq = -1.0;
    double acc = dq * r;
    auto step = [&](double di, double e) {
        double qn = fma(-e, r, di - s);
        const double dqn = fma(e * dq, r * r, -1.0);
        if (fabs(qn) < pivmin) qn = -pivmin;
        r = recip(qn);
        dq = dqn;
        c += qn < 0.0;
        acc = fma(dq, r, acc);
    };
    int i = 1;
    for (; i + 4 <= d; i += 4) {  // the loads of four steps ahead of the recurrence
        double dv[4], ev[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            dv[u] = dd[i + u];
            ev[u] = e2[i + u - 1];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) step(dv[u], ev[u]);
    }
    for (; i < d; i++) step(dd[i], e2[i - 1]);
    S = acc;
    return c;
}

// Inverse-iteration helpers on one column x (stride d) of the work area, with
// the LU factors of T - sg I in F (5 planes of stride d^2, same column).
struct Column {
    double* x;
    double* F;
    size_t ld, plane;  // d, d^2
    __device__ double& X(int i) const { return x[(size_t)i * ld]; }
    __device__ double& f(int q, int i) const { return F[q * plane + (size_t)i * ld]; }
};

// x <- L^{-1} P (scale x); loads run kChunk rows ahead of the recurrence.
__device__ void apply_pl(const Column& c, int d, double scale) {
    double bc = c.X(0) * scale;
    for (int i0 = 0; i0 + 1 < d; i0 += kChunk) {
        double xn[kChunk], f0[kChunk], sw[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            const int i = min(i0 + u, d - 2);
            xn[u] = c.X(i + 1);
            f0[u] = c.f(0, i);
            sw[u] = c.f(4, i);
        }
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            const int i = i0 + u;
            if (i + 1 < d) {
                const double bn = xn[u] * scale;
                const bool s = sw[u] != 0.0;
                c.X(i) = s ? bn : bc;
                bc = s ? fma(-f0[u], bn, bc) : fma(-f0[u], bc, bn);
            }
        }
    }
    c.X(d - 1) = bc;
}

// x <- U^{-1} x; returns ||x||^2.
__device__ double apply_u(const Column& c, int d) {
    double x1 = 0.0, x2 = 0.0, s2 = 0.0;
    for (int i0 = d - 1; i0 >= 0; i0 -= kChunk) {
        double xb[kChunk], f1[kChunk], f2[kChunk], f3[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            const int i = max(i0 - u, 0);
            xb[u] = c.X(i);
            f1[u] = c.f(1, i);
            f2[u] = c.f(2, i);
            f3[u] = c.f(3, i);
        }
#pragma unroll
        for (int u = 0; u < kChunk; u++) {
            const int i = i0 - u;
            if (i >= 0) {
                const double xi = fma(-f3[u], x2, fma(-f2[u], x1, xb[u])) * f1[u];
                c.X(i) = xi;
                s2 = fma(xi, xi, s2);
                x2 = x1;
                x1 = xi;
            }
        }
    }
    return s2;
}

// x <- x - sum_m (y_m . x / ||y_m||^2) y_m over the cluster's earlier columns
// (twice); returns the new ||x||^2.
__device__ double orthogonalise(const Column& c, int d, const double* Sv, int l0, int l, const double* xs) {
    for (int pass = 0; pass < 2; pass++)
        for (int m = l0; m < l; m++) {
            const double* y = Sv + m;
            double dot = 0.0;
            for (int i0 = 0; i0 < d; i0 += kChunk) {
                double yv[kChunk], xv[kChunk];
#pragma unroll
                for (int u = 0; u < kChunk; u++) {
                    const int i = min(i0 + u, d - 1);
                    yv[u] = y[(size_t)i * d];
                    xv[u] = c.X(i);
                }
#pragma unroll
                for (int u = 0; u < kChunk; u++)
                    if (i0 + u < d) dot = fma(yv[u], xv[u], dot);
            }
            const double f = dot / xs[m];
            for (int i0 = 0; i0 < d; i0 += kChunk) {
                double yv[kChunk], xv[kChunk];
#pragma unroll
                for (int u = 0; u < kChunk; u++) {
                    const int i = min(i0 + u, d - 1);
                    yv[u] = y[(size_t)i * d];
                    xv[u] = c.X(i);
                }
#pragma unroll
                for (int u = 0; u < kChunk; u++)
                    if (i0 + u < d) c.X(i0 + u) = fma(-f, yv[u], xv[u]);
            }
        }
    double s2 = 0.0;
    for (int i0 = 0; i0 < d; i0 += kChunk) {
        double xv[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; u++) xv[u] = c.X(min(i0 + u, d - 1));
#pragma unroll
        for (int u = 0; u < kChunk; u++)
            if (i0 + u < d) s2 = fma(xv[u], xv[u], s2);
    }
    return s2;
}

// The block's share of one thread: RI x CJ elements.  Three homes:
//   kTileReg   all in registers;
//   kTileSplit the first kSplitRegs in registers, the rest in LDS laid out
//              [slot][thread] (blocks just above the register budget);
//   kTileGlobal in global memory laid out [wavefront][r][c][lane]: a
//              wavefront's accesses coalesce (the largest blocks).
constexpr int kTileReg = 0, kTileGlobal = 1, kTileSplit = 2;
constexpr int kSplitRegs = 33;

template <int NT, int RI, int CJ, int MODE>
struct Tile;
template <int NT, int RI, int CJ>
struct Tile<NT, RI, CJ, kTileReg> {
    double v[RI][CJ];
    __device__ __forceinline__ double& at(int r, int c) { return v[r][c]; }
    __device__ __forceinline__ double col(int r, int ck) {
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < CJ; c++)
            if (c == ck) x = v[r][c];
        return x;
    }
};
template <int NT, int RI, int CJ>
struct Tile<NT, RI, CJ, kTileSplit> {
    double v[kSplitRegs];
    double* lds;  // (RI CJ - kSplitRegs) x NT
    // r, c are compile-time constants wherever the tile is swept (unrolled loops)
    __device__ __forceinline__ double& at(int r, int c) {
        const int e = r * CJ + c;
        return e < kSplitRegs ? v[e] : lds[(e - kSplitRegs) * NT + threadIdx.x];
    }
    __device__ __forceinline__ double col(int r, int ck) {
        double x = 0.0;
#pragma unroll
        for (int c = 0; c < CJ; c++)
            if (c == ck) x = at(r, c);
        return x;
    }
};
template <int NT, int RI, int CJ>
struct Tile<NT, RI, CJ, kTileGlobal> {
    double* p;
    __device__ __forceinline__ double& at(int r, int c) {
        return p[((size_t)(threadIdx.x >> 6) * RI * CJ + r * CJ + c) * 64 + (threadIdx.x & 63)];
    }
    __device__ __forceinline__ double col(int r, int ck) { return at(r, ck); }
};

// Eigen-decomposition of the tridiagonal T (diagonal dd, off-diagonal ee, in
// LDS), one thread per eigenvalue: eigenvalues ascending in lamv (Sturm counts
// and bracketed Newton steps, warm-started from lprev, which receives the new
// ones), unnormalised eigenvectors in the columns of Sv (inverse iteration
// with the LU factors in fac), xs[l] = ||S(:, l)||^2.  Returns the number of
// eigenvalues <= 0; tnorm = ||T|| estimate, zero_block when T = 0.
template <int NT>
__device__ __forceinline__ int tri_eigen(int d, const double* dd, const double* ee, double* e2, double* lamv,
                                         double* xs, double* red, double* lprev, double* Sv, double* fac,
                                         int* status, double& tnorm_out, bool& zero_block_out) {
    const int t = threadIdx.x;
    double glo = INFINITY, ghi = -INFINITY, em = 0.0;
    if (t < d) {
        const double el = t > 0 ? fabs(ee[t - 1]) : 0.0;
        const double er = t + 1 < d ? fabs(ee[t]) : 0.0;
        glo = dd[t] - el - er;
        ghi = dd[t] + el + er;
        if (t + 1 < d) {
            e2[t] = ee[t] * ee[t];
            em = e2[t];
        }
    }
    glo = block_min<NT>(glo, red);
    ghi = block_max<NT>(ghi, red);
    em = block_max<NT>(em, red);
    const double tnorm = fmax(fabs(glo), fabs(ghi));
    tnorm_out = tnorm;
    const double pivmin = DBL_MIN * fmax(1.0, em);
    const double gl = glo - 2.0 * kEps * tnorm * d - 2.0 * pivmin;
    const double gu = ghi + 2.0 * kEps * tnorm * d + 2.0 * pivmin;
    const double abstol = 2.0 * kEps * tnorm;
    const double dprev = lprev[d];
    const bool zero_block = !(tnorm > 0.0);
    zero_block_out = zero_block;
    double change = 0.0;
    bool bad = false;
    if (t < d) {
        const int l = t;
        double lam = 0.0;
        if (!zero_block) {
            double lo = gl, hi = gu;
            int clo = 0, chi = d;
            int passes = 0;
            double s = 0.5 * (lo + hi), S = 0.0;
            bool haveS = false;
            auto eval = [&](double x, double& Sx) {
                const int c = sturm(x, d, dd, e2, pivmin, Sx);
                passes++;
                if (c <= l) {
                    lo = x;
                    clo = c;
                } else {
                    hi = x;
                    chi = c;
                }
                return c;
            };
            const double lp = lprev[l];
            if (dprev >= 0.0 && isfinite(lp) && isfinite(dprev)) {
                // expanding search around the previous eigenvalue of this block
                double del = fmax(2.0 * dprev, 8.0 * abstol);
                s = fmin(fmax(lp, lo), hi);
                const int c0 = eval(s, S);
                haveS = true;
                for (int e = 0; e < 8; e++) {
                    const double x = (c0 <= l) ? s + del : s - del;
                    if (!(x > lo && x < hi)) break;
                    double Sx;
                    const int cx = eval(x, Sx);
                    if ((c0 <= l) != (cx <= l)) break;
                    del *= 4.0;
                }
            }
            lam = 0.5 * (lo + hi);
            double last = INFINITY;
            while (passes < kMaxPasses) {
                if (hi - lo <= abstol) {
                    lam = 0.5 * (lo + hi);
                    break;
                }
                double sn = 0.0;
                bool newton = false;
                if (clo == l && chi == l + 1 && haveS) {
                    const double step = -1.0 / S;
                    sn = s + step;
                    if (fabs(step) <= 2.0 * abstol) {
                        lam = fmin(fmax(sn, lo), hi);
                        break;
                    }
                    newton = isfinite(sn) && sn > lo && sn < hi && fabs(step) <= 0.5 * last;
                    if (newton) last = fabs(step);
                }
                if (!newton) {
                    sn = 0.5 * (lo + hi);
                    last = INFINITY;
                }
                s = sn;
                eval(s, S);
                haveS = true;
                lam = s;
            }
            if (passes >= kMaxPasses) bad = true;
            if (isfinite(lp)) change = fabs(lam - lp);
        }
        lamv[l] = lam;
        xs[l] = 1.0;
    }
    change = block_max<NT>(change, red);  // also publishes lamv
    if (t < d) lprev[t] = lamv[t];
    if (t == 0) lprev[d] = (isfinite(dprev) && dprev >= 0.0) ? change : 1e-2 * tnorm;

    // inverse iteration, one thread per cluster of close eigenvalues; the
    // vectors stay unnormalised, xs[l] = ||S(:, l)||^2
    if (!zero_block && t < d) {
        auto linked = [&](int l) {
            return l > 0 && (lamv[l] - lamv[l - 1]) <= kClusterRel * fmax(fabs(lamv[l]), fabs(lamv[l - 1])) +
                                                           kClusterAbs * tnorm;
        };
        if (!linked(t)) {
            const int l0 = t;
            int l1 = l0;
            while (l1 + 1 < d && linked(l1 + 1)) l1++;
            const double ptol = kEps * tnorm;
            double sprev = -INFINITY;
            for (int l = l0; l <= l1; l++) {
                double sg = lamv[l];
                const double pert = 10.0 * kEps * fmax(fabs(sg), 1e-3 * tnorm);
                if (l > l0 && sg <= sprev + pert) sg = sprev + pert;
                sprev = sg;
                const Column col{Sv + l, fac + l, (size_t)d, (size_t)d * d};
                // T - sg I = P L U with partial pivoting ([q][i]: 0 multiplier,
                // 1 reciprocal pivot, 2-3 superdiagonals of U, 4 interchange),
                // fused with L^{-1} P applied to the start vector
                double dcur = dd[0] - sg, ucur = d > 1 ? ee[0] : 0.0;
                double bc = start_value((uint32_t)l, 0u);
                for (int i = 0; i + 1 < d; i++) {
                    const double sub = ee[i];
                    double dnext = dd[i + 1] - sg;
                    double unext = (i + 2 < d) ? ee[i + 1] : 0.0;
                    const double bn = start_value((uint32_t)l, (uint32_t)(i + 1));
                    double f, u0, u1, u2, sw, xi;
                    if (fabs(dcur) >= fabs(sub)) {
                        if (fabs(dcur) < ptol) dcur = (dcur < 0.0) ? -ptol : ptol;
                        f = sub * recip(dcur);
                        u0 = dcur;
                        u1 = ucur;
                        u2 = 0.0;
                        sw = 0.0;
                        dnext = fma(-f, ucur, dnext);
                        xi = bc;
                        bc = fma(-f, bc, bn);
                    } else {
                        f = dcur * recip(sub);
                        u0 = sub;
                        u1 = dnext;
                        u2 = unext;
                        sw = 1.0;
                        dnext = fma(-f, dnext, ucur);
                        unext = -f * unext;
                        xi = bn;
                        bc = fma(-f, bn, bc);
                    }
                    col.f(0, i) = f;
                    col.f(1, i) = recip(u0);
                    col.f(2, i) = u1;
                    col.f(3, i) = u2;
                    col.f(4, i) = sw;
                    col.X(i) = xi;
                    dcur = dnext;
                    ucur = unext;
                }
                if (fabs(dcur) < ptol) dcur = (dcur < 0.0) ? -ptol : ptol;
                col.f(1, d - 1) = recip(dcur);
                col.f(2, d - 1) = 0.0;
                col.f(3, d - 1) = 0.0;
                col.X(d - 1) = bc;
                double s2 = apply_u(col, d);
                if (l > l0) s2 = orthogonalise(col, d, Sv, l0, l, xs);
                apply_pl(col, d, 1.0 / sqrt(s2));
                s2 = apply_u(col, d);
                if (l > l0) s2 = orthogonalise(col, d, Sv, l0, l, xs);
                if (!(s2 > 0.0) || !isfinite(s2)) {
                    bad = true;
                    s2 = INFINITY;  // drops the vector from T+
                }
                xs[l] = s2;
            }
        }
    }
    if (bad) atomicOr(status, (int)kStatusJacobiNotConverged);
    // eigenvalues are ascending: the first ps are <= 0 (a count, so that the
    // range below stays inside [0, d] whatever the values)
    return d - (int)block_sum<NT>((t < d && !(lamv[t] <= 0.0)) ? 1.0 : 0.0, red);  // also completes Sv, xs
}

// Waves per SIMD the compiler must leave room for: two workgroups of the
// 512-thread class share a CU (its hot loops stay spill-free at 128 VGPRs).
template <int NT>
constexpr int kTriWavesPerEU = NT == 512 ? 4 : 1;

template <int NT, int RI, int CJ, int MODE>
__global__ __launch_bounds__(NT, kTriWavesPerEU<NT>) void k_project_tri(BatchDev B, const ProjItem* items,
                                                                      double* work, double* lamst,
                                                                      const int* route, int want) {
    constexpr bool GM = MODE == kTileGlobal;
    constexpr int PR = NT / 32;
    constexpr int DMAX = PR * RI;
    constexpr int kStage = MODE == kTileSplit ? 4 : kStageMax;  // LDS left for staging eigenvectors
    constexpr int NLDS = MODE == kTileSplit ? RI * CJ - kSplitRegs : 0;
    static_assert(MODE != kTileSplit || NLDS > 0, "split tile needs an LDS part");
    static_assert(RI <= 8, "reduce_scatter8 handles at most 8 rows per thread");
    static_assert(DMAX <= NT, "one eigenvalue per thread");
    static_assert(PR * RI == 32 * CJ, "square register tiling");
    __shared__ double colb[2][DMAX];
    __shared__ double pbuf[DMAX];
    __shared__ double vbuf[DMAX], wbuf[DMAX];  // rank-2 update not yet applied to the tile
    __shared__ double dd[DMAX], ee[DMAX], e2[DMAX], tau[DMAX], lamv[DMAX], xs[DMAX];
    __shared__ double stage[kStage][DMAX];
    __shared__ double red[NT / 64];
    __shared__ double tile_lds[NLDS > 0 ? NLDS * NT : 1];

    const ProjItem it = items[blockIdx.x];
    InstState& st = B.st[it.inst];
    if (!st.active || !in_pgroup(B, it.inst)) return;
    if (route && route[it.gblock] != want) return;  // not this launch's share of the refined classes
    const InstDev I = B.inst[it.inst];
    const int d = it.d;
    const int t = threadIdx.x, tr = t >> 5, tc = t & 31, lane = t & 63;
    // the same indices, recomputed at their use inside the Householder steps
    const int wv_s = __builtin_amdgcn_readfirstlane(t >> 6);
    auto tr_now = [&]() { return 2 * wv_s + (lane_now() >> 5); };
    auto tc_now = [&]() { return lane_now() & 31; };
    const double mu = st.mu;
    const double* U = B.u + I.poff + it.off;
    double* hv = work + it.work_off;        // d x d, row k = reflector k
    double* Sv = hv + (size_t)d * d;        // d x d, column l = eigenvector l (unnormalised)
    double* fac = Sv + (size_t)d * d;       // 5 x d x d LU factors, [q][i][l]
    double* lprev = lamst + it.lam_off;     // d eigenvalues + last largest change

    for (int i = t; i < 2 * DMAX; i += NT) (&colb[0][0])[i] = 0.0;
    for (int i = t; i < DMAX; i += NT) vbuf[i] = wbuf[i] = 0.0;
    Tile<NT, RI, CJ, MODE> A;
    if constexpr (GM) A.p = fac + 5 * (size_t)d * d;  // NT RI CJ doubles after the factors
    if constexpr (MODE == kTileSplit) A.lds = tile_lds;
#pragma unroll
    for (int r = 0; r < RI; r++) {
        const int i = tr + PR * r;
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            const int j = tc + 32 * c;
            const int ic = min(i, d - 1), jc = min(j, d - 1);
            const double s = 0.5 * (U[(size_t)ic * d + jc] + U[(size_t)jc * d + ic]);
            A.at(r, c) = (i < d && j < d) ? s : 0.0;
        }
        row_fence();
    }
    __syncthreads();

    // column k of the register matrix (below the diagonal) to colb[buf], diagonal to dd[k]
    auto put_col = [&](int k, int buf) {
        if (tc == (k & 31)) {
            const int ck = k >> 5;
#pragma unroll
            for (int r = 0; r < RI; r++) {
                const int i = tr + PR * r;
                const double v = A.col(r, ck);
                if (i < d) {
                    if (i > k) colb[buf][i] = v;
                    else if (i == k) dd[k] = v;
                }
            }
        }
    };
    // pbuf[i] = tk (A v)_i for every row i < DMAX
    auto matvec = [&](const double (&vc)[CJ], double tk) {
        double s8[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            double s = 0.0;
            if (r < RI) {
#pragma unroll
                for (int c = 0; c < CJ; c++) s = fma(A.at(r, c), vc[c], s);
            }
            s8[r] = s;
            row_fence();
        }
        const double ps = reduce_scatter8(s8);
        const int ln = lane_now();
        const int rr = ((ln >> 4) & 1) * 4 + ((ln >> 3) & 1) * 2 + ((ln >> 2) & 1);  // reduce_scatter8 entry
        if ((ln & 3) == 0 && rr < RI) pbuf[tr_now() + PR * rr] = tk * ps;
    };
    // A -= v w^T + w v^T, w = p - kk v, over rows/columns i with lo < i < d:
    // v_i = vbuf[i] * vscale except v_one = 1; p in pbuf.
    auto rank2 = [&](const double* vbuf, const double (&vc)[CJ], double vscale, int lo, int one, double kk) {
        double wc[CJ];
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            const int j = tc + 32 * c;
            const double pj = pbuf[j];
            wc[c] = (j > lo && j < d) ? fma(-kk, vc[c], pj) : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RI; r++) {
            const int i = tr + PR * r;
            const double vraw = vbuf[i], pi = pbuf[i];
            const bool in = i > lo && i < d;
            const double vi = in ? (i == one ? 1.0 : vraw * vscale) : 0.0;
            const double wi = in ? fma(-kk, vi, pi) : 0.0;
#pragma unroll
            for (int c = 0; c < CJ; c++) A.at(r, c) = fma(-wi, vc[c], fma(-vi, wc[c], A.at(r, c)));
            row_fence();
        }
    };

    // Column k of the register matrix to colb[buf]: rows below the diagonal,
    // zero elsewhere; the diagonal to dd[k].
    // Row tests are made on the slot index r (a constant after unrolling)
    // against per-step bounds, so that no row index has to live in a register.
    auto slot_of = [&](int i, int trr) { return (i >= trr && (i - trr) % PR == 0) ? (i - trr) / PR : -1; };
    auto put_col_z = [&](int k, int buf) {
        if (tc_now() == (k & 31)) {
            const int ck = k >> 5, trr = tr_now();
            const int r_k = slot_of(k, trr), r_lo = k >= trr ? (k - trr) / PR + 1 : 0;  // rows > k from r_lo
            const int r_e = d > trr ? (d - 1 - trr) / PR + 1 : 0;                     // rows < d below r_e
#pragma unroll
            for (int r = 0; r < RI; r++) {
                const double v = A.col(r, ck);
                colb[buf][trr + PR * r] = (r >= r_lo && r < r_e) ? v : 0.0;
                if (r == r_k) dd[k] = v;
            }
        }
    };
    // A -= v w^T + w v^T with v_i = vsrc[i] * vscale (v_one = 1) and
    // w = p - kk v, p in pbuf.  vsrc is zero outside the rows that the step
    // reduces or transforms, so no row or column needs a mask: rows and
    // columns outside the trailing matrix receive finite garbage that no later
    // step reads (they only meet zero entries of v), and padding rows stay zero.
    auto rank2_z = [&](const double* vsrc, const double (&vc)[CJ], double vscale, int one, double kk) {
        double wc[CJ];
        const int tcc = tc_now(), trr = tr_now();
#pragma unroll
        for (int c = 0; c < CJ; c++) wc[c] = fma(-kk, vc[c], pbuf[tcc + 32 * c]);
        const int r_one = one >= 0 ? slot_of(one, trr) : -1;
        // the LDS loads of row r + 1 are issued before the updates of row r
        double vn = vsrc[trr], pn = pbuf[trr];
#pragma unroll
        for (int r = 0; r < RI; r++) {
            const double vr = vn, pr = pn;
            if (r + 1 < RI) {
                vn = vsrc[trr + PR * (r + 1)];
                pn = pbuf[trr + PR * (r + 1)];
            }
            const double vi = (r == r_one) ? 1.0 : vr * vscale;
            const double wi = fma(-kk, vi, pr);
#pragma unroll
            for (int c = 0; c < CJ; c++) A.at(r, c) = fma(-wi, vc[c], fma(-vi, wc[c], A.at(r, c)));
            row_fence();
        }
    };
    // sum over the rows i < DMAX of f(i), on every lane
    auto rows_sum = [&](auto&& f) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < (DMAX + 63) / 64; m++) {
            const int i = lane + 64 * m;
            if (DMAX % 64 == 0 || i < DMAX) s += f(i);
        }
        return wsum(s);
    };

    // One pass over the tile: apply the pending rank-2 update
    // A -= v w^T + w v^T (v, w in vbuf / wbuf, zero where it does not apply)
    // and accumulate A v_new; pbuf[i] = tk (A v_new)_i for every row i < DMAX.
    auto fused = [&](const double (&vn)[CJ], double tk) {
        double vo[CJ], wo[CJ];
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            vo[c] = vbuf[tc + 32 * c];
            wo[c] = wbuf[tc + 32 * c];
        }
        double s8[8];
#pragma unroll
        for (int r = 0; r < 8; r++) {
            double sr = 0.0;
            if (r < RI) {
                const int i = tr + PR * r;
                const double vi = vbuf[i], wi = wbuf[i];
#pragma unroll
                for (int c = 0; c < CJ; c++) {
                    const double a = fma(-wi, vo[c], fma(-vi, wo[c], A.at(r, c)));
                    A.at(r, c) = a;
                    sr = fma(a, vn[c], sr);
                }
            }
            s8[r] = sr;
            row_fence();
        }
        const double ps = reduce_scatter8(s8);
        const int ln = lane_now();
        const int rr = ((ln >> 4) & 1) * 4 + ((ln >> 3) & 1) * 2 + ((ln >> 2) & 1);
        if ((ln & 3) == 0 && rr < RI) pbuf[tr_now() + PR * rr] = tk * ps;
    };
    // apply the pending update without a product
    auto apply_pending = [&]() {
        double vo[CJ], wo[CJ];
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            vo[c] = vbuf[tc + 32 * c];
            wo[c] = wbuf[tc + 32 * c];
        }
#pragma unroll
        for (int r = 0; r < RI; r++) {
            const int i = tr + PR * r;
            const double vi = vbuf[i], wi = wbuf[i];
#pragma unroll
            for (int c = 0; c < CJ; c++) A.at(r, c) = fma(-wi, vo[c], fma(-vi, wo[c], A.at(r, c)));
            row_fence();
        }
    };

    // ---------------------------------------------------------------- U = Q T Q^T
    if constexpr (!GM) {
    put_col_z(0, 0);
    for (int k = 0; k + 2 < d; k++) {
        lds_sync();
        const double* col = colb[k & 1];  // zero outside rows k+1 .. d-1
        const double sg = rows_sum([&](int i) { const double ci = col[i]; return i > k + 1 ? ci * ci : 0.0; });
        const double alpha = col[k + 1];
        double tk = 0.0, sc = 0.0, beta = alpha;
        if (sg != 0.0) {
            beta = -copysign(sqrt(fma(alpha, alpha, sg)), alpha);
            tk = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
        }
        if (t == 0) {
            ee[k] = beta;
            tau[k] = tk;
            hv[(size_t)k * d + k] = tk;  // tau of reflector k, read back when the eigenvectors are formed
        }
        if (tk != 0.0) {  // uniform: every wavefront computed the same sums
            const int tn = 64 * wv_s + lane_now();
            if (tn > k && tn < d) hv[(size_t)k * d + tn] = (tn == k + 1) ? 1.0 : col[tn] * sc;
            double vc[CJ];
            const int tcc = tc_now();
            const int c_one = (((k + 1 - tcc) & 31) == 0 && k + 1 >= tcc) ? (k + 1 - tcc) >> 5 : -1;
#pragma unroll
            for (int c = 0; c < CJ; c++) vc[c] = (c == c_one) ? 1.0 : col[tcc + 32 * c] * sc;
            matvec(vc, tk);
            lds_sync();
            const double kk =
                0.5 * tk * rows_sum([&](int i) { return ((i == k + 1) ? 1.0 : col[i] * sc) * pbuf[i]; });
            rank2_z(col, vc, sc, k + 1, kk);
        }
        put_col_z(k + 1, (k + 1) & 1);
    }
    if (d >= 2) put_col_z(d - 1, (d - 1) & 1);  // dd[d-1]; keeps column d-2 in the other buffer
    __syncthreads();                  // also completes the reflector stores
    if (t == 0 && d >= 2) ee[d - 2] = colb[(d - 2) & 1][d - 1];
    __syncthreads();

    } else {
    // Step k reduces column k.  The rank-2 update of step k is applied to the
    // tile during the product of step k + 1 (one pass per step); only the
    // owners of column k + 1 apply it at once, to publish that column.
    put_col(0, 0);
    for (int k = 0; k + 2 < d; k++) {
        lds_sync();  // column k, the pending update of step k - 1
        const double* col = colb[k & 1];
        double sg = 0.0;
        for (int i = k + 2 + lane; i < DMAX; i += 64) sg = fma(col[i], col[i], sg);  // zero beyond d
        sg = wsum(sg);
        const double alpha = col[k + 1];
        double tk = 0.0, sc = 0.0, beta = alpha;
        if (sg != 0.0) {
            beta = -copysign(sqrt(fma(alpha, alpha, sg)), alpha);
            tk = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
        }
        if (t == 0) {
            ee[k] = beta;
            tau[k] = tk;
            hv[(size_t)k * d + k] = tk;  // tau of reflector k, read back when the eigenvectors are formed
        }
        // v_k = (0.., 1, col[k+2..] sc), zero when no reflection is needed
        auto vk = [&](int i) {
            const double ci = col[i];
            return tk == 0.0 ? 0.0 : (i == k + 1 ? 1.0 : ((i > k + 1 && i < d) ? ci * sc : 0.0));
        };
        if (tk != 0.0)
            for (int i = k + 1 + t; i < d; i += NT) hv[(size_t)k * d + i] = vk(i);
        double vn[CJ];
#pragma unroll
        for (int c = 0; c < CJ; c++) vn[c] = vk(tc + 32 * c);
        fused(vn, tk);
        lds_sync();  // pbuf; every wavefront is done with the old vbuf / wbuf
        double kk = 0.0;
        for (int i = k + 1 + lane; i < d; i += 64) kk = fma(vk(i), pbuf[i], kk);
        kk = 0.5 * tk * wsum(kk);
        auto wk = [&](int i) { return (i > k && i < d) ? fma(-kk, vk(i), pbuf[i]) : 0.0; };
        for (int i = t; i < DMAX; i += NT) {
            vbuf[i] = vk(i);
            wbuf[i] = wk(i);
        }
        // column k + 1 of the reduced matrix, from the owners' tile
        if (tc == ((k + 1) & 31)) {
            const int ck = (k + 1) >> 5;
            const double vj = vk(k + 1), wj = wk(k + 1);
#pragma unroll
            for (int r = 0; r < RI; r++) {
                const int i = tr + PR * r;
                if (i > k && i < d) {
                    const double a = fma(-wk(i), vj, fma(-vk(i), wj, A.col(r, ck)));
                    if (i > k + 1) colb[(k + 1) & 1][i] = a;
                    else dd[k + 1] = a;
                }
            }
        }
    }
    __syncthreads();  // last column, pending update, reflector stores
    if (d >= 2) {
        // the last diagonal element, with the pending update applied
        const int k = d - 1;
        if (tc == (k & 31)) {
            const int ck = k >> 5;
#pragma unroll
            for (int r = 0; r < RI; r++)
                if (tr + PR * r == k) {
                    const double v = vbuf[k], w = wbuf[k];
                    dd[k] = fma(-w, v, fma(-v, w, A.col(r, ck)));
                }
        }
        if (t == 0) ee[d - 2] = colb[(d - 2) & 1][d - 1];
    }
    __syncthreads();

    }
    // ---------------------------------------------------------------- T = S L S^T
    double tnorm;
    bool zero_block;
    const int ps = tri_eigen<NT>(d, dd, ee, e2, lamv, xs, red, lprev, Sv, fac, &st.status, tnorm, zero_block);
    if (zero_block)  // T = 0: any orthonormal basis is an eigenbasis; keep one for the eigenvector formation
        for (int e = t; e < d * d; e += NT) Sv[e] = (e / d == e % d) ? 1.0 : 0.0;

    // ---------------------------------------------------------------- T+ in registers
    const bool use_pos = zero_block || (d - ps) <= ps;
    const int lb = use_pos ? ps : 0, le = use_pos ? d : ps;
    const double sgn = use_pos ? 1.0 : -1.0;
#pragma unroll
    for (int r = 0; r < RI; r++) {
        const int i = tr + PR * r;
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            const int j = tc + 32 * c;
            double v = 0.0;
            if (!use_pos && i < d && j < d) {
                if (i == j) v = dd[i];
                else if (i == j + 1) v = ee[j];
                else if (j == i + 1) v = ee[i];
            }
            A.at(r, c) = v;
        }
    }
    if (!zero_block) {
        for (int l0 = lb; l0 < le; l0 += kStage) {
            lds_sync();
            for (int e = t; e < kStage * DMAX; e += NT) {
                const int i = e / kStage, m = e - i * kStage;
                stage[m][i] = (l0 + m < le && i < d) ? Sv[(size_t)i * d + l0 + m] : 0.0;
            }
            lds_sync();
            const int nm = (le - l0 < kStage) ? le - l0 : kStage;
            for (int m = 0; m < nm; m++) {
                const double wl = sgn * lamv[l0 + m] / xs[l0 + m];
                double sj[CJ];
#pragma unroll
                for (int c = 0; c < CJ; c++) sj[c] = stage[m][tc + 32 * c];
#pragma unroll
                for (int r = 0; r < RI; r++) {
                    const double ar = wl * stage[m][tr + PR * r];
#pragma unroll
                    for (int c = 0; c < CJ; c++) A.at(r, c) = fma(ar, sj[c], A.at(r, c));
                    row_fence();
                }
            }
        }
    }

    // ---------------------------------------------------------------- U+ = Q T+ Q^T
    if constexpr (!GM) {
    if (d >= 3) {
        auto load_v = [&](int k, int buf) {
            if (t < DMAX) colb[buf][t] = (t > k && t < d) ? hv[(size_t)k * d + t] : 0.0;
        };
        load_v(d - 3, (d - 3) & 1);
        for (int k = d - 3; k >= 0; k--) {
            lds_sync();
            const double tk = tau[k];
            if (k > 0) load_v(k - 1, (k - 1) & 1);
            if (tk != 0.0) {
                const double* v = colb[k & 1];  // zero outside rows k+1 .. d-1
                double vc[CJ];
                const int tcc = tc_now();
#pragma unroll
                for (int c = 0; c < CJ; c++) vc[c] = v[tcc + 32 * c];
                matvec(vc, tk);
                lds_sync();
                const double kk = 0.5 * tk * rows_sum([&](int i) { return v[i] * pbuf[i]; });
                rank2_z(v, vc, 1.0, -1, kk);
            }
        }
    }

    } else {
    // M <- H_k M H_k for k = d-3 .. 0; as in the reduction, each step's
    // rank-2 update is applied during the next step's product.
    if (d >= 3) {
        auto load_v = [&](int k, int buf) {
            for (int i = t; i < DMAX; i += NT) colb[buf][i] = (i > k && i < d) ? hv[(size_t)k * d + i] : 0.0;
        };
        lds_sync();  // the tile's last readers of vbuf / wbuf are done
        for (int i = t; i < DMAX; i += NT) vbuf[i] = wbuf[i] = 0.0;
        load_v(d - 3, (d - 3) & 1);
        for (int k = d - 3; k >= 0; k--) {
            lds_sync();  // v_k, the pending update
            const double tk = tau[k];
            if (k > 0) load_v(k - 1, (k - 1) & 1);
            if (tk != 0.0) {
                const double* v = colb[k & 1];
                double vn[CJ];
#pragma unroll
                for (int c = 0; c < CJ; c++) vn[c] = v[tc + 32 * c];
                fused(vn, tk);
                lds_sync();
                double kk = 0.0;
                for (int i = k + 1 + lane; i < d; i += 64) kk = fma(v[i], pbuf[i], kk);
                kk = 0.5 * tk * wsum(kk);
                for (int i = t; i < DMAX; i += NT) {
                    vbuf[i] = v[i];
                    wbuf[i] = i < d ? fma(-kk, v[i], pbuf[i]) : 0.0;
                }
            }
        }
        lds_sync();
        apply_pending();
    }

    }
    // ---------------------------------------------------------------- BPSDP block update
    double ex = 0.0, dr = 0.0;
    double* X = B.x + I.poff + it.off;
    double* Z = B.z + I.poff + it.off;
    double* V1 = B.v1 + I.poff + it.off;
    const double* Cc = B.c + I.poff + it.off;
    const double inv_mu = 1.0 / mu;
#pragma unroll
    for (int r = 0; r < RI; r++) {
        const int i = tr + PR * r;
#pragma unroll
        for (int c = 0; c < CJ; c++) {
            const int j = tc + 32 * c;
            if (i < d && j < d) {
                const size_t e = (size_t)i * d + j;
                const double up = A.at(r, c);
                const double uij = U[e];
                const double sym = 0.5 * (uij + U[(size_t)j * d + i]);
                const double xn = up * inv_mu;
                const double zn = up - sym;
                const double xo = X[e];
                const double res = (uij - mu * xo) + zn;
                X[e] = xn;
                Z[e] = zn;
                const double ce = Cc[e];
                V1[e] = ce - zn;
                ex += ce * xn;
                dr += res * res;
            }
        }
        row_fence();
    }
    ex = block_sum<NT>(ex, red);
    dr = block_sum<NT>(dr, red);
    if (t == 0) {
        st.pending = 1;
        B.proj_part[2 * (size_t)it.gblock] = ex;
        B.proj_part[2 * (size_t)it.gblock + 1] = dr;
    }
}

// ---- One wavefront per block (blocks of up to 64 rows).  Lane j holds column
// j of the block in registers, so the products of the Householder reduction
// and of the back-transform are lane-local and the rank-2 updates need no
// exchange; the vectors they read (v, w, the column being reduced) are
// broadcast through LDS.  Rows are visited in groups of eight, and groups
// that the current step cannot change are skipped (uniform branches).
constexpr int kW = 64;

__global__ __launch_bounds__(kW) void k_project_wave(BatchDev B, const ProjItem* items, double* work, double* lamst) {
    __shared__ double colb[kW];  // column k of the reduced matrix, rows >= k
    __shared__ double vbuf[kW], wbuf[kW];
    __shared__ double dd[kW], ee[kW], e2[kW], tau[kW], lamv[kW], xs[kW];
    __shared__ double stage[kStageMax][kW];
    __shared__ double red[1];

    const ProjItem it = items[blockIdx.x];
    InstState& st = B.st[it.inst];
    if (!st.active || !in_pgroup(B, it.inst)) return;
    const InstDev I = B.inst[it.inst];
    const int d = it.d;
    const int j = threadIdx.x;
    const int jc = min(j, d - 1);
    const int ng = (d + 7) >> 3;  // groups of eight rows that hold rows < d
    const double mu = st.mu;
    const double* U = B.u + I.poff + it.off;
    double* hv = work + it.work_off;     // d x d, row k = reflector k
    double* Sv = hv + (size_t)d * d;     // d x d, column l = eigenvector l (unnormalised)
    double* fac = Sv + (size_t)d * d;    // 5 x d x d LU factors, [q][i][l]
    double* lprev = lamst + it.lam_off;  // d eigenvalues + last largest change

    double a[kW];
#pragma unroll
    for (int g = 0; g < kW / 8; g++) {
#pragma unroll
        for (int u = 0; u < 8; u++) a[8 * g + u] = 0.0;
        if (g < ng) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = 8 * g + u, ic = min(i, d - 1);
                const double sv = 0.5 * (U[(size_t)ic * d + jc] + U[(size_t)jc * d + ic]);
                a[i] = (i < d && j < d) ? sv : 0.0;
            }
        }
        row_fence();
    }

    // lane k writes column k (rows >= k, by groups) to colb
    auto publish = [&](int k) {
        if (j == k) {
#pragma unroll
            for (int g = 0; g < kW / 8; g++)
                if (g < ng && 8 * g + 7 >= k) {
#pragma unroll
                    for (int u = 0; u < 8; u++) colb[8 * g + u] = a[8 * g + u];
                }
        }
    };
    // tk (M v)_j over the rows i > lo (v in vbuf, zero elsewhere)
    // tk (M v)_j over the rows i > lo (v in vbuf, zero elsewhere)
    auto product = [&](int lo, double tk) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int g = 0; g < kW / 8; g++) {
            if (g < ng && 8 * g + 7 > lo) {
                const int i = 8 * g;
                s0 = fma(a[i + 0], vbuf[i + 0], s0);
                s1 = fma(a[i + 1], vbuf[i + 1], s1);
                s2 = fma(a[i + 2], vbuf[i + 2], s2);
                s3 = fma(a[i + 3], vbuf[i + 3], s3);
                s0 = fma(a[i + 4], vbuf[i + 4], s0);
                s1 = fma(a[i + 5], vbuf[i + 5], s1);
                s2 = fma(a[i + 6], vbuf[i + 6], s2);
                s3 = fma(a[i + 7], vbuf[i + 7], s3);
            }
            row_fence();
        }
        return tk * ((s0 + s1) + (s2 + s3));
    };
    // M -= v w^T + w v^T on column j over the rows i > lo
    auto update = [&](int lo, double vj, double wj) {
#pragma unroll
        for (int g = 0; g < kW / 8; g++) {
            if (g < ng && 8 * g + 7 > lo) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int i = 8 * g + u;
                    a[i] = fma(-wbuf[i], vj, fma(-vbuf[i], wj, a[i]));
                }
            }
            row_fence();
        }
    };

    // ---------------------------------------------------------------- U = Q T Q^T
    for (int k = 0; k + 2 < d; k++) {
        publish(k);
        __syncthreads();
        const double alpha = colb[k + 1];
        const double xj = (j > k + 1 && j < d) ? colb[j] : 0.0;
        const double sg = wsum(xj * xj);
        double tk = 0.0, sc = 0.0, beta = alpha;
        if (sg != 0.0) {
            beta = -copysign(sqrt(fma(alpha, alpha, sg)), alpha);
            tk = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
        }
        if (j == 0) {
            dd[k] = colb[k];
            ee[k] = beta;
            tau[k] = tk;
        }
        if (tk != 0.0) {  // uniform
            const double vj = (j == k + 1) ? 1.0 : xj * sc;
            if (j > k && j < d) hv[(size_t)k * d + j] = vj;
            vbuf[j] = vj;
            __syncthreads();
            const double pj = product(k, tk);
            const double kk = 0.5 * tk * wsum(vj * pj);
            const double wj = (j > k && j < d) ? fma(-kk, vj, pj) : 0.0;
            wbuf[j] = wj;
            __syncthreads();
            update(k, vj, wj);
        }
        __syncthreads();  // colb, vbuf, wbuf are rewritten by the next step
    }
    for (int k = d > 2 ? d - 2 : 0; k < d; k++) {  // the last two columns
        publish(k);
        __syncthreads();
        if (j == 0) {
            dd[k] = colb[k];
            if (k + 1 < d) ee[k] = colb[k + 1];
        }
        __syncthreads();
    }

    // ---------------------------------------------------------------- T = S L S^T
    double tnorm;
    bool zero_block;
    const int ps = tri_eigen<kW>(d, dd, ee, e2, lamv, xs, red, lprev, Sv, fac, &st.status, tnorm, zero_block);

    // ---------------------------------------------------------------- T+ in registers
    const bool use_pos = zero_block || (d - ps) <= ps;
    const int lb = use_pos ? ps : 0, le = use_pos ? d : ps;
    const double sgn = use_pos ? 1.0 : -1.0;
    {
        const double tjj = dd[jc], tj1 = (j + 1 < d) ? ee[j] : 0.0, tjm = (j > 0 && j < d) ? ee[j - 1] : 0.0;
        const bool keep_t = !use_pos && j < d;
#pragma unroll
        for (int i = 0; i < kW; i++) {
            double v = 0.0;
            if (i == j) v = tjj;
            else if (i == j + 1) v = tj1;
            else if (i + 1 == j) v = tjm;
            a[i] = (keep_t && i < d) ? v : 0.0;
        }
    }
    if (!zero_block) {
        for (int l0 = lb; l0 < le; l0 += kStageMax) {
            __syncthreads();
            for (int e = j; e < kStageMax * kW; e += kW) {
                const int i = e / kStageMax, m = e - i * kStageMax;
                stage[m][i] = (l0 + m < le && i < d) ? Sv[(size_t)i * d + l0 + m] : 0.0;
            }
            __syncthreads();
            const int nm = (le - l0 < kStageMax) ? le - l0 : kStageMax;
            for (int m = 0; m < nm; m++) {
                const double cj = sgn * lamv[l0 + m] / xs[l0 + m] * stage[m][j];
#pragma unroll
                for (int g = 0; g < kW / 8; g++) {
                    if (g < ng) {
#pragma unroll
                        for (int u = 0; u < 8; u++) a[8 * g + u] = fma(stage[m][8 * g + u], cj, a[8 * g + u]);
                    }
                    row_fence();
                }
            }
        }
    }

    // ---------------------------------------------------------------- U+ = Q T+ Q^T
    if (d >= 3) {
        double vnext = (j > d - 3 && j < d) ? hv[(size_t)(d - 3) * d + j] : 0.0;
        for (int k = d - 3; k >= 0; k--) {
            const double tk = tau[k];
            const double vj = vnext;
            __syncthreads();  // the readers of the previous vbuf / wbuf are done
            vbuf[j] = vj;
            if (k > 0) vnext = (j > k - 1 && j < d) ? hv[(size_t)(k - 1) * d + j] : 0.0;
            __syncthreads();
            if (tk != 0.0) {
                const double pj = product(k, tk);
                const double kk = 0.5 * tk * wsum(vj * pj);
                const double wj = j < d ? fma(-kk, vj, pj) : 0.0;
                wbuf[j] = wj;
                __syncthreads();
                update(-1, vj, wj);
            }
        }
    }

    // ---------------------------------------------------------------- BPSDP block update
    double ex = 0.0, dr = 0.0;
    if (j < d) {
        double* X = B.x + I.poff + it.off;
        double* Z = B.z + I.poff + it.off;
        double* V1 = B.v1 + I.poff + it.off;
        const double* Cc = B.c + I.poff + it.off;
        const double inv_mu = 1.0 / mu;
#pragma unroll
        for (int g = 0; g < kW / 8; g++) {
            if (g < ng) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int i = 8 * g + u;
                    if (i < d) {
                        const size_t e = (size_t)i * d + j;
                        const double up = a[i];
                        const double uij = U[e];
                        const double sym = 0.5 * (uij + U[(size_t)j * d + i]);
                        const double xn = up * inv_mu;
                        const double zn = up - sym;
                        const double xo = X[e];
                        const double res = (uij - mu * xo) + zn;
                        X[e] = xn;
                        Z[e] = zn;
                        const double ce = Cc[e];
                        V1[e] = ce - zn;
                        ex += ce * xn;
                        dr += res * res;
                    }
                }
            }
            row_fence();
        }
    }
    ex = wsum(ex);
    dr = wsum(dr);
    if (j == 0) {
        st.pending = 1;
        B.proj_part[2 * (size_t)it.gblock] = ex;
        B.proj_part[2 * (size_t)it.gblock + 1] = dr;
    }
}

}  // namespace

int project_tri_max_d(int cls) {
    switch (cls) {
        case 0: return 64;
        case 1: return 128;
        case 2: return 192;
        case 3: return 224;
        case 4: return 256;
    }
    return 0;
}

void launch_project_tri(int cls, const BatchDev& B, const ProjItem* items, int nitems, double* work, double* lam,
                        const int* route, int want, hipStream_t s) {
    if (nitems <= 0) return;
    switch (cls) {
        case 0:
            hipLaunchKernelGGL(k_project_wave, dim3(nitems), dim3(kW), 0, s, B, items, work, lam);
            break;
        case 1:
            hipLaunchKernelGGL((k_project_tri<512, 8, 4, kTileReg>), dim3(nitems), dim3(512), 0, s, B, items, work, lam, route, want);
            break;
        case 2:
            hipLaunchKernelGGL((k_project_tri<768, 8, 6, kTileReg>), dim3(nitems), dim3(768), 0, s, B, items, work,
                               lam, route, want);
            break;
        case 3:
            hipLaunchKernelGGL((k_project_tri<1024, 7, 7, kTileSplit>), dim3(nitems), dim3(1024), 0, s, B, items,
                               work, lam, route, want);
            break;
        case 4:
            hipLaunchKernelGGL((k_project_tri<1024, 8, 8, kTileGlobal>), dim3(nitems), dim3(1024), 0, s, B, items,
                               work, lam, route, want);
            break;
        default:
            break;
    }
}

}  // namespace v2h
