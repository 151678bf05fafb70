// dpm_device.hpp -- device helpers shared by the gfx950 kernels.
//
// All arithmetic is FP64 and the library is built with -ffp-contract=off, so
// every expression below is the plain IEEE sequence written (no FMA
// contraction), which is what makes GPU and CPU-oracle results bit-identical
// on the disk path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpm {

constexpr int WAVE = 64;     // CDNA wavefront width
constexpr int BLOCK = 256;   // particles per workgroup in the per-particle kernels
constexpr int CHUNK = 64;    // width of the canonical reduction chunk (one wavefront)

// Minimum-image difference, include/particles/base/kernels.cuh:65-72:
//   d - L*round(d/L), round = half away from zero.  Used where exactness near the
//   half-box matters (neighbour criteria).
__device__ __forceinline__ double pbc_div(double a, double b, double L) {
    double d = a - b;
    return d - L * round(d / L);
}

// Minimum image used by the force, stress, contact and list-build kernels.
// With `fast` set (the host sets it when L/2 exceeds every contact and list
// cutoff), |d| < L/2 returns d itself and other d use d*(1/L) for the rounding.
// Both are bit-identical to pbc_div for every pair closer than L/2 minus a few
// ulps; the only pairs where they can differ have |dx| within ulps of L/2, which
// is beyond every cutoff, so forces, energies, stresses, contacts and neighbour
// sets are unchanged.  Without `fast` the reference's division form is used.
__device__ __forceinline__ double pbc(double a, double b, double L, double invL, double halfL,
                                      int fast) {
    double d = a - b;
    if (fast) {
        if (fabs(d) < halfL) return d;
        return d - L * round(d * invL);
    }
    return d - L * round(d / L);
}

struct PairParams {
    double Lx, Ly, invLx, invLy, halfLx, halfLy;
    double e_c, n_c, ncm1;  // ncm1 = n_c - 1
    double inv2nc;          // 1/(2 n_c) when 2 n_c is a power of two (exact), else 0
    double two_nc;
    int fast;
    int fastdiv;            // n_c == 2 and e_c, every contact radius within [2^-100, 2^100]
};

// Correctly rounded square root and quotient for operands known to be normal numbers
// of moderate size.  These are the refinement sequences the compiler itself emits for
// IEEE sqrt and division on this target (hardware reciprocal-square-root / reciprocal
// seed, Newton refinement of the seed, residual corrections with FMA) without its
// operand pre-scaling and special-value fix-up stages.  Those stages change the result
// only for zero, infinite, NaN or subnormal operands, arguments below 2^-767 (sqrt),
// or quotients within 2^-768 of the overflow or underflow thresholds; pair_core uses
// these only when 2^-700 <= d2 and the host has bounded e_c and the radii, so every
// call returns the same bits as the IEEE operation.
__device__ __forceinline__ double sqrt_rn_normal(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double e = fma(-g, g, x);
    g = fma(e, h, g);
    e = fma(-g, g, x);
    return fma(e, h, g);
}

__device__ __forceinline__ double div_rn_normal(double a, double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    double q = a * y;
    double r = fma(-b, q, a);
    return fma(r, y, q);
}

// overlap^(n_c-1): n_c == 2 gives exactly overlap.
__device__ __forceinline__ double pow_ov(double ov, const PairParams& p) {
    return p.ncm1 == 1.0 ? ov : pow(ov, p.ncm1);
}

// calcPointPointInteraction, include/particles/base/kernels.cuh:154-183, for a
// given minimum-image separation.  Returns the half pair energy; writes the force
// on the first point.
__device__ __forceinline__ double pair_core(double dx, double dy, double r, double orad,
                                            const PairParams& p, double& fx, double& fy) {
    double rs = r + orad;
    double d2 = dx * dx + dy * dy;
    if (d2 >= rs * rs) {
        fx = 0.0;
        fy = 0.0;
        return 0.0;
    }
    double ov, op, fm;
    if (p.fastdiv && d2 >= 0x1p-700) {
        double d = sqrt_rn_normal(d2);
        ov = 1.0 - div_rn_normal(d, rs);
        op = ov;
        fm = div_rn_normal(p.e_c * op, rs * d);
    } else {
        double d = sqrt(d2);
        ov = 1.0 - (d / rs);
        op = pow_ov(ov, p);
        fm = p.e_c * op / (rs * d);
    }
    double num = p.e_c * ov * op;
    double e = p.inv2nc != 0.0 ? num * p.inv2nc : num / p.two_nc;
    fx = fm * dx;
    fy = fm * dy;
    return e;
}

__device__ __forceinline__ double pair_law(double px, double py, double r, double ox, double oy,
                                           double orad, const PairParams& p, double& fx,
                                           double& fy) {
    double dx = pbc(px, ox, p.Lx, p.invLx, p.halfLx, p.fast);
    double dy = pbc(py, oy, p.Ly, p.invLy, p.halfLy, p.fast);
    return pair_core(dx, dy, r, orad, p, fx, fy);
}

// Minimum image without a branch: d - L*rint(d/L) (round half to even).  It equals the
// reference's d - L*round(d/L) (half away from zero) except when d/L lies within an
// ulp of a half-integer, where both results have magnitude L/2, beyond every contact
// and list cutoff whenever the caller may use it (PairParams::fast).
__device__ __forceinline__ double pbc_rint(double a, double b, double L, double invL) {
    double d = a - b;
    double k = __builtin_rint(d * invL);
    return d - L * k;
}

// Pair force (and, with ENERGY, half pair energy) of one listed neighbour for the hot
// loops: calcPointPointInteraction (include/particles/base/kernels.cuh:154-183) with
// the branch-free minimum image, the contact test folded into selects and the short
// sqrt/division sequences; bit-identical to pair_law.  `valid` false (a slot past the
// row end) contributes +0.0, which leaves the running sums unchanged.  Needs
// p.fast && p.fastdiv (EXACT_IMAGE: p.fastdiv only, with the reference's minimum image);
// pair_law covers the other cases.
template <bool ENERGY, bool EXACT_IMAGE = false>
__device__ __forceinline__ double pair_hot(double4 qi, double4 q, bool valid, const PairParams& p,
                                           double& fx, double& fy) {
    double dx = EXACT_IMAGE ? pbc(qi.x, q.x, p.Lx, p.invLx, p.halfLx, 0)
                            : pbc_rint(qi.x, q.x, p.Lx, p.invLx);
    double dy = EXACT_IMAGE ? pbc(qi.y, q.y, p.Ly, p.invLy, p.halfLy, 0)
                            : pbc_rint(qi.y, q.y, p.Ly, p.invLy);
    double rs = qi.z + q.z;
    double d2 = dx * dx + dy * dy;
    bool contact = valid && d2 < rs * rs;
    double d = sqrt_rn_normal(d2);
    double ov = 1.0 - div_rn_normal(d, rs);
    double fm = div_rn_normal(p.e_c * ov, rs * d);
    if (contact && d2 < 0x1p-700) {  // separations the short sequences do not cover
        d = sqrt(d2);
        ov = 1.0 - (d / rs);
        fm = p.e_c * ov / (rs * d);
    }
    fx = contact ? fm * dx : 0.0;
    fy = contact ? fm * dy : 0.0;
    if (!ENERGY) return 0.0;
    double num = p.e_c * ov * ov;
    return contact ? num * p.inv2nc : 0.0;
}

// Pairwise-halving tree over the 64 lanes (lane t += lane t+s, s = 32..1);
// lane 0 receives the canonical chunk sum.
__device__ __forceinline__ double wave_tree_sum(double v) {
#pragma unroll
    for (int s = WAVE / 2; s > 0; s >>= 1) v = v + __shfl_down(v, s, WAVE);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int s = WAVE / 2; s > 0; s >>= 1) v = fmax(v, __shfl_down(v, s, WAVE));
    return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
#pragma unroll
    for (int s = WAVE / 2; s > 0; s >>= 1) v = max(v, __shfl_down(v, s, WAVE));
    return v;
}

// Block-wide maxima of two values at once (the two shuffle chains interleave);
// results valid in thread 0.
__device__ __forceinline__ void block_max2(double& a, double& b) {
    __shared__ double sm2[2][BLOCK / WAVE];
#pragma unroll
    for (int s = WAVE / 2; s > 0; s >>= 1) {
        double ta = __shfl_down(a, s, WAVE), tb = __shfl_down(b, s, WAVE);
        a = fmax(a, ta);
        b = fmax(b, tb);
    }
    int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    if (l == 0) {
        sm2[0][w] = a;
        sm2[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x / WAVE); ++k) {
            a = fmax(a, sm2[0][k]);
            b = fmax(b, sm2[1][k]);
        }
    }
    __syncthreads();
}

// Block-wide max over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ double block_max(double v) {
    __shared__ double sm[BLOCK / WAVE];
    v = wave_max(v);
    int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    if (l == 0) sm[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sm[0];
        for (int k = 1; k < (int)(blockDim.x / WAVE); ++k) r = fmax(r, sm[k]);
    }
    __syncthreads();
    return r;
}

// Block-wide max of an int (all threads must call); result valid in thread 0.
__device__ __forceinline__ int block_max_int(int v) {
    __shared__ int smi[BLOCK / WAVE];
    v = wave_max_i(v);
    int w = threadIdx.x / WAVE, l = threadIdx.x % WAVE;
    if (l == 0) smi[w] = v;
    __syncthreads();
    int r = 0;
    if (threadIdx.x == 0) {
        r = smi[0];
        for (int k = 1; k < (int)(blockDim.x / WAVE); ++k) r = max(r, smi[k]);
    }
    __syncthreads();
    return r;
}

// Final stage of the canonical sum over nc chunk partials, done by one wave:
// lane t sums P[t], P[t+64], ... in order starting from 0.0, then the lanes are
// tree-reduced.  Result valid in lane 0.
// The loads of eight rows are issued before their additions (the partials were just
// written by other XCDs, so each load is a long-latency trip); the additions keep
// the order.
constexpr int FS_ROWS = 8;

__device__ __forceinline__ double wave_final_sum(const double* P, int64_t nc) {
    int t = threadIdx.x % WAVE;
    double s = 0.0;
    int64_t k = t;
    for (; k + (FS_ROWS - 1) * WAVE < nc; k += FS_ROWS * WAVE) {
        double v[FS_ROWS];
#pragma unroll
        for (int u = 0; u < FS_ROWS; ++u) v[u] = P[k + u * WAVE];
#pragma unroll
        for (int u = 0; u < FS_ROWS; ++u) s = s + v[u];
    }
    for (; k < nc; k += WAVE) s = s + P[k];
    return wave_tree_sum(s);
}

// wave_final_sum of four arrays at once (all loads of a row batch in flight together).
__device__ __forceinline__ void wave_final_sum4(const double* A, const double* B,
                                                const double* C, const double* D, int64_t nc,
                                                double& ra, double& rb, double& rc,
                                                double& rd) {
    int t = threadIdx.x % WAVE;
    double sa = 0.0, sb = 0.0, sc = 0.0, sd = 0.0;
    int64_t k = t;
    constexpr int R = FS_ROWS / 2;
    for (; k + (R - 1) * WAVE < nc; k += R * WAVE) {
        double va[R], vb[R], vc[R], vd[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            va[u] = A[k + u * WAVE];
            vb[u] = B[k + u * WAVE];
            vc[u] = C[k + u * WAVE];
            vd[u] = D[k + u * WAVE];
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            sa = sa + va[u];
            sb = sb + vb[u];
            sc = sc + vc[u];
            sd = sd + vd[u];
        }
    }
    for (; k < nc; k += WAVE) {
        sa = sa + A[k];
        sb = sb + B[k];
        sc = sc + C[k];
        sd = sd + D[k];
    }
    ra = wave_tree_sum(sa);
    rb = wave_tree_sum(sb);
    rc = wave_tree_sum(sc);
    rd = wave_tree_sum(sd);
}

// XCD-aware tile index.  Workgroups are dispatched round-robin over the 8
// XCDs (block b runs on XCD b % 8); this bijection gives XCD x the contiguous
// tile range [x*q + min(x,r), ...), so the spatially sorted particles a block
// gathers from stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int NX = 8;
    int q = nblocks / NX, r = nblocks % NX;
    int x = b % NX, k = b / NX;
    return x * q + min(x, r) + k;
}

}  // namespace dpm
