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
    int pad;
};

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
    double d = sqrt(d2);
    double ov = 1.0 - (d / rs);
    double op = pow_ov(ov, p);
    double num = p.e_c * ov * op;
    double e = p.inv2nc != 0.0 ? num * p.inv2nc : num / p.two_nc;
    double fm = p.e_c * op / (rs * d);
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

// Final stage of the canonical sum over nc chunk partials, done by one wave:
// lane t sums P[t], P[t+64], ... in order starting from 0.0, then the lanes are
// tree-reduced.  Result valid in lane 0.
__device__ __forceinline__ double wave_final_sum(const double* P, int64_t nc) {
    int t = threadIdx.x % WAVE;
    double s = 0.0;
    for (int64_t k = t; k < nc; k += WAVE) s = s + P[k];
    return wave_tree_sum(s);
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
