// dpm_core.hip -- gfx950 kernels and host control for one particle system
// (disk path, neighbour lists, reductions, NVE / FIRE / Adam loops, jamming).
//
// Layout in HBM (one system):
//   P0, P1      : double4[N] {x, y, radius, mass}; P0 is current between calls,
//                 fused NVE chunks ping-pong positions P0 <-> P1 (one 32 B gather
//                 per neighbour)
//   vel, frc, lref, cref : double2[N];  pe, ke : double[N]
//   nbr         : int32 ELL, column-major nbr[k*N + i] (lanes coalesce); nnum[N]
//   sidx, cell_of : int32[N];  cell_start : int32[ncells+1] counting-sort offsets
//   pmax_nd/cd  : double[ceil(N/256)] per-block displacement maxima
//   part        : double[4*ceil(N/64) + ...] canonical-sum chunk partials
//
// Execution model of the hot loops (NVE, FIRE): chunks of steps are captured in a
// hipGraph.  A step whose drift crossed the rebuild threshold halts itself (and the
// rest of the chunk); a gated rebuild chain at the end of every chunk performs the
// rebuild and finishes the halted step on the device, so the host only polls a
// pinned copy of the control block per chunk.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <type_traits>

#include "dpm_kernels.hpp"

namespace dpm {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

// =========================================================================
//                                  kernels
// =========================================================================

__device__ __forceinline__ int wrap_cell(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// The list row of particle i (kernelUpdateCellNeighborList, base/kernels.cu:440-505):
// 3x3 stencil, dy outer, dx inner, ascending sorted index within a cell, written to
// the column-major ELL.  `pos(j)` supplies the positions the list is built from.
template <class Pos>
__device__ __forceinline__ int stencil_build(const Dev& d, const NbrParams& nb, int64_t i,
                                             double2 xi, Pos pos) {
    const int ncd = nb.ncd;
    int cid = d.cell_of[i];
    int cx = cid % ncd, cy = cid / ncd;
    int s[9], e[9];
#pragma unroll
    for (int oy = -1; oy <= 1; ++oy) {
        int yy = wrap_cell(cy + oy, ncd);
#pragma unroll
        for (int ox = -1; ox <= 1; ++ox) {
            int c = wrap_cell(cx + ox, ncd) + yy * ncd;
            int t = (oy + 1) * 3 + (ox + 1);
            s[t] = d.cell_start[c];
            e[t] = d.cell_start[c + 1];
        }
    }
    int cnt = 0;
    const int cap = d.cap;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        for (int j = s[t]; j < e[t]; ++j) {
            if (j == i) continue;
            double2 xj = pos(j);
            double dx = pbc(xi.x, xj.x, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
            double dy = pbc(xi.y, xj.y, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
            if (dx * dx + dy * dy < nb.rc2) {  // isWithinCutoffSquared, kernels.cuh:249-260
                if (cnt < cap) d.nbr[(int64_t)cnt * d.N + i] = j;
                cnt++;
            }
        }
    }
    d.nnum[i] = cnt;
    return cnt;
}

// Sum of harmonic pair forces (and energies) over the ELL row of particle il, in row
// order: kernelCalcDiskForces, src/particles/disk/kernels.cu:11-57.  Every lane of the
// wave calls this (nn = 0 for lanes without a particle); the wave walks its longest
// row in batches of ROW_BATCH neighbours, gathering a batch's records together and
// loading the next batch's indices while the current batch is summed, strictly in
// row order.  Index loads use a column clamped into the ELL (every slot holds a valid
// particle index: the array starts zeroed and is only written with indices); slots
// past a lane's row end contribute +0.0, which leaves its sums unchanged.
constexpr int ROW_BATCH = 4;

// HOT selects pair_hot (the host passes it when p.fast && p.fastdiv); each kernel is
// instantiated for both so the hot loop carries no registers for the general path.
template <bool ENERGY, bool HOT>
__device__ __forceinline__ void row_forces(const Dev& d, const double4* P, const PairParams& p,
                                           int64_t il, double4 qi, int nn, double& ax,
                                           double& ay, double& e) {
    const int cap = d.cap;
    const int* col = d.nbr + il;
    const int64_t N = d.N;
    const int wmax = __builtin_amdgcn_readfirstlane(wave_max_i(nn));
    ax = 0.0;
    ay = 0.0;
    e = 0.0;
    int jn[ROW_BATCH];
#pragma unroll
    for (int t = 0; t < ROW_BATCH; ++t) jn[t] = col[(int64_t)min(t, cap - 1) * N];
    for (int k = 0; k < wmax; k += ROW_BATCH) {
        double4 q[ROW_BATCH];
#pragma unroll
        for (int t = 0; t < ROW_BATCH; ++t) q[t] = P[jn[t]];
        if (k + ROW_BATCH < wmax) {  // next batch's indices (wave-uniform test)
#pragma unroll
            for (int t = 0; t < ROW_BATCH; ++t)
                jn[t] = col[(int64_t)min(k + ROW_BATCH + t, cap - 1) * N];
        }
#pragma unroll
        for (int t = 0; t < ROW_BATCH; ++t) {
            if (k + t < wmax) {
                const bool valid = k + t < nn;
                double fx = 0.0, fy = 0.0, ie = 0.0;
                if (HOT) {
                    ie = pair_hot<ENERGY>(qi, q[t], valid, p, fx, fy);
                } else if (valid) {
                    ie = pair_law(qi.x, qi.y, qi.z, q[t].x, q[t].y, q[t].z, p, fx, fy);
                }
                ax += fx;
                ay += fy;
                if (ENERGY) e += ie;
            }
        }
    }
}

// row_forces for the fused NVE step, with the first ROW_PF index columns loaded by
// the caller ahead of everything that waits on the control block, batches of ROW_PF
// records gathered together, and the wave's slot bound taken from ballots instead of
// a shuffle reduction.  Same slots, same order, same arithmetic as row_forces.
constexpr int ROW_PF = 2;

__device__ __forceinline__ bool wave_any(bool x) {
    return __builtin_amdgcn_ballot_w64(x) != 0;
}

template <bool HOT>
__device__ __forceinline__ void row_forces_pf(const Dev& d, const double4* P, const PairParams& p,
                                              int64_t il, double4 qi, int nn, int (&jn)[ROW_PF],
                                              double& ax, double& ay) {
    const int cap = d.cap;
    const int* col = d.nbr + il;
    const int64_t N = d.N;
    ax = 0.0;
    ay = 0.0;
    for (int k = 0; wave_any(k < nn); k += ROW_PF) {
        double4 q[ROW_PF];
#pragma unroll
        for (int t = 0; t < ROW_PF; ++t) q[t] = P[jn[t]];  // every slot holds a valid index
        if (wave_any(k + ROW_PF < nn)) {  // next batch's indices
#pragma unroll
            for (int t = 0; t < ROW_PF; ++t) jn[t] = col[(int64_t)min(k + ROW_PF + t, cap - 1) * N];
        }
#pragma unroll
        for (int t = 0; t < ROW_PF; ++t) {
            if (wave_any(k + t < nn)) {
                const bool valid = k + t < nn;
                double fx = 0.0, fy = 0.0;
                if (HOT) {
                    pair_hot<false>(qi, q[t], valid, p, fx, fy);
                } else if (valid) {
                    pair_law(qi.x, qi.y, qi.z, q[t].x, q[t].y, q[t].z, p, fx, fy);
                }
                ax += fx;
                ay += fy;
            }
        }
    }
}

// ---- NVE (src/integrator/nve.cu:15-25) ---------------------------------------
// One launch per step: zero + forces + closing kick(dt/2) of step s, then the
// opening kick(dt/2) + drift(dt) of step s+1 (kernelCalcDiskForces,
// kernelUpdateVelocities twice, kernelUpdatePositions; base/kernels.cu:48-139).
// Positions are read from P[cur] and written to P[1-cur], so neighbours are read
// at their old positions.  The forces and energies of these intermediate steps are
// consumed in registers and never stored: the step that ends a run (k_nve_force)
// writes the arrays the caller can observe.
template <bool HOT>
__global__ __launch_bounds__(BLOCK) void k_nve_fused(Dev d, PairParams p, NbrParams nb,
                                                     double hdt, double dt, int cur, int pos,
                                                     int last) {
    // the control-block words this step reads, all at once
    Ctl* c = d.ctl;
    const int halted = c->halt;
    const long long e = c->epoch;
    const long long cfl = c->cflag[pos & 1], lfl = c->flag[pos & 1], ovf = c->ovf;
    const double4* Pc = cur ? d.P1 : d.P0;
    double4* Pn = cur ? d.P0 : d.P1;
    int tile;
    const int64_t i = particle_of_thread(tile);
    const bool live = i < d.N;
    const int64_t il = live ? i : 0;
    // loads that do not depend on the control block go first: their latency overlaps
    // the control-block read and each other
    const double4 qi = Pc[il];
    int nraw = d.nnum[il];
    const double2 v0 = d.vel[il];
    const double2 lc = d.cref[il];
    double2 ln = d.lref[il];
    int jn[ROW_PF];
#pragma unroll
    for (int t = 0; t < ROW_PF; ++t) jn[t] = d.nbr[(int64_t)min(t, d.cap - 1) * d.N + il];
    // compiler-only fence: every load above is issued before the branch below (none is
    // waited for here)
    __asm__ __volatile__("" ::: "memory");
    if (halted) return;
    // pending_rebuild and ovf_pending on the words read above
    const long long st0 = stamp(e, pos);
    const bool ovf_other = ovf != 0 && ovf != st0;
    int need = cfl == st0 ? 2 : (lfl == st0 ? 1 : (ovf_other ? 3 : 0));
    // list rebuilds of a cell list are done here, row by row; the rest halts the chunk
    bool in_kernel = need == 1 && nb.method == DPM_NBR_CELL && !ovf_other;
    if (need != 0 && !in_kernel) {
        if (threadIdx.x == 0) {
            c->halt = 1;
            c->halt_pos = pos;
            c->halt_mode = H_NVE_FUSED;
            c->halt_kind = need;
        }
        return;
    }
    int cnt = 0;
    if (in_kernel) {
        if (live) {  // rebuild this particle's row, then sum it like any other row
            ln = make_double2(qi.x, qi.y);
            cnt = stencil_build(d, nb, i, ln, [&](int j) {
                double4 q = Pc[j];
                return make_double2(q.x, q.y);
            });
            d.lref[i] = ln;
            if (cnt > d.cap) c->ovf = stamp(e, pos);
            nraw = cnt;
#pragma unroll
            for (int t = 0; t < ROW_PF; ++t) jn[t] = d.nbr[(int64_t)min(t, d.cap - 1) * d.N + il];
        }
    }
    const int nn = live ? min(nraw, d.cap) : 0;
    double ax, ay;
    row_forces_pf<HOT>(d, Pc, p, il, qi, nn, jn, ax, ay);
    double nd = 0.0, cd = 0.0;
    if (live) {
        double fx = 0.0 + ax, fy = 0.0 + ay;
        double dtm = hdt / qi.w;
        double2 v = v0;
        v.x += fx * dtm;  // closing half-kick of step s
        v.y += fy * dtm;
        v.x += fx * dtm;  // opening half-kick of step s+1
        v.y += fy * dtm;
        double x = qi.x + v.x * dt, y = qi.y + v.y * dt;
        d.vel[i] = v;
        Pn[i] = make_double4(x, y, qi.z, qi.w);
        double dxn = x - ln.x, dyn = y - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = x - lc.x, dyc = y - lc.y;
        cd = dxc * dxc + dyc * dyc;
    }
    if (in_kernel) {  // row-length statistics of the rebuild, one atomic per block
        int m = block_max_int(cnt);
        if (threadIdx.x == 0) atomicMax(&c->max_nnum_all, m);
        if (blockIdx.x == 0 && threadIdx.x == 0) c->n_rebuilds += 1;
    }
    long long st = last ? stamp(e + 1, 0) : stamp(e, pos + 1);
    if (last)  // the maxima of the chunk's final state, for the checks and queries after it
        disp_epilogue(d, nb, tile, nd, cd, (pos + 1) & 1, st);
    else
        disp_flags(d, nb, nd, cd, (pos + 1) & 1, st);
    if (blockIdx.x == 0 && threadIdx.x == 0) c->progress = pos + 1;
}

// Zero + forces + closing kick(dt/2) on P0 (the end of a run, the first half of a
// split step, or the resumption of a halted step).
template <bool HOT>
__device__ __forceinline__ void nve_force_body(const Dev& d, const PairParams& p, double hdt,
                                               int pos, int hmode, int gate) {
    Ctl* c = d.ctl;
    bool closes = hmode == H_NVE_FORCE;  // the closing force of a run
    if (gate == G_NORMAL) {
        if (c->halt || (closes && c->closed)) return;  // a run is closed once
        int need = pending_rebuild(c, c->epoch, pos);
        if (need) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                c->halt = 1;
                c->halt_pos = pos;
                c->halt_mode = hmode;
                c->halt_kind = need;
            }
            return;
        }
    } else if (gate == G_RESUME) {
        if (!c->halt || !resumable(d) || c->halt_mode == H_FIRE) return;
        closes = c->halt_mode == H_NVE_FORCE;
    }
    if (closes && blockIdx.x == 0 && threadIdx.x == 0) c->closed = 1;
    int tile;
    const int64_t i = particle_of_thread(tile);
    const bool live = i < d.N;
    const int64_t il = live ? i : 0;
    const double4 qi = d.P0[il];
    const int nn = live ? min(d.nnum[il], d.cap) : 0;
    double ax, ay, en;
    row_forces<true, HOT>(d, d.P0, p, il, qi, nn, ax, ay, en);
    if (!live) return;
    double fx = 0.0 + ax, fy = 0.0 + ay;
    d.frc[i] = make_double2(fx, fy);
    d.pe[i] = 0.0 + en;
    double dtm = hdt / qi.w;
    double2 v = d.vel[i];
    v.x += fx * dtm;
    v.y += fy * dtm;
    d.vel[i] = v;
}

template <bool HOT>
__global__ __launch_bounds__(BLOCK) void k_nve_force(Dev d, PairParams p, double hdt, int pos,
                                                     int hmode, int gate) {
    nve_force_body<HOT>(d, p, hdt, pos, hmode, gate);
}

// End of a launch unit that changes no loop state (the closing force of a run): publishes
// the control block (one wave).
__global__ __launch_bounds__(WAVE) void k_unit_end(Dev d) { publish_ctl(d); }

// Opening kick(dt/2) + drift(dt) in place on P0.  (wr_pos, epoch + wr_eoff) names
// the step that must check the rebuild request; progress >= 0 is recorded (a resumed
// step adds one to the count).
__global__ __launch_bounds__(BLOCK) void k_nve_kick_drift(Dev d, NbrParams nb, double hdt,
                                                          double dt, int wr_pos, int wr_eoff,
                                                          int progress, int gate) {
    Ctl* c = d.ctl;
    if (gate == G_NORMAL) {
        if (c->halt) return;
    } else if (gate == G_RESUME) {
        if (!c->halt || !resumable(d) || c->halt_mode != H_NVE_FUSED) return;
    }
    long long e = c->epoch;
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double4 q = d.P0[i];
        double2 v = d.vel[i], f = d.frc[i];
        double dtm = hdt / q.w;
        v.x += f.x * dtm;
        v.y += f.y * dtm;
        q.x += v.x * dt;
        q.y += v.y * dt;
        d.vel[i] = v;
        d.P0[i] = q;
        disp_of(d, i, q.x, q.y, nd, cd);
    }
    disp_epilogue(d, nb, tile, nd, cd, wr_pos & 1, stamp(e + wr_eoff, wr_pos));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (gate == G_RESUME) c->progress += 1;  // the finished halted step
        else if (progress >= 0) c->progress = progress;
    }
}

// ---- stand-alone pieces of the Particle interface -------------------------------
// Disk::calculateForces as called stand-alone: forces are accumulated (+=), exactly
// as the reference kernel does into caller-zeroed arrays.
template <bool HOT>
__global__ __launch_bounds__(BLOCK) void k_disk_forces_acc(Dev d, PairParams p) {
    int tile;
    const int64_t i = particle_of_thread(tile);
    const bool live = i < d.N;
    const int64_t il = live ? i : 0;
    const double4 qi = d.P0[il];
    const int nn = live ? min(d.nnum[il], d.cap) : 0;
    double ax, ay, e;
    row_forces<true, HOT>(d, d.P0, p, il, qi, nn, ax, ay, e);
    if (!live || nn == 0) return;  // the reference kernel skips empty rows
    double2 f = d.frc[i];
    f.x += ax;
    f.y += ay;
    d.frc[i] = f;
    d.pe[i] += e;
}

__global__ void k_zero_forces(Dev d) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    d.frc[i] = make_double2(0.0, 0.0);
    d.pe[i] = 0.0;
    if (d.torque) d.torque[i] = 0.0;
}

__global__ void k_zero_velocities(Dev d) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    d.vel[i] = make_double2(0.0, 0.0);
    if (d.omega) d.omega[i] = 0.0;
}

// kernelUpdateVelocities, src/particles/base/kernels.cu:117-139
__global__ void k_kick(Dev d, double dt) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i], f = d.frc[i];
    double dtm = dt / d.P0[i].w;
    v.x += f.x * dtm;
    v.y += f.y * dtm;
    d.vel[i] = v;
}

// kernelUpdatePositions, src/particles/base/kernels.cu:48-81
__global__ __launch_bounds__(BLOCK) void k_drift(Dev d, NbrParams nb, double dt) {
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double4 q = d.P0[i];
        double2 v = d.vel[i];
        q.x += v.x * dt;
        q.y += v.y * dt;
        d.P0[i] = q;
        disp_of(d, i, q.x, q.y, nd, cd);
    }
    disp_epilogue(d, nb, tile, nd, cd, -1, 0);
}

// kernelMixVelocitiesAndForces, src/particles/base/kernels.cu:91-112
__global__ void k_mix(Dev d, double alpha) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    mix_one(v, d.frc[i], alpha);
    d.vel[i] = v;
}

// kernelCalculateTranslationalKineticEnergy (base/kernels.cu:163-180) and the
// rigid version (rigid_bumpy/kernels.cu:254-269)
__global__ void k_kinetic(Dev d, int rigid_rot) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    double m = d.P0[i].w;
    double v2 = v.x * v.x + v.y * v.y;
    if (rigid_rot) {
        double w = d.omega[i];
        d.ke[i] = 0.5 * m * v2 + 0.5 * d.moi[i] * w * w;
    } else {
        d.ke[i] = 0.5 * m * v2;
    }
}

__device__ __forceinline__ double sum_value(const Dev& d, int what, int64_t i) {
    switch (what) {
        case S_PE: return d.pe[i];
        case S_KE: return d.ke[i];
        case S_FXVX: return d.frc[i].x * d.vel[i].x;
        case S_FYVY: return d.frc[i].y * d.vel[i].y;
        case S_FX: return d.frc[i].x;
        case S_FY: return d.frc[i].y;
        case S_SXX: return d.sxx[i];
        case S_SYY: return d.syy[i];
        case S_SXY: return d.sxy[i];
        case S_SYX: return d.syx[i];
        case S_VX: return d.vel[i].x;
        case S_VY: return d.vel[i].y;
        case S_R2: {
            double r = d.P0[i].z;
            return r * r;
        }
        case S_TQW: return d.torque[i] * d.omega[i];
        case S_COUNT: return (double)d.contacts[i];
    }
    return 0.0;
}

// First stage of the canonical sum: one wave per 64-chunk.
__global__ __launch_bounds__(BLOCK) void k_chunk_sum(Dev d, int what, double* out) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v = (i < d.N) ? sum_value(d, what, i) : 0.0;
    v = wave_tree_sum(v);
    int64_t c = i / CHUNK;
    if ((threadIdx.x % WAVE) == 0 && c * CHUNK < d.N) out[c] = v;
}

__global__ void k_final_sum(const double* part, int64_t nc, double* out) {
    double v = wave_final_sum(part, nc);
    if (threadIdx.x == 0) *out = v;
}

__device__ __forceinline__ double max_value(const Dev& d, int what, int64_t i) {
    switch (what) {
        case M_F2: {
            double2 f = d.frc[i];
            return f.x * f.x + f.y * f.y;
        }
        case M_T2: {
            double t = d.torque[i];
            return t * t;
        }
    }
    return 0.0;
}

__global__ __launch_bounds__(BLOCK) void k_block_max(Dev d, int what, double* out) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double v = (i < d.N) ? max_value(d, what, i) : 0.0;
    double b = block_max(v);
    if (threadIdx.x == 0) out[blockIdx.x] = b;
}

__global__ __launch_bounds__(BLOCK) void k_final_max(const double* part, int64_t n, double* out) {
    double m = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += BLOCK) m = fmax(m, part[k]);
    m = block_max(m);
    if (threadIdx.x == 0) *out = m;
}

// kernelCalcDiskStressTensor, src/particles/disk/kernels.cu:203-251
__global__ __launch_bounds__(BLOCK) void k_disk_stress(Dev d, PairParams p, double area) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double4 qi = d.P0[i];
    double2 v = d.vel[i];
    double m = qi.w;
    double axx = m * v.x * v.x, axy = m * v.x * v.y, ayx = axy, ayy = m * v.y * v.y;
    int nn = min(d.nnum[i], d.cap);
    for (int k = 0; k < nn; ++k) {
        int j = d.nbr[(int64_t)k * d.N + i];
        double4 qj = d.P0[j];
        double xd = pbc(qi.x, qj.x, p.Lx, p.invLx, p.halfLx, p.fast);
        double yd = pbc(qi.y, qj.y, p.Ly, p.invLy, p.halfLy, p.fast);
        double fx, fy;
        pair_law(qi.x, qi.y, qi.z, qj.x, qj.y, qj.z, p, fx, fy);
        axx += xd * fx / 2;
        axy += xd * fy / 2;
        ayx += yd * fx / 2;
        ayy += yd * fy / 2;
    }
    d.sxx[i] = axx / area;
    d.sxy[i] = axy / area;
    d.syx[i] = ayx / area;
    d.syy[i] = ayy / area;
}

// kernelCountDiskContacts, src/particles/disk/kernels.cu:177-201
__global__ __launch_bounds__(BLOCK) void k_disk_contacts(Dev d, PairParams p) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double4 qi = d.P0[i];
    int c = 0;
    int nn = min(d.nnum[i], d.cap);
    for (int k = 0; k < nn; ++k) {
        int j = d.nbr[(int64_t)k * d.N + i];
        double4 qj = d.P0[j];
        double dx = pbc(qi.x, qj.x, p.Lx, p.invLx, p.halfLx, p.fast);
        double dy = pbc(qi.y, qj.y, p.Ly, p.invLy, p.halfLy, p.fast);
        double dist = sqrt(dx * dx + dy * dy);
        if (dist < qi.z + qj.z) c++;
    }
    d.contacts[i] = c;
}

// positions.scale (Particle::scalePositionsFull, particle.cu:1070-1072); rigid:
// kernelScalePositions, rigid_bumpy/kernels.cu:1308-1333
__global__ void k_scale_positions(Dev d, double f) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double4 q = d.P0[i];
    if (d.vpos) {
        double nx = q.x * f, ny = q.y * f;
        int st = d.vstart[i], nv = d.nverts[i];
        for (int v = st; v < st + nv; ++v) {
            double2 w = d.vpos[v];
            w.x += (nx - q.x);
            w.y += (ny - q.y);
            d.vpos[v] = w;
        }
        q.x = nx;
        q.y = ny;
    } else {
        q.x *= f;
        q.y *= f;
    }
    d.P0[i] = q;
}

__global__ void k_add_velocity(Dev d, double ax, double ay, double scale) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    if (scale != 0.0) {
        v.x *= scale;
        v.y *= scale;
    } else {
        v.x -= ax;
        v.y -= ay;
    }
    d.vel[i] = v;
}

// kernelAdamStep, src/particles/base/kernels.cu:550-611
__global__ __launch_bounds__(BLOCK) void k_adam_update(Dev d, NbrParams nb, double alpha,
                                                       double b1, double b2, double omb1,
                                                       double omb2, double eps) {
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double2 f = d.frc[i], m1 = d.m1[i], m2 = d.m2[i];
        m1.x = fma(b1, m1.x, (b1 - 1) * f.x);
        m1.y = fma(b1, m1.y, (b1 - 1) * f.y);
        m2.x = fma(b2, m2.x, (1 - b2) * f.x * f.x);
        m2.y = fma(b2, m2.y, (1 - b2) * f.y * f.y);
        double mhx = m1.x / omb1, mhy = m1.y / omb1;
        double vhx = m2.x / omb2, vhy = m2.y / omb2;
        double ux = -alpha * mhx / (sqrt(vhx) + eps);
        double uy = -alpha * mhy / (sqrt(vhy) + eps);
        double4 q = d.P0[i];
        q.x += ux;
        q.y += uy;
        d.P0[i] = q;
        disp_of(d, i, q.x, q.y, nd, cd);
        d.m1[i] = m1;
        d.m2[i] = m2;
    }
    disp_epilogue(d, nb, tile, nd, cd, -1, 0);
}

// ------------------------------------------------------------- neighbours
// Every rebuild kernel takes `gated`: when set it runs only if the chain chose a
// rebuild that includes it: the cell kernels for kind 2 (cells + list), the list
// kernels for kinds 1 (list), 2 and 3 (list from the stored build positions).  The cell
// kernels also bin the grid that Verlet rows are built from; there they run with every
// list rebuild (gated == 2).
__device__ __forceinline__ bool cell_gate(const Ctl* c, int gated) {
    return gated == 0 || (gated == 2 ? c->kind >= 1 : c->kind == 2);
}
#define GATE_CELL \
    if (!cell_gate(d.ctl, gated)) return;
#define GATE_LIST \
    if (gated && d.ctl->kind < 1) return;

__global__ void k_fill_int(int* a, int v, int64_t n, const Ctl* gate, int gated) {
    if (gate && !cell_gate(gate, gated)) return;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}

__device__ __forceinline__ int cell_coord(double pos, double cs, int n) {
    // getPBCCellIndex, include/particles/base/kernels.cuh:285-301
    long long idx = (long long)floor(pos / cs);
    return (int)(idx >= 0 ? idx % n : (idx % n + n) % n);
}

// kernelGetCellIndexForParticle (base/kernels.cu:309-330) + counting-sort histogram
__global__ void k_cell_bin(Dev d, NbrParams nb, int gated) {
    GATE_CELL;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double4 q = d.P0[i];
    long long c = (long long)cell_coord(q.x, nb.cs, nb.ncd) +
                  (long long)cell_coord(q.y, nb.cs, nb.ncd) * nb.ncd;
    if (!(c >= 0 && c < nb.ncells)) {  // non-finite position: keep the index in range, report
        c = 0;
        d.ctl->bad_pos = 1;
    }
    d.cell_tmp[i] = (int)c;
    atomicAdd(&d.cell_cnt[c], 1);
}

// Exclusive scan of the histogram into cell_start, in three parallel passes
// (block sums, scan of the block sums, per-block scan + offset).  Replaces
// kernelGetFirstParticleIndexForCell's O(n_cells*N) search (base/kernels.cu:332-438);
// empty cells get start == end instead of -1, which the stencil loop handles
// identically (base/kernels.cu:470-478).
constexpr int SCAN_TILE = 1024;  // cells per block (4 per thread)

__global__ __launch_bounds__(BLOCK) void k_scan_a(Dev d, long long nc, int gated) {
    GATE_CELL;
    __shared__ int sm[BLOCK / WAVE];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    int s = 0;
    for (int t = 0; t < 4; ++t)
        if (base + t < nc) s += d.cell_cnt[base + t];
    for (int o = WAVE / 2; o > 0; o >>= 1) s += __shfl_down(s, o, WAVE);
    if (threadIdx.x % WAVE == 0) sm[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) d.scan_bsum[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

__global__ __launch_bounds__(1024) void k_scan_b(Dev d, long long nb_blocks, int gated) {
    GATE_CELL;
    __shared__ int sums[1024];
    int t = threadIdx.x;
    long long per = (nb_blocks + 1023) / 1024;
    long long lo = t * per, hi = min(lo + per, nb_blocks);
    int loc = 0;
    for (long long b = lo; b < hi; ++b) loc += d.scan_bsum[b];
    sums[t] = loc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        int v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int run = sums[t] - loc;
    for (long long b = lo; b < hi; ++b) {
        int x = d.scan_bsum[b];
        d.scan_bsum[b] = run;
        run += x;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_c(Dev d, long long nc, int gated) {
    GATE_CELL;
    __shared__ int sm[BLOCK];
    long long base = (long long)blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    int v[4], s = 0;
    for (int t = 0; t < 4; ++t) {
        v[t] = (base + t < nc) ? d.cell_cnt[base + t] : 0;
        s += v[t];
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
        int x = ((int)threadIdx.x >= off) ? sm[threadIdx.x - off] : 0;
        __syncthreads();
        sm[threadIdx.x] += x;
        __syncthreads();
    }
    int run = d.scan_bsum[blockIdx.x] + sm[threadIdx.x] - s;
    for (int t = 0; t < 4; ++t) {
        if (base + t < nc) {
            d.cell_start[base + t] = run;
            d.cell_cur[base + t] = run;
        }
        run += v[t];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) d.cell_start[nc] = (int)d.N;
}

__global__ void k_cell_scatter(Dev d, int gated) {
    GATE_CELL;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    int c = d.cell_tmp[i];
    int p = atomicAdd(&d.cell_cur[c], 1);
    d.perm[p] = (int)i;
}

// Restore stability inside each cell: ascending previous index, which makes the
// order identical to thrust::sort_by_key's stable radix sort (particle.cu:1293).
__global__ void k_cell_sort(Dev d, long long nc, int gated) {
    GATE_CELL;
    long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= nc) return;
    int s = d.cell_start[c], e = d.cell_start[c + 1];
    for (int a = s + 1; a < e; ++a) {
        int key = d.perm[a];
        int b = a - 1;
        while (b >= s && d.perm[b] > key) {
            d.perm[b + 1] = d.perm[b];
            --b;
        }
        d.perm[b + 1] = key;
    }
}

// kernelReorderParticleData, base/kernels.cu:508-543 (+ Adam moments, static ids,
// rigid-body particle arrays of kernelReorderRigidBumpyParticleData,
// rigid_bumpy/kernels.cu:796-842)
__global__ void k_reorder_gather(Dev d, Spare sp, int gated) {
    GATE_CELL;
    int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= d.N) return;
    int o = d.perm[k];
    double4 q = d.P0[o];
    sp.P[k] = q;
    sp.vel[k] = d.vel[o];
    sp.frc[k] = d.frc[o];
    sp.sidx[k] = d.sidx[o];
    d.cref[k] = make_double2(q.x, q.y);
    d.cell_of[k] = d.cell_tmp[o];
    if (d.m1) {
        sp.m1[k] = d.m1[o];
        sp.m2[k] = d.m2[o];
    }
    if (d.ma) sp.ma[k] = d.ma[o];
    if (d.vpos) {
        sp.angle[k] = d.angle[o];
        sp.omega[k] = d.omega[o];
        sp.torque[k] = d.torque[o];
        sp.moi[k] = d.moi[o];
        sp.nverts[k] = d.nverts[o];
        sp.old_to_new[o] = (int)k;
    }
}

__global__ __launch_bounds__(BLOCK) void k_reorder_copyback(Dev d, Spare sp, int gated) {
    GATE_CELL;
    int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (threadIdx.x == 0) d.pmax_cd[blockIdx.x] = 0.0;  // the grid has one block per tile
    if (k >= d.N) return;
    double4 q = sp.P[k];
    d.P0[k] = q;
    d.P1[k] = q;
    d.vel[k] = sp.vel[k];
    d.frc[k] = sp.frc[k];
    d.sidx[k] = sp.sidx[k];
    if (d.m1) {
        d.m1[k] = sp.m1[k];
        d.m2[k] = sp.m2[k];
    }
    if (d.ma) d.ma[k] = sp.ma[k];
    if (d.vpos) {
        d.angle[k] = sp.angle[k];
        d.omega[k] = sp.omega[k];
        d.torque[k] = sp.torque[k];
        d.moi[k] = sp.moi[k];
        d.nverts[k] = sp.nverts[k];
    }
}

__device__ __forceinline__ bool within(double ax, double ay, double bx, double by,
                                       const NbrParams& nb, double c2) {
    // isWithinCutoffSquared, include/particles/base/kernels.cuh:249-260
    double dx = pbc(ax, bx, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
    double dy = pbc(ay, by, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
    return dx * dx + dy * dy < c2;
}

// kernelUpdateCellNeighborList, base/kernels.cu:440-505: 3x3 stencil in the
// reference's order (dy outer, dx inner, ascending sorted index), written to the
// column-major ELL.  The nine cell ranges are loaded up front.

// kind 3 rebuilds the list of the last in-kernel rebuild from its build positions
// (lref) without touching the displacement state.
__global__ __launch_bounds__(BLOCK) void k_build_cell(Dev d, NbrParams nb, int gated) {
    GATE_LIST;
    const bool from_ref = d.ctl->kind == 3;
    int tile;
    int64_t i = particle_of_thread(tile);
    if (threadIdx.x == 0 && !from_ref) d.pmax_nd[tile] = 0.0;
    int cnt = 0;
    if (i < d.N) {
        if (from_ref) {
            cnt = stencil_build(d, nb, i, d.lref[i], [&](int j) { return d.lref[j]; });
        } else {
            double4 q = d.P0[i];
            double2 xi = make_double2(q.x, q.y);
            cnt = stencil_build(d, nb, i, xi, [&](int j) {
                double4 qj = d.P0[j];
                return make_double2(qj.x, qj.y);
            });
            d.lref[i] = xi;
        }
    }
    // row-length statistics: one atomic per block (same-address atomics serialise)
    int m = block_max_int(cnt);
    if (threadIdx.x == 0) {
        atomicMax(&d.ctl->max_nnum, m);
        atomicMax(&d.ctl->max_nnum_all, m);
    }
}


// kernelUpdateNeighborList, base/kernels.cu:256-293: all pairs, ascending j,
// j positions staged through LDS.
__global__ __launch_bounds__(BLOCK) void k_build_verlet(Dev d, NbrParams nb, int gated) {
    GATE_LIST;
    __shared__ double2 tp[BLOCK];
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (threadIdx.x == 0) d.pmax_nd[blockIdx.x] = 0.0;
    double4 qi = (i < d.N) ? d.P0[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    int cnt = 0;
    const int cap = d.cap;
    for (int64_t base = 0; base < d.N; base += BLOCK) {
        __syncthreads();
        if (base + threadIdx.x < d.N) {
            double4 q = d.P0[base + threadIdx.x];
            tp[threadIdx.x] = make_double2(q.x, q.y);
        }
        __syncthreads();
        int lim = (int)min((int64_t)BLOCK, d.N - base);
        if (i < d.N) {
            for (int t = 0; t < lim; ++t) {
                int64_t j = base + t;
                if (j == i) continue;
                if (within(qi.x, qi.y, tp[t].x, tp[t].y, nb, nb.rc2)) {
                    if (cnt < cap) d.nbr[(int64_t)cnt * d.N + i] = (int)j;
                    cnt++;
                }
            }
        }
    }
    if (i < d.N) {
        d.nnum[i] = cnt;
        d.lref[i] = make_double2(qi.x, qi.y);
    }
    int m = block_max_int(i < d.N ? cnt : 0);
    if (threadIdx.x == 0) {
        atomicMax(&d.ctl->max_nnum, m);
        atomicMax(&d.ctl->max_nnum_all, m);
    }
}

// kernelUpdateNeighborList's rows (base/kernels.cu:256-293: all pairs, ascending j) built
// from a cell grid whose cells are at least the cutoff wide: every particle within the
// cutoff of i lies in the 3 x 3 cells around i's, each cell lists its particles in
// ascending index (stable counting sort of unmoved particles), and the nine lists are
// merged in ascending order, so each row holds the same indices in the same order as
// the all-pairs scan.  Particles are binned by k_cell_bin from P0 and are not reordered.
__global__ __launch_bounds__(BLOCK) void k_build_verlet_grid(Dev d, NbrParams nb, int gated) {
    GATE_LIST;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (threadIdx.x == 0) d.pmax_nd[blockIdx.x] = 0.0;
    int cnt = 0;
    if (i < d.N) {
        const double4 qi = d.P0[i];
        const int ncd = nb.ncd;
        const int cid = d.cell_tmp[i];
        const int cx = cid % ncd, cy = cid / ncd;
        int s[9], e[9], h[9];  // cursor, end and head index of each cell's list
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy) {
            const int yy = wrap_cell(cy + oy, ncd);
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox) {
                const int c = wrap_cell(cx + ox, ncd) + yy * ncd;
                const int t = (oy + 1) * 3 + (ox + 1);
                s[t] = d.cell_start[c];
                e[t] = d.cell_start[c + 1];
                h[t] = s[t] < e[t] ? d.perm[s[t]] : INT_MAX;
            }
        }
        const int cap = d.cap;
        for (;;) {
            int best = INT_MAX, bt = -1;
#pragma unroll
            for (int t = 0; t < 9; ++t)
                if (h[t] < best) {
                    best = h[t];
                    bt = t;
                }
            if (bt < 0) break;
#pragma unroll
            for (int t = 0; t < 9; ++t)
                if (t == bt) {
                    s[t]++;
                    h[t] = s[t] < e[t] ? d.perm[s[t]] : INT_MAX;
                }
            if (best == i) continue;
            const double4 qj = d.P0[best];
            if (within(qi.x, qi.y, qj.x, qj.y, nb, nb.rc2)) {
                if (cnt < cap) d.nbr[(int64_t)cnt * d.N + i] = best;
                cnt++;
            }
        }
        d.nnum[i] = cnt;
        d.lref[i] = make_double2(qi.x, qi.y);
    }
    int m = block_max_int(i < d.N ? cnt : 0);
    if (threadIdx.x == 0) {
        atomicMax(&d.ctl->max_nnum, m);
        atomicMax(&d.ctl->max_nnum_all, m);
    }
}

// The per-block displacement maxima of the current positions (P0) against the list and
// cell reference positions: what disp_epilogue leaves behind, rebuilt after a loop whose
// steps decide their rebuilds by wave ballots (disp_flags), for the checks and queries
// that follow it.
__global__ __launch_bounds__(BLOCK) void k_disp_partials(Dev d, NbrParams nb) {
    int tile;
    const int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double4 q = d.P0[i];
        disp_of(d, i, q.x, q.y, nd, cd);
    }
    disp_epilogue(d, nb, tile, nd, cd, -1, 0);
}

// ------------------------------------------------------------- rebuild chain
// mode 0: end of a chunk (acts only if a step halted); 1: check (decide from the
// current displacement maxima); 2: forced full rebuild (initialisation, restore).
__global__ __launch_bounds__(BLOCK) void k_chain_decide(Dev d, NbrParams nb, int mode) {
    Ctl* c = d.ctl;
    if (c->overflow || (mode == 0 && !c->halt)) {
        if (threadIdx.x == 0) c->kind = 0;
        return;
    }
    int64_t nblk = (d.N + BLOCK - 1) / BLOCK;
    double mn = 0.0, mc = 0.0;
    for (int64_t k = threadIdx.x; k < nblk; k += BLOCK) {
        mn = fmax(mn, d.pmax_nd[k]);
        mc = fmax(mc, d.pmax_cd[k]);
    }
    mn = block_max(mn);
    mc = block_max(mc);
    if (threadIdx.x != 0) return;
    int kind = 0;
    if (mode == 0) {
        kind = c->halt_kind;
    } else if (mode == 2) {
        kind = nb.method == DPM_NBR_CELL ? 2 : 1;
    } else if (nb.method == DPM_NBR_CELL) {
        if (3.0 * mc > nb.cthr) kind = 2;
        else if (3.0 * mn > nb.thr) kind = 1;
    } else if (nb.method == DPM_NBR_VERLET) {
        if (3.0 * mn > nb.thr) kind = 1;
    }
    c->kind = kind;
    if (kind >= 1) {
        c->max_nnum = 0;
        c->max_vnum = 0;
    }
    if (kind == 1 || kind == 2) c->n_rebuilds += 1;
    if (kind == 2) c->n_cell_rebuilds += 1;
}

// A fused NVE step halted at an odd position left the current positions in P1.
__global__ void k_chain_copy_cur(Dev d) {
    const Ctl* c = d.ctl;
    if (!c->halt || c->halt_mode != H_NVE_FUSED || !(c->halt_pos & 1)) return;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double4 a = d.P1[i], b = d.P0[i];
    d.P0[i] = make_double4(a.x, a.y, b.z, b.w);
}

// One wave.  mode 0 ends a launch unit (a chunk's chain or a recovery) and publishes the
// control block; modes 1 and 2 end a stand-alone rebuild.
__global__ __launch_bounds__(WAVE) void k_chain_end(Dev d, int mode) {
    Ctl* c = d.ctl;
    if (threadIdx.x == 0) {
        bool over = c->max_nnum > d.cap || (d.vnbr != nullptr && c->max_vnum > d.vcap);
        if (mode != 0) {
            if (c->kind >= 1 && over) c->overflow = 1;
        } else {
            if (c->halt) {
                if (c->overflow || over) {
                    c->overflow = 1;
                } else {
                    c->halt = 0;
                    c->ovf = 0;  // the whole list was rebuilt with enough room
                }
            }
            c->last_progress = c->progress;
            c->progress = 0;
            c->epoch += 1;
        }
    }
    if (mode != 0) return;
    __syncthreads();
    publish_ctl(d);
}

// End of an NVE chunk (one wave): the chunk's completed steps are published and the next
// epoch starts.  A halted step stays halted; the host finishes it (rebuild chain) when it
// reads the status.
__global__ __launch_bounds__(WAVE) void k_chunk_end(Dev d) {
    Ctl* c = d.ctl;
    if (threadIdx.x == 0) {
        c->last_progress = c->progress;
        c->progress = 0;
        c->epoch += 1;
    }
    __syncthreads();
    publish_ctl(d);
}

// Host fix after a capacity overflow: rebuild only the list, without counting a
// new rebuild (the reference regrows inside the same update call).
__global__ void k_overflow_begin(Dev d) {
    Ctl* c = d.ctl;
    c->overflow = 0;
    if (c->kind != 3) c->kind = 1;
    c->max_nnum = 0;
    c->max_vnum = 0;
}

__global__ void k_ctl_reset(Dev d, int config) {
    Ctl* c = d.ctl;
    c->halt = 0;
    c->progress = 0;
    c->last_progress = 0;
    c->done = 0;
    c->overflow = 0;
    c->kind = 0;
    c->closed = 0;
    c->pub_seq = 0;
    c->epoch += 1;  // requests stamped before the reset never match a later step
    if (config) {  // a new neighbour configuration restarts the capacity statistics
        c->max_nnum = 0;
        c->max_nnum_all = 0;
        c->max_vnum = 0;
        c->max_vnum_all = 0;
    }
}

// ------------------------------------------------------------- FIRE loop
// One FIRE iteration (src/integrator/fire.cu:13-46) is three launches:
//   k_fire_move   : [uphill back-step, v=0] + kick(dt/2) + mix + drift(dt)
//   k_fire_force  : zero + forces + kick(dt/2) + chunk partials of P, PE, |F|max
//   k_fire_ctl    : canonical sums, minimizeFire's stop test for this step and
//                   Fire::step's decision for the next (one wave)
__global__ __launch_bounds__(BLOCK) void k_fire_move(Dev d, NbrParams nb, int pos) {
    Ctl* c = d.ctl;
    if (c->halt) return;
    const FireCtl* F = d.fire;
    if (F->done) return;
    long long e = c->epoch;
    double dt = F->dt, alpha = F->alpha;
    int uphill = F->uphill;
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double4 q = d.P0[i];
        double2 v = d.vel[i], f = d.frc[i];
        if (uphill) {  // updatePositions(-dt/2) with the decreased dt; setVelocitiesToZero
            double bdt = -dt / 2.0;
            q.x += v.x * bdt;
            q.y += v.y * bdt;
            v.x = 0.0;
            v.y = 0.0;
        }
        double dtm = (dt / 2.0) / q.w;
        v.x += f.x * dtm;
        v.y += f.y * dtm;
        mix_one(v, f, alpha);
        q.x += v.x * dt;
        q.y += v.y * dt;
        d.P0[i] = q;
        d.vel[i] = v;
        disp_of(d, i, q.x, q.y, nd, cd);
    }
    (void)tile;
    disp_flags(d, nb, nd, cd, pos & 1, stamp(e, pos));  // maxima: k_disp_partials at the end
}

// kick: 0 = forces only (minimizeFire's initial evaluation), 1 = + closing kick
template <bool HOT>
__global__ __launch_bounds__(BLOCK) void k_fire_force(Dev d, PairParams p, int pos, int kick,
                                                      int gate) {
    Ctl* c = d.ctl;
    const FireCtl* F = d.fire;
    if (gate == G_NORMAL) {
        if (c->halt || F->done) return;
        int need = pending_rebuild(c, c->epoch, pos);
        if (need) {
            if (threadIdx.x == 0) {
                c->halt = 1;
                c->halt_pos = pos;
                c->halt_mode = H_FIRE;
                c->halt_kind = need;
            }
            return;
        }
    } else if (gate == G_RESUME) {
        if (!c->halt || !resumable(d) || c->halt_mode != H_FIRE) return;
    }
    double half = F->dt / 2.0;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    const bool live = i < d.N;
    const int64_t il = live ? i : 0;
    const double4 qi = d.P0[il];
    const int nn = live ? min(d.nnum[il], d.cap) : 0;
    double ax, ay, e;
    row_forces<true, HOT>(d, d.P0, p, il, qi, nn, ax, ay, e);
    double pxv = 0.0, pyv = 0.0, pe = 0.0, f2 = 0.0;
    if (live) {
        double fx = 0.0 + ax, fy = 0.0 + ay;
        d.frc[i] = make_double2(fx, fy);
        pe = 0.0 + e;
        d.pe[i] = pe;
        double2 v = d.vel[i];
        if (kick) {
            double dtm = half / qi.w;
            v.x += fx * dtm;
            v.y += fy * dtm;
            d.vel[i] = v;
        }
        pxv = fx * v.x;
        pyv = fy * v.y;
        f2 = fx * fx + fy * fy;
    }
    int64_t nc = (d.N + CHUNK - 1) / CHUNK;
    double* P = d.part;
    double s0 = wave_tree_sum(pxv);
    double s1 = wave_tree_sum(pyv);
    double s2 = wave_tree_sum(pe);
    int64_t ch = i / CHUNK;
    if ((threadIdx.x % WAVE) == 0 && ch < nc) {
        P[ch] = s0;
        P[nc + ch] = s1;
        P[2 * nc + ch] = s2;
    }
    double b = block_max(f2);
    if (threadIdx.x == 0) P[3 * nc + blockIdx.x] = b;
}

__device__ __forceinline__ void fire_finish(const Dev& d, FireCtl* F) {
    F->done = 1;
    d.ctl->done = 1;
}

// init: 1 = minimizeFire's prologue (records the first average energy)
__global__ void k_fire_ctl(Dev d, int init, int pos, int gate) {
    Ctl* c = d.ctl;
    FireCtl* F = d.fire;
    if (gate == G_NORMAL) {
        if (c->halt || F->done) return;
    } else if (gate == G_RESUME) {
        if (!c->halt || !resumable(d) || c->halt_mode != H_FIRE) return;
    }
    int64_t nc = (d.N + CHUNK - 1) / CHUNK;
    int64_t nblk = (d.N + BLOCK - 1) / BLOCK;
    const double* P = d.part;
    double px, py, pe, ptw;
    // (the fourth array is the rigid spin power; a disk system sums px's array again)
    wave_final_sum4(P, P + nc, P + 2 * nc, d.vpos ? P + 3 * nc + nblk : P, nc, px, py, pe, ptw);
    if (!d.vpos) ptw = 0.0;
    double fm = 0.0;
    {
        int64_t k = threadIdx.x;
        for (; k + (FS_ROWS - 1) * WAVE < nblk; k += FS_ROWS * WAVE) {
            double v[FS_ROWS];
#pragma unroll
            for (int u = 0; u < FS_ROWS; ++u) v[u] = P[3 * nc + k + u * WAVE];
#pragma unroll
            for (int u = 0; u < FS_ROWS; ++u) fm = fmax(fm, v[u]);
        }
        for (; k < nblk; k += WAVE) fm = fmax(fm, P[3 * nc + k]);
    }
    fm = wave_max(fm);
    if (threadIdx.x != 0) return;
    if (!init) {
        if (gate == G_RESUME) c->progress += 1;  // the finished halted step
        else c->progress = pos + 1;
    }
    double avg = pe / F->dof;
    F->fmax_sq = fm;
    F->avg_pe = avg;
    if (init) {
        F->last_avg_pe = avg;
        if (F->max_steps <= 0) {
            fire_finish(d, F);
            return;
        }
    } else {
        // minimizeFire stop test (src/routines/minimization.cu:99-105)
        double diff = fabs(avg / F->last_avg_pe - 1);
        F->last_avg_pe = avg;
        bool fok = F->fmax_target_sq > 0.0 && fm < F->fmax_target_sq;
        F->step += 1;
        if (diff < F->avg_pe_diff_target || avg < F->avg_pe_target || fok) {
            F->converged = 1;
            fire_finish(d, F);
            return;
        }
        if (F->step >= F->max_steps) {
            fire_finish(d, F);
            return;
        }
    }
    // Fire::step decision for the next iteration (src/integrator/fire.cu:14-33);
    // RigidBumpy::getPowerFire adds sum(tau*w) (rigid_bumpy.cu:1181-1191)
    double power = px + py;
    if (d.vpos) power = power + ptw;
    F->power = power;
    if (power > 0) {
        F->n_good++;
        F->n_bad = 0;
        if (F->n_good > 5) {
            F->dt = fmin(F->dt * 1.1, F->dt_max);
            F->alpha *= 0.99;
        }
        F->uphill = 0;
    } else {
        F->n_good = 0;
        F->n_bad++;
        if (F->n_bad > 10) {
            F->stopped = 1;
            F->step += 1;  // the stopped step counts as executed
            fire_finish(d, F);
            return;
        }
        F->dt = fmax(F->dt * 0.5, F->dt_min);
        F->alpha = F->alpha_init;
        F->uphill = 1;
    }
}

}  // namespace dpm

// =========================================================================
//                             host-side helpers
// =========================================================================
// Launches the HOT instantiation of a pair-force kernel when its PairParams allow the
// short path (the PairParams argument must be named p at the call site).
#define LAUNCH_HOT(kern, ...)                                  \
    do {                                                       \
        if (p.fast && p.fastdiv)                               \
            hipLaunchKernelGGL((kern<true>), __VA_ARGS__);     \
        else                                                   \
            hipLaunchKernelGGL((kern<false>), __VA_ARGS__);    \
    } while (0)

namespace dpm {

// Graph captures, device allocations and frees of all systems are serialised: several
// systems may be driven from several host threads at once, and an allocation or free
// must not fall inside another thread's capture.
std::mutex& capture_mutex() {
    static std::mutex m;
    return m;
}

// Fills and copies go through the system's own stream (never the legacy stream, which
// would tie them to graph captures running for other systems in other threads).
int dalloc_bytes(dpm_system* s, void** p, size_t bytes) {
    void* q = nullptr;
    {
        std::lock_guard<std::mutex> g(capture_mutex());
        DPM_HIP(hipMalloc(&q, bytes));
    }
    DPM_HIP(hipMemsetAsync(q, 0, bytes, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    s->allocs.push_back(q);
    *p = q;
    return 0;
}

void dfree(dpm_system* s, void* p) {
    if (!p) return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
    if (it != s->allocs.end()) s->allocs.erase(it);
    std::lock_guard<std::mutex> g(capture_mutex());
    (void)hipFree(p);
}

int64_t pow2_at_least(int64_t m) {
    return (int64_t)std::pow(2.0, std::ceil(std::log2((double)m)));
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kernel launch: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

// Host<->device copies ordered on the system stream (the stream is non-blocking,
// so the legacy null stream would not wait for kernels queued on it).
int scopy(dpm_system* s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    DPM_HIP(hipMemcpyAsync(dst, src, bytes, kind, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

// The fast minimum image is exact when L/2 exceeds every contact and list cutoff.
static bool fast_ok(const dpm_system* s) {
    double lim = 2.0 * s->rmax;
    if (s->method == DPM_NBR_ALL) return false;  // its cutoff is 10 box diagonals
    if (s->method == DPM_NBR_CELL || s->method == DPM_NBR_VERLET) lim = std::max(lim, s->rc);
    if (s->kind == DPM_BUMPY) lim = std::max({lim, 2.0 * s->rv, s->rc_v, s->pcut_v});
    double half = 0.5 * std::min(s->box[0], s->box[1]);
    return half > lim * (1.0 + 1e-9);
}

// The short correctly rounded sqrt/division sequences apply (see sqrt_rn_normal) when
// n_c == 2 and e_c and every contact radius lie within [2^-100, 2^100].
static bool fastdiv_ok(const dpm_system* s) {
    const double lo = 0x1p-100, hi = 0x1p100;
    double rmin = s->kind == DPM_BUMPY ? s->rv : s->rmin;
    double rmax = s->kind == DPM_BUMPY ? s->rv : s->rmax;
    return s->n_c == 2.0 && s->e_c >= lo && s->e_c <= hi && rmin >= lo && rmax <= hi;
}

PairParams params_of(const dpm_system* s) {
    PairParams p = pair_params(s, s->e_c, s->n_c);
    p.fast = fast_ok(s) ? 1 : 0;
    p.fastdiv = fastdiv_ok(s) ? 1 : 0;
    if (s->method == DPM_NBR_ALL)  // pair forces still only need L/2 > largest contact
        p.fast = 0.5 * std::min(s->box[0], s->box[1]) > 2.0 * s->rmax * (1.0 + 1e-9) ? 1 : 0;
    return p;
}

NbrParams nbr_params(const dpm_system* s) {
    NbrParams nb{};
    nb.method = s->method;
    nb.ncd = (int)std::max<int64_t>(s->ncd, 1);
    nb.ncells = s->ncells;
    nb.fast = fast_ok(s) ? 1 : 0;
    nb.rc2 = s->rc * s->rc;
    nb.thr = s->thr_sq;
    nb.cthr = s->cell_thr_sq;
    nb.cs = s->cell_size;
    nb.Lx = s->box[0];
    nb.Ly = s->box[1];
    nb.invLx = 1.0 / s->box[0];
    nb.invLy = 1.0 / s->box[1];
    nb.halfLx = 0.5 * s->box[0];
    nb.halfLy = 0.5 * s->box[1];
    nb.vc2 = s->rc_v * s->rc_v;
    nb.vp2 = s->pcut_v * s->pcut_v;
    if (s->method == DPM_NBR_VERLET && s->vg_ncd >= 3) {  // the grid of the Verlet rows
        nb.ncd = (int)s->vg_ncd;
        nb.ncells = s->vg_ncd * s->vg_ncd;
        nb.cs = s->box[0] / (double)s->vg_ncd;
    }
    return nb;
}

int launch_chunk_sum(dpm_system* s, int what, double* out_dev) {
    hipLaunchKernelGGL(k_chunk_sum, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, what,
                       s->d.part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(WAVE), 0, s->stream, s->d.part, s->nchunk,
                       out_dev);
    return check_launch();
}

int sum_to_host(dpm_system* s, int what, double* out) {
    double* dv = &s->d.ctl->scal[0];
    if (launch_chunk_sum(s, what, dv)) return -1;
    return scopy(s, out, dv, sizeof(double), hipMemcpyDeviceToHost);
}

int max_to_host(dpm_system* s, int what, double* out) {
    double* dv = &s->d.ctl->scal[1];
    const double* src = nullptr;
    int64_t n = s->nblk;
    if (what == M_ND) src = s->d.pmax_nd;
    else if (what == M_CD) src = s->d.pmax_cd;
    else {
        double* tmp = s->d.part;
        hipLaunchKernelGGL(k_block_max, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, what, tmp);
        src = tmp;
    }
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLOCK), 0, s->stream, src, n, dv);
    if (check_launch()) return -1;
    return scopy(s, out, dv, sizeof(double), hipMemcpyDeviceToHost);
}

// =========================================================================
//                 timing (HIP events per launch or per chunk)
// =========================================================================
// Events come from a free list; a pair returns to it once its time is collected.
static hipEvent_t pool_event(dpm_system* s) {
    if (s->ev_free.empty()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        s->ev_pool.push_back(e);
        return e;
    }
    hipEvent_t e = s->ev_free.back();
    s->ev_free.pop_back();
    return e;
}

Timed::Timed(dpm_system* s_, int cls_) : s(s_), cls(cls_) {
    if (s->timing != 1) return;
    a = pool_event(s);
    b = pool_event(s);
    (void)hipEventRecord(a, s->stream);
}

Timed::~Timed() {
    if (s->timing != 1) return;
    (void)hipEventRecord(b, s->stream);
    s->ev_pending.push_back({cls, {a, b}});
}

static void account(dpm_system* s, const std::pair<int, std::pair<hipEvent_t, hipEvent_t>>& pe) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
        s->ktime_ms[pe.first] += ms;
        s->klaunch[pe.first] += 1;
    }
    s->ev_free.push_back(pe.second.first);
    s->ev_free.push_back(pe.second.second);
}

// Collects every pending pair (waits for the stream).
void collect_timing(dpm_system* s) {
    if (s->ev_pending.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto& pe : s->ev_pending) account(s, pe);
    s->ev_pending.clear();
}

// Collects the pairs whose end event has completed, without waiting (chunk timing).
static void collect_finished(dpm_system* s) {
    size_t k = 0;
    while (k < s->ev_pending.size() && hipEventQuery(s->ev_pending[k].second.second) == hipSuccess)
        account(s, s->ev_pending[k++]);
    s->ev_pending.erase(s->ev_pending.begin(), s->ev_pending.begin() + k);
}

// =========================================================================
//                 neighbour management (chain, rebuilds, capacity)
// =========================================================================
int ensure_ell(dpm_system* s, int64_t cap) {
    if (cap <= s->phys_cap) return 0;
    dfree(s, s->d.nbr);
    s->d.nbr = nullptr;
    if (dalloc(s, &s->d.nbr, s->N * cap)) return -1;
    s->phys_cap = cap;
    s->d.cap = (int)cap;
    s->gen++;
    return 0;
}

// The rebuild kernels (cell binning, scan, scatter, stable fix-up, reorder, list
// build) in stream order; with `gated` each one consults the chain's decision.
// The particle list build of a rebuild (after the cell chain for the cell method).
// Counting sort of the particles into the cells of `nb` (histogram, scan, scatter, stable
// fix-up), leaving cell_start and the cell lists in perm; cg is the cell kernels' gate.
static void enqueue_cell_sort(dpm_system* s, const NbrParams& nb, int cg) {
    const Dev& d = s->d;
    hipStream_t st = s->stream;
    long long nc = nb.ncells;
    long long nsb = (nc + SCAN_TILE - 1) / SCAN_TILE;
    const Ctl* g = cg ? d.ctl : nullptr;
    hipLaunchKernelGGL(k_fill_int, pgrid(nc), dim3(BLOCK), 0, st, d.cell_cnt, 0, (int64_t)nc, g,
                       cg);
    hipLaunchKernelGGL(k_cell_bin, pgrid(s->N), dim3(BLOCK), 0, st, d, nb, cg);
    hipLaunchKernelGGL(k_scan_a, dim3((unsigned)nsb), dim3(BLOCK), 0, st, d, nc, cg);
    hipLaunchKernelGGL(k_scan_b, dim3(1), dim3(1024), 0, st, d, nsb, cg);
    hipLaunchKernelGGL(k_scan_c, dim3((unsigned)nsb), dim3(BLOCK), 0, st, d, nc, cg);
    hipLaunchKernelGGL(k_cell_scatter, pgrid(s->N), dim3(BLOCK), 0, st, d, cg);
    hipLaunchKernelGGL(k_cell_sort, pgrid(nc), dim3(BLOCK), 0, st, d, nc, cg);
}

void enqueue_list_build(dpm_system* s, int gated) {
    NbrParams nb = nbr_params(s);
    if (s->method == DPM_NBR_CELL) {
        hipLaunchKernelGGL(k_build_cell, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, nb, gated);
    } else if (s->method == DPM_NBR_VERLET && s->vg_ncd >= 3) {
        enqueue_cell_sort(s, nb, gated ? 2 : 0);
        hipLaunchKernelGGL(k_build_verlet_grid, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, nb,
                           gated);
    } else {
        hipLaunchKernelGGL(k_build_verlet, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, nb,
                           gated);
    }
}

static void enqueue_rebuild(dpm_system* s, int gated) {
    const Dev& d = s->d;
    NbrParams nb = nbr_params(s);
    hipStream_t st = s->stream;
    if (s->method == DPM_NBR_CELL) {
        enqueue_cell_sort(s, nb, gated ? 1 : 0);
        hipLaunchKernelGGL(k_reorder_gather, pgrid(s->N), dim3(BLOCK), 0, st, d, s->sp, gated);
        if (s->kind == DPM_BUMPY) bumpy_enqueue_vertex_reorder(s, gated);
        hipLaunchKernelGGL(k_reorder_copyback, pgrid(s->N), dim3(BLOCK), 0, st, d, s->sp, gated);
    }
    enqueue_list_build(s, gated);
    if (s->kind == DPM_BUMPY) bumpy_enqueue_vertex_build(s, gated);
}

// The halted step's remaining work. resume: 0 none, 1 NVE, 2 FIRE, 3 rigid bumpy FIRE.
static void enqueue_resume(dpm_system* s, int resume, double dt) {
    const Dev& d = s->d;
    NbrParams nb = nbr_params(s);
    PairParams p = params_of(s);
    hipStream_t st = s->stream;
    if (resume == 1) {
        LAUNCH_HOT(k_nve_force, pgrid(s->N), dim3(BLOCK), 0, st, d, p, 0.5 * dt, 0, 0,
                           (int)G_RESUME);
        hipLaunchKernelGGL(k_nve_kick_drift, pgrid(s->N), dim3(BLOCK), 0, st, d, nb, 0.5 * dt, dt,
                           0, 1, -1, (int)G_RESUME);
    } else if (resume == 2) {
        LAUNCH_HOT(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, st, d, p, 0, 1,
                           (int)G_RESUME);
        hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, st, d, 0, 0, (int)G_RESUME);
    } else if (resume == 3) {  // rigid bumpy FIRE
        bumpy_enqueue_fire_forces(s, 1, G_RESUME);
        hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, st, d, 0, 0, (int)G_RESUME);
    }
}

// End of a chunk: decide, rebuild and finish a halted step, all on the device.
static void enqueue_chain(dpm_system* s, int resume, double dt) {
    const Dev& d = s->d;
    hipStream_t st = s->stream;
    if (s->method != DPM_NBR_CELL && s->method != DPM_NBR_VERLET) {
        hipLaunchKernelGGL(k_chain_end, dim3(1), dim3(WAVE), 0, st, d, 0);
        return;
    }
    Timed t(s, DPM_K_REBUILD);
    hipLaunchKernelGGL(k_chain_decide, dim3(1), dim3(BLOCK), 0, st, d, nbr_params(s), 0);
    if (resume == 1) hipLaunchKernelGGL(k_chain_copy_cur, pgrid(s->N), dim3(BLOCK), 0, st, d);
    enqueue_rebuild(s, 1);
    enqueue_resume(s, resume, dt);
    hipLaunchKernelGGL(k_chain_end, dim3(1), dim3(WAVE), 0, st, d, 0);
}

// Host-side fix after a capacity overflow: regrow the ELL to 2^ceil(log2 max) (the
// reference's rule, particle.cu:1122-1130), rebuild the list and (for a halted
// loop) finish the halted step.
static int fix_overflow(dpm_system* s, int resume, double dt) {
    for (int attempt = 0; attempt < 8; ++attempt) {
        Ctl c;
        if (scopy(s, &c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost)) return -1;
        if (!c.overflow) return 0;
        if (c.max_nnum > s->phys_cap && ensure_ell(s, pow2_at_least(c.max_nnum))) return -1;
        if (s->kind == DPM_BUMPY && c.max_vnum > s->vphys_cap &&
            bumpy_ensure_vcap(s, pow2_at_least(c.max_vnum)))
            return -1;
        const Dev& d = s->d;
        NbrParams nb = nbr_params(s);
        hipStream_t st = s->stream;
        hipLaunchKernelGGL(k_overflow_begin, dim3(1), dim3(1), 0, st, d);
        enqueue_list_build(s, 1);
        if (s->kind == DPM_BUMPY) bumpy_enqueue_vertex_build(s, 1);
        enqueue_resume(s, resume, dt);
        // halted loops finish like a chunk end; stand-alone rebuilds only check
        hipLaunchKernelGGL(k_chain_end, dim3(1), dim3(WAVE), 0, st, d, resume ? 0 : 1);
        if (check_launch()) return -1;
    }
    set_error("neighbour capacity did not converge");
    return -1;
}

// Stand-alone rebuild: mode 1 = Particle::checkForNeighborUpdate, 2 = forced.
// kind_out: 0 none, 1 list, 2 cells + list.
int device_rebuild(dpm_system* s, int mode, int* kind_out) {
    if (kind_out) *kind_out = 0;
    if (s->method == DPM_NBR_NONE) return 0;
    if (s->method == DPM_NBR_ALL && mode == 1) return 0;  // checkForAllToAllUpdate: no-op
    hipLaunchKernelGGL(k_chain_decide, dim3(1), dim3(BLOCK), 0, s->stream, s->d, nbr_params(s),
                       mode);
    enqueue_rebuild(s, 1);
    hipLaunchKernelGGL(k_chain_end, dim3(1), dim3(WAVE), 0, s->stream, s->d, mode);
    if (check_launch()) return -1;
    Ctl c;
    if (scopy(s, &c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost)) return -1;
    if (c.overflow && fix_overflow(s, 0, 0.0)) {
        s->nbr_ready = false;
        return -1;
    }
    if (kind_out) *kind_out = c.kind;
    return 0;
}

int check_and_rebuild(dpm_system* s, int* kind_out) { return device_rebuild(s, 1, kind_out); }

// =========================================================================
//                         chunked device-resident loops
// =========================================================================
typedef void (*ChunkFn)(dpm_system*, double, int64_t);

static int ensure_graph(dpm_system* s, std::map<int64_t, Graph>& cache, int64_t steps, double dt,
                        ChunkFn fill) {
    Graph& g = cache[steps];
    if (g.exec && g.gen == s->gen && g.dt == dt) return 0;
    if (g.exec) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
    }
    std::lock_guard<std::mutex> lk(capture_mutex());
    DPM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    fill(s, dt, steps);
    DPM_HIP(hipStreamEndCapture(s->stream, &g.graph));
    DPM_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    g.gen = s->gen;
    g.dt = dt;
    return 0;
}

// c NVE steps (positions 0..c-1) + the chain.  An odd c splits its last step into
// force + kick-drift so every chunk ends with the positions in P0.
static void nve_chunk(dpm_system* s, double dt, int64_t c) {
    const Dev& d = s->d;
    PairParams p = params_of(s);
    NbrParams nb = nbr_params(s);
    hipStream_t st = s->stream;
    int64_t fused = (c % 2) ? c - 1 : c;
    for (int64_t k = 0; k < fused; ++k) {
        Timed t(s, DPM_K_FORCE);
        LAUNCH_HOT(k_nve_fused, pgrid(s->N), dim3(BLOCK), 0, st, d, p, nb, 0.5 * dt, dt,
                           (int)(k & 1), (int)k, (int)(k == c - 1));
    }
    if (c % 2) {
        int k = (int)(c - 1);
        {
            Timed t(s, DPM_K_FORCE);
            LAUNCH_HOT(k_nve_force, pgrid(s->N), dim3(BLOCK), 0, st, d, p, 0.5 * dt, k,
                               (int)H_NVE_FUSED, (int)G_NORMAL);
        }
        Timed t(s, DPM_K_DRIFT);
        hipLaunchKernelGGL(k_nve_kick_drift, pgrid(s->N), dim3(BLOCK), 0, st, d, nb, 0.5 * dt, dt,
                           0, 1, k + 1, (int)G_NORMAL);
    }
    hipLaunchKernelGGL(k_chunk_end, dim3(1), dim3(WAVE), 0, st, d);
}

static void fire_chunk(dpm_system* s, double, int64_t c) {
    const Dev& d = s->d;
    PairParams p = params_of(s);
    NbrParams nb = nbr_params(s);
    hipStream_t st = s->stream;
    for (int64_t k = 0; k < c; ++k) {
        {
            Timed t(s, DPM_K_DRIFT);
            hipLaunchKernelGGL(k_fire_move, pgrid(s->N), dim3(BLOCK), 0, st, d, nb, (int)k);
        }
        {
            Timed t(s, DPM_K_FORCE);
            LAUNCH_HOT(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, st, d, p,
                               (int)k, 1, (int)G_NORMAL);
        }
        hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, st, d, 0, (int)k, (int)G_NORMAL);
    }
    enqueue_chain(s, 2, 0.0);
}

static void bumpy_fire_chunk(dpm_system* s, double, int64_t c) {
    for (int64_t k = 0; k < c; ++k) {
        bumpy_enqueue_fire_step(s, (int)k);
        hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, s->stream, s->d, 0, (int)k,
                           (int)G_NORMAL);
    }
    enqueue_chain(s, 3, 0.0);
}

static void recovery_fill(dpm_system* s, double dt, int64_t) { enqueue_chain(s, 1, dt); }

// The closing force + kick of an NVE run (the last step's second half); it records in
// the control block that it ran (`closed`), or halts when that step needs a rebuild.
static void nve_close(dpm_system* s, double dt) {
    {
        Timed t(s, DPM_K_FORCE);
        PairParams p = params_of(s);
        LAUNCH_HOT(k_nve_force, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, p, 0.5 * dt, 0,
                   (int)H_NVE_FORCE, (int)G_NORMAL);
    }
    hipLaunchKernelGGL(k_unit_end, dim3(1), dim3(WAVE), 0, s->stream, s->d);
}
typedef void (*TailFn)(dpm_system*, double);

// The chunk lengths a graph-launched loop uses: every length up to 32 (a run's last
// chunk is its remainder) and powers of two up to the configured chunk.  A longer
// remainder that is not a power of two is run as power-of-two chunks.
constexpr int64_t EXACT_LEN_MAX = 32;
static int64_t graph_len(int64_t c) {
    if (c <= EXACT_LEN_MAX) return c;
    int64_t p = 1;
    while (2 * p <= c) p *= 2;
    return p;
}

// Captures every chunk length a run can launch (and the recovery chain) ahead of the
// run, so that no capture falls inside a run whose length was not seen before.
static int ensure_all_graphs(dpm_system* s, std::map<int64_t, Graph>& cache, double dt,
                             ChunkFn chunk_fn) {
    // a marker entry (no graph) records the set last completed: a run with the same
    // arguments checks it alone (the run's first kernels are already on the device)
    Graph& mark = cache[INT64_MIN];
    if (mark.gen == s->gen && mark.dt == dt && mark.chunk == s->chunk) return 0;
    for (int64_t c = 1; c <= std::min<int64_t>(s->chunk, EXACT_LEN_MAX); ++c)
        if (ensure_graph(s, cache, c, dt, chunk_fn)) return -1;
    for (int64_t c = 2 * EXACT_LEN_MAX; c <= s->chunk; c *= 2)
        if (ensure_graph(s, cache, c, dt, chunk_fn)) return -1;
    if (s->chunk > EXACT_LEN_MAX && ensure_graph(s, cache, s->chunk, dt, chunk_fn)) return -1;
    if (ensure_graph(s, cache, -1, dt, recovery_fill)) return -1;
    mark.gen = s->gen;  // (std::map entries stay where they are as the map grows)
    mark.dt = dt;
    mark.chunk = s->chunk;
    return 0;
}

// Runs `total` steps in chunks with up to two launch units in flight; each unit ends
// with a copy of the control block to pinned memory, from which the host counts the
// completed steps.  FIRE chunks carry the gated rebuild chain themselves
// (host_recovery false).  NVE chunks, whose list rebuilds happen inside the step
// kernel and halt only for cell rebuilds, end with k_chunk_end; when one reports a
// halted step the host enqueues the rebuild chain (captured once as a graph) as a
// recovery unit behind the chunk already in flight, without waiting for the device.
// A capacity overflow is fixed on the host.  With `stop_on_done` the loop ends when
// the device reports completion (FIRE).  The chunk length adapts: after a chunk that
// halted at step p the next one is about 2(p + 1) steps long (a halted chunk's later
// steps are no-ops), and each chunk that runs through doubles it again, up to the
// configured length.  Chunk boundaries do not change results.
//
// `tail` (NVE: the closing force) is a unit of its own, enqueued as soon as every step
// has been launched, before the host waits for them: the loop ends once the device
// reports it applied (`closed`).  A tail launched behind a halted step does nothing and
// is enqueued again after the recovery; a tail that halts itself is finished by the
// recovery chain, which then closes the run.
static int run_chunks(dpm_system* s, int64_t total, double dt, int resume, ChunkFn chunk_fn,
                      std::map<int64_t, Graph>& cache, bool stop_on_done, bool host_recovery,
                      TailFn tail = nullptr) {
    enum { U_CHUNK = 0, U_RECOVERY = 1, U_TAIL = 2 };
    struct Unit {
        int64_t c;       // steps planned (0 for a recovery or tail unit)
        int seq;         // publication sequence number (mirror slot seq mod 4)
        int kind;
        bool covered;    // launched before a recovery that finishes its halted step
    };
    const bool graphs = s->use_graph && s->timing != 1;
    if (graphs && tail && ensure_all_graphs(s, cache, dt, chunk_fn)) return -1;
    std::deque<Unit> q;
    int64_t done = 0, planned = 0;
    int next_slot = 0;  // publication sequence of the next unit (k_ctl_reset starts it at 0)
    bool need_recovery = false;
    bool closed = tail == nullptr, tail_queued = false;
    int64_t cur = s->chunk;
    auto recovering = [&q] {
        for (auto& x : q)
            if (x.kind == U_RECOVERY) return true;
        return false;
    };
    while (done < total || need_recovery || !closed) {
        // the tail goes behind the last step, never behind a recovery that may close
        while (q.size() < 2 && (need_recovery || planned < total ||
                                (!closed && !tail_queued && !recovering()))) {
            Unit u{0, 0, U_CHUNK, false};
            hipEvent_t ua = nullptr, ub = nullptr;
            const bool timed = s->timing == 2 && (need_recovery || planned < total);
            if (timed) {
                ua = pool_event(s);
                ub = pool_event(s);
                DPM_HIP(hipEventRecord(ua, s->stream));
            }
            if (need_recovery) {
                if (graphs) {
                    if (ensure_graph(s, cache, -1, dt, recovery_fill)) return -1;
                    DPM_HIP(hipGraphLaunch(cache[-1].exec, s->stream));
                } else {
                    enqueue_chain(s, resume, dt);
                    if (check_launch()) return -1;
                }
                for (auto& x : q) x.covered = true;
                u.kind = U_RECOVERY;
                planned += 1;
                need_recovery = false;
            } else if (planned < total) {
                int64_t c = std::min<int64_t>(cur, total - planned);
                if (graphs && (tail || c == cur)) {
                    c = graph_len(c);
                    if (ensure_graph(s, cache, c, dt, chunk_fn)) return -1;
                    DPM_HIP(hipGraphLaunch(cache[c].exec, s->stream));
                } else {
                    chunk_fn(s, dt, c);
                    if (check_launch()) return -1;
                }
                u.c = c;
                planned += c;
            } else {
                tail(s, dt);
                if (check_launch()) return -1;
                u.kind = U_TAIL;
                tail_queued = true;
            }
            if (timed) {
                DPM_HIP(hipEventRecord(ub, s->stream));
                s->ev_pending.push_back(
                    {u.kind == U_RECOVERY ? DPM_K_REBUILD : DPM_K_FORCE, {ua, ub}});
            }
            // the unit's last kernel publishes the control block into mirror slot seq mod 4
            u.seq = next_slot++;
            DPM_HIP(hipEventRecord(s->poll_ev[u.seq & 3], s->stream));
            q.push_back(u);
        }
        if (q.empty()) break;
        Unit f = q.front();
        q.pop_front();
        DPM_HIP(hipEventSynchronize(s->poll_ev[f.seq & 3]));
        if (s->timing == 1) collect_timing(s);
        else if (s->timing == 2) collect_finished(s);
        Ctl hc;
        std::memcpy(&hc, (const void*)&s->hctl[f.seq & 3], sizeof(Ctl));
        if (hc.pub_seq != f.seq) {  // the copy carries the sequence number it was written for
            set_error("launch loop: status of unit " + std::to_string(f.seq) +
                      " not published (mirror holds " + std::to_string(hc.pub_seq) + ")");
            (void)hipStreamSynchronize(s->stream);
            return -1;
        }
        if (f.kind == U_TAIL) {
            tail_queued = false;
            closed = hc.closed != 0;
            // a tail that halted itself: the recovery applies the closing force
            if (!closed && hc.halt && !f.covered) need_recovery = true;
        } else {
            done += hc.last_progress;
            planned -= (f.kind == U_RECOVERY ? 1 : f.c) - hc.last_progress;
            if (tail && f.kind == U_RECOVERY) closed = hc.closed != 0;
        }
        if (hc.overflow) {
            // the halted step is unfinished and every later unit in flight was a no-op
            DPM_HIP(hipStreamSynchronize(s->stream));
            if (fix_overflow(s, resume, dt)) return -1;
            Ctl after;
            if (scopy(s, &after, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost)) return -1;
            next_slot = after.pub_seq;  // the fix published its own unit end
            done += after.last_progress;
            q.clear();
            planned = done;
            need_recovery = false;
            tail_queued = false;
            if (tail) closed = after.closed != 0;
            if (stop_on_done && after.done) break;
            continue;
        }
        if (host_recovery && hc.halt && f.kind == U_CHUNK && !f.covered) need_recovery = true;
        if (stop_on_done && hc.done) break;
        if (f.kind == U_CHUNK && f.c > 0) {
            if (hc.last_progress < f.c)
                cur = std::min<int64_t>(s->chunk, pow2_at_least(2 * (hc.last_progress + 1)));
            else
                cur = std::min<int64_t>(s->chunk, 2 * cur);
            cur = std::max<int64_t>(cur, 2);
        }
    }
    DPM_HIP(hipStreamSynchronize(s->stream));
    if (s->timing) collect_timing(s);
    return 0;
}

int nve_run(dpm_system* s, double dt, int64_t n) {
    if (s->kind == DPM_BUMPY) return bumpy_nve_run(s, dt, n);
    if (!s->nbr_ready) {
        set_error("nve_run: neighbour list not initialised");
        return -1;
    }
    if (n <= 0) return 0;
    const Dev& d = s->d;
    hipStream_t st = s->stream;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, st, d, 0);
    {  // opening kick + drift of the first step; its request is checked at position 0
        Timed t(s, DPM_K_DRIFT);
        hipLaunchKernelGGL(k_nve_kick_drift, pgrid(s->N), dim3(BLOCK), 0, st, d, nbr_params(s),
                           0.5 * dt, dt, 0, 0, -1, (int)G_NORMAL);
    }
    if (check_launch()) return -1;
    // n - 1 fused steps, then the closing force + kick of the last step
    return run_chunks(s, n - 1, dt, 1, nve_chunk, s->nve_graphs, false, true, nve_close);
}

int minimize_fire(dpm_system* s, const dpm_fire_params* prm, dpm_fire_result* out) {
    const bool rigid = s->kind == DPM_BUMPY;
    if (!s->nbr_ready) {
        set_error("minimize_fire: neighbour list not initialised");
        return -1;
    }
    // minimizeFire prologue (src/routines/minimization.cu:76-94)
    FireCtl f{};
    f.alpha = prm->alpha_init;
    f.alpha_init = prm->alpha_init;
    f.dt = prm->dt;
    f.dt_max = 10.0 * prm->dt;
    f.dt_min = 0.001 * prm->dt;
    f.avg_pe_target = prm->avg_pe_target;
    f.avg_pe_diff_target = prm->avg_pe_diff_target;
    f.fmax_target_sq = prm->fmax_target > 0 ? prm->fmax_target * prm->fmax_target : 0.0;
    f.dof = (double)s->N;
    f.max_steps = prm->max_steps;
    const Dev& d = s->d;
    hipStream_t st = s->stream;
    DPM_HIP(hipMemcpyAsync(d.fire, &f, sizeof(FireCtl), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, st, d, 0);
    hipLaunchKernelGGL(k_zero_velocities, pgrid(s->N), dim3(BLOCK), 0, st, d);
    if (check_and_rebuild(s, nullptr)) return -1;
    PairParams p = params_of(s);
    if (rigid) {
        bumpy_enqueue_fire_forces(s, 0, G_ALWAYS);
    } else {
        LAUNCH_HOT(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, st, d, p, 0, 0,
                   (int)G_ALWAYS);
    }
    hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, st, d, 1, 0, (int)G_ALWAYS);
    if (check_launch()) return -1;
    if (prm->max_steps > 0 &&
        run_chunks(s, prm->max_steps, 0.0, rigid ? 3 : 2, rigid ? bumpy_fire_chunk : fire_chunk,
                   s->fire_graphs, true, false))
        return -1;
    hipLaunchKernelGGL(k_disp_partials, pgrid(s->N), dim3(BLOCK), 0, st, d, nbr_params(s));
    hipLaunchKernelGGL(k_zero_velocities, pgrid(s->N), dim3(BLOCK), 0, st, d);
    if (scopy(s, &f, d.fire, sizeof(FireCtl), hipMemcpyDeviceToHost)) return -1;
    if (out) {
        out->steps = std::min<int64_t>(f.step, prm->max_steps);
        out->stopped = f.stopped;
        out->converged = f.converged;
        out->avg_pe = f.avg_pe;
        out->dt = f.dt;
        out->alpha = f.alpha;
        double f2 = 0.0;
        if (dpm_reduce(s, DPM_RED_FMAX_SQ, &f2)) return -1;
        out->fmax_sq = f2;
    }
    return check_launch();
}

void reset_ctl(dpm_system* s, int config) {
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, s->stream, s->d, config);
}

// Cell arrays for at least nc cells (the cell method's grid, or the Verlet rows' grid).
int ensure_cells(dpm_system* s, int64_t nc) {
    if (nc <= s->ncells_alloc) return 0;
    DPM_HIP(hipStreamSynchronize(s->stream));  // a queued rebuild may still use the old ones
    dfree(s, s->d.cell_start);
    dfree(s, s->d.cell_cnt);
    dfree(s, s->d.cell_cur);
    dfree(s, s->d.scan_bsum);
    s->d.cell_start = s->d.cell_cnt = s->d.cell_cur = s->d.scan_bsum = nullptr;
    int64_t nsb = (nc + SCAN_TILE - 1) / SCAN_TILE;
    if (dalloc(s, &s->d.cell_start, nc + 1) || dalloc(s, &s->d.cell_cnt, nc) ||
        dalloc(s, &s->d.cell_cur, nc) || dalloc(s, &s->d.scan_bsum, nsb + 1))
        return -1;
    s->ncells_alloc = nc;
    s->gen++;
    return 0;
}

// The grid the Verlet method builds its rows from (k_build_verlet_grid): cells at least
// rc (1 + 1e-9) wide, at most about 2 per particle, three or more a side, in a square box;
// systems under VGRID_MIN_N particles (where one all-pairs pass costs less than the grid's
// launches) and other boxes use the all-pairs scan.  Either way the rows are identical.
constexpr int64_t VGRID_MIN_N = 1024;
int update_verlet_grid(dpm_system* s) {
    int64_t ncd = 0;
    if (s->method == DPM_NBR_VERLET && s->N >= VGRID_MIN_N && s->box[0] == s->box[1] &&
        s->rc > 0) {
        double c = std::floor(s->box[0] / (s->rc * (1.0 + 1e-9)));
        c = std::min(c, std::floor(std::sqrt(2.0 * (double)s->N)));
        if (c >= 3) ncd = (int64_t)c;
    }
    if (ncd == s->vg_ncd) return 0;
    if (ncd && ensure_cells(s, ncd * ncd)) return -1;
    s->vg_ncd = ncd;
    s->gen++;
    return 0;
}

}  // namespace dpm

// =========================================================================
//                                 C ABI
// =========================================================================
using namespace dpm;

namespace {

int fetch_sidx(dpm_system* s, std::vector<int>& sidx) {
    sidx.resize(s->N);
    return scopy(s, sidx.data(), s->d.sidx, sizeof(int) * s->N, hipMemcpyDeviceToHost);
}

int require_nbr(dpm_system* s, const char* who) {
    if (!s->nbr_ready) {
        set_error(std::string(who) + ": neighbour list not initialised");
        return -1;
    }
    return 0;
}

int ensure_adam(dpm_system* s) {
    if (s->d.m1) return 0;
    if (dalloc(s, &s->d.m1, s->N) || dalloc(s, &s->d.m2, s->N) || dalloc(s, &s->sp.m1, s->N) ||
        dalloc(s, &s->sp.m2, s->N))
        return -1;
    // RigidBumpy::initAdamVariables adds the angle moments (rigid_bumpy.cu:1461-1465)
    if (s->kind == DPM_BUMPY && (dalloc(s, &s->d.ma, s->N) || dalloc(s, &s->sp.ma, s->N)))
        return -1;
    s->gen++;
    return 0;
}

int ensure_stress(dpm_system* s) {
    if (s->d.sxx) return 0;
    if (dalloc(s, &s->d.sxx, s->N) || dalloc(s, &s->d.sxy, s->N) || dalloc(s, &s->d.syx, s->N) ||
        dalloc(s, &s->d.syy, s->N) || dalloc(s, &s->d.contacts, s->N) ||
        dalloc(s, &s->d.vcontacts, s->N))
        return -1;
    s->gen++;
    return 0;
}

int64_t ndof_of(dpm_system* s) {
    if (s->kind == DPM_BUMPY) return s->rotation ? 3 * s->N : 2 * s->N;
    return 2 * s->N;
}

// Per-particle fields: 'P' = component of P0 (radius and mass also mirrored in P1),
// 'V'/'F' = component of vel/frc, 'D' = a plain double array.
bool field_slot(dpm_system* s, int f, char* kind, int* comp, double** plain) {
    *plain = nullptr;
    *comp = 0;
    switch (f) {
        case DPM_F_X: *kind = 'P'; *comp = 0; return true;
        case DPM_F_Y: *kind = 'P'; *comp = 1; return true;
        case DPM_F_RAD: *kind = 'P'; *comp = 2; return true;
        case DPM_F_MASS: *kind = 'P'; *comp = 3; return true;
        case DPM_F_VX: *kind = 'V'; *comp = 0; return true;
        case DPM_F_VY: *kind = 'V'; *comp = 1; return true;
        case DPM_F_FX: *kind = 'F'; *comp = 0; return true;
        case DPM_F_FY: *kind = 'F'; *comp = 1; return true;
        case DPM_F_PE: *kind = 'D'; *plain = s->d.pe; return true;
        case DPM_F_KE: *kind = 'D'; *plain = s->d.ke; return true;
        case DPM_F_STRESS_XX: *kind = 'D'; *plain = s->d.sxx; return *plain != nullptr;
        case DPM_F_STRESS_XY: *kind = 'D'; *plain = s->d.sxy; return *plain != nullptr;
        case DPM_F_STRESS_YX: *kind = 'D'; *plain = s->d.syx; return *plain != nullptr;
        case DPM_F_STRESS_YY: *kind = 'D'; *plain = s->d.syy; return *plain != nullptr;
    }
    return false;
}

}  // namespace

namespace dpm {
// Largest radius in P0 (host copy), used to validate the fast minimum image.
int refresh_rmax(dpm_system* s) {
    std::vector<double4> h(s->N);
    if (scopy(s, h.data(), s->d.P0, sizeof(double4) * s->N, hipMemcpyDeviceToHost)) return -1;
    double m = 0.0, mn = HUGE_VAL;
    for (const auto& q : h) {
        m = std::max(m, q.z);
        mn = std::min(mn, q.z);
    }
    s->rmax = m;
    s->rmin = mn;
    s->gen++;
    return 0;
}
}  // namespace dpm

extern "C" {

const char* dpm_last_error(void) { return dpm::g_err.c_str(); }

int dpm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

dpm_system* dpm_create(int kind, int64_t n, int64_t nv) {
    if (n <= 0 || n >= (int64_t)1 << 30 || (kind != DPM_DISK && kind != DPM_BUMPY) ||
        (kind == DPM_BUMPY && (nv <= 0 || nv >= (int64_t)1 << 30))) {
        set_error("dpm_create: bad arguments");
        return nullptr;
    }
    if (dpm_device_count() <= 0) {
        set_error("dpm_create: no HIP device");
        return nullptr;
    }
    auto* s = new dpm_system();
    s->kind = kind;
    s->N = n;
    s->Nv = kind == DPM_BUMPY ? nv : 0;
    s->nblk = (n + BLOCK - 1) / BLOCK;
    s->nchunk = (n + CHUNK - 1) / CHUNK;
    (void)hipGetDevice(&s->device);
    bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
    Dev& d = s->d;
    d.N = n;
    d.Nv = s->Nv;
    ok = ok && !dalloc(s, &d.P0, n) && !dalloc(s, &d.P1, n) && !dalloc(s, &d.vel, n) &&
         !dalloc(s, &d.frc, n) && !dalloc(s, &d.lref, n) && !dalloc(s, &d.cref, n) &&
         !dalloc(s, &d.pe, n) && !dalloc(s, &d.ke, n) && !dalloc(s, &d.sidx, n) &&
         !dalloc(s, &d.nnum, n) && !dalloc(s, &d.cell_of, n) && !dalloc(s, &d.cell_tmp, n) &&
         !dalloc(s, &d.perm, n) && !dalloc(s, &d.pmax_nd, s->nblk) &&
         !dalloc(s, &d.pmax_cd, s->nblk) && !dalloc(s, &d.part, 4 * s->nchunk + s->nblk + 64) &&
         !dalloc(s, &d.ctl, 1) && !dalloc(s, &d.fire, 1) && !dalloc(s, &s->sp.P, n) &&
         !dalloc(s, &s->sp.vel, n) && !dalloc(s, &s->sp.frc, n) && !dalloc(s, &s->sp.sidx, n);
    if (ok) {
        std::vector<int> id(n);
        for (int64_t k = 0; k < n; ++k) id[k] = (int)k;
        ok = scopy(s, d.sidx, id.data(), sizeof(int) * n, hipMemcpyHostToDevice) == 0;
        Ctl c{};
        c.epoch = 1;
        ok = ok && scopy(s, d.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice) == 0;
    }
    // fine-grained (coherent) pinned memory: kernels write the status into it directly
    ok = ok && hipHostMalloc((void**)&s->hctl, 4 * sizeof(Ctl),
                             hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess;
    ok = ok && hipHostGetDevicePointer((void**)&s->d.hpub, s->hctl, 0) == hipSuccess;
    for (int k = 0; ok && k < 4; ++k)
        ok = hipEventCreateWithFlags(&s->poll_ev[k], hipEventDisableTiming) == hipSuccess;
    if (ok && kind == DPM_BUMPY) ok = bumpy_alloc(s) == 0;
    if (!ok) {
        std::string e = dpm::g_err;
        dpm_destroy(s);
        set_error("dpm_create: allocation failed: " + e);
        return nullptr;
    }
    return s;
}

void dpm_destroy(dpm_system* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto* cache : {&s->nve_graphs, &s->fire_graphs})
        for (auto& kv : *cache)
            if (kv.second.exec) {
                (void)hipGraphExecDestroy(kv.second.exec);
                (void)hipGraphDestroy(kv.second.graph);
            }
    {
        std::lock_guard<std::mutex> g(capture_mutex());
        for (void* p : s->allocs) (void)hipFree(p);
        for (void* p : s->snap_allocs) (void)hipFree(p);
    }
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    for (auto e : s->poll_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->hctl) (void)hipHostFree(s->hctl);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int dpm_set_box(dpm_system* s, double lx, double ly) {
    if (!(lx > 0) || !(ly > 0)) {
        set_error("dpm_set_box: box sides must be positive");
        return -1;
    }
    s->box[0] = lx;
    s->box[1] = ly;
    if (s->method == DPM_NBR_CELL) s->cell_size = lx / s->ncd;  // Particle::updateCellSize
    s->gen++;
    return update_verlet_grid(s);
}

int dpm_get_box(dpm_system* s, double* lx, double* ly) {
    *lx = s->box[0];
    *ly = s->box[1];
    return 0;
}

int dpm_set_energy(dpm_system* s, double e_c, double n_c) {
    s->e_c = e_c;
    s->n_c = n_c;
    s->gen++;
    return 0;
}

static int upload_field(dpm_system* s, int f, const void* data, int64_t count);

int dpm_upload(dpm_system* s, int f, const void* data, int64_t count) {
    if (upload_field(s, f, data, count)) return -1;
    // rigid bumpy: whether the vertices still sit on their rings (the vertex list build's
    // scan gate assumes it)
    if (s->kind == DPM_BUMPY && s->nv_max > 0 &&
        (f == DPM_F_X || f == DPM_F_Y || f == DPM_F_RAD || f == DPM_F_VERTEX_X ||
         f == DPM_F_VERTEX_Y))
        return bumpy_check_ring(s);
    return 0;
}

static int upload_field(dpm_system* s, int f, const void* data, int64_t count) {
    std::vector<int> sidx;
    if (fetch_sidx(s, sidx)) return -1;
    if (s->kind == DPM_BUMPY && ((f >= DPM_F_ANGLE && f <= DPM_F_MOI) || f >= DPM_F_VERTEX_X))
        return bumpy_upload(s, f, data, count, sidx);
    if (count != s->N) {
        set_error("dpm_upload: count mismatch");
        return -1;
    }
    const double* src = (const double*)data;
    char kind;
    int comp;
    double* plain;
    if (!field_slot(s, f, &kind, &comp, &plain) || f == DPM_F_KE) {
        set_error("dpm_upload: field not uploadable");
        return -1;
    }
    if (kind == 'D') {
        std::vector<double> h(s->N);
        for (int64_t k = 0; k < s->N; ++k) h[k] = src[sidx[k]];
        return scopy(s, plain, h.data(), sizeof(double) * s->N, hipMemcpyHostToDevice);
    }
    if (kind == 'P') {
        std::vector<double4> h(s->N);
        if (scopy(s, h.data(), s->d.P0, sizeof(double4) * s->N, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->N; ++k) (&h[k].x)[comp] = src[sidx[k]];
        if (f == DPM_F_RAD) {
            double m = 0.0, mn = HUGE_VAL;
            for (int64_t k = 0; k < s->N; ++k) {
                m = std::max(m, h[k].z);
                mn = std::min(mn, h[k].z);
            }
            s->rmax = m;
            s->rmin = mn;
            s->gen++;
        }
        if (scopy(s, s->d.P0, h.data(), sizeof(double4) * s->N, hipMemcpyHostToDevice)) return -1;
        return scopy(s, s->d.P1, h.data(), sizeof(double4) * s->N, hipMemcpyHostToDevice);
    }
    double2* p2 = kind == 'V' ? s->d.vel : s->d.frc;
    std::vector<double2> h(s->N);
    if (scopy(s, h.data(), p2, sizeof(double2) * s->N, hipMemcpyDeviceToHost)) return -1;
    for (int64_t k = 0; k < s->N; ++k) (comp ? h[k].y : h[k].x) = src[sidx[k]];
    return scopy(s, p2, h.data(), sizeof(double2) * s->N, hipMemcpyHostToDevice);
}

int dpm_download(dpm_system* s, int f, void* data, int64_t count) {
    std::vector<int> sidx;
    if (fetch_sidx(s, sidx)) return -1;
    if (s->kind == DPM_BUMPY && ((f >= DPM_F_ANGLE && f <= DPM_F_MOI) || f >= DPM_F_VERTEX_X))
        return bumpy_download(s, f, data, count, sidx);
    if (count != s->N) {
        set_error("dpm_download: count mismatch");
        return -1;
    }
    if (f == DPM_F_STATIC_INDEX) {
        int64_t* dst = (int64_t*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[k] = sidx[k];
        return 0;
    }
    if (f == DPM_F_NUM_NEIGHBORS || f == DPM_F_CONTACTS) {
        int* src = f == DPM_F_NUM_NEIGHBORS ? s->d.nnum : s->d.contacts;
        if (!src) {
            set_error("dpm_download: field not computed");
            return -1;
        }
        std::vector<int> h(s->N);
        if (scopy(s, h.data(), src, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
        int64_t* dst = (int64_t*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    char kind;
    int comp;
    double* plain;
    if (!field_slot(s, f, &kind, &comp, &plain)) {
        set_error("dpm_download: unknown or unallocated field");
        return -1;
    }
    double* dst = (double*)data;
    if (kind == 'D') {
        std::vector<double> h(s->N);
        if (scopy(s, h.data(), plain, sizeof(double) * s->N, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    if (kind == 'P') {
        std::vector<double4> h(s->N);
        if (scopy(s, h.data(), s->d.P0, sizeof(double4) * s->N, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = (&h[k].x)[comp];
        return 0;
    }
    double2* p2 = kind == 'V' ? s->d.vel : s->d.frc;
    std::vector<double2> h(s->N);
    if (scopy(s, h.data(), p2, sizeof(double2) * s->N, hipMemcpyDeviceToHost)) return -1;
    for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = comp ? h[k].y : h[k].x;
    return 0;
}

int dpm_neighbor_config(dpm_system* s, int method, double cutoff, double thr_sq, int64_t ncd,
                        double cell_thr_sq) {
    if (method != DPM_NBR_ALL && method != DPM_NBR_VERLET && method != DPM_NBR_CELL) {
        set_error("dpm_neighbor_config: invalid method");
        return -1;
    }
    s->method = method;
    s->rc = cutoff;
    s->thr_sq = thr_sq;
    s->nbr_ready = false;
    // Particle::initParticleIndex (particle.cu:1202-1207): the current order becomes static
    {
        std::vector<int> id(s->N);
        for (int64_t k = 0; k < s->N; ++k) id[k] = (int)k;
        if (scopy(s, s->d.sidx, id.data(), sizeof(int) * s->N, hipMemcpyHostToDevice)) return -1;
    }
    if (method == DPM_NBR_CELL) {
        if (ncd < 3) {
            set_error("dpm_neighbor_config: need at least 3 cells per side");
            return -1;
        }
        s->ncd = ncd;
        s->ncells = ncd * ncd;
        s->cell_thr_sq = cell_thr_sq;
        s->cell_size = s->box[0] / ncd;
        if (ensure_cells(s, s->ncells)) return -1;
    }
    s->vg_ncd = 0;
    if (update_verlet_grid(s)) return -1;
    int64_t cap = 16;
    if (method == DPM_NBR_ALL) {
        // Particle::initAllToAllList (particle.cu:1186-1200)
        cap = std::max<int64_t>(1, s->N - 1);
        double diag = std::sqrt(s->box[0] * s->box[0] + s->box[1] * s->box[1]);
        s->rc = 10.0 * diag;
        s->thr_sq = s->rc * s->rc;
    }
    if (ensure_ell(s, cap)) return -1;
    reset_ctl(s, 1);
    s->gen++;
    return check_launch();
}

int dpm_neighbor_init(dpm_system* s) {
    if (s->method == DPM_NBR_NONE) {
        set_error("dpm_neighbor_init: no method configured");
        return -1;
    }
    // initVerletListVariables (particle.cu:1171-1178): references and displacements zeroed
    hipStream_t st = s->stream;
    DPM_HIP(hipMemsetAsync(s->d.lref, 0, sizeof(double2) * s->N, st));
    DPM_HIP(hipMemsetAsync(s->d.cref, 0, sizeof(double2) * s->N, st));
    DPM_HIP(hipMemsetAsync(s->d.pmax_nd, 0, sizeof(double) * s->nblk, st));
    DPM_HIP(hipMemsetAsync(s->d.pmax_cd, 0, sizeof(double) * s->nblk, st));
    reset_ctl(s, 0);
    if (device_rebuild(s, 2, nullptr)) return -1;
    s->nbr_ready = true;
    DPM_HIP(hipStreamSynchronize(st));
    return 0;
}

int dpm_neighbor_check(dpm_system* s, int* rebuilt) {
    if (require_nbr(s, "dpm_neighbor_check")) return -1;
    return check_and_rebuild(s, rebuilt);
}

int dpm_neighbor_pairs(dpm_system* s, int64_t* oi, int64_t* oj, int64_t cap, int64_t* count) {
    if (require_nbr(s, "dpm_neighbor_pairs")) return -1;
    std::vector<int> sidx, nn(s->N), nb((size_t)(s->N * s->phys_cap));
    if (fetch_sidx(s, sidx)) return -1;
    if (scopy(s, nn.data(), s->d.nnum, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
    if (scopy(s, nb.data(), s->d.nbr, sizeof(int) * s->N * s->phys_cap, hipMemcpyDeviceToHost))
        return -1;
    int64_t c = 0;
    for (int64_t i = 0; i < s->N; ++i)
        for (int64_t k = 0; k < nn[i] && k < s->phys_cap; ++k) {
            int j = nb[k * s->N + i];
            if (c < cap) {
                oi[c] = sidx[i];
                oj[c] = sidx[j];
            }
            c++;
        }
    *count = c;
    if (c > cap) {
        set_error("dpm_neighbor_pairs: capacity too small");
        return -1;
    }
    return 0;
}

int dpm_neighbor_stats(dpm_system* s, int64_t* ma, int64_t* tot, int64_t* nr, int64_t* ncr) {
    Ctl c;
    if (scopy(s, &c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost)) return -1;
    if (ma) {
        // Particle::updateVerletList's capacity rule (particle.cu:1122-1130): starts at
        // 4 and becomes 2^ceil(log2 max) whenever a build exceeds it
        if (s->method == DPM_NBR_ALL) *ma = std::max<int64_t>(1, s->N - 1);
        else *ma = c.max_nnum_all > 4 ? pow2_at_least(c.max_nnum_all) : 4;
    }
    if (tot) {
        std::vector<int> nn(s->N);
        if (scopy(s, nn.data(), s->d.nnum, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
        int64_t t = 0;
        for (int v : nn) t += v;
        *tot = t;
    }
    if (nr) *nr = c.n_rebuilds;
    if (ncr) *ncr = c.n_cell_rebuilds;
    return 0;
}

int dpm_zero_forces(dpm_system* s) {
    hipLaunchKernelGGL(k_zero_forces, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d);
    if (s->kind == DPM_BUMPY) return bumpy_forces(s, -1);
    return check_launch();
}

int dpm_compute_forces(dpm_system* s) {
    if (require_nbr(s, "dpm_compute_forces")) return -1;
    if (s->kind == DPM_BUMPY) return bumpy_forces(s, 1);
    PairParams p = params_of(s);
    LAUNCH_HOT(k_disk_forces_acc, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, p);
    return check_launch();
}

int dpm_kick(dpm_system* s, double dt) {
    if (s->kind == DPM_BUMPY) return bumpy_kick(s, dt);
    hipLaunchKernelGGL(k_kick, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, dt);
    return check_launch();
}

int dpm_drift(dpm_system* s, double dt) {
    if (s->kind == DPM_BUMPY) return bumpy_drift(s, dt);
    hipLaunchKernelGGL(k_drift, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, nbr_params(s), dt);
    return check_launch();
}

int dpm_mix(dpm_system* s, double alpha) {
    if (s->kind == DPM_BUMPY) return bumpy_mix(s, alpha);
    hipLaunchKernelGGL(k_mix, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, alpha);
    return check_launch();
}

int dpm_zero_velocities(dpm_system* s) {
    hipLaunchKernelGGL(k_zero_velocities, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d);
    return check_launch();
}

int dpm_nve_run(dpm_system* s, double dt, int64_t n) {
    if (n < 0) {
        set_error("dpm_nve_run: negative step count");
        return -1;
    }
    return nve_run(s, dt, n);
}

int dpm_reduce(dpm_system* s, int what, double* out) {
    switch (what) {
        case DPM_RED_PE: return sum_to_host(s, S_PE, out);
        case DPM_RED_KE:
        case DPM_RED_TEMPERATURE: {
            hipLaunchKernelGGL(k_kinetic, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d,
                               (int)(s->kind == DPM_BUMPY && s->rotation));
            double ke;
            if (sum_to_host(s, S_KE, &ke)) return -1;
            *out = what == DPM_RED_KE ? ke : ke * 2.0 / ndof_of(s);
            return 0;
        }
        case DPM_RED_POWER: {
            double a, b;
            if (sum_to_host(s, S_FXVX, &a) || sum_to_host(s, S_FYVY, &b)) return -1;
            double p = a + b;
            if (s->kind == DPM_BUMPY) {
                double c;
                if (sum_to_host(s, S_TQW, &c)) return -1;
                p = p + c;
            }
            *out = p;
            return 0;
        }
        case DPM_RED_FMAX_SQ: {
            double a;
            if (max_to_host(s, M_F2, &a)) return -1;
            if (s->kind == DPM_BUMPY) {
                double b;
                if (max_to_host(s, M_T2, &b)) return -1;
                a = std::max(a, b);
            }
            *out = a;
            return 0;
        }
        case DPM_RED_FORCE_BALANCE: {
            double a, b;
            if (sum_to_host(s, S_FX, &a) || sum_to_host(s, S_FY, &b)) return -1;
            *out = std::sqrt(a * a + b * b);
            return 0;
        }
        case DPM_RED_PRESSURE: {
            if (dpm_compute_stress(s)) return -1;
            double a, b;
            if (sum_to_host(s, S_SXX, &a) || sum_to_host(s, S_SYY, &b)) return -1;
            *out = (a + b) / 2;
            return 0;
        }
        case DPM_RED_STRESS_XX:
        case DPM_RED_STRESS_XY:
        case DPM_RED_STRESS_YX:
        case DPM_RED_STRESS_YY: {
            if (dpm_compute_stress(s)) return -1;
            int w = what == DPM_RED_STRESS_XX   ? S_SXX
                    : what == DPM_RED_STRESS_XY ? S_SXY
                    : what == DPM_RED_STRESS_YX ? S_SYX
                                                : S_SYY;
            return sum_to_host(s, w, out);
        }
        case DPM_RED_CONTACTS:
        case DPM_RED_VERTEX_CONTACTS: {
            if (dpm_count_contacts(s)) return -1;
            std::vector<int> h(s->N);
            int* src = what == DPM_RED_CONTACTS ? s->d.contacts : s->d.vcontacts;
            if (scopy(s, h.data(), src, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
            int64_t t = 0;
            for (int v : h) t += v;
            *out = (double)t;
            return 0;
        }
        case DPM_RED_PACKING_FRACTION: {
            double a;
            if (sum_to_host(s, S_R2, &a)) return -1;
            const double PI = 3.141592653589793238462643383279502884197;
            *out = a * PI / (s->box[0] * s->box[1]);
            return 0;
        }
        case DPM_RED_MAX_NEIGH_DISP_SQ: return max_to_host(s, M_ND, out);
        case DPM_RED_MAX_CELL_DISP_SQ: return max_to_host(s, M_CD, out);
    }
    set_error("dpm_reduce: unknown quantity");
    return -1;
}

int dpm_minimize_fire(dpm_system* s, const dpm_fire_params* p, dpm_fire_result* out) {
    return minimize_fire(s, p, out);
}

int dpm_fire_state_init(dpm_fire_state* st, double a, double dt) {
    st->alpha = a;
    st->alpha_init = a;
    st->dt = dt;
    st->dt_max = 10.0 * dt;
    st->dt_min = 0.001 * dt;
    st->n_good = 0;
    st->n_bad = 0;
    st->stopped = 0;
    st->power = 0.0;
    return 0;
}

// Fire::step (src/integrator/fire.cu:13-46) with the state on the host.
int dpm_fire_step(dpm_system* s, dpm_fire_state* st) {
    if (require_nbr(s, "dpm_fire_step")) return -1;
    double power;
    if (dpm_reduce(s, DPM_RED_POWER, &power)) return -1;
    st->power = power;
    if (power > 0) {
        st->n_good++;
        st->n_bad = 0;
        if (st->n_good > 5) {
            st->dt = std::min(st->dt * 1.1, st->dt_max);
            st->alpha *= 0.99;
        }
    } else {
        st->n_good = 0;
        st->n_bad++;
        if (st->n_bad > 10) {
            st->stopped = 1;
            return 0;
        }
        st->dt = std::max(st->dt * 0.5, st->dt_min);
        st->alpha = st->alpha_init;
        if (dpm_drift(s, -st->dt / 2.0) || dpm_zero_velocities(s)) return -1;
    }
    if (dpm_kick(s, st->dt / 2.0) || dpm_mix(s, st->alpha) || dpm_drift(s, st->dt) ||
        dpm_zero_forces(s) || check_and_rebuild(s, nullptr) || dpm_compute_forces(s) ||
        dpm_kick(s, st->dt / 2.0))
        return -1;
    return 0;
}

int dpm_adam_step(dpm_system* s, int64_t step, double alpha, double b1, double b2, double eps) {
    if (require_nbr(s, "dpm_adam_step")) return -1;
    if (ensure_adam(s)) return -1;
    // Adam::minimize (src/integrator/adam.cu:11-18)
    if (dpm_zero_forces(s) || check_and_rebuild(s, nullptr) || dpm_compute_forces(s)) return -1;
    double omb1 = 1 - std::pow(b1, step + 1);
    double omb2 = 1 - std::pow(b2, step + 1);
    if (s->kind == DPM_BUMPY) return bumpy_adam_update(s, alpha, b1, b2, omb1, omb2, eps);
    hipLaunchKernelGGL(k_adam_update, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d,
                       nbr_params(s), alpha, b1, b2, omb1, omb2, eps);
    return check_launch();
}

int dpm_adam_reset(dpm_system* s) {
    if (ensure_adam(s)) return -1;
    // Particle::initAdamVariables (particle.cu:319-322)
    DPM_HIP(hipMemsetAsync(s->d.m1, 0, sizeof(double2) * s->N, s->stream));
    DPM_HIP(hipMemsetAsync(s->d.m2, 0, sizeof(double2) * s->N, s->stream));
    if (s->d.ma) DPM_HIP(hipMemsetAsync(s->d.ma, 0, sizeof(double2) * s->N, s->stream));
    return 0;
}

int dpm_minimize_adam(dpm_system* s, double alpha, double b1, double b2, double eps, double pe_t,
                      double diff_t, int64_t max_steps, int64_t* steps_out) {
    if (require_nbr(s, "dpm_minimize_adam")) return -1;
    if (ensure_adam(s)) return -1;
    // minimizeAdam (src/routines/minimization.cu:3-33)
    if (dpm_adam_reset(s)) return -1;
    int64_t step = 0;
    double dof = (double)ndof_of(s), last = 0.0;
    while (step < max_steps) {
        if (dpm_adam_step(s, step, alpha, b1, b2, eps)) return -1;
        double pe;
        if (sum_to_host(s, S_PE, &pe)) return -1;
        double avg = pe / dof;
        double diff = std::abs(avg / last - 1);
        last = avg;
        if (diff < diff_t || avg < pe_t) break;
        step++;
    }
    if (steps_out) *steps_out = step;
    return 0;
}

int dpm_compute_stress(dpm_system* s) {
    if (require_nbr(s, "dpm_compute_stress")) return -1;
    if (ensure_stress(s)) return -1;
    if (s->kind == DPM_BUMPY) return bumpy_stress(s);
    hipLaunchKernelGGL(k_disk_stress, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, params_of(s),
                       s->box[0] * s->box[1]);
    return check_launch();
}

int dpm_count_contacts(dpm_system* s) {
    if (require_nbr(s, "dpm_count_contacts")) return -1;
    if (ensure_stress(s)) return -1;
    if (s->kind == DPM_BUMPY) return bumpy_contacts(s);
    hipLaunchKernelGGL(k_disk_contacts, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d,
                       params_of(s));
    return check_launch();
}

int dpm_scale_positions(dpm_system* s, double f) {
    hipLaunchKernelGGL(k_scale_positions, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, f);
    return check_launch();
}

// Particle::scaleBox (particle.cu:1057-1064).  The reference's pow(area, 1/2) is taken
// as the correctly rounded square root (the batched engine computes the same on the
// device, where only sqrt is correctly rounded).
int dpm_scale_box(dpm_system* s, double f) {
    if (dpm_scale_positions(s, f)) return -1;
    double side = std::sqrt(s->box[0] * s->box[1]);
    double ns = side * f;
    return dpm_set_box(s, ns, ns);
}

// Particle::scaleToPackingFractionFull (particle.cu:1047-1055)
int dpm_scale_to_packing_fraction(dpm_system* s, double phi) {
    double a;
    if (sum_to_host(s, S_R2, &a)) return -1;
    const double PI = 3.141592653589793238462643383279502884197;
    double area = a * PI;
    double ns = std::pow(area / phi, 1.0 / 2);
    double side = std::pow(s->box[0] * s->box[1], 1.0 / 2);
    if (dpm_scale_positions(s, ns / side)) return -1;
    return dpm_set_box(s, ns, ns);
}

static int snap_alloc(dpm_system* s) {
    if (!s->snap_allocs.empty()) return 0;
    auto al = [&](auto** p, int64_t n) {
        size_t bytes = sizeof(**p) * (size_t)std::max<int64_t>(n, 1);
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) return -1;
        s->snap_allocs.push_back(q);
        *p = (std::remove_reference_t<decltype(*p)>)q;
        return 0;
    };
    int64_t n = s->N;
    if (al(&s->snap.P, n) || al(&s->snap.vel, n) || al(&s->snap.frc, n) || al(&s->snap.sidx, n)) {
        set_error("snapshot allocation failed");
        return -1;
    }
    if (s->kind == DPM_BUMPY) {
        int64_t v = s->Nv;
        if (al(&s->snap.angle, n) || al(&s->snap.omega, n) || al(&s->snap.torque, n) ||
            al(&s->snap.moi, n) || al(&s->snap.nverts, n) || al(&s->snap.vstart, n) ||
            al(&s->snap.vpos, v) || al(&s->snap.vmass, v) || al(&s->snap.vpart, v) ||
            al(&s->snap.svidx, v)) {
            set_error("snapshot allocation failed");
            return -1;
        }
    }
    return 0;
}

// Particle::setLastState (particle.cu:956-974), held in HBM instead of host memory.
int dpm_snapshot_save(dpm_system* s) {
    if (snap_alloc(s)) return -1;
    auto cp = [&](void* dst, const void* src, size_t b) {
        return hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s->stream);
    };
    int64_t n = s->N;
    DPM_HIP(cp(s->snap.P, s->d.P0, sizeof(double4) * n));
    DPM_HIP(cp(s->snap.vel, s->d.vel, sizeof(double2) * n));
    DPM_HIP(cp(s->snap.frc, s->d.frc, sizeof(double2) * n));
    DPM_HIP(cp(s->snap.sidx, s->d.sidx, sizeof(int) * n));
    if (s->kind == DPM_BUMPY && bumpy_snapshot(s, true)) return -1;
    s->snap_box[0] = s->box[0];
    s->snap_box[1] = s->box[1];
    s->snap_valid = true;
    return 0;
}

// Particle::revertToLastState (particle.cu:977-991) + a full neighbour rebuild.
int dpm_snapshot_restore(dpm_system* s) {
    if (!s->snap_valid) {
        set_error("dpm_snapshot_restore: no snapshot");
        return -1;
    }
    auto cp = [&](void* dst, const void* src, size_t b) {
        return hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s->stream);
    };
    int64_t n = s->N;
    DPM_HIP(cp(s->d.P0, s->snap.P, sizeof(double4) * n));
    DPM_HIP(cp(s->d.P1, s->snap.P, sizeof(double4) * n));
    DPM_HIP(cp(s->d.vel, s->snap.vel, sizeof(double2) * n));
    DPM_HIP(cp(s->d.frc, s->snap.frc, sizeof(double2) * n));
    DPM_HIP(cp(s->d.sidx, s->snap.sidx, sizeof(int) * n));
    if (s->kind == DPM_BUMPY && bumpy_snapshot(s, false)) return -1;
    if (dpm_set_box(s, s->snap_box[0], s->snap_box[1])) return -1;
    if (s->nbr_ready && device_rebuild(s, 2, nullptr)) return -1;
    return 0;
}

// Particle::removeMeanVelocities (particle.cu:865-869)
int dpm_remove_mean_velocity(dpm_system* s) {
    double a, b;
    if (sum_to_host(s, S_VX, &a) || sum_to_host(s, S_VY, &b)) return -1;
    double mx = a / s->N, my = b / s->N;
    hipLaunchKernelGGL(k_add_velocity, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, mx, my, 0.0);
    return check_launch();
}

// Particle::scaleVelocitiesToTemperature (particle.cu:899-906)
int dpm_scale_to_temperature(dpm_system* s, double T) {
    double cur;
    if (dpm_reduce(s, DPM_RED_TEMPERATURE, &cur)) return -1;
    if (cur <= 0.0) {
        set_error("dpm_scale_to_temperature: current temperature is zero");
        return -1;
    }
    double f = std::sqrt(T / cur);
    hipLaunchKernelGGL(k_add_velocity, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, 0.0, 0.0, f);
    return check_launch();
}
// Pressure-bisection jamming, scripts/runArthurJamWithFire.cpp:36-106.
int dpm_jam(dpm_system* s, const dpm_jam_params* P, dpm_jam_result* out) {
    if (require_nbr(s, "dpm_jam")) return -1;
    double P_low = P->pressure_target, P_high = P->pressure_target * (1.0 + P->window_rel);
    int64_t fire_total = 0;
    dpm_fire_result fr;
    for (int it = 0; it < 1000; ++it) {
        if (dpm_minimize_fire(s, &P->fire, &fr)) return -1;
        fire_total += fr.steps;
        double pr;
        if (dpm_reduce(s, DPM_RED_PRESSURE, &pr)) return -1;
        if (pr < P_low) break;
        if (dpm_scale_box(s, P->decompress_factor)) return -1;
    }
    double L_low = -1, L_high = -1, L_last = -1;
    double rate = 1.0 - P->strain_step;
    int64_t cstep = 0;
    int status = 1;
    double pr = 0.0, nc = 0.0;
    if (dpm_snapshot_save(s)) return -1;
    while (cstep < P->max_compression_steps) {
        if (dpm_minimize_fire(s, &P->fire, &fr)) return -1;
        fire_total += fr.steps;
        if (dpm_reduce(s, DPM_RED_PRESSURE, &pr)) return -1;
        if (dpm_reduce(s, DPM_RED_CONTACTS, &nc)) return -1;
        double L = std::sqrt(s->box[0] * s->box[1]);
        double L_prior = L;
        if (pr < P_low) {
            L_low = L;
            if (dpm_snapshot_save(s)) return -1;
            L_last = L;
            if (L_high != -1) {
                L = (L_high + L_low) / 2;
                L_high = -1;
            } else {
                L = L * rate;
            }
        } else {
            if (pr > P_high) {
                L_high = L;
                if (dpm_snapshot_restore(s)) return -1;
                L_prior = L_last;
                L = (L_high + L_low) / 2;
            } else if (nc > P->min_contacts_per_particle * (double)s->N) {
                status = 0;
                break;
            }
        }
        if (L_low > 0 && L_high > 0 && std::abs(L_low / L_high - 1) < P->length_tol) {
            status = 2;
            break;
        }
        if (dpm_scale_box(s, L / L_prior)) return -1;
        cstep++;
    }
    if (out) {
        out->compression_steps = cstep;
        out->fire_steps_total = fire_total;
        out->status = status;
        out->pressure = pr;
        double phi, pe;
        if (dpm_reduce(s, DPM_RED_PACKING_FRACTION, &phi) || dpm_reduce(s, DPM_RED_PE, &pe))
            return -1;
        out->packing_fraction = phi;
        out->pe = pe;
        out->box_length = s->box[0];
        out->contacts = (int64_t)nc;
    }
    return 0;
}

int dpm_set_option(dpm_system* s, int opt, int64_t v) {
    switch (opt) {
        case DPM_OPT_GRAPH: s->use_graph = v != 0; return 0;
        case DPM_OPT_CHUNK:
            if (v < 2 || v > 4096 || (v % 2)) {
                set_error("dpm_set_option: chunk must be even and in [2, 4096]");
                return -1;
            }
            s->chunk = v;
            return 0;
        case DPM_OPT_TIMING:
            if (v < 0 || v > 2) {
                set_error("dpm_set_option: timing must be 0, 1 or 2");
                return -1;
            }
            s->timing = (int)v;
            // the events of a few launch units, created now rather than inside a timed run
            while (v && s->ev_pool.size() < 8) {
                hipEvent_t e;
                DPM_HIP(hipEventCreate(&e));
                s->ev_pool.push_back(e);
                s->ev_free.push_back(e);
            }
            return 0;
    }
    set_error("dpm_set_option: unknown option");
    return -1;
}

int dpm_kernel_time(dpm_system* s, int cls, double* ms, int64_t* n) {
    if (cls < 0 || cls > 2) {
        set_error("dpm_kernel_time: bad class");
        return -1;
    }
    collect_timing(s);
    *ms = s->ktime_ms[cls];
    *n = s->klaunch[cls];
    return 0;
}

int dpm_reset_kernel_time(dpm_system* s) {
    collect_timing(s);
    for (int k = 0; k < 3; ++k) {
        s->ktime_ms[k] = 0.0;
        s->klaunch[k] = 0;
    }
    return 0;
}

int dpm_synchronize(dpm_system* s) {
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

}  // extern "C"
