// dpm_batch.hip -- batched jamming engine: many independent systems of uneven N (disks
// and rigid bumpy particles side by side) in one device arena, all advanced by the same
// per-phase kernels, which a hipGraph replays; each system carries its own FIRE state,
// jamming controller and mode, so a system that has finished drops out (its waves exit
// at once) while the others go on.
//
// Per system the arithmetic is that of one dpm_system with the Verlet method running
// dpm_jam (the pressure bisection of scripts/runArthurJamWithFire.cpp:36-106 around
// minimizeFire, src/routines/minimization.cu:75-112): the same pair law and row order
// (ascending index), the same FIRE update, the same canonical 64-element reduction
// tree, the same controller steps and the same list rebuilds (a full rebuild after a
// restore; the prologue's test on the displacements at the end of the previous
// minimisation), so every system's result is bit-identical to that system jammed alone
// (alone or in any mix).  A rigid bumpy system follows the rigid path of dpm_system
// (vertex rows from the particle rows, vertex forces summed per particle in vertex
// order, the rigid FIRE moves).
//
// Arena layout: system s owns slots [off_s, off_s + n_s) with off_s a multiple of 64,
// so a wavefront never spans two systems and a wave's 64 lanes are one canonical
// chunk.  Padding slots belong to their system but never hold a particle.
//   P, snap   : double4 {x, y, radius, mass}      vel, frc, lref : double2
//   pe        : double                            nnum, sys_of  : int32
//   nbr       : int32 ELL, column-major nbr[k*S + slot] (S = arena slots)
//   part      : per chunk, 10 partial sums / maxima (below)
// A rigid system also owns vertex slots [voff_s, voff_s + nvs_s) (voff_s a multiple of
// 64; a padding vertex has vpart = -1), its particles' vertices contiguous in particle
// order (vst = first vertex slot, nvt = count):
//   vpos, vsnap : double2                 vpart, vsys_of, vnum : int32
//   vnbr      : int32 ELL, column-major vnbr[k*V + vslot] (V = vertex slots)
//   ang, asnap, omg, tq, moi : double per particle slot; lpos, mpos, rot0, rot1 : the
//   centres and rotations of the step's moves, read by the vertex pass
//
// One batched step is a fixed sequence of launches:
//   k_b_prep_move  per slot: snapshot save/restore and box scaling requested by the
//               controller (for a rigid particle with its vertices), zero velocities
//               (minimizeFire prologue); then the FIRE move of systems inside a
//               minimisation (Fire::step first half) and its displacement flags
//   k_b_build   per (system, 256-slot tile): O(n^2) row build through LDS tiles, for the
//               systems whose Verlet test (3 max d^2 > thr^2, b_rebuild_now) asks for it
//   k_b_build_cells  systems of 128+ particles: the same rows from a cell grid in LDS
//   k_b_force   forces (+ closing kick) and the chunk partials of P, PE, max F^2
//   k_b_ctl     per system: minimizeFire's stop test and Fire::step's decision; clears
//               the controller's applied requests and counts the step
//   k_b_eval    after a minimisation: zero v, stress, contacts, r^2 partials and the
//               displacement flags the next prologue tests
//   k_b_jam     per system: the jamming controller (next minimisation, or done)
// and for rigid systems (launched when the arena has vertex slots):
//   k_rb_vmove  per vertex: the rotations of the step's rigid moves
//   k_rb_vbuild per particle (a half-wave over its vertices): vertex rows from the
//               particle rows (after the builds)
//   k_rb_force  per 64-particle chunk: vertex forces, torques and energies from the vertex
//               rows into LDS, their sums over each particle's vertices in vertex order,
//               the closing kick and the chunk partials
//   k_rb_eval   per particle (a half-wave over its vertices), before k_b_eval: the
//               stress and contacts of a rigid system that finished a minimisation
#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "dpm_kernels.hpp"
#include "dpm_rigid.hpp"

namespace dpm {

// BM_RESTORE: the step that restores the snapshot and rebuilds the lists at the restored
// positions (dpm_snapshot_restore's full rebuild) before the box scaling and the
// minimisation that follow it
enum BMode {
    BM_PROLOGUE = 0, BM_FIRE = 1, BM_EVAL = 2, BM_DONE = 3, BM_EMPTY = 4, BM_RESTORE = 5
};
enum BKind { BK_DISK = 0, BK_RIGID = 1 };

// dpm_jam_params of one system (minimizeFire's and the controller's constants)
struct BJam {
    double alpha0, dt0, pe_t, pe_diff_t, fmax2;
    long long max_steps;
    double P_low, P_high, rate, dfac, ltol, zmin;
    long long max_csteps;
};
// BJ_MIN: one minimisation and nothing else (a minimisation request)
enum BPhase { BJ_DECOMP = 0, BJ_COMP = 1, BJ_MIN = 2 };
enum BPart {
    BP_PX = 0, BP_PY, BP_PE, BP_F2, BP_SXX, BP_SYY, BP_R2, BP_DMAX, BP_Z, BP_TW, BP_COUNT
};

struct BSys {
    long long off;  // first arena slot (multiple of 64)
    int n, nc;      // particles, 64-slot chunks
    double Lx, Ly;
    double rmax;
    // minimizeFire / Fire state (the fields of FireCtl)
    double alpha, dt, dt_max, dt_min, last_avg_pe, avg_pe, power, fmax_sq;
    long long n_good, n_bad, step;
    int stopped, done, converged, uphill;
    int mode;      // BMode of the next batched step
    int rebuild;   // Verlet rebuild before the next force evaluation
    int snap_op;   // pending for b_prep_slot: 1 save, 2 restore
    int pad0;
    double scale_f;  // pending position scale (0: none)
    // jamming controller (dpm_jam)
    int jphase, dec_it, status, pad1;
    double L_low, L_high, L_last, snapLx, snapLy, pr;
    long long cstep, fire_total, contacts;
    double pe_final, phi_final;
    long long done_step;  // batched step in which the system finished
    BJam J;
    // per-system constants: list cutoff, Verlet threshold, the half-box limit of the
    // branch-free minimum image, contact energy; rigid systems also their vertex slots,
    // vertex radius and vertex cutoffs (vc2: vertex-vertex, vp2: vertex-particle)
    int kind, rotation;
    int nv, nvs;
    long long voff;
    double rc, rc2, thr2, lim, e_c;
    double rv, vc2, vp2;
    double pend_f;  // BM_RESTORE: the box scaling applied after the restore's rebuild
};

struct BDev {
    long long nsys, nslot, ntiles;
    int cap;
    BSys* sys;
    int* sys_of;
    double4 *P, *snap;
    double2 *vel, *frc, *lref;
    double* pe;
    int *nnum, *nbr;
    double* part;      // BP_COUNT arrays of nslot/64 chunk values
    int2* tiles;       // (system, first slot) of each build tile
    int2* ctiles;      // (system, part) of each cell-grid build workgroup
    long long nctiles; // capacity of ctiles (the cell-grid build's grid size)
    int* flags;        // [0] capacity overflow, [1] largest row, [2] systems finished,
                       // [3] batched steps executed, [4] build tiles in use,
                       // [5] cell-grid workgroups in use, [6] largest vertex row,
                       // [9] rigid systems entering their evaluation (evlist entries)
    int* evlist;       // those systems, appended by k_b_ctl
    int* done_list;    // ids of finished systems, in order (ring of nsys entries)
    // common parameters
    double n_c, ncm1, two_nc, inv2nc;
    double cell_scale;  // cell-grid cells: at least cell_scale * rc wide (>= 1)
    int fastdiv, pad;
    // rigid bumpy arena (nvslot = 0: none)
    long long nvslot;
    int vcap, pad2;
    int *vsys_of, *vpart, *vnum, *vnbr, *vst, *nvt;
    double2 *vpos, *vsnap;
    double *ang, *asnap, *omg, *tq, *moi;
    double2 *lpos, *mpos, *rot0, *rot1;
    double2* evs;  // k_rb_eval's per-particle stress (xx, yy) for k_b_eval
    int* evz;      // and contact count
};

__device__ __forceinline__ double* bpart(const BDev& b, int which) {
    return b.part + (long long)which * (b.nslot / WAVE);
}

__device__ __forceinline__ PairParams bparams(const BDev& b, const BSys& S) {
    PairParams p{};
    p.Lx = S.Lx;
    p.Ly = S.Ly;
    p.invLx = 1.0 / S.Lx;
    p.invLy = 1.0 / S.Ly;
    p.halfLx = 0.5 * S.Lx;
    p.halfLy = 0.5 * S.Ly;
    p.e_c = S.e_c;
    p.n_c = b.n_c;
    p.ncm1 = b.ncm1;
    p.two_nc = b.two_nc;
    p.inv2nc = b.inv2nc;
    p.fast = 0.5 * fmin(S.Lx, S.Ly) > S.lim * (1.0 + 1e-9) ? 1 : 0;
    p.fastdiv = b.fastdiv;
    return p;
}

__device__ __forceinline__ void fire_reset(const BDev& b, BSys& S) {
    S.alpha = S.J.alpha0;
    S.dt = S.J.dt0;
    S.dt_max = 10.0 * S.J.dt0;
    S.dt_min = 0.001 * S.J.dt0;
    S.n_good = 0;
    S.n_bad = 0;
    S.step = 0;
    S.stopped = 0;
    S.done = 0;
    S.converged = 0;
    S.uphill = 0;
    S.power = 0.0;
}

// ------------------------------------------------------------------ kernels
// The controller's pending operations on one particle slot, in the controller's order:
// revertToLastState (2) or setLastState (1), then scaleBox's position scaling
// (kernelScalePositions); op 3 (leaving a full decompression loop) saves after the
// scaling.  A rigid particle's angle goes with its position.
__device__ __forceinline__ double4 b_slot_ops(const BDev& b, const BSys& S, long long i,
                                              bool rigid) {
    double4 q = b.P[i];
    double a = rigid ? b.ang[i] : 0.0;
    if (S.snap_op == 2) {
        q = b.snap[i];
        if (rigid) a = b.asnap[i];
    }
    if (S.snap_op == 1) {
        b.snap[i] = q;
        if (rigid) b.asnap[i] = a;
    }
    if (S.scale_f != 0.0) {
        q.x *= S.scale_f;
        q.y *= S.scale_f;
    }
    if (S.snap_op == 3) {
        b.snap[i] = q;
        if (rigid) b.asnap[i] = a;
    }
    b.P[i] = q;
    if (rigid) b.ang[i] = a;
    return q;
}

// The same operations on vertex slot v of particle p: the vertex snapshot, and the
// scaling's translation of the vertex with its centre (kernelScalePositions,
// rigid_bumpy/kernels.cu:1308-1333).  Reads the centre before b_slot_ops moves it.
__device__ __forceinline__ void rb_vertex_ops(const BDev& b, const BSys& S, long long v, int p) {
    double2 x = b.vpos[v];
    if (S.snap_op == 2) x = b.vsnap[v];
    if (S.snap_op == 1) b.vsnap[v] = x;
    if (S.scale_f != 0.0) {
        const double4 q = S.snap_op == 2 ? b.snap[p] : b.P[p];
        const double nx = q.x * S.scale_f, ny = q.y * S.scale_f;
        x.x += (nx - q.x);
        x.y += (ny - q.y);
    }
    if (S.snap_op == 3) b.vsnap[v] = x;
    b.vpos[v] = x;
}

// Each slot's controller work (b_prep_slot) and then its FIRE move (b_move_slot), in
// one kernel: both touch only their own slot, so the order per slot is the order of
// two separate passes.  Every condition that guards a ballot is uniform over a wave
// (one system per wave).
// The prologue's rebuild test reads the displacement flags k_b_eval recorded at the end of
// the previous minimisation, as dpm_minimize_fire's prologue reads the displacement
// maxima of the previous one's closing k_disp_partials: the controller's scaling in
// between is not measured (a restore rebuilds in its own step).
__device__ __forceinline__ void b_prep_slot(const BDev& b, const BSys& S, long long i,
                                            bool rigid) {
    const bool start = S.mode == BM_PROLOGUE;
    if (!start && S.snap_op == 0) return;  // wave-uniform: one system per wave
    if (i - S.off >= S.n) return;
    if (rigid) {  // the particle's vertices first: they read the centre before it moves
        const int st = b.vst[i], nv = b.nvt[i];
        for (int k = 0; k < nv; ++k) rb_vertex_ops(b, S, st + k, (int)i);
    }
    b_slot_ops(b, S, i, rigid);
    if (start) {
        b.vel[i] = make_double2(0.0, 0.0);  // minimizeFire: setVelocitiesToZero
        if (rigid) b.omg[i] = 0.0;
    }
}

// Fire::step first half (src/integrator/fire.cu:14-40): [uphill back-step with the
// decreased dt, zero v] + kick(dt/2) + mix + drift(dt)
__device__ __forceinline__ void b_move_slot(const BDev& b, const BSys& S, long long i) {
    if (S.mode != BM_FIRE) return;
    const bool live = i - S.off < S.n;
    double d2 = 0.0;
    if (live) {
        const double dt = S.dt, alpha = S.alpha;
        double4 q = b.P[i];
        double2 v = b.vel[i], f = b.frc[i];
        if (S.uphill) {
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
        b.P[i] = q;
        b.vel[i] = v;
        double2 l = b.lref[i];
        double dx = q.x - l.x, dy = q.y - l.y;
        d2 = dx * dx + dy * dy;
    }
    const bool far = __builtin_amdgcn_ballot_w64(3.0 * d2 > S.thr2) != 0;
    if ((threadIdx.x % WAVE) == 0) bpart(b, BP_DMAX)[i / WAVE] = far ? 1.0 : 0.0;
}

// The rigid Fire::step moves (rigid_fire_move); the centres and the sines and cosines of
// the two rotations are left for k_rb_vmove (k_bfire_move's arithmetic).
__device__ __forceinline__ void rb_move_slot(const BDev& b, const BSys& S, long long i) {
    if (S.mode != BM_FIRE) return;
    const bool live = i - S.off < S.n;
    double d2 = 0.0;
    if (live) {
        double4 q = b.P[i];
        double2 v = b.vel[i];
        double a = b.ang[i], w = b.omg[i];
        double2 old, mid;
        double dth0, dth;
        rigid_fire_move(q, v, a, w, b.frc[i], b.tq[i], b.moi[i], S.dt, S.alpha, S.uphill,
                        S.rotation, old, mid, dth0, dth);
        b.lpos[i] = old;
        b.mpos[i] = mid;
        if (S.uphill) {
            double s0, c0;
            dpm_sincos(dth0, &s0, &c0);
            b.rot0[i] = make_double2(s0, c0);
        }
        double s1, c1;
        dpm_sincos(dth, &s1, &c1);
        b.rot1[i] = make_double2(s1, c1);
        b.P[i] = q;
        b.ang[i] = a;
        b.vel[i] = v;
        b.omg[i] = w;
        double2 l = b.lref[i];
        double dx = q.x - l.x, dy = q.y - l.y;
        d2 = dx * dx + dy * dy;
    }
    const bool far = __builtin_amdgcn_ballot_w64(3.0 * d2 > S.thr2) != 0;
    if ((threadIdx.x % WAVE) == 0) bpart(b, BP_DMAX)[i / WAVE] = far ? 1.0 : 0.0;
}

__global__ __launch_bounds__(BLOCK) void k_b_prep_move(BDev b) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= b.nslot) return;
    const int s = b.sys_of[i];
    if (s < 0) return;  // free arena slots (64-aligned: whole waves)
    const BSys& S = b.sys[s];
    const bool rigid = S.kind == BK_RIGID;
    b_prep_slot(b, S, i, rigid);
    if (rigid) rb_move_slot(b, S, i);
    else b_move_slot(b, S, i);
}

// Vertex slots: the rotations of the step's rigid moves (k_bfire_vertices).
__global__ __launch_bounds__(BLOCK) void k_rb_vmove(BDev b) {
    const long long v = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= b.nvslot) return;
    const int s = b.vsys_of[v];
    if (s < 0) return;
    const BSys& S = b.sys[s];
    if (S.mode != BM_FIRE) return;
    const int p = b.vpart[v];
    if (p < 0) return;
    const int uphill = S.uphill;
    const double2 r0 = uphill ? b.rot0[p] : make_double2(0.0, 1.0), r1 = b.rot1[p];
    const double4 nq = b.P[p];
    b.vpos[v] = bfire_vertex_one(b.vpos[v], b.lpos[p], b.mpos[p], make_double2(nq.x, nq.y), r0.x,
                                 r0.y, r1.x, r1.y, uphill);
}

// Whether system S rebuilds its rows in this step (called by a whole wave): the initial
// build, a restore step (dpm_snapshot_restore's full rebuild), or
// Particle::checkForVerletListUpdate (particle.cu:1141-1147) on the displacement flags of
// the step (the move's, or at a prologue those of the last minimisation's end): some
// 3 d^2 > thr^2.
__device__ __forceinline__ bool b_rebuild_now(const BDev& b, const BSys& S) {
    if (S.mode == BM_DONE || S.mode == BM_EVAL || S.mode == BM_EMPTY) return false;
    if (S.rebuild || S.mode == BM_RESTORE) return true;
    const double* D = bpart(b, BP_DMAX) + S.off / WAVE;  // per wave: some 3 d^2 > thr
    bool far = false;
    for (int c = threadIdx.x % WAVE; c < S.nc; c += WAVE) far = far || D[c] != 0.0;
    return __builtin_amdgcn_ballot_w64(far) != 0;
}

// Systems of BC_MIN_N particles and more get their rows from a cell grid instead of the
// all-pairs scan (one work item per BC_PER particles, listed in ctiles; the workgroup
// holding an item bins the whole system into its LDS and builds the rows of every
// parts-th block of 1024 particles).  Cells are at least rc (1 + 1e-9) wide,
// so every particle within rc of i lies in the 3 x 3 cells around i's, and the rows
// hold the same indices in the same ascending order as the all-pairs scan.
constexpr int BC_MIN_N = 128;      // systems from this size up use the cell grid
constexpr int BC_MAX_N = 16384;    // particles per system the LDS lists hold
constexpr int BC_MAX_D = 96;       // cells per side at most
constexpr int BC_CELLS = BC_MAX_D * BC_MAX_D;
constexpr int BC_BLOCK = 1024;
constexpr int BC_PER = 2048;       // particles per workgroup (each bins the whole system
                                   // itself, then builds the rows of its part)

struct BCGrid {
    int nx, ny;
    double csx, csy;
};

__device__ __forceinline__ bool bc_grid(const BDev& b, const BSys& S, BCGrid& g) {
    if (S.n < BC_MIN_N || S.n > BC_MAX_N) return false;
    const double w = S.rc * (1.0 + 1e-9) * b.cell_scale;
    g.nx = (int)fmin((double)BC_MAX_D, floor(S.Lx / w));
    g.ny = (int)fmin((double)BC_MAX_D, floor(S.Ly / w));
    if (g.nx < 3 || g.ny < 3) return false;
    g.csx = S.Lx / g.nx;
    g.csy = S.Ly / g.ny;
    return true;
}

// the cell of one coordinate: wrapped into [0, L], clamped to the grid
__device__ __forceinline__ int bc_coord(double x, double L, double cs, int n) {
    double w = x - L * floor(x / L);
    int c = (int)(w / cs);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// One work item (system, part): every early return is uniform over the workgroup.
// cend: counts, then (after the fill) each cell's end; lst: the system's particles
// (system-local) grouped by cell.
__device__ __forceinline__ void bc_build_item(const BDev& b, int2 ct, int* cend, int* lst,
                                              int* wsum) {
    const BSys& S = b.sys[ct.x];
    if (!b_rebuild_now(b, S)) return;  // the same verdict in every wave of the workgroup
    BCGrid g;
    if (!bc_grid(b, S, g)) return;
    const int part = ct.y, parts = (S.n + BC_PER - 1) / BC_PER;
    const int tid = threadIdx.x, n = S.n, ncell = g.nx * g.ny;
    const long long off = S.off;
    for (int c = tid; c < ncell; c += BC_BLOCK) cend[c] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += BC_BLOCK) {
        double4 q = b.P[off + i];
        int c = bc_coord(q.x, S.Lx, g.csx, g.nx) + g.nx * bc_coord(q.y, S.Ly, g.csy, g.ny);
        atomicAdd(&cend[c], 1);
    }
    __syncthreads();
    // exclusive scan of the counts: each thread owns a run of consecutive cells
    const int per = (ncell + BC_BLOCK - 1) / BC_BLOCK;
    const int c0 = min(tid * per, ncell), c1 = min(c0 + per, ncell);
    int local = 0;
    for (int c = c0; c < c1; ++c) local += cend[c];
    const int lane = tid % WAVE, wv = tid / WAVE;
    int incl = local;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        int o = __shfl_up(incl, d, WAVE);
        if (lane >= d) incl += o;
    }
    if (lane == WAVE - 1) wsum[wv] = incl;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wv; ++w) before += wsum[w];
    int run = before + incl - local;
    for (int c = c0; c < c1; ++c) {
        int t = cend[c];
        cend[c] = run;
        run += t;
    }
    __syncthreads();
    for (int i = tid; i < n; i += BC_BLOCK) {
        double4 q = b.P[off + i];
        int c = bc_coord(q.x, S.Lx, g.csx, g.nx) + g.nx * bc_coord(q.y, S.Ly, g.csy, g.ny);
        lst[atomicAdd(&cend[c], 1)] = i;
    }
    __syncthreads();
    // each cell's members in ascending order (a few per cell)
    for (int c = tid; c < ncell; c += BC_BLOCK) {
        const int a = c == 0 ? 0 : cend[c - 1], e = cend[c];
        for (int u = a + 1; u < e; ++u) {
            int v = lst[u], k = u - 1;
            while (k >= a && lst[k] > v) {
                lst[k + 1] = lst[k];
                --k;
            }
            lst[k + 1] = v;
        }
    }
    __syncthreads();
    const PairParams p = bparams(b, S);
    constexpr int NONE = 0x7fffffff;
    int mx = 0;
    for (int il = part * BC_BLOCK + tid; il < n; il += parts * BC_BLOCK) {
        const long long i = off + il;
        const double4 qi = b.P[i];
        const int cx = bc_coord(qi.x, S.Lx, g.csx, g.nx), cy = bc_coord(qi.y, S.Ly, g.csy, g.ny);
        // the nine cells' ranges, merged in ascending particle order
        int pa[9], pe[9], h[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            int x = cx + t % 3 - 1, y = cy + t / 3 - 1;
            x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
            y = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
            const int c = x + g.nx * y;
            pa[t] = c == 0 ? 0 : cend[c - 1];
            pe[t] = cend[c];
            h[t] = pa[t] < pe[t] ? lst[pa[t]] : NONE;
        }
        int cnt = 0;
        while (true) {
            int best = h[0], bt = 0;
#pragma unroll
            for (int t = 1; t < 9; ++t)
                if (h[t] < best) {
                    best = h[t];
                    bt = t;
                }
            if (best == NONE) break;
#pragma unroll
            for (int t = 0; t < 9; ++t)
                if (t == bt) {
                    pa[t]++;
                    h[t] = pa[t] < pe[t] ? lst[pa[t]] : NONE;
                }
            if (best == il) continue;
            const long long j = off + best;
            const double4 qj = b.P[j];
            double dx, dy;
            if (p.fast) {
                dx = pbc_rint(qi.x, qj.x, p.Lx, p.invLx);
                dy = pbc_rint(qi.y, qj.y, p.Ly, p.invLy);
            } else {
                dx = pbc(qi.x, qj.x, p.Lx, p.invLx, p.halfLx, 0);
                dy = pbc(qi.y, qj.y, p.Ly, p.invLy, p.halfLy, 0);
            }
            if (dx * dx + dy * dy < S.rc2) {
                if (cnt < b.cap) b.nbr[(long long)cnt * b.nslot + i] = (int)j;
                cnt++;
            }
        }
        b.nnum[i] = cnt;
        b.lref[i] = make_double2(qi.x, qi.y);
        mx = max(mx, cnt);
    }
    mx = wave_max_i(mx);
    if (lane == 0) {
        atomicMax(&b.flags[1], mx);
        if (mx > b.cap) atomicMax(&b.flags[0], 1);
    }
}

// A fixed grid of BC_GRID workgroups (the LDS holds one per CU) walks the work list, so
// the graph-baked launch stays small whatever the list's capacity.
constexpr int BC_GRID = 512;
__global__ __launch_bounds__(BC_BLOCK) void k_b_build_cells(BDev b) {
    __shared__ int cend[BC_CELLS];
    __shared__ int lst[BC_MAX_N];
    __shared__ int wsum[BC_BLOCK / WAVE];
    const int count = b.flags[5];
    for (int w = blockIdx.x; w < count; w += gridDim.x) {
        bc_build_item(b, b.ctiles[w], cend, lst, wsum);
        __syncthreads();  // the next item reuses the LDS
    }
}

// kernelUpdateNeighborList (base/kernels.cu:256-293) for the systems that need it:
// every other particle of the system with d^2 < rc^2, ascending index.  One block per
// 256-slot tile of a system; the system's positions pass through LDS.
__global__ __launch_bounds__(BLOCK) void k_b_build(BDev b) {
    if ((int)blockIdx.x >= b.flags[4]) return;
    const int2 t = b.tiles[blockIdx.x];
    const BSys& S = b.sys[t.x];
    if (!b_rebuild_now(b, S)) return;  // the same verdict in every wave of the block
    BCGrid g;
    if (bc_grid(b, S, g)) return;  // k_b_build_cells builds this system
    __shared__ double2 tp[BLOCK];
    const long long i = (long long)t.y + threadIdx.x;
    const long long end = S.off + S.n;
    const bool live = i < end;
    double4 qi = live ? b.P[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    // when half the box exceeds every cutoff the branch-free minimum image gives the
    // same verdict for every pair (it differs only where |d| is half the box)
    const PairParams p = bparams(b, S);
    int cnt = 0;
    for (long long base = S.off; base < end; base += BLOCK) {
        __syncthreads();
        if (base + threadIdx.x < end) {
            double4 q = b.P[base + threadIdx.x];
            tp[threadIdx.x] = make_double2(q.x, q.y);
        }
        __syncthreads();
        int lim = (int)min((long long)BLOCK, end - base);
        if (live) {
            for (int u = 0; u < lim; ++u) {
                long long j = base + u;
                if (j == i) continue;
                double dx, dy;
                if (p.fast) {
                    dx = pbc_rint(qi.x, tp[u].x, p.Lx, p.invLx);
                    dy = pbc_rint(qi.y, tp[u].y, p.Ly, p.invLy);
                } else {
                    dx = pbc(qi.x, tp[u].x, p.Lx, p.invLx, p.halfLx, 0);
                    dy = pbc(qi.y, tp[u].y, p.Ly, p.invLy, p.halfLy, 0);
                }
                if (dx * dx + dy * dy < S.rc2) {
                    if (cnt < b.cap) b.nbr[(long long)cnt * b.nslot + i] = (int)j;
                    cnt++;
                }
            }
        }
    }
    if (live) {
        b.nnum[i] = cnt;
        b.lref[i] = make_double2(qi.x, qi.y);
    }
    int m = block_max_int(live ? cnt : 0);
    if (threadIdx.x == 0) {
        atomicMax(&b.flags[1], m);
        if (m > b.cap) atomicMax(&b.flags[0], 1);
    }
}

// zero + calculateForces + (closing) updateVelocities(dt/2), chunk partials of
// F.v, PE and max |F|^2 (k_fire_force's arithmetic)
// FASTDIV (the batch's pair law allows the short sqrt/division sequences): pair_hot, with
// the reference's minimum image for a box too small for the branch-free one; otherwise
// pair_law.  The two instantiations keep pair_law's registers out of the usual one.
template <bool FASTDIV>
__global__ __launch_bounds__(BLOCK) void k_b_force(BDev b) {
    // XCD-aware block order: a system's slots (and the neighbours they gather) stay in
    // one XCD's L2
    // (any workgroup size that is a multiple of 64 up to BLOCK: the pass is per slot)
    const long long i =
        (long long)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= b.nslot) return;
    const int s = b.sys_of[i];
    if (s < 0) return;  // free arena slots (64-aligned: whole waves)
    const BSys& S = b.sys[s];
    if (S.mode != BM_PROLOGUE && S.mode != BM_FIRE) return;
    if (S.kind == BK_RIGID) return;  // k_rb_force
    const int kick = S.mode == BM_FIRE;
    const PairParams p = bparams(b, S);
    const bool live = i - S.off < S.n;
    const long long il = live ? i : S.off;
    const double4 qi = b.P[il];
    const int nn = live ? min(b.nnum[il], b.cap) : 0;
    const int* col = b.nbr + il;
    double ax = 0.0, ay = 0.0, e = 0.0;
    const bool fast = p.fast;
    // two neighbours per batch, the next batch's indices loaded ahead; every ELL slot
    // holds a valid arena slot (zero-filled at allocation, only written with slots), and
    // a slot past the row end adds +0.0, which leaves the sums unchanged
    const int last = b.cap - 1;
    int jn[2] = {col[0], col[(long long)min(1, last) * b.nslot]};
    for (int k = 0; __builtin_amdgcn_ballot_w64(k < nn) != 0; k += 2) {
        const double4 q2[2] = {b.P[jn[0]], b.P[jn[1]]};
        if (__builtin_amdgcn_ballot_w64(k + 2 < nn) != 0) {
            jn[0] = col[(long long)min(k + 2, last) * b.nslot];
            jn[1] = col[(long long)min(k + 3, last) * b.nslot];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const bool valid = k + t < nn;
            const double4 q = q2[t];
            double fx = 0.0, fy = 0.0, ie = 0.0;
            if constexpr (FASTDIV) {
                if (fast) ie = pair_hot<true>(qi, q, valid, p, fx, fy);
                else ie = pair_hot<true, true>(qi, q, valid, p, fx, fy);
            } else if (valid) {
                ie = pair_law(qi.x, qi.y, qi.z, q.x, q.y, q.z, p, fx, fy);
            }
            ax += fx;
            ay += fy;
            e += ie;
        }
    }
    double pxv = 0.0, pyv = 0.0, pe = 0.0, f2 = 0.0;
    if (live) {
        double fx = 0.0 + ax, fy = 0.0 + ay;
        b.frc[i] = make_double2(fx, fy);
        pe = 0.0 + e;
        b.pe[i] = pe;
        double2 v = b.vel[i];
        if (kick) {
            double dtm = (S.dt / 2.0) / qi.w;
            v.x += fx * dtm;
            v.y += fy * dtm;
            b.vel[i] = v;
        }
        pxv = fx * v.x;
        pyv = fy * v.y;
        f2 = fx * fx + fy * fy;
    }
    double s0 = wave_tree_sum(pxv);
    double s1 = wave_tree_sum(pyv);
    double s2 = wave_tree_sum(pe);
    double m = wave_max(f2);
    if ((threadIdx.x % WAVE) == 0) {
        long long ch = i / WAVE;
        bpart(b, BP_PX)[ch] = s0;
        bpart(b, BP_PY)[ch] = s1;
        bpart(b, BP_PE)[ch] = s2;
        bpart(b, BP_F2)[ch] = m;
    }
}

// kernelUpdateVertexNeighborList (rigid_bumpy/kernels.cu:870-922) for the rigid systems
// whose particle rows were rebuilt in this step: for vertex v of particle p, over p's
// particle row in order, the particles q whose centre lies within vp of v, then q's
// vertices in order within vc of v (k_bumpy_vertex_build's row order).  A half-wave per
// particle slot, its lanes over the particle's vertices: the row entries and q's vertex
// positions are the same address for the 32 lanes (one load serves them all).  Two
// consecutive slots share a wave, and a 64-slot chunk belongs to one system, so every
// early return is uniform over the wave.  q's vertices are scanned only when v lies
// within ring_reach of q's centre (vertex_ring_near): the rows are unchanged.
constexpr int RB_HALF = 32;
__global__ __launch_bounds__(BLOCK) void k_rb_vbuild(BDev b) {
    const long long t = (long long)blockIdx.x * BLOCK + threadIdx.x;
    const long long i = t / RB_HALF;
    const int lane = (int)(t % RB_HALF);
    if (i >= b.nslot) return;
    const int s = b.sys_of[i];
    if (s < 0) return;  // free arena slots (64-aligned: whole waves)
    const BSys& S = b.sys[s];
    if (S.kind != BK_RIGID || !b_rebuild_now(b, S)) return;
    const PairParams pp = bparams(b, S);
    const double vc = sqrt(S.vc2);
    const bool shift = fmin(pp.halfLx, pp.halfLy) > vertex_shift_limit(S.rmax, S.rv, vc);
    const int hs = (int)(threadIdx.x % WAVE) / RB_HALF * RB_HALF;  // the half's first lane
    const unsigned lt = (1u << lane) - 1u;  // the half's lanes below this one
    int mx = 0;
    if (i - S.off < S.n) {  // uniform over the half-wave
        const int st = b.vst[i], nv = b.nvt[i];
        const int nn = min(b.nnum[i], b.cap);
        // the particle row, one entry per lane (nn <= cap = RB_HALF): the neighbour's
        // record and vertex range, loaded at once and passed round by cross-lane reads
        // instead of two dependent loads per entry
        int rf = 0, rn = 0;
        double rx = 0.0, ry = 0.0, rr = 0.0;
        if (lane < nn) {
            const int q = b.nbr[(long long)lane * b.nslot + i];
            const double4 c = b.P[q];
            rx = c.x;
            ry = c.y;
            rr = c.z;
            rf = b.vst[q];
            rn = b.nvt[q];
        }
        for (int a0 = 0; a0 < nv; a0 += RB_HALF) {
            // lane: vertex a0 + lane of this particle; the rows of the half's vertices grow
            // together, q by q, each q's vertices tested by all 32 lanes at once for one
            // gated vertex after another
            const bool has = a0 + lane < nv;
            const double2 xv = has ? b.vpos[st + a0 + lane] : make_double2(0.0, 0.0);
            int cnt = 0;
            for (int k = 0; k < nn; ++k) {
                const double4 cq = make_double4(__shfl(rx, k, RB_HALF), __shfl(ry, k, RB_HALF),
                                                __shfl(rr, k, RB_HALF), 0.0);
                const int f0 = __shfl(rf, k, RB_HALF), nq = __shfl(rn, k, RB_HALF);
                const double kx = image_index(xv.x - cq.x, pp.Lx, pp.invLx, pp.halfLx, pp.fast);
                const double ky = image_index(xv.y - cq.y, pp.Ly, pp.invLy, pp.halfLy, pp.fast);
                const double dx = pbc_known(xv.x, cq.x, pp.Lx, kx, pp.fast);  // = pbc(xv, cq)
                const double dy = pbc_known(xv.y, cq.y, pp.Ly, ky, pp.fast);
                const double d2 = dx * dx + dy * dy;
                const bool gate = has && d2 < S.vp2 && vertex_ring_near(d2, cq.z, S.rv, vc);
                unsigned gm = (unsigned)(__builtin_amdgcn_ballot_w64(gate) >> hs);
                if (gm == 0) continue;
                for (int u0 = 0; u0 < nq; u0 += RB_HALF) {
                    const bool uin = u0 + lane < nq;
                    const int u = f0 + u0 + lane;
                    const double2 xu = uin ? b.vpos[u] : make_double2(0.0, 0.0);
                    for (unsigned m = gm; m != 0; m &= m - 1) {  // gated vertices, in order
                        const int g = __builtin_ctz(m);
                        const double gx = __shfl(xv.x, g, RB_HALF), gy = __shfl(xv.y, g, RB_HALF);
                        double ex, ey;
                        if (shift) {
                            ex = pbc_known(gx, xu.x, pp.Lx, __shfl(kx, g, RB_HALF), pp.fast);
                            ey = pbc_known(gy, xu.y, pp.Ly, __shfl(ky, g, RB_HALF), pp.fast);
                        } else {
                            ex = pbc(gx, xu.x, pp.Lx, pp.invLx, pp.halfLx, pp.fast);
                            ey = pbc(gy, xu.y, pp.Ly, pp.invLy, pp.halfLy, pp.fast);
                        }
                        const bool hit = uin && ex * ex + ey * ey < S.vc2;
                        const unsigned hm = (unsigned)(__builtin_amdgcn_ballot_w64(hit) >> hs);
                        if (hm == 0) continue;
                        // ascending u: lane order; vertex g's row continues at its count
                        const int pos = __shfl(cnt, g, RB_HALF) + __builtin_popcount(hm & lt);
                        if (hit && pos < b.vcap)
                            b.vnbr[(long long)pos * b.nvslot + st + a0 + g] = u;
                        if (lane == g) cnt += __builtin_popcount(hm);
                    }
                }
            }
            if (has) {
                b.vnum[st + a0 + lane] = cnt;
                mx = max(mx, cnt);
            }
        }
    }
    mx = wave_max_i(mx);
    if ((threadIdx.x % WAVE) == 0) {
        atomicMax(&b.flags[6], mx);
        if (mx > b.vcap) atomicMax(&b.flags[0], 1);
    }
}

// The force phase of a rigid FIRE step for one 64-particle chunk per workgroup:
// kernelCalcRigidBumpyForces1 (rigid_bumpy/kernels.cu:366-417) for every vertex of the
// chunk's particles -- force, torque about the centre (calcTorque,
// include/particles/base/kernels.cuh:237-241) and energy from the vertex row, in row
// order -- into LDS, a piece of RB_PIECE vertices at a time; their per-particle sums in
// vertex order (kernelCalcRigidBumpyParticleForces1, :419-442; four lanes per particle,
// one per quantity, the running sums carried across pieces); then the closing
// updateVelocities(dt/2) of a FIRE step (kernelUpdateRigidVelocities) and the chunk
// partials of F.v, tau.w, PE and max(F^2, tau^2) (k_bfire_forces + k_bfire_partials'
// arithmetic).  The per-vertex values never leave LDS.
// 1,024 threads and a piece of 1,792 vertices (64 particles of up to 28 vertices in one
// piece, 57 KB of LDS): the chunks are few (one per 64 particles), so each carries as many
// waves as a workgroup holds (3.34 packings/s on 64 rigid systems against 3.19 with 512
// threads and pieces of 1,024)
constexpr int RB_FT = 1024;             // threads per chunk
constexpr int RB_PIECE = 1792;          // vertices per LDS piece

// Force, torque and energy of vertex slot v (pair_law's path: the batch's pair law does not
// allow the short sequences)
__device__ __forceinline__ void rb_vforce_law(const BDev& b, const PairParams& p, double rv,
                                              long long v, double& ax, double& ay,
                                              double& e) {
    const double2 xv = b.vpos[v];
    const int nn = min(b.vnum[v], b.vcap);
    ax = 0.0;
    ay = 0.0;
    e = 0.0;
    for (int k = 0; k < nn; ++k) {
        const double2 xu = b.vpos[b.vnbr[(long long)k * b.nvslot + v]];
        double fx, fy;
        const double ie = pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
        ax += fx;
        ay += fy;
        e += ie;
    }
}

template <bool FASTDIV>
__global__ __launch_bounds__(RB_FT) void k_rb_force(BDev b) {
    const long long c0 = (long long)xcd_tile(blockIdx.x, gridDim.x) * WAVE;  // chunk's slot
    if (c0 >= b.nslot) return;
    const int s = b.sys_of[c0];
    if (s < 0) return;  // a free chunk (64-aligned)
    const BSys& S = b.sys[s];
    if (S.kind != BK_RIGID || (S.mode != BM_PROLOGUE && S.mode != BM_FIRE)) return;
    const PairParams p = bparams(b, S);
    const long long end = min(c0 + WAVE, S.off + S.n);  // live slots [c0, end)
    if (end <= c0) return;
    __shared__ double sv[4][RB_PIECE];
    __shared__ double sres[4][WAVE];
    const int tid = threadIdx.x;
    const int j = tid >> 2, comp = tid & 3;  // the summing lane: particle c0 + j, quantity
    const bool sums = tid < 4 * WAVE && c0 + j < end;
    const int jst = sums ? b.vst[c0 + j] : 0, jnv = sums ? b.nvt[c0 + j] : 0;
    const long long lo = b.vst[c0], hi = (long long)b.vst[end - 1] + b.nvt[end - 1];
    const double rv = S.rv;
    const int last = b.vcap - 1;
    const long long row1 = (long long)min(1, last) * b.nvslot;
    double acc = 0.0;
    for (long long base = lo; base < hi; base += RB_PIECE) {
        const long long top = min(base + RB_PIECE, hi);
        for (long long v = base + tid; v < top; v += RB_FT) {
            double ax = 0.0, ay = 0.0, e = 0.0;
            const double2 xv = b.vpos[v];
            if constexpr (FASTDIV) {
                const double4 qv = make_double4(xv.x, xv.y, rv, 0.0);
                const int nn = min(b.vnum[v], b.vcap);
                const int* col = b.vnbr + v;
                // two neighbours per batch, the next batch's indices loaded ahead; every
                // ELL slot holds a valid vertex slot (zero-filled at allocation, only
                // written with slots), and a slot past the row end adds +0.0
                int u0 = col[0], u1 = col[row1];
                for (int k = 0; __builtin_amdgcn_ballot_w64(k < nn) != 0; k += 2) {
                    const double2 x0 = b.vpos[u0], x1 = b.vpos[u1];
                    if (__builtin_amdgcn_ballot_w64(k + 2 < nn) != 0) {
                        u0 = col[(long long)min(k + 2, last) * b.nvslot];
                        u1 = col[(long long)min(k + 3, last) * b.nvslot];
                    }
                    const double4 q0 = make_double4(x0.x, x0.y, rv, 0.0);
                    const double4 q1 = make_double4(x1.x, x1.y, rv, 0.0);
                    double fx, fy, ie;
                    if (p.fast) ie = pair_hot<true>(qv, q0, k < nn, p, fx, fy);
                    else ie = pair_hot<true, true>(qv, q0, k < nn, p, fx, fy);
                    ax += fx;
                    ay += fy;
                    e += ie;
                    if (p.fast) ie = pair_hot<true>(qv, q1, k + 1 < nn, p, fx, fy);
                    else ie = pair_hot<true, true>(qv, q1, k + 1 < nn, p, fx, fy);
                    ax += fx;
                    ay += fy;
                    e += ie;
                }
            } else {
                rb_vforce_law(b, p, rv, v, ax, ay, e);
            }
            // calcTorque (include/particles/base/kernels.cuh:237-241) about the centre
            const double4 cq = b.P[b.vpart[v]];
            const double dx = xv.x - cq.x, dy = xv.y - cq.y;
            const double t = ay * dx - ax * dy;
            const int l = (int)(v - base);
            sv[0][l] = 0.0 + ax;
            sv[1][l] = 0.0 + ay;
            sv[2][l] = 0.0 + t;
            sv[3][l] = 0.0 + e;
        }
        __syncthreads();
        if (sums) {  // this particle's vertices inside the piece, in vertex order
            const long long a0 = max((long long)jst, base), a1 = min((long long)jst + jnv, top);
            const double* src = sv[comp];
            for (long long a = a0; a < a1; ++a) acc += src[a - base];
        }
        __syncthreads();  // the next piece reuses the LDS
    }
    if (sums) sres[comp][j] = 0.0 + acc;
    __syncthreads();
    if (tid >= WAVE) return;  // wave 0: one lane per particle slot of the chunk
    const long long i = c0 + tid;
    double pxv = 0.0, pyv = 0.0, ptw = 0.0, pe = 0.0, f2 = 0.0;
    if (i < end) {
        const double fx = sres[0][tid], fy = sres[1][tid], t = sres[2][tid];
        pe = sres[3][tid];
        b.frc[i] = make_double2(fx, fy);
        b.tq[i] = t;
        b.pe[i] = pe;
        double2 v = b.vel[i];
        double w = b.omg[i];
        if (S.mode == BM_FIRE) {
            const double hdt = S.dt / 2.0;
            if (S.rotation) w += t * hdt / b.moi[i];
            else w = 0.0;
            const double m = b.P[i].w;
            v.x += fx * hdt / m;
            v.y += fy * hdt / m;
            b.vel[i] = v;
            b.omg[i] = w;
        }
        pxv = fx * v.x;
        pyv = fy * v.y;
        ptw = t * w;
        f2 = fmax(fx * fx + fy * fy, t * t);
    }
    const double s0 = wave_tree_sum(pxv);
    const double s1 = wave_tree_sum(pyv);
    const double s2 = wave_tree_sum(pe);
    const double s3 = wave_tree_sum(ptw);
    const double m = wave_max(f2);
    if (tid == 0) {
        const long long ch = c0 / WAVE;
        bpart(b, BP_PX)[ch] = s0;
        bpart(b, BP_PY)[ch] = s1;
        bpart(b, BP_PE)[ch] = s2;
        bpart(b, BP_TW)[ch] = s3;
        bpart(b, BP_F2)[ch] = m;
    }
}

// Vertex ring, vertex masses and moment of inertia of the particles of one rigid system
// (k_bumpy_init_vertices' arithmetic: kernelInitializeVerticesOnParticles,
// rigid_bumpy/kernels.cu:1137-1162, kernelGetMomentsOfInertia, :1171-1189).
__global__ __launch_bounds__(BLOCK) void k_rb_init(BDev b, long long off, int n, double rv) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= n) return;
    const long long p = off + k;
    const double4 cq = b.P[p];
    const double a = b.ang[p];
    const int nv = b.nvt[p], st = b.vst[p];
    const double2 c = make_double2(cq.x, cq.y);
    for (int i = 0; i < nv; ++i) b.vpos[st + i] = ring_vertex(c, cq.z, a, nv, rv, i);
    const double vm = cq.w / static_cast<double>(nv);
    double I = 0.0;
    for (int i = 0; i < nv; ++i) {
        const double rx = b.vpos[st + i].x - c.x, ry = b.vpos[st + i].y - c.y;
        I += vm * (rx * rx + ry * ry);
    }
    b.moi[p] = I;
}

// Particle::scaleBox's box update (particle.cu:1057-1064): side = sqrt(area) * f
__device__ __forceinline__ void scale_box_of(BSys& S, double f) {
    double side = sqrt(S.Lx * S.Ly);
    double ns = side * f;
    S.Lx = ns;
    S.Ly = ns;
    S.scale_f = f;
}

// the end of a minimisation: the system is evaluated in this step (k_rb_eval lists the
// rigid ones)
__device__ __forceinline__ void enter_eval(const BDev& b, BSys& S, int s) {
    S.mode = BM_EVAL;
    if (S.kind == BK_RIGID) b.evlist[atomicAdd(&b.flags[9], 1)] = s;
}

// one wave per system: minimizeFire's stop test (minimization.cu:96-105, with the
// optional |F|max criterion) and Fire::step's decision for the next iteration
// (fire.cu:14-33) -- k_fire_ctl's arithmetic on the system's chunk partials
__global__ __launch_bounds__(WAVE) void k_b_ctl(BDev b) {
    BSys& S = b.sys[blockIdx.x];
    const int mode = S.mode;
    if (threadIdx.x == 0) {
        if (blockIdx.x == 0) b.flags[3] += 1;  // one thread of the step counts it
        // the controller's requests of the last step are applied (k_b_prep_move)
        S.snap_op = 0;
        S.scale_f = 0.0;
    }
    if (mode == BM_RESTORE) {  // the lists are current for the restored state: scaleBox
        if (threadIdx.x == 0) {
            scale_box_of(S, S.pend_f);
            S.pend_f = 0.0;
            S.rebuild = 0;
            S.mode = BM_PROLOGUE;
        }
        return;
    }
    if (mode != BM_PROLOGUE && mode != BM_FIRE) return;
    const bool rigid = S.kind == BK_RIGID;
    const long long c0 = S.off / WAVE;
    // (the fourth sum is the rigid spin power; a disk system sums its PE array again)
    double px, py, pe, ptw;
    wave_final_sum4(bpart(b, BP_PX) + c0, bpart(b, BP_PY) + c0, bpart(b, BP_PE) + c0,
                    bpart(b, rigid ? BP_TW : BP_PE) + c0, S.nc, px, py, pe, ptw);
    double fm = 0.0;
    for (int k = threadIdx.x; k < S.nc; k += WAVE) fm = fmax(fm, bpart(b, BP_F2)[c0 + k]);
    fm = wave_max(fm);
    if (threadIdx.x != 0) return;
    S.rebuild = 0;  // the list this step used is current
    const double dof = (double)S.n;
    double avg = pe / dof;
    S.fmax_sq = fm;
    S.avg_pe = avg;
    if (mode == BM_PROLOGUE) {
        S.last_avg_pe = avg;
        S.mode = BM_FIRE;
        if (S.J.max_steps <= 0) {
            S.done = 1;
            enter_eval(b, S, blockIdx.x);
            return;
        }
    } else {
        double diff = fabs(avg / S.last_avg_pe - 1);
        S.last_avg_pe = avg;
        bool fok = S.J.fmax2 > 0.0 && fm < S.J.fmax2;
        S.step += 1;
        if (diff < S.J.pe_diff_t || avg < S.J.pe_t || fok) {
            S.converged = 1;
            S.done = 1;
            enter_eval(b, S, blockIdx.x);
            return;
        }
        if (S.step >= S.J.max_steps) {
            S.done = 1;
            enter_eval(b, S, blockIdx.x);
            return;
        }
    }
    // RigidBumpy::getPowerFire adds sum(tau*w) (rigid_bumpy.cu:1181-1191)
    double power = px + py;
    if (rigid) power = power + ptw;
    S.power = power;
    if (power > 0) {
        S.n_good++;
        S.n_bad = 0;
        if (S.n_good > 5) {
            S.dt = fmin(S.dt * 1.1, S.dt_max);
            S.alpha *= 0.99;
        }
        S.uphill = 0;
    } else {
        S.n_good = 0;
        S.n_bad++;
        if (S.n_bad > 10) {
            S.stopped = 1;
            S.step += 1;
            S.done = 1;
            enter_eval(b, S, blockIdx.x);
            return;
        }
        S.dt = fmax(S.dt * 0.5, S.dt_min);
        S.alpha = S.J.alpha0;
        S.uphill = 1;
    }
}

// after a minimisation: setVelocitiesToZero (minimizeFire's end), then
// calculateStressTensor (kernelCalcDiskStressTensor, disk/kernels.cu:203-251) and
// countContacts (kernelCountDiskContacts, disk/kernels.cu:177-201) per particle, with
// chunk partials of sigma_xx, sigma_yy, r^2 (packing fraction) and the contact count
__device__ __forceinline__ void b_eval_disk(const BDev& b, const BSys& S, const PairParams& p,
                                            long long i, double& sxx, double& syy,
                                            double& z) {
    b.vel[i] = make_double2(0.0, 0.0);
    const double4 qi = b.P[i];
    const double2 v = make_double2(0.0, 0.0);
    const double m = qi.w;
    double axx = m * v.x * v.x, ayy = m * v.y * v.y;
    int c = 0;
    const int nn = min(b.nnum[i], b.cap);
    for (int k = 0; k < nn; ++k) {
        int j = b.nbr[(long long)k * b.nslot + i];
        double4 qj = b.P[j];
        double xd = pbc(qi.x, qj.x, p.Lx, p.invLx, p.halfLx, p.fast);
        double yd = pbc(qi.y, qj.y, p.Ly, p.invLy, p.halfLy, p.fast);
        double fx, fy;
        pair_law(qi.x, qi.y, qi.z, qj.x, qj.y, qj.z, p, fx, fy);
        axx += xd * fx / 2;
        ayy += yd * fy / 2;
        double dist = sqrt(xd * xd + yd * yd);
        if (dist < qi.z + qj.z) c++;
    }
    double area = S.Lx * S.Ly;
    sxx = axx / area;
    syy = ayy / area;
    z = (double)c;
}

// The rigid forms: setVelocitiesToZero (v and omega), kernelCalcRigidBumpyStressTensor
// (rigid_bumpy/kernels.cu:705-789: the vertex kinetic terms, here of zero velocity, and
// the vertex-pair virial) and kernelCountRigidBumpyContacts (rigid_bumpy/kernels.cu:
// 924-997: a neighbour q counts once when a vertex of i touches a vertex of q), the
// arithmetic of k_bumpy_stress and the count of k_bumpy_contacts.
__device__ __forceinline__ void rb_eval(const BDev& b, const BSys& S, const PairParams& p,
                                        long long i, double& sxx, double& syy, double& z) {
    b.vel[i] = make_double2(0.0, 0.0);
    b.omg[i] = 0.0;
    const int st = b.vst[i], nv = b.nvt[i];
    const double rv = S.rv;
    // one pass over the vertex rows for both the stress and the contacts.  The kinetic
    // terms of the vertices are vm * 0 * 0 = +0.0 (zero velocity and spin), which leave
    // the sums unchanged (a sum that starts at +0.0 never becomes -0.0), so they are not
    // added; the pair terms go in the reference's order (vertices, then row entries).
    // A neighbour q of i counts when some vertex of i touches some vertex of q; every
    // vertex-row entry belongs to a particle of i's row (the rows are built from it), so
    // the touching entries mark row positions in a mask and the count is its population.
    double axx = 0.0, ayy = 0.0;
    const double vrs = 2 * rv;
    const int nn = min(b.nnum[i], b.cap);
    unsigned long long seen = 0;  // row positions (the row capacity is at most 64)
    long long lastq = -1;         // touching partners repeat along the ring: looked up once
    for (int a = 0; a < nv; ++a) {
        const int v = st + a;
        const double2 xv = b.vpos[v];
        const int vn = min(b.vnum[v], b.vcap);
        for (int k2 = 0; k2 < vn; ++k2) {
            const int u = b.vnbr[(long long)k2 * b.nvslot + v];
            const int q = b.vpart[u];
            const double2 xu = b.vpos[u];
            const double xd = pbc(xv.x, xu.x, p.Lx, p.invLx, p.halfLx, p.fast);
            const double yd = pbc(xv.y, xu.y, p.Ly, p.invLy, p.halfLy, p.fast);
            if (q != i) {
                double fx, fy;
                pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
                axx += xd * fx / 2;
                ayy += yd * fy / 2;
            }
            const double dist = sqrt(xd * xd + yd * yd);
            if (!(dist < vrs) || q == lastq) continue;
            lastq = q;
            // q's position in the row, four entries per probe (loads issued together)
            for (int k = 0; k < nn; k += 4) {
                int r[4];
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    r[t] = k + t < nn ? b.nbr[(long long)(k + t) * b.nslot + i] : -1;
                int hit = -1;
#pragma unroll
                for (int t = 3; t >= 0; --t)
                    if (r[t] == q) hit = k + t;
                if (hit >= 0) {
                    seen |= 1ull << hit;
                    break;
                }
            }
        }
    }
    const double area = S.Lx * S.Ly;
    sxx = axx / area;
    syy = ayy / area;
    const int nc = __popcll(seen);
    z = (double)nc;
}

// The evaluation of rigid particles at the end of a minimisation (rb_eval's results)
// with a half-wave per particle, lanes over its vertices: each lane walks its vertex row
// (the gathers and pair laws of all the particle's vertices at once) and keeps the row's
// nonzero stress terms, in row order, in LDS; then one lane adds them up in the order of
// the serial walk, vertex after vertex.  Terms that are +-0.0 (pairs not in contact) are
// left out: the running sums start at +0.0 and never become -0.0, so adding them changes
// nothing.  The contact mask is the union of the lanes' masks.  A vertex with more than
// RE_CAP nonzero terms sends its particle down the serial walk (rb_eval).
constexpr int RE_CAP = 8;
constexpr int RE_GRID = 1024;  // workgroups of the launch (at most)
__device__ __forceinline__ void rb_eval_half(const BDev& b, const BSys& S, long long i,
                                             double2 (*sterm)[RE_CAP]) {
    const int lane = (int)(threadIdx.x % RB_HALF);
    const PairParams p = bparams(b, S);
    const int st = b.vst[i], nv = b.nvt[i];
    const double rv = S.rv, vrs = 2 * rv;
    const int nn = min(b.nnum[i], b.cap);
    const int hs = (int)(threadIdx.x % WAVE) / RB_HALF * RB_HALF;  // the half's first lane
    double2* mine = sterm[threadIdx.x];
    double axx = 0.0, ayy = 0.0;        // lane 0's running sums
    unsigned long long seen = 0;         // this lane's touching partners' row positions
    bool spill = false;
    for (int a0 = 0; a0 < nv; a0 += RB_HALF) {
        const int a = a0 + lane;
        int cnt = 0;
        if (a < nv) {
            const int v = st + a;
            const double2 xv = b.vpos[v];
            const int vn = min(b.vnum[v], b.vcap);
            long long lastq = -1;
            for (int k2 = 0; k2 < vn; ++k2) {
                const int u = b.vnbr[(long long)k2 * b.nvslot + v];
                const int q = b.vpart[u];
                const double2 xu = b.vpos[u];
                const double xd = pbc(xv.x, xu.x, p.Lx, p.invLx, p.halfLx, p.fast);
                const double yd = pbc(xv.y, xu.y, p.Ly, p.invLy, p.halfLy, p.fast);
                if (q != i) {
                    double fx, fy;
                    pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
                    const double tx = xd * fx / 2, ty = yd * fy / 2;
                    if (tx != 0.0 || ty != 0.0) {
                        if (cnt < RE_CAP) mine[cnt] = make_double2(tx, ty);
                        cnt++;
                    }
                }
                const double dist = sqrt(xd * xd + yd * yd);
                if (!(dist < vrs) || q == lastq) continue;
                lastq = q;
                for (int k = 0; k < nn; ++k)
                    if (b.nbr[(long long)k * b.nslot + i] == q) {
                        seen |= 1ull << k;
                        break;
                    }
            }
        }
        spill = spill || cnt > RE_CAP;
        // lane 0 adds the half's terms in vertex order (sterm is written by the lanes of
        // this half only; a wave's LDS accesses complete in order)
        __syncwarp();
        for (int l = 0; l < RB_HALF && a0 + l < nv; ++l) {
            const int cl = __shfl(cnt, l, RB_HALF);
            if (lane == 0) {
                const double2* src = sterm[threadIdx.x + l];
                for (int k = 0; k < min(cl, RE_CAP); ++k) {
                    axx += src[k].x;
                    ayy += src[k].y;
                }
            }
        }
        __syncwarp();  // the next vertex group rewrites sterm
    }
    const bool any_spill = (__builtin_amdgcn_ballot_w64(spill) >> hs) & 0xffffffffull;
    // the union of the lanes' contact masks
#pragma unroll
    for (int o = RB_HALF / 2; o > 0; o >>= 1) {
        const unsigned lo = __shfl_xor((unsigned)seen, o, RB_HALF);
        const unsigned hi = __shfl_xor((unsigned)(seen >> 32), o, RB_HALF);
        seen |= ((unsigned long long)hi << 32) | lo;
    }
    if (lane != 0) return;
    double sxx, syy, z;
    if (any_spill) {
        rb_eval(b, S, p, i, sxx, syy, z);
    } else {
        const double area = S.Lx * S.Ly;
        sxx = axx / area;
        syy = ayy / area;
        z = (double)__popcll(seen);
    }
    b.evs[i] = make_double2(sxx, syy);
    b.evz[i] = (int)z;
}

// the systems k_b_ctl listed in this step, a half-wave per particle over the grid (a step
// in which no rigid system finishes a minimisation returns at once)
__global__ __launch_bounds__(BLOCK) void k_rb_eval(BDev b) {
    __shared__ double2 sterm[BLOCK][RE_CAP];
    const int count = b.flags[9];
    const int halves = BLOCK / RB_HALF;
    const long long h0 = (long long)blockIdx.x * halves + threadIdx.x / RB_HALF;
    for (int e = 0; e < count; ++e) {
        const BSys& S = b.sys[b.evlist[e]];
        for (long long h = h0; h < S.n; h += (long long)gridDim.x * halves)
            rb_eval_half(b, S, S.off + h, sterm);  // uniform over the half-wave
    }
}

__global__ __launch_bounds__(BLOCK) void k_b_eval(BDev b) {
    const long long i = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (i == 0) b.flags[9] = 0;  // k_rb_eval has read this step's list
    if (i >= b.nslot) return;
    const int s = b.sys_of[i];
    if (s < 0) return;  // free arena slots (64-aligned: whole waves)
    const BSys& S = b.sys[s];
    if (S.mode != BM_EVAL) return;
    const PairParams p = bparams(b, S);
    const bool live = i - S.off < S.n;
    double sxx = 0.0, syy = 0.0, r2 = 0.0, z = 0.0, d2 = 0.0;
    if (live) {
        if (S.kind == BK_RIGID) {  // k_rb_eval's results
            b.vel[i] = make_double2(0.0, 0.0);
            b.omg[i] = 0.0;
            const double2 ev = b.evs[i];
            sxx = ev.x;
            syy = ev.y;
            z = (double)b.evz[i];
        } else {
            b_eval_disk(b, S, p, i, sxx, syy, z);
        }
        const double4 q = b.P[i];
        r2 = q.z * q.z;
        // the closing displacement maxima of dpm_minimize_fire (k_disp_partials), which
        // the next minimisation's prologue tests: 3 max d^2 > thr holds exactly when
        // 3 d^2 > thr holds for some slot (3x is monotone under rounding), so one ballot
        // per wave replaces the wave maximum
        const double2 l = b.lref[i];
        const double dx = q.x - l.x, dy = q.y - l.y;
        d2 = dx * dx + dy * dy;
    }
    const bool far = __builtin_amdgcn_ballot_w64(3.0 * d2 > S.thr2) != 0;
    if ((threadIdx.x % WAVE) == 0) bpart(b, BP_DMAX)[i / WAVE] = far ? 1.0 : 0.0;
    double s0 = wave_tree_sum(sxx);
    double s1 = wave_tree_sum(syy);
    double s2 = wave_tree_sum(r2);
    double s3 = wave_tree_sum(z);
    if ((threadIdx.x % WAVE) == 0) {
        long long ch = i / WAVE;
        bpart(b, BP_SXX)[ch] = s0;
        bpart(b, BP_SYY)[ch] = s1;
        bpart(b, BP_R2)[ch] = s2;
        bpart(b, BP_Z)[ch] = s3;
    }
}

__device__ __forceinline__ void start_minimization(const BDev& b, BSys& S) {
    fire_reset(b, S);
    S.mode = BM_PROLOGUE;
}

// The jamming controller of dpm_jam after a minimisation, for lane 0 of the system's
// wave; returns whether the system has finished.
__device__ __forceinline__ bool b_jam_decide(const BDev& b, BSys& S, double sxx, double syy,
                                             double z) {
    S.fire_total += min(S.step, S.J.max_steps);
    const double pr = (sxx + syy) / 2;
    const double P_low = S.J.P_low, P_high = S.J.P_high;
    bool finished = false;
    if (S.jphase == BJ_MIN) {  // minimizeFire alone: its end state is the result
        S.pr = pr;
        S.contacts = (long long)z;
        S.status = S.converged ? 0 : 1;
        finished = true;
    } else if (S.jphase == BJ_DECOMP) {
        // for (it < 1000) { minimize; if (P < P_low) break; scaleBox(f); }
        bool leave = pr < P_low;
        int op = 1;
        if (!leave) {
            scale_box_of(S, S.J.dfac);
            S.dec_it += 1;
            leave = S.dec_it >= 1000;
            op = 3;  // the snapshot below is taken after this scaling
        }
        if (leave) {  // L_* = -1, setLastState, then the compression loop
            S.jphase = BJ_COMP;
            S.L_low = S.L_high = S.L_last = -1;
            S.cstep = 0;
            S.status = 1;
            S.pr = 0.0;
            S.contacts = 0;
            S.snap_op = op;
            S.snapLx = S.Lx;
            S.snapLy = S.Ly;
            if (S.cstep >= S.J.max_csteps) finished = true;
            else start_minimization(b, S);
        } else {
            start_minimization(b, S);
        }
    } else {
        S.pr = pr;
        S.contacts = (long long)z;
        double L = sqrt(S.Lx * S.Ly);
        double L_prior = L;
        bool stop = false;
        int restore = 0;
        if (pr < P_low) {
            S.L_low = L;
            S.snap_op = 1;
            S.snapLx = S.Lx;
            S.snapLy = S.Ly;
            S.L_last = L;
            if (S.L_high != -1) {
                L = (S.L_high + S.L_low) / 2;
                S.L_high = -1;
            } else {
                L = L * S.J.rate;
            }
        } else {
            if (pr > P_high) {
                S.L_high = L;
                restore = 1;
                L_prior = S.L_last;
                L = (S.L_high + S.L_low) / 2;
            } else if (z > S.J.zmin * (double)S.n) {
                S.status = 0;
                stop = true;
            }
        }
        if (!stop && S.L_low > 0 && S.L_high > 0 && fabs(S.L_low / S.L_high - 1) < S.J.ltol) {
            S.status = 2;
            stop = true;
        }
        if (restore) {  // revertToLastState: positions (b_prep_slot) and box
            S.snap_op = 2;
            S.Lx = S.snapLx;
            S.Ly = S.snapLy;
        }
        if (stop) {
            finished = true;
        } else if (restore && S.cstep + 1 < S.J.max_csteps) {
            // dpm_snapshot_restore's full rebuild at the restored state runs in a step of
            // its own (BM_RESTORE); k_b_ctl then scales the box
            S.cstep += 1;
            start_minimization(b, S);
            S.mode = BM_RESTORE;
            S.pend_f = L / L_prior;
        } else {
            scale_box_of(S, L / L_prior);
            S.cstep += 1;
            if (S.cstep >= S.J.max_csteps) finished = true;
            else start_minimization(b, S);
        }
    }
    return finished;
}

// one wave per system: the jamming controller of dpm_jam after a minimisation.  A
// system that finishes gets the controller's last snapshot / scaling operations at once
// (its slots see no further step).
__global__ __launch_bounds__(WAVE) void k_b_jam(BDev b) {
    BSys& S = b.sys[blockIdx.x];
    if (S.mode != BM_EVAL) return;
    const long long c0 = S.off / WAVE;
    double sxx, syy, pe, r2;
    wave_final_sum4(bpart(b, BP_SXX) + c0, bpart(b, BP_SYY) + c0, bpart(b, BP_PE) + c0,
                    bpart(b, BP_R2) + c0, S.nc, sxx, syy, pe, r2);
    double z = 0.0;
    for (int k = threadIdx.x; k < S.nc; k += WAVE) z += bpart(b, BP_Z)[c0 + k];
    z = wave_tree_sum(z);  // integer counts: exact in any order
    __shared__ int apply;
    if (threadIdx.x == 0) {
        const bool finished = b_jam_decide(b, S, sxx, syy, z);
        apply = finished && (S.snap_op != 0 || S.scale_f != 0.0);
        if (finished) {
            S.done_step = b.flags[3];
            S.pe_final = pe;
            const double PI = 3.141592653589793238462643383279502884197;
            S.phi_final = r2 * PI / (S.Lx * S.Ly);
            S.mode = BM_DONE;
        }
    }
    __syncthreads();
    if (apply) {
        const bool rigid = S.kind == BK_RIGID;
        if (rigid)  // vertices first: they read the centres before the particle pass
            for (long long v = S.voff + threadIdx.x; v < S.voff + S.nvs; v += WAVE) {
                const int p = b.vpart[v];
                if (p >= 0) rb_vertex_ops(b, S, v, p);
            }
        __syncthreads();
        for (long long i = S.off + threadIdx.x; i < S.off + S.n; i += WAVE)
            b_slot_ops(b, S, i, rigid);
        __syncthreads();
    }
    if (threadIdx.x == 0 && S.mode == BM_DONE) {
        S.snap_op = 0;
        S.scale_f = 0.0;
        int k = atomicAdd(&b.flags[2], 1);
        b.done_list[k % b.nsys] = (int)blockIdx.x;
    }
}

}  // namespace dpm

// =========================================================================
//                               host side
// =========================================================================
struct dpm_batch {
    int64_t nsys = 0, nslot = 0, ntiles = 0;
    int nct = 0;  // cell-grid build workgroups of a batch (ctiles entries)
    int device = 0;
    hipStream_t stream = nullptr;
    dpm::BDev d{};
    std::vector<void*> allocs;
    std::vector<long long> off;
    std::vector<int> n;
    std::vector<dpm::BSys> hsys;
    std::vector<double> rmax, rmin, box0;  // per system: radii range, box side as set
    // rigid bumpy systems: vertices (0: a disk system), vertex slots, first vertex slot,
    // vertex radius, contact energy, rotation; nvslot = vertex slots of the arena
    int64_t nvslot = 0;
    std::vector<int64_t> nv, nvs, voff;
    std::vector<double> rv, ec;
    std::vector<int> rot;
    // the state as set (autotuning restores it): positions, rigid angles and vertices
    double4* P_init = nullptr;
    double* A_init = nullptr;
    double2* V_init = nullptr;
    double rc = 1.7, thr_sq = 3.0 * 0.15 * 0.15, e_c = 1.0, n_c = 2.0;
    int64_t steps_per_graph = 32;
    int force_block = 256;  // k_b_force workgroup size (64, 128, 256)
    int cell_pct = 100;     // cell-grid build: cell width in % of the cutoff (>= 100)
    bool disks = true;      // the arena can hold disk systems (the disk force pass runs)
    bool use_graph = true;
    bool ran = false;  // a dpm_batch_jam / dpm_batch_minimize has started
    int* hflags = nullptr;  // pinned: [0] overflow, [1] largest row, [2] systems done
    hipEvent_t poll[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t graph_steps = 0;
    // DPM_OPT_TIMING: HIP events around every k_b_force launch (eager launches)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    std::vector<hipEvent_t> ev_free, ev_all;
    double force_ms = 0.0;
    int64_t force_launches = 0;
};

namespace dpm {

// The cell-grid build workgroups of a system: one per BC_PER particles when the system
// is in the cell-grid size range (k_b_build_cells re-checks the grid for the box of the
// step and leaves small boxes to k_b_build).
static void cell_tiles(std::vector<int2>& out, int s, int n) {
    if (n < BC_MIN_N || n > BC_MAX_N) return;
    for (int part = 0; part < (n + BC_PER - 1) / BC_PER; ++part) out.push_back(make_int2(s, part));
}

// Copies and fills are ordered on the batch's own stream (never the legacy stream, which
// would tie them to graph captures of other threads' systems).
static hipError_t bsync(dpm_batch* b, hipError_t e) {
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(b->stream);
}
#define BCOPY(dst, src, bytes, kind) \
    DPM_HIP(bsync(b, hipMemcpyAsync((dst), (src), (bytes), (kind), b->stream)))

static int balloc(dpm_batch* b, void** p, size_t bytes) {
    void* q = nullptr;
    {
        std::lock_guard<std::mutex> g(capture_mutex());
        DPM_HIP(hipMalloc(&q, std::max<size_t>(bytes, 64)));
    }
    DPM_HIP(bsync(b, hipMemsetAsync(q, 0, std::max<size_t>(bytes, 64), b->stream)));
    b->allocs.push_back(q);
    *p = q;
    return 0;
}

template <class T>
static int balloc_n(dpm_batch* b, T** p, int64_t n) {
    void* q = nullptr;
    if (balloc(b, &q, sizeof(T) * (size_t)std::max<int64_t>(n, 1))) return -1;
    *p = (T*)q;
    return 0;
}

// Stream, arena and tables for up to nsys systems in nslot slots and nvslot vertex slots
// (sys_of, vsys_of, vpart = -1: free).
static int batch_alloc(dpm_batch* b, int64_t nsys, int64_t nslot, int64_t tiles_cap,
                       int64_t nvslot) {
    (void)hipGetDevice(&b->device);
    BDev& d = b->d;
    d.nsys = nsys;
    d.nslot = nslot;
    d.ntiles = tiles_cap;
    // parts = ceil(n / BC_PER) <= n / BC_PER + n / BC_MIN_N for n >= BC_MIN_N, and the
    // systems share the nslot slots
    d.nctiles = nslot / BC_PER + nslot / BC_MIN_N + 1;
    d.cap = 32;  // at most 32: k_rb_vbuild holds a row entry per half-wave lane (and
                 // rb_eval marks row positions in a 64-bit mask)
    DPM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    if (balloc_n(b, &d.sys, nsys) || balloc_n(b, &d.sys_of, nslot) || balloc_n(b, &d.P, nslot) ||
        balloc_n(b, &d.snap, nslot) || balloc_n(b, &d.vel, nslot) || balloc_n(b, &d.frc, nslot) ||
        balloc_n(b, &d.lref, nslot) || balloc_n(b, &d.pe, nslot) || balloc_n(b, &d.nnum, nslot) ||
        balloc_n(b, &d.nbr, nslot * d.cap) || balloc_n(b, &d.part, (nslot / WAVE) * BP_COUNT) ||
        balloc_n(b, &d.tiles, tiles_cap) || balloc_n(b, &d.ctiles, d.nctiles) ||
        balloc_n(b, &d.flags, 10) || balloc_n(b, &d.evlist, nsys) ||
        balloc_n(b, &d.done_list, nsys))
        return -1;
    DPM_HIP(bsync(b, hipMemsetAsync(d.sys_of, 0xff, sizeof(int) * nslot, b->stream)));  // -1: free
    d.nvslot = nvslot;
    d.vcap = 32;
    b->nvslot = nvslot;
    if (nvslot > 0) {
        if (balloc_n(b, &d.vsys_of, nvslot) || balloc_n(b, &d.vpart, nvslot) ||
            balloc_n(b, &d.vnum, nvslot) || balloc_n(b, &d.vnbr, nvslot * d.vcap) ||
            balloc_n(b, &d.vpos, nvslot) || balloc_n(b, &d.vsnap, nvslot) ||
            balloc_n(b, &d.vst, nslot) ||
            balloc_n(b, &d.nvt, nslot) || balloc_n(b, &d.ang, nslot) ||
            balloc_n(b, &d.asnap, nslot) || balloc_n(b, &d.omg, nslot) ||
            balloc_n(b, &d.tq, nslot) || balloc_n(b, &d.moi, nslot) ||
            balloc_n(b, &d.lpos, nslot) || balloc_n(b, &d.mpos, nslot) ||
            balloc_n(b, &d.rot0, nslot) || balloc_n(b, &d.rot1, nslot) ||
            balloc_n(b, &d.evs, nslot) || balloc_n(b, &d.evz, nslot))
            return -1;
        DPM_HIP(bsync(b, hipMemsetAsync(d.vsys_of, 0xff, sizeof(int) * nvslot, b->stream)));
        DPM_HIP(bsync(b, hipMemsetAsync(d.vpart, 0xff, sizeof(int) * nvslot, b->stream)));
    }
    DPM_HIP(hipHostMalloc((void**)&b->hflags, 2 * 4 * sizeof(int), hipHostMallocDefault));
    for (int k = 0; k < 2; ++k)
        DPM_HIP(hipEventCreateWithFlags(&b->poll[k], hipEventDisableTiming));
    return 0;
}

// The numbers every system shares, baked into the kernels' arguments.
static void batch_common(dpm_batch* b, bool fastdiv) {
    BDev& d = b->d;
    d.n_c = b->n_c;
    d.ncm1 = b->n_c - 1.0;
    d.two_nc = 2.0 * b->n_c;
    int ex;
    double m = frexp(d.two_nc, &ex);
    d.inv2nc = (m == 0.5) ? 1.0 / d.two_nc : 0.0;
    d.fastdiv = fastdiv ? 1 : 0;
    d.cell_scale = b->cell_pct / 100.0;
}

// The short correctly rounded sqrt/division sequences apply (fastdiv_ok of dpm_system):
// n_c == 2, e_c and every contact radius (a rigid system's: the vertex radius) within
// [2^-100, 2^100].
static bool fastdiv_range(double n_c, double e_c, double rlo, double rhi) {
    const double lo = 0x1p-100, hi = 0x1p100;
    return n_c == 2.0 && e_c >= lo && e_c <= hi && rlo >= lo && rhi <= hi;
}

// A system's constants.  Disks: the batch's cutoff, threshold and contact energy.  Rigid
// bumpy systems: the rules of RigidBumpy::setNeighborSize (rigid_bumpy.cu:1278-1300)
// with the standard multipliers -- list cutoff 1.5 d_max, threshold (0.2 r_v)^2, vertex
// cutoff 1.5 (2 r_v), vertex-particle cutoff d_max -- and the half-box limit over every
// contact and cutoff (fast_ok of dpm_system).
static void sys_constants(const dpm_batch* b, int64_t s, double rmax, double rv, double e_c,
                          int rotation, int64_t nv, int64_t nvs, long long voff, BSys& f) {
    if (nv > 0) {
        const double dmax = 2.0 * rmax;
        f.kind = BK_RIGID;
        f.rc = 1.5 * dmax;
        f.thr2 = (0.2 * rv) * (0.2 * rv);
        const double vc = 1.5 * 2.0 * rv;
        f.vc2 = vc * vc;
        f.vp2 = dmax * dmax;
        f.lim = std::max({2.0 * rmax, f.rc, 2.0 * rv, vc, dmax});
        f.rv = rv;
        f.e_c = e_c;
        f.rotation = rotation;
        f.nv = (int)nv;
        f.nvs = (int)nvs;
        f.voff = voff;
    } else {
        f.kind = BK_DISK;
        f.rc = b->rc;
        f.thr2 = b->thr_sq;
        f.lim = std::max(2.0 * rmax, b->rc);
        f.e_c = b->e_c;
    }
    f.rc2 = f.rc * f.rc;
    (void)s;
}

static BJam jam_constants(const dpm_jam_params* P) {
    BJam J{};
    J.alpha0 = P->fire.alpha_init;
    J.dt0 = P->fire.dt;
    J.pe_t = P->fire.avg_pe_target;
    J.pe_diff_t = P->fire.avg_pe_diff_target;
    J.fmax2 = P->fire.fmax_target > 0 ? P->fire.fmax_target * P->fire.fmax_target : 0.0;
    J.max_steps = P->fire.max_steps;
    J.P_low = P->pressure_target;
    J.P_high = P->pressure_target * (1.0 + P->window_rel);
    J.rate = 1.0 - P->strain_step;
    J.dfac = P->decompress_factor;
    J.ltol = P->length_tol;
    J.zmin = P->min_contacts_per_particle;
    J.max_csteps = P->max_compression_steps;
    return J;
}

static hipEvent_t batch_event(dpm_batch* b) {
    if (b->ev_free.empty()) {
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        b->ev_all.push_back(e);
        return e;
    }
    hipEvent_t e = b->ev_free.back();
    b->ev_free.pop_back();
    return e;
}

// Adds the elapsed time of every recorded k_b_force launch (waits for the stream).
static void batch_collect_timing(dpm_batch* b) {
    if (b->ev_pending.empty()) return;
    (void)hipStreamSynchronize(b->stream);
    for (auto& pr : b->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            b->force_ms += ms;
            b->force_launches += 1;
        }
        b->ev_free.push_back(pr.first);
        b->ev_free.push_back(pr.second);
    }
    b->ev_pending.clear();
}

static void enqueue_batch_step(dpm_batch* b) {
    const BDev& d = b->d;
    hipStream_t st = b->stream;
    dim3 slots((unsigned)((b->nslot + BLOCK - 1) / BLOCK));
    const bool rigid = b->nvslot > 0;
    dim3 vslots((unsigned)((b->nvslot + BLOCK - 1) / BLOCK));
    hipLaunchKernelGGL(k_b_prep_move, slots, dim3(BLOCK), 0, st, d);
    if (rigid) hipLaunchKernelGGL(k_rb_vmove, vslots, dim3(BLOCK), 0, st, d);
    hipLaunchKernelGGL(k_b_build, dim3((unsigned)b->ntiles), dim3(BLOCK), 0, st, d);
    hipLaunchKernelGGL(k_b_build_cells, dim3((unsigned)std::min<long long>(d.nctiles, BC_GRID)),
                       dim3(BC_BLOCK), 0, st, d);
    if (rigid)
        hipLaunchKernelGGL(k_rb_vbuild, dim3((unsigned)(b->nslot * RB_HALF / BLOCK)), dim3(BLOCK),
                           0, st, d);
    if (b->disks) {  // a batch of rigid systems only leaves the disk force pass out
        hipEvent_t ea = nullptr, eb = nullptr;
        if (b->timing) {
            ea = batch_event(b);
            eb = batch_event(b);
            (void)hipEventRecord(ea, st);
        }
        const int fb = b->force_block;
        dim3 fslots((unsigned)((b->nslot + fb - 1) / fb));
        if (d.fastdiv)
            hipLaunchKernelGGL(k_b_force<true>, fslots, dim3(fb), 0, st, d);
        else
            hipLaunchKernelGGL(k_b_force<false>, fslots, dim3(fb), 0, st, d);
        if (b->timing) {
            (void)hipEventRecord(eb, st);
            b->ev_pending.push_back({ea, eb});
        }
    }
    if (rigid) {
        // DPM_OPT_TIMING times this pass instead when the batch holds rigid systems only
        hipEvent_t ra = nullptr, rb = nullptr;
        if (b->timing && !b->disks) {
            ra = batch_event(b);
            rb = batch_event(b);
            (void)hipEventRecord(ra, st);
        }
        dim3 chunks((unsigned)(b->nslot / WAVE));
        if (d.fastdiv)
            hipLaunchKernelGGL(k_rb_force<true>, chunks, dim3(RB_FT), 0, st, d);
        else
            hipLaunchKernelGGL(k_rb_force<false>, chunks, dim3(RB_FT), 0, st, d);
        if (ra) {
            (void)hipEventRecord(rb, st);
            b->ev_pending.push_back({ra, rb});
        }
    }
    hipLaunchKernelGGL(k_b_ctl, dim3((unsigned)b->nsys), dim3(WAVE), 0, st, d);
    if (rigid)
        hipLaunchKernelGGL(k_rb_eval,
                           dim3((unsigned)std::min<long long>(b->nslot * RB_HALF / BLOCK,
                                                              RE_GRID)),
                           dim3(BLOCK), 0, st, d);
    hipLaunchKernelGGL(k_b_eval, slots, dim3(BLOCK), 0, st, d);
    hipLaunchKernelGGL(k_b_jam, dim3((unsigned)b->nsys), dim3(WAVE), 0, st, d);
}

static void drop_graph(dpm_batch* b) {
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
    }
}

static int ensure_batch_graph(dpm_batch* b, int64_t k) {
    if (b->exec && b->graph_steps == k) return 0;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
    }
    std::lock_guard<std::mutex> lk(capture_mutex());
    DPM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int64_t s = 0; s < k; ++s) enqueue_batch_step(b);
    DPM_HIP(hipStreamEndCapture(b->stream, &b->graph));
    DPM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_steps = k;
    return 0;
}

}  // namespace dpm

using namespace dpm;

extern "C" {

}  // extern "C"

namespace dpm {

// n_vertices == nullptr: every system is a disk system.
static dpm_batch* batch_create(int64_t n_systems, const int64_t* n_particles,
                               const int64_t* n_vertices) {
    if (n_systems <= 0 || n_systems >= (int64_t)1 << 24 || !n_particles) {
        set_error("dpm_batch_create: bad arguments");
        return nullptr;
    }
    if (dpm_device_count() <= 0) {
        set_error("dpm_batch_create: no HIP device");
        return nullptr;
    }
    auto* b = new dpm_batch();
    b->nsys = n_systems;
    b->off.resize(n_systems);
    b->n.resize(n_systems);
    b->rmax.assign(n_systems, 0.0);
    b->rmin.assign(n_systems, 0.0);
    b->box0.assign(n_systems, 0.0);
    b->nv.assign(n_systems, 0);
    b->nvs.assign(n_systems, 0);
    b->voff.assign(n_systems, 0);
    b->rv.assign(n_systems, 0.0);
    b->ec.assign(n_systems, 0.0);
    b->rot.assign(n_systems, 1);
    long long vacc = 0;
    for (int64_t s = 0; s < n_systems && n_vertices; ++s) {
        int64_t v = n_vertices[s];
        if (v < 0 || v >= (int64_t)1 << 30 || (v > 0 && v < n_particles[s])) {
            set_error("dpm_batch_create: a rigid system needs a vertex per particle at least");
            delete b;
            return nullptr;
        }
        b->nv[s] = v;
        b->nvs[s] = (v + WAVE - 1) / WAVE * WAVE;
        b->voff[s] = vacc;
        vacc += b->nvs[s];
    }
    b->disks = !n_vertices ||
               std::any_of(b->nv.begin(), b->nv.end(), [](int64_t v) { return v == 0; });
    if (vacc >= (int64_t)1 << 31) {
        set_error("dpm_batch_create: vertex arena exceeds 2^31 slots");
        delete b;
        return nullptr;
    }
    long long acc = 0;
    std::vector<int2> tiles;
    for (int64_t s = 0; s < n_systems; ++s) {
        int64_t ns = n_particles[s];
        if (ns <= 0 || ns >= (int64_t)1 << 24) {
            set_error("dpm_batch_create: every system needs 1 .. 2^24 particles");
            delete b;
            return nullptr;
        }
        b->off[s] = acc;
        b->n[s] = (int)ns;
        for (int64_t t = 0; t < ns; t += BLOCK) tiles.push_back(make_int2((int)s, (int)(acc + t)));
        acc += (ns + WAVE - 1) / WAVE * WAVE;
    }
    if (acc >= (int64_t)1 << 31) {
        set_error("dpm_batch_create: arena exceeds 2^31 slots");
        delete b;
        return nullptr;
    }
    b->nslot = acc;
    b->ntiles = (int64_t)tiles.size();
    bool ok = batch_alloc(b, n_systems, acc, b->ntiles, vacc) == 0;
    BDev& d = b->d;
    if (ok) {
        std::vector<int> so(acc);
        for (int64_t s = 0; s < n_systems; ++s) {
            long long end = s + 1 < n_systems ? b->off[s + 1] : acc;
            for (long long k = b->off[s]; k < end; ++k) so[k] = (int)s;
        }
        ok = bsync(b, hipMemcpyAsync(d.sys_of, so.data(), sizeof(int) * acc,
                                     hipMemcpyHostToDevice, b->stream)) == hipSuccess &&
             bsync(b, hipMemcpyAsync(d.tiles, tiles.data(), sizeof(int2) * tiles.size(),
                                     hipMemcpyHostToDevice, b->stream)) == hipSuccess;
        if (ok && vacc > 0) {
            std::vector<int> vso(vacc);
            for (int64_t s = 0; s < n_systems; ++s)
                for (long long k = 0; k < b->nvs[s]; ++k) vso[b->voff[s] + k] = (int)s;
            ok = bsync(b, hipMemcpyAsync(d.vsys_of, vso.data(), sizeof(int) * vacc,
                                         hipMemcpyHostToDevice, b->stream)) == hipSuccess;
        }
        std::vector<int2> ct;
        for (int64_t s = 0; s < n_systems; ++s) cell_tiles(ct, (int)s, b->n[s]);
        b->nct = (int)ct.size();
        ok = ok && (long long)ct.size() <= d.nctiles;  // always holds (bound of batch_alloc)
        if (ok && b->nct)
            ok = bsync(b, hipMemcpyAsync(d.ctiles, ct.data(), sizeof(int2) * ct.size(),
                                         hipMemcpyHostToDevice, b->stream)) == hipSuccess;
    }
    b->hsys.assign(n_systems, BSys{});
    for (int64_t s = 0; s < n_systems; ++s) {
        b->hsys[s].off = b->off[s];
        b->hsys[s].n = b->n[s];
        b->hsys[s].nc = (b->n[s] + WAVE - 1) / WAVE;
        b->hsys[s].Lx = b->hsys[s].Ly = 1.0;
    }
    if (!ok) {
        set_error("dpm_batch_create: device allocation failed");
        dpm_batch_destroy(b);
        return nullptr;
    }
    return b;
}

}  // namespace dpm

extern "C" {

dpm_batch* dpm_batch_create(int64_t n_systems, const int64_t* n_particles) {
    return batch_create(n_systems, n_particles, nullptr);
}

dpm_batch* dpm_batch_create_mixed(int64_t n_systems, const int64_t* n_particles,
                                  const int64_t* n_vertices) {
    if (!n_vertices) {
        set_error("dpm_batch_create_mixed: null vertex counts");
        return nullptr;
    }
    return batch_create(n_systems, n_particles, n_vertices);
}

void dpm_batch_destroy(dpm_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    {
        std::lock_guard<std::mutex> g(capture_mutex());
        for (void* p : b->allocs) (void)hipFree(p);
    }
    if (b->hflags) (void)hipHostFree(b->hflags);
    for (auto e : b->poll)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b->ev_all) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int dpm_batch_set_system(dpm_batch* b, int64_t s, const double* x, const double* y,
                         const double* rad, const double* mass, double box_side) {
    if (s < 0 || s >= b->nsys || !(box_side > 0)) {
        set_error("dpm_batch_set_system: bad system index or box side");
        return -1;
    }
    if (b->nv[s] > 0) {
        set_error("dpm_batch_set_system: system " + std::to_string(s) +
                  " is a rigid bumpy system (dpm_batch_set_rigid_system)");
        return -1;
    }
    int ns = b->n[s];
    std::vector<double4> h(ns);
    double rm = 0.0, rn = 1e300;
    for (int k = 0; k < ns; ++k) {
        h[k] = make_double4(x[k], y[k], rad[k], mass[k]);
        rm = std::max(rm, rad[k]);
        rn = std::min(rn, rad[k]);
        if (!(rad[k] > 0) || !(mass[k] > 0)) {
            set_error("dpm_batch_set_system: radii and masses must be positive");
            return -1;
        }
    }
    BCOPY(b->d.P + b->off[s], h.data(), sizeof(double4) * ns, hipMemcpyHostToDevice);
    b->rmax[s] = rm;
    b->rmin[s] = rn;
    b->box0[s] = box_side;
    b->hsys[s].Lx = b->hsys[s].Ly = box_side;
    b->hsys[s].rmax = rm;
    return 0;
}

int dpm_batch_get_system(dpm_batch* b, int64_t s, double* x, double* y, double* box_side) {
    if (s < 0 || s >= b->nsys) {
        set_error("dpm_batch_get_system: bad system index");
        return -1;
    }
    DPM_HIP(hipStreamSynchronize(b->stream));
    int ns = b->n[s];
    std::vector<double4> h(ns);
    BCOPY(h.data(), b->d.P + b->off[s], sizeof(double4) * ns, hipMemcpyDeviceToHost);
    for (int k = 0; k < ns; ++k) {
        if (x) x[k] = h[k].x;
        if (y) y[k] = h[k].y;
    }
    if (box_side) {  // the box as set until a run has started
        BSys S;
        BCOPY(&S, b->d.sys + s, sizeof(BSys), hipMemcpyDeviceToHost);
        *box_side = b->ran ? S.Lx : b->box0[s];
    }
    return 0;
}

int dpm_batch_set_rigid_system(dpm_batch* b, int64_t s, const double* x, const double* y,
                               const double* rad, const double* mass, const double* angle,
                               const int64_t* vertices_per_particle, double vertex_radius,
                               double e_c, int32_t rotation, double box_side) {
    if (s < 0 || s >= b->nsys || !(box_side > 0)) {
        set_error("dpm_batch_set_rigid_system: bad system index or box side");
        return -1;
    }
    if (b->nv[s] == 0) {
        set_error("dpm_batch_set_rigid_system: system " + std::to_string(s) +
                  " was created without vertices");
        return -1;
    }
    if (!x || !y || !rad || !mass || !angle || !vertices_per_particle ||
        !(vertex_radius >= 0x1p-100 && vertex_radius <= 0x1p100) || !(e_c > 0)) {
        set_error("dpm_batch_set_rigid_system: bad arguments");
        return -1;
    }
    const int ns = b->n[s];
    const long long off = b->off[s], voff = b->voff[s];
    std::vector<double4> h(ns);
    std::vector<double> a(ns);
    std::vector<int> vst(ns), nvt(ns), vpart(b->nvs[s], -1);
    double rm = 0.0;
    long long acc = 0;
    for (int k = 0; k < ns; ++k) {
        if (!(rad[k] > vertex_radius) || !(mass[k] > 0) || vertices_per_particle[k] <= 0) {
            set_error("dpm_batch_set_rigid_system: radii must exceed the vertex radius, "
                      "masses be positive, and every particle have a vertex");
            return -1;
        }
        h[k] = make_double4(x[k], y[k], rad[k], mass[k]);
        a[k] = angle[k];
        rm = std::max(rm, rad[k]);
        vst[k] = (int)(voff + acc);
        nvt[k] = (int)vertices_per_particle[k];
        for (int64_t u = 0; u < vertices_per_particle[k] && acc + u < b->nv[s]; ++u)
            vpart[acc + u] = (int)(off + k);
        acc += vertices_per_particle[k];
    }
    if (acc != b->nv[s]) {
        set_error("dpm_batch_set_rigid_system: vertex counts do not sum to the system's "
                  "vertices");
        return -1;
    }
    BDev& d = b->d;
    BCOPY(d.P + off, h.data(), sizeof(double4) * ns, hipMemcpyHostToDevice);
    BCOPY(d.ang + off, a.data(), sizeof(double) * ns, hipMemcpyHostToDevice);
    BCOPY(d.vst + off, vst.data(), sizeof(int) * ns, hipMemcpyHostToDevice);
    BCOPY(d.nvt + off, nvt.data(), sizeof(int) * ns, hipMemcpyHostToDevice);
    BCOPY(d.vpart + voff, vpart.data(), sizeof(int) * vpart.size(), hipMemcpyHostToDevice);
    hipLaunchKernelGGL(k_rb_init, pgrid(ns), dim3(BLOCK), 0, b->stream, d, off, ns,
                       vertex_radius);
    if (check_launch()) return -1;
    DPM_HIP(hipStreamSynchronize(b->stream));
    b->rmax[s] = rm;
    b->rmin[s] = rm;  // marks the system as set
    b->rv[s] = vertex_radius;
    b->ec[s] = e_c;
    b->rot[s] = rotation != 0;
    b->box0[s] = box_side;
    b->hsys[s].Lx = b->hsys[s].Ly = box_side;
    b->hsys[s].rmax = rm;
    return 0;
}

int dpm_batch_get_rigid_system(dpm_batch* b, int64_t s, double* x, double* y, double* angle,
                               double* vx, double* vy, double* box_side) {
    if (s < 0 || s >= b->nsys || b->nv[s] == 0) {
        set_error("dpm_batch_get_rigid_system: bad system index or not a rigid system");
        return -1;
    }
    if (dpm_batch_get_system(b, s, x, y, box_side)) return -1;
    const int ns = b->n[s];
    if (angle) BCOPY(angle, b->d.ang + b->off[s], sizeof(double) * ns, hipMemcpyDeviceToHost);
    if (vx || vy) {
        std::vector<double2> h(b->nv[s]);
        BCOPY(h.data(), b->d.vpos + b->voff[s], sizeof(double2) * h.size(), hipMemcpyDeviceToHost);
        for (size_t k = 0; k < h.size(); ++k) {
            if (vx) vx[k] = h[k].x;
            if (vy) vy[k] = h[k].y;
        }
    }
    return 0;
}

int dpm_batch_config(dpm_batch* b, double e_c, double n_c, double cutoff, double disp_thresh_sq) {
    if (!(e_c > 0) || !(n_c > 0) || !(cutoff > 0) || !(disp_thresh_sq > 0)) {
        set_error("dpm_batch_config: parameters must be positive");
        return -1;
    }
    b->e_c = e_c;
    b->n_c = n_c;
    b->rc = cutoff;
    b->thr_sq = disp_thresh_sq;
    return 0;
}

int dpm_batch_set_option(dpm_batch* b, int option, int64_t value) {
    if (option == DPM_OPT_GRAPH) {
        b->use_graph = value != 0;
        return 0;
    }
    if (option == DPM_OPT_CHUNK) {
        if (value < 1 || value > 1024) {
            set_error("dpm_batch_set_option: steps per launch must be in [1, 1024]");
            return -1;
        }
        b->steps_per_graph = value;
        return 0;
    }
    if (option == DPM_OPT_FORCE_BLOCK || option == DPM_OPT_CELL_PCT) {
        const bool blk = option == DPM_OPT_FORCE_BLOCK;
        if (blk ? !(value == 64 || value == 128 || value == 256) : !(value >= 100 && value <= 400)) {
            set_error(blk ? "dpm_batch_set_option: force block must be 64, 128 or 256"
                          : "dpm_batch_set_option: cell width must be 100-400 % of the cutoff");
            return -1;
        }
        (blk ? b->force_block : b->cell_pct) = (int)value;
        drop_graph(b);
        return 0;
    }
    if (option == DPM_OPT_TIMING) {
        batch_collect_timing(b);
        b->timing = value != 0;
        b->force_ms = 0.0;
        b->force_launches = 0;
        return 0;
    }
    set_error("dpm_batch_set_option: unknown option");
    return -1;
}

int dpm_batch_kernel_time(dpm_batch* b, double* force_ms, int64_t* launches) {
    batch_collect_timing(b);
    if (force_ms) *force_ms = b->force_ms;
    if (launches) *launches = b->force_launches;
    return 0;
}

int dpm_batch_vertex_row_entries(dpm_batch* b, int64_t* total) {
    if (!total) {
        set_error("dpm_batch_vertex_row_entries: null output");
        return -1;
    }
    DPM_HIP(hipStreamSynchronize(b->stream));
    int64_t t = 0;
    if (b->nvslot > 0) {
        std::vector<int> nn(b->nvslot), own(b->nvslot);
        BCOPY(nn.data(), b->d.vnum, sizeof(int) * b->nvslot, hipMemcpyDeviceToHost);
        BCOPY(own.data(), b->d.vpart, sizeof(int) * b->nvslot, hipMemcpyDeviceToHost);
        for (int64_t k = 0; k < b->nvslot; ++k)
            if (own[k] >= 0) t += std::min(nn[k], b->d.vcap);
    }
    *total = t;
    return 0;
}

int dpm_batch_row_entries(dpm_batch* b, int64_t* total) {
    if (!total) {
        set_error("dpm_batch_row_entries: null output");
        return -1;
    }
    DPM_HIP(hipStreamSynchronize(b->stream));
    std::vector<int> nn(b->nslot), own(b->nslot);
    BCOPY(nn.data(), b->d.nnum, sizeof(int) * b->nslot, hipMemcpyDeviceToHost);
    BCOPY(own.data(), b->d.sys_of, sizeof(int) * b->nslot, hipMemcpyDeviceToHost);
    int64_t t = 0;
    for (int64_t k = 0; k < b->nslot; ++k)
        if (own[k] >= 0) t += std::min(nn[k], b->d.cap);
    *total = t;
    return 0;
}

}  // extern "C"

namespace dpm {

// dpm_batch_jam, or with min_only dpm_batch_minimize: every system runs its jam (its
// minimisation) from the state as set, in one batched launch sequence.
static int batch_run(dpm_batch* b, const dpm_jam_params* P, bool min_only,
                     int64_t max_batched_steps, dpm_jam_result* out, int64_t* steps_done) {
    if (!P || !out || max_batched_steps < 0) {
        set_error("dpm_batch_jam: bad arguments");
        return -1;
    }
    BDev& d = b->d;
    // common parameters (pair_params's derived numbers, dpm_jam's loop constants)
    bool fastdiv = true;
    for (int64_t s = 0; s < b->nsys; ++s) {
        if (b->rmin[s] <= 0) {
            set_error("dpm_batch_jam: system " + std::to_string(s) + " has no particles set");
            return -1;
        }
        fastdiv = fastdiv && (b->nv[s] > 0
                                  ? fastdiv_range(b->n_c, b->ec[s], b->rv[s], b->rv[s])
                                  : fastdiv_range(b->n_c, b->e_c, b->rmin[s], b->rmax[s]));
    }
    batch_common(b, fastdiv);
    const BJam J = jam_constants(P);
    // every system starts minimizeFire (decompression phase) with a fresh list
    for (int64_t s = 0; s < b->nsys; ++s) {
        BSys& S = b->hsys[s];
        BSys f{};
        f.off = S.off;
        f.n = S.n;
        f.nc = S.nc;
        f.Lx = b->box0[s];
        f.Ly = b->box0[s];
        f.rmax = b->rmax[s];
        sys_constants(b, s, b->rmax[s], b->rv[s], b->ec[s], b->rot[s], b->nv[s], b->nvs[s],
                      b->voff[s], f);
        f.J = J;
        f.alpha = J.alpha0;
        f.dt = J.dt0;
        f.dt_max = 10.0 * J.dt0;
        f.dt_min = 0.001 * J.dt0;
        f.mode = BM_PROLOGUE;
        f.rebuild = 1;
        f.jphase = min_only ? BJ_MIN : BJ_DECOMP;
        f.status = 1;
        f.L_low = f.L_high = f.L_last = -1;
        S = f;
    }
    DPM_HIP(hipMemcpyAsync(d.sys, b->hsys.data(), sizeof(BSys) * b->nsys, hipMemcpyHostToDevice,
                           b->stream));
    b->ran = true;
    DPM_HIP(hipMemsetAsync(d.flags, 0, 10 * sizeof(int), b->stream));
    const int nt[2] = {(int)b->ntiles, b->nct};  // live until the synchronisation below
    DPM_HIP(hipMemcpyAsync(d.flags + 4, nt, 2 * sizeof(int), hipMemcpyHostToDevice, b->stream));
    DPM_HIP(hipStreamSynchronize(b->stream));
    if (b->exec) {  // the parameters are baked into the graph's kernel arguments
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
    }
    // up to two launch units in flight; each ends with a copy of the flags
    int64_t launched = 0, finished_steps = 0;
    int inflight = 0, slot = 0;
    int64_t unit_steps[2] = {0, 0};
    bool all_done = false;
    while (!all_done) {
        while (inflight < 2 && launched < max_batched_steps) {
            int64_t k = std::min<int64_t>(b->steps_per_graph, max_batched_steps - launched);
            if (b->use_graph && !b->timing && k == b->steps_per_graph) {
                if (ensure_batch_graph(b, k)) return -1;
                DPM_HIP(hipGraphLaunch(b->exec, b->stream));
            } else {
                for (int64_t s = 0; s < k; ++s) enqueue_batch_step(b);
                if (check_launch()) return -1;
            }
            int ws = (slot + inflight) % 2;
            DPM_HIP(hipMemcpyAsync(b->hflags + 4 * ws, d.flags, 4 * sizeof(int),
                                   hipMemcpyDeviceToHost, b->stream));
            DPM_HIP(hipEventRecord(b->poll[ws], b->stream));
            unit_steps[ws] = k;
            launched += k;
            inflight++;
        }
        if (inflight == 0) break;
        DPM_HIP(hipEventSynchronize(b->poll[slot]));
        finished_steps += unit_steps[slot];
        const int* f = b->hflags + 4 * slot;
        inflight--;
        slot = (slot + 1) % 2;
        if (f[0]) {
            DPM_HIP(hipStreamSynchronize(b->stream));
            set_error("dpm_batch_jam: a neighbour row exceeded the row capacity");
            return -1;
        }
        if (f[2] >= b->nsys) all_done = true;
    }
    DPM_HIP(hipStreamSynchronize(b->stream));
    batch_collect_timing(b);
    if (steps_done) *steps_done = finished_steps;
    BCOPY(b->hsys.data(), d.sys, sizeof(BSys) * b->nsys, hipMemcpyDeviceToHost);
    for (int64_t s = 0; s < b->nsys; ++s) {
        const BSys& S = b->hsys[s];
        dpm_jam_result& r = out[s];
        r.compression_steps = S.cstep;
        r.fire_steps_total = S.fire_total;
        r.status = S.mode == BM_DONE ? S.status : -1;  // -1: still running
        r.pressure = S.pr;
        r.packing_fraction = S.phi_final;
        r.pe = S.pe_final;
        r.box_length = S.Lx;
        r.contacts = S.contacts;
    }
    return 0;
}

}  // namespace dpm

extern "C" {

int dpm_batch_jam(dpm_batch* b, const dpm_jam_params* P, int64_t max_batched_steps,
                  dpm_jam_result* out, int64_t* steps_done) {
    return batch_run(b, P, false, max_batched_steps, out, steps_done);
}

int dpm_batch_minimize(dpm_batch* b, const dpm_fire_params* F, int64_t max_batched_steps,
                       dpm_jam_result* out, int64_t* steps_done) {
    if (!F) {
        set_error("dpm_batch_minimize: bad arguments");
        return -1;
    }
    dpm_jam_params P{};
    P.fire = *F;
    return batch_run(b, &P, true, max_batched_steps, out, steps_done);
}

}  // extern "C"

namespace dpm {

// The state as set (positions, and the rigid systems' angles and vertices) saved to, or
// restored from, the autotuner's copies.
static hipError_t tune_state(dpm_batch* b, bool save) {
    auto cp = [&](void* dst, void* src, size_t n) {
        return save ? hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, b->stream)
                    : hipMemcpyAsync(src, dst, n, hipMemcpyDeviceToDevice, b->stream);
    };
    hipError_t e = cp(b->P_init, b->d.P, sizeof(double4) * b->nslot);
    if (e == hipSuccess && b->nvslot > 0) e = cp(b->A_init, b->d.ang, sizeof(double) * b->nslot);
    if (e == hipSuccess && b->nvslot > 0)
        e = cp(b->V_init, b->d.vpos, sizeof(double2) * b->nvslot);
    return e;
}

static int tune_begin(dpm_batch* b) {
    if (!b->P_init && balloc_n(b, &b->P_init, b->nslot)) return -1;
    if (b->nvslot > 0 && !b->A_init &&
        (balloc_n(b, &b->A_init, b->nslot) || balloc_n(b, &b->V_init, b->nvslot)))
        return -1;
    DPM_HIP(tune_state(b, true));
    return 0;
}

// One trial with the batch's current settings: an untimed launch unit from the state as
// set (graph capture and instantiation), then trial_steps timed batched steps from it
// again; ms = wall milliseconds per batched step.
static int tune_trial(dpm_batch* b, const dpm_jam_params* P, int64_t trial_steps, double* ms) {
    std::vector<dpm_jam_result> tmp(b->nsys);
    drop_graph(b);
    DPM_HIP(tune_state(b, false));
    int64_t done = 0;
    if (batch_run(b, P, false, b->steps_per_graph, tmp.data(), &done)) return -1;
    DPM_HIP(tune_state(b, false));
    DPM_HIP(hipStreamSynchronize(b->stream));
    auto t0 = std::chrono::steady_clock::now();
    if (batch_run(b, P, false, trial_steps, tmp.data(), &done)) return -1;
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
              .count() /
          (double)std::max<int64_t>(done, 1);
    return 0;
}

static int tune_end(dpm_batch* b) {
    drop_graph(b);
    DPM_HIP(tune_state(b, false));
    DPM_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

}  // namespace dpm

extern "C" {

// Autotuner: runs `trial_steps` batched steps of the jam from the configured state for
// every (cutoff, steps per launch) pair, restores the state after each trial, and keeps
// the fastest pair.  Neither choice changes any result bit (rows only add listed
// non-contacts, which contribute +0.0; launch grouping is not visible to the systems).
int dpm_batch_autotune(dpm_batch* b, const dpm_jam_params* P, int64_t trial_steps,
                       const double* cutoffs, int n_cutoffs, const int64_t* chunks,
                       int n_chunks, double* best_cutoff, int64_t* best_chunk,
                       double* best_ms_per_step) {
    if (!P || trial_steps <= 0 || n_cutoffs <= 0 || n_chunks <= 0 || !cutoffs || !chunks) {
        set_error("dpm_batch_autotune: bad arguments");
        return -1;
    }
    for (int i = 0; i < n_cutoffs; ++i)
        if (!(cutoffs[i] > 0)) {
            set_error("dpm_batch_autotune: bad candidate");
            return -1;
        }
    for (int j = 0; j < n_chunks; ++j)
        if (chunks[j] < 1 || chunks[j] > 1024) {
            set_error("dpm_batch_autotune: bad candidate");
            return -1;
        }
    if (tune_begin(b)) return -1;
    // the cutoff candidates apply to the disk systems (rigid systems keep their own)
    double rmax_disk = 0.0;
    for (int64_t s = 0; s < b->nsys; ++s)
        if (b->nv[s] == 0) rmax_disk = std::max(rmax_disk, b->rmax[s]);
    const double rc0 = b->rc, thr0 = b->thr_sq;
    const int64_t ch0 = b->steps_per_graph;
    double best = 1e300, bc = rc0;
    int64_t bk = ch0;
    for (int i = 0; i < n_cutoffs; ++i) {
        for (int j = 0; j < n_chunks; ++j) {
            // rebuild when a particle moved half the skin (the threshold keeps its form)
            double skin = cutoffs[i] - 2.0 * rmax_disk;
            if (!(skin > 0)) continue;
            b->rc = cutoffs[i];
            b->thr_sq = 3.0 * (0.5 * skin) * (0.5 * skin);
            b->steps_per_graph = chunks[j];
            double ms;
            if (tune_trial(b, P, trial_steps, &ms)) return -1;
            if (ms < best) {
                best = ms;
                bc = cutoffs[i];
                bk = chunks[j];
            }
        }
    }
    if (best == 1e300) {
        b->rc = rc0;
        b->thr_sq = thr0;
        b->steps_per_graph = ch0;
        set_error("dpm_batch_autotune: no cutoff exceeds the largest contact distance");
        return -1;
    }
    double skin = bc - 2.0 * rmax_disk;
    b->rc = bc;
    b->thr_sq = 3.0 * (0.5 * skin) * (0.5 * skin);
    b->steps_per_graph = bk;
    if (tune_end(b)) return -1;
    if (best_cutoff) *best_cutoff = bc;
    if (best_chunk) *best_chunk = bk;
    if (best_ms_per_step) *best_ms_per_step = best;
    return 0;
}

// Launch shape: the force pass's workgroup size (64, 128 or 256 slots) and the cell-grid
// build's cell width (100, 125 or 150 % of the cutoff), tried one after the other with
// the current cutoff and steps per launch.  Neither changes a result bit: the force pass
// is per slot, and cells at least the cutoff wide give the same rows.
int dpm_batch_autotune_shape(dpm_batch* b, const dpm_jam_params* P, int64_t trial_steps,
                             int64_t* best_block, int64_t* best_cell_pct,
                             double* best_ms_per_step) {
    if (!P || trial_steps <= 0) {
        set_error("dpm_batch_autotune_shape: bad arguments");
        return -1;
    }
    if (tune_begin(b)) return -1;
    double best = 1e300;
    int bb = b->force_block, bp = b->cell_pct;
    for (int blk : {64, 128, 256}) {
        b->force_block = blk;
        double ms;
        if (tune_trial(b, P, trial_steps, &ms)) return -1;
        if (ms < best) {
            best = ms;
            bb = blk;
        }
    }
    b->force_block = bb;
    for (int pct : {100, 125, 150}) {
        b->cell_pct = pct;
        double ms;
        if (tune_trial(b, P, trial_steps, &ms)) return -1;
        if (ms < best) {
            best = ms;
            bp = pct;
        }
    }
    b->cell_pct = bp;
    if (tune_end(b)) return -1;
    if (best_block) *best_block = bb;
    if (best_cell_pct) *best_cell_pct = bp;
    if (best_ms_per_step) *best_ms_per_step = best;
    return 0;
}

int dpm_batch_finish_steps(dpm_batch* b, int64_t* steps_out) {
    if (!steps_out) {
        set_error("dpm_batch_finish_steps: null output");
        return -1;
    }
    DPM_HIP(hipStreamSynchronize(b->stream));
    BCOPY(b->hsys.data(), b->d.sys, sizeof(BSys) * b->nsys, hipMemcpyDeviceToHost);
    for (int64_t s = 0; s < b->nsys; ++s)
        steps_out[s] = b->hsys[s].mode == BM_DONE ? b->hsys[s].done_step : -1;
    return 0;
}

}  // extern "C"

// =========================================================================
//           serving loop (continuous batching of jam and minimisation requests)
// =========================================================================
// Requests (disk or rigid bumpy systems, each with its own jamming parameters) are
// queued on submission and admitted while the device budget allows: a request takes a
// free region of the batched engine's arena (and, for a rigid bumpy system, of its
// vertex arena) and a system entry, and joins the running batch at the next launch.  A
// finished system leaves the arena at once, so its space and entry serve the next
// request.  Every result is the one the request's system gives when jammed alone.
#include <deque>
#include <map>
#include <mutex>

struct dpm_server_req {
    int64_t id = 0;
    int kind = DPM_DISK;
    int mode = DPM_REQ_JAM;
    int64_t n = 0, nv = 0;
    std::vector<double> x, y, r, m, ang;
    std::vector<int64_t> vcounts;
    double rv = 0.0, side = 0.0, e_c = 1.0;
    dpm_jam_params jam{};
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    std::chrono::steady_clock::time_point t_submit, t_done;
    dpm_jam_result res{};
    std::vector<double> rx, ry;
    double rside = 0.0;
    std::string err;
    int sys = -1;
    long long off = -1, slots = 0, voff = -1, vslots = 0;
    int64_t bytes = 0;
};

// Locking: `mu` guards the request table, the queue and every counter; the batch (arena,
// stream, graph) is driven only by the thread calling dpm_server_progress.
struct dpm_server {
    dpm_batch* b = nullptr;
    int64_t budget = 0, used = 0;
    std::mutex mu;
    std::map<int64_t, dpm_server_req> reqs;
    std::deque<int64_t> queue;                    // submitted, not admitted
    std::vector<int> free_sys;
    std::map<long long, long long> free_ranges;   // first slot -> length
    std::map<long long, long long> vfree_ranges;  // first vertex slot -> length
    std::vector<int64_t> sys_req;                 // system entry -> request id (0: free)
    int64_t next_id = 1, seen_done = 0, finished = 0, active = 0;
    std::vector<int2> tiles_host;
};

namespace dpm {

// device bytes of one arena slot and of one vertex slot (arrays of batch_alloc)
constexpr int64_t SLOT_BYTES = 32 + 32 + 16 + 16 + 16 + 8 + 4 + 4 + 32 * 4 + 8;
constexpr int64_t RIGID_SLOT_BYTES = 4 + 4 + 8 * 5 + 16 * 4 + 16 + 4;
constexpr int64_t VSLOT_BYTES = 4 + 4 + 4 + 32 * 4 + 16 * 2;

static void serve_upload_tiles(dpm_server* sv) {
    dpm_batch* b = sv->b;
    sv->tiles_host.clear();
    for (size_t s = 0; s < sv->sys_req.size(); ++s) {
        if (!sv->sys_req[s]) continue;
        const dpm_server_req& q = sv->reqs.at(sv->sys_req[s]);
        for (int64_t t = 0; t < q.n; t += BLOCK)
            sv->tiles_host.push_back(make_int2((int)s, (int)(q.off + t)));
    }
    std::vector<int2> ct;
    for (size_t s = 0; s < sv->sys_req.size(); ++s)
        if (sv->sys_req[s]) cell_tiles(ct, (int)s, (int)sv->reqs.at(sv->sys_req[s]).n);
    // the bound of batch_alloc: ct never exceeds d.nctiles entries
    int nt[2] = {(int)sv->tiles_host.size(), (int)std::min<long long>(ct.size(), b->d.nctiles)};
    if (nt[0])
        (void)bsync(b, hipMemcpyAsync(b->d.tiles, sv->tiles_host.data(), sizeof(int2) * nt[0],
                                      hipMemcpyHostToDevice, b->stream));
    if (nt[1])
        (void)bsync(b, hipMemcpyAsync(b->d.ctiles, ct.data(), sizeof(int2) * nt[1],
                                      hipMemcpyHostToDevice, b->stream));
    (void)bsync(b, hipMemcpyAsync(b->d.flags + 4, nt, 2 * sizeof(int), hipMemcpyHostToDevice,
                                  b->stream));
}

// first fit in a map of free ranges (first slot -> length)
static bool take_range(std::map<long long, long long>& fr, long long slots, long long* off) {
    for (auto it = fr.begin(); it != fr.end(); ++it) {
        if (it->second < slots) continue;
        *off = it->first;
        long long rest = it->second - slots, start = it->first + slots;
        fr.erase(it);
        if (rest) fr[start] = rest;
        return true;
    }
    return false;
}

static void give_range(std::map<long long, long long>& fr, long long off, long long slots) {
    auto it = fr.emplace(off, slots).first;
    auto nx = std::next(it);
    if (nx != fr.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        fr.erase(nx);
    }
    if (it != fr.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            fr.erase(it);
        }
    }
}

// Systems that finished since the last look leave the arena.  The stream is idle.
static int serve_collect(dpm_server* sv) {
    dpm_batch* b = sv->b;
    int fl[8];
    BCOPY(fl, b->d.flags, sizeof(fl), hipMemcpyDeviceToHost);
    if (fl[0]) {
        set_error("dpm_server: a neighbour row exceeded the row capacity");
        return -1;
    }
    int64_t total = fl[2];
    if (total == sv->seen_done) return 0;
    std::vector<int> ids(b->nsys);
    BCOPY(ids.data(), b->d.done_list, sizeof(int) * b->nsys, hipMemcpyDeviceToHost);
    bool changed = false;
    for (int64_t k = sv->seen_done; k < total; ++k) {
        int s = ids[k % b->nsys];
        int64_t id = sv->sys_req[s];
        if (!id) continue;
        dpm_server_req& q = sv->reqs.at(id);
        BSys S;
        BCOPY(&S, b->d.sys + s, sizeof(BSys), hipMemcpyDeviceToHost);
        std::vector<double4> h(q.n);
        BCOPY(h.data(), b->d.P + q.off, sizeof(double4) * q.n, hipMemcpyDeviceToHost);
        q.rx.resize(q.n);
        q.ry.resize(q.n);
        for (int64_t i = 0; i < q.n; ++i) {
            q.rx[i] = h[i].x;
            q.ry[i] = h[i].y;
        }
        q.res.compression_steps = S.cstep;
        q.res.fire_steps_total = S.fire_total;
        q.res.status = S.status;
        q.res.pressure = S.pr;
        q.res.packing_fraction = S.phi_final;
        q.res.pe = S.pe_final;
        q.res.box_length = S.Lx;
        q.res.contacts = S.contacts;
        q.rside = S.Lx;
        q.state = 2;
        q.t_done = std::chrono::steady_clock::now();
        // the system entry and its slots become free
        S.mode = BM_EMPTY;
        BCOPY(b->d.sys + s, &S, sizeof(BSys), hipMemcpyHostToDevice);
        std::vector<int> neg(std::max(q.slots, q.vslots), -1);
        BCOPY(b->d.sys_of + q.off, neg.data(), sizeof(int) * q.slots, hipMemcpyHostToDevice);
        give_range(sv->free_ranges, q.off, q.slots);
        if (q.vslots) {
            BCOPY(b->d.vsys_of + q.voff, neg.data(), sizeof(int) * q.vslots,
                  hipMemcpyHostToDevice);
            BCOPY(b->d.vpart + q.voff, neg.data(), sizeof(int) * q.vslots,
                  hipMemcpyHostToDevice);
            give_range(sv->vfree_ranges, q.voff, q.vslots);
        }
        sv->free_sys.push_back(s);
        sv->sys_req[s] = 0;
        sv->used -= q.bytes;
        sv->finished++;
        sv->active--;
        changed = true;
    }
    sv->seen_done = total;
    if (changed) serve_upload_tiles(sv);
    return 0;
}

// The request's particles (and a rigid system's angles, vertex layout and vertex ring)
// into its arena ranges, then its system entry (minimizeFire, decompression phase).
static int serve_place(dpm_server* sv, dpm_server_req& q, int s) {
    dpm_batch* b = sv->b;
    BDev& d = b->d;
    const long long off = q.off;
    std::vector<double4> h(q.n);
    double rm = 0.0;
    for (int64_t i = 0; i < q.n; ++i) {
        h[i] = make_double4(q.x[i], q.y[i], q.r[i], q.m[i]);
        rm = std::max(rm, q.r[i]);
    }
    BCOPY(d.P + off, h.data(), sizeof(double4) * q.n, hipMemcpyHostToDevice);
    std::vector<int> own(std::max(q.slots, q.vslots), s);
    BCOPY(d.sys_of + off, own.data(), sizeof(int) * q.slots, hipMemcpyHostToDevice);
    if (q.kind == DPM_BUMPY) {
        std::vector<int> vst(q.n), nvt(q.n), vpart(q.vslots, -1);
        long long acc = 0;
        for (int64_t i = 0; i < q.n; ++i) {
            vst[i] = (int)(q.voff + acc);
            nvt[i] = (int)q.vcounts[i];
            for (int64_t u = 0; u < q.vcounts[i]; ++u) vpart[acc + u] = (int)(off + i);
            acc += q.vcounts[i];
        }
        BCOPY(d.ang + off, q.ang.data(), sizeof(double) * q.n, hipMemcpyHostToDevice);
        BCOPY(d.vst + off, vst.data(), sizeof(int) * q.n, hipMemcpyHostToDevice);
        BCOPY(d.nvt + off, nvt.data(), sizeof(int) * q.n, hipMemcpyHostToDevice);
        BCOPY(d.vpart + q.voff, vpart.data(), sizeof(int) * q.vslots, hipMemcpyHostToDevice);
        BCOPY(d.vsys_of + q.voff, own.data(), sizeof(int) * q.vslots, hipMemcpyHostToDevice);
        hipLaunchKernelGGL(k_rb_init, pgrid(q.n), dim3(BLOCK), 0, b->stream, d, off, (int)q.n,
                           q.rv);
        if (check_launch()) return -1;
        DPM_HIP(hipStreamSynchronize(b->stream));
    }
    BSys f{};
    f.off = off;
    f.n = (int)q.n;
    f.nc = (int)((q.n + WAVE - 1) / WAVE);
    f.Lx = f.Ly = q.side;
    f.rmax = rm;
    const bool rigid = q.kind == DPM_BUMPY;
    sys_constants(b, s, rm, q.rv, q.e_c, 1, rigid ? q.nv : 0, q.vslots, q.voff, f);
    f.J = jam_constants(&q.jam);
    f.alpha = f.J.alpha0;
    f.dt = f.J.dt0;
    f.dt_max = 10.0 * f.J.dt0;
    f.dt_min = 0.001 * f.J.dt0;
    f.mode = BM_PROLOGUE;
    f.rebuild = 1;
    f.jphase = q.mode == DPM_REQ_MINIMIZE ? BJ_MIN : BJ_DECOMP;
    f.status = 1;
    f.L_low = f.L_high = f.L_last = -1;
    BCOPY(d.sys + s, &f, sizeof(BSys), hipMemcpyHostToDevice);
    return 0;
}

// Queued requests that fit join the arena.
static int serve_admit(dpm_server* sv) {
    bool tiles_changed = false;
    for (auto it = sv->queue.begin(); it != sv->queue.end();) {
        dpm_server_req& q = sv->reqs.at(*it);
        if (sv->used + q.bytes > sv->budget || sv->free_sys.empty()) {
            ++it;
            continue;
        }
        long long off, voff = -1;
        if (!take_range(sv->free_ranges, q.slots, &off)) {
            ++it;
            continue;
        }
        if (q.vslots && !take_range(sv->vfree_ranges, q.vslots, &voff)) {
            give_range(sv->free_ranges, off, q.slots);
            ++it;
            continue;
        }
        int s = sv->free_sys.back();
        sv->free_sys.pop_back();
        q.sys = s;
        q.off = off;
        q.voff = voff;
        q.state = 1;
        sv->used += q.bytes;
        sv->sys_req[s] = q.id;
        sv->active++;
        if (serve_place(sv, q, s)) return -1;
        // the request's inputs now live in the arena
        for (auto* v : {&q.x, &q.y, &q.r, &q.m, &q.ang}) std::vector<double>().swap(*v);
        std::vector<int64_t>().swap(q.vcounts);
        tiles_changed = true;
        it = sv->queue.erase(it);
    }
    if (tiles_changed) serve_upload_tiles(sv);
    return 0;
}

}  // namespace dpm

extern "C" {

// Every server entry point turns a C++ exception (allocation failure) into an error
// return: nothing is thrown across the C ABI.
#define SERVER_GUARD(fail_value)                                   \
    catch (const std::exception& e) {                             \
        set_error(std::string("dpm_server: ") + e.what());         \
        return fail_value;                                         \
    }                                                              \
    catch (...) {                                                  \
        set_error("dpm_server: unknown exception");                \
        return fail_value;                                         \
    }

dpm_server* dpm_server_create(int64_t max_systems, int64_t arena_slots, int64_t budget_bytes,
                              int64_t vertex_slots) {
    if (max_systems <= 0 || max_systems >= (int64_t)1 << 24 || arena_slots < WAVE ||
        arena_slots >= (int64_t)1 << 31 || budget_bytes <= 0 || vertex_slots < 0 ||
        vertex_slots >= (int64_t)1 << 31) {
        set_error("dpm_server_create: bad arguments");
        return nullptr;
    }
    if (dpm_device_count() <= 0) {
        set_error("dpm_server_create: no HIP device");
        return nullptr;
    }
    dpm_server* sv = nullptr;
    try {
        sv = new dpm_server();
        sv->b = new dpm_batch();
        int64_t slots = arena_slots / WAVE * WAVE;
        int64_t vslots = vertex_slots / WAVE * WAVE;
        dpm_batch* b = sv->b;
        b->nsys = max_systems;
        b->nslot = slots;
        b->ntiles = slots / WAVE;
        std::string err;
        if (batch_alloc(b, max_systems, slots, b->ntiles, vslots)) {
            err = dpm_last_error();
        } else {
            std::vector<BSys> empty(max_systems);
            for (auto& S : empty) S.mode = BM_EMPTY;
            if (bsync(b, hipMemcpyAsync(b->d.sys, empty.data(), sizeof(BSys) * max_systems,
                                        hipMemcpyHostToDevice, b->stream)) != hipSuccess)
                err = "device copy failed";
        }
        if (err.empty()) {
            sv->budget = budget_bytes;
            sv->sys_req.assign(max_systems, 0);
            for (int64_t s = max_systems - 1; s >= 0; --s) sv->free_sys.push_back((int)s);
            sv->free_ranges[0] = slots;
            if (vslots) sv->vfree_ranges[0] = vslots;
            batch_common(b, true);  // every request's pair law is checked at submission
            if (b->use_graph && ensure_batch_graph(b, b->steps_per_graph)) err = dpm_last_error();
        }
        if (!err.empty()) {
            dpm_batch_destroy(sv->b);
            delete sv;
            set_error("dpm_server_create: " + err);
            return nullptr;
        }
        return sv;
    } catch (const std::exception& e) {
        if (sv) {
            if (sv->b) dpm_batch_destroy(sv->b);
            delete sv;
        }
        set_error(std::string("dpm_server_create: ") + e.what());
        return nullptr;
    }
}

void dpm_server_destroy(dpm_server* sv) {
    if (!sv) return;
    dpm_batch_destroy(sv->b);
    delete sv;
}

int64_t dpm_server_submit(dpm_server* sv, const dpm_request* r) try {
    if (!r || r->n_particles <= 0 || r->n_particles >= (int64_t)1 << 24 || !r->x || !r->y ||
        !r->radius || !r->mass || !(r->box_side > 0) ||
        (r->kind != DPM_DISK && r->kind != DPM_BUMPY) ||
        (r->mode != DPM_REQ_JAM && r->mode != DPM_REQ_MINIMIZE)) {
        set_error("dpm_server_submit: bad request");
        return -1;
    }
    dpm_server_req q;
    q.kind = r->kind;
    q.mode = r->mode;
    q.n = r->n_particles;
    q.x.assign(r->x, r->x + q.n);
    q.y.assign(r->y, r->y + q.n);
    q.r.assign(r->radius, r->radius + q.n);
    q.m.assign(r->mass, r->mass + q.n);
    q.side = r->box_side;
    q.e_c = r->e_c;
    q.jam = r->jam;
    for (int64_t i = 0; i < q.n; ++i)
        if (!(q.r[i] >= 0x1p-100 && q.r[i] <= 0x1p100) || !(q.m[i] > 0)) {
            set_error("dpm_server_submit: radii must lie in [2^-100, 2^100], masses > 0");
            return -1;
        }
    q.slots = (q.n + WAVE - 1) / WAVE * WAVE;
    q.bytes = q.slots * SLOT_BYTES;
    if (q.slots > sv->b->nslot) {
        set_error("dpm_server_submit: system larger than the arena");
        return -1;
    }
    if (q.kind == DPM_DISK) {
        if (q.e_c != sv->b->e_c) {
            set_error("dpm_server_submit: disk requests share the server's contact energy");
            return -1;
        }
    } else {
        if (!r->angle || !r->vertices_per_particle ||
            !(r->vertex_radius >= 0x1p-100 && r->vertex_radius <= 0x1p100) ||
            r->n_vertices < q.n || !(q.e_c >= 0x1p-100 && q.e_c <= 0x1p100)) {
            set_error("dpm_server_submit: rigid bumpy requests need angles, vertex counts, a "
                      "vertex radius in [2^-100, 2^100] and a contact energy in that range");
            return -1;
        }
        if (sv->b->nvslot == 0) {
            set_error("dpm_server_submit: this server takes no rigid bumpy requests");
            return -1;
        }
        q.nv = r->n_vertices;
        q.ang.assign(r->angle, r->angle + q.n);
        q.vcounts.assign(r->vertices_per_particle, r->vertices_per_particle + q.n);
        q.rv = r->vertex_radius;
        int64_t acc = 0;
        for (int64_t i = 0; i < q.n; ++i) {
            if (q.vcounts[i] <= 0 || !(q.r[i] > q.rv)) {
                set_error("dpm_server_submit: every particle needs a vertex and a radius above "
                          "the vertex radius");
                return -1;
            }
            acc += q.vcounts[i];
        }
        if (acc != q.nv) {
            set_error("dpm_server_submit: vertex counts do not sum to n_vertices");
            return -1;
        }
        q.vslots = (q.nv + WAVE - 1) / WAVE * WAVE;
        if (q.vslots > sv->b->nvslot) {
            set_error("dpm_server_submit: system larger than the vertex arena");
            return -1;
        }
        q.bytes += q.slots * RIGID_SLOT_BYTES + q.vslots * VSLOT_BYTES;
    }
    if (q.bytes > sv->budget) {
        set_error("dpm_server_submit: request exceeds the device budget");
        return -1;
    }
    std::lock_guard<std::mutex> g(sv->mu);
    q.id = sv->next_id++;
    q.t_submit = std::chrono::steady_clock::now();
    int64_t id = q.id;
    sv->reqs.emplace(id, std::move(q));
    sv->queue.push_back(id);
    return id;
}
SERVER_GUARD(-1)

int dpm_server_progress(dpm_server* sv, int64_t batched_steps, int64_t* finished_total,
                        int64_t* in_flight, int64_t* queued) try {
    dpm_batch* b = sv->b;
    int64_t active = 0;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        if (serve_collect(sv) || serve_admit(sv)) return -1;
        active = sv->active;
    }
    if (active > 0 && batched_steps > 0) {
        int64_t done = 0;
        while (done < batched_steps) {
            int64_t k = std::min<int64_t>(b->steps_per_graph, batched_steps - done);
            if (b->use_graph && k == b->steps_per_graph) {
                if (ensure_batch_graph(b, k)) return -1;
                DPM_HIP(hipGraphLaunch(b->exec, b->stream));
            } else {
                for (int64_t s = 0; s < k; ++s) enqueue_batch_step(b);
                if (check_launch()) return -1;
            }
            done += k;
        }
        DPM_HIP(hipStreamSynchronize(b->stream));
        std::lock_guard<std::mutex> g(sv->mu);
        if (serve_collect(sv) || serve_admit(sv)) return -1;
    }
    std::lock_guard<std::mutex> g(sv->mu);
    if (finished_total) *finished_total = sv->finished;
    if (in_flight) *in_flight = sv->active;
    if (queued) *queued = (int64_t)sv->queue.size();
    return 0;
}
SERVER_GUARD(-1)

int dpm_server_result(dpm_server* sv, int64_t id, int32_t* state, dpm_jam_result* out,
                      double* latency_s, double* x, double* y, double* box_side) try {
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) {
        set_error("dpm_server_result: unknown request");
        return -1;
    }
    const dpm_server_req& q = it->second;
    if (state) *state = q.state;
    if (q.state == 3) {
        set_error("dpm_server_result: request failed: " + q.err);
        return -1;
    }
    if (q.state != 2) return 0;
    if (out) *out = q.res;
    if (latency_s)
        *latency_s = std::chrono::duration<double>(q.t_done - q.t_submit).count();
    if (x) std::copy(q.rx.begin(), q.rx.end(), x);
    if (y) std::copy(q.ry.begin(), q.ry.end(), y);
    if (box_side) *box_side = q.rside;
    return 0;
}
SERVER_GUARD(-1)

int dpm_server_release(dpm_server* sv, int64_t id) try {
    std::lock_guard<std::mutex> g(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) {
        set_error("dpm_server_release: unknown request");
        return -1;
    }
    if (it->second.state != 2 && it->second.state != 3) {
        set_error("dpm_server_release: request still queued or running");
        return -1;
    }
    sv->reqs.erase(it);
    return 0;
}
SERVER_GUARD(-1)

}  // extern "C"
