// dpm_core.hip -- gfx950 kernels and host control for one particle system
// (disk path, neighbour lists, reductions, NVE / FIRE / Adam loops, jamming).
//
// Reference correspondence is cited per kernel.  Layout in HBM (one system):
//   pos, vel, frc, lref, cref : double2[N]   (x,y interleaved: one 16 B gather)
//   rad, mass, pe, ke         : double[N]
//   nbr                       : int32 ELL, column-major nbr[k*N + i] (lanes coalesce)
//   nnum, sidx, cell_of       : int32[N]
//   cell_start                : int32[ncells+1] counting-sort offsets
//   pmax_nd, pmax_cd          : double[ceil(N/256)] per-block displacement maxima
//   part                      : double[4*ceil(N/64)] canonical-sum chunk partials
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>

#include "dpm_kernels.hpp"

namespace dpm {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

// =========================================================================
//                                  kernels
// =========================================================================

__device__ __forceinline__ int64_t particle_of_thread(int& tile) {
    tile = xcd_tile(blockIdx.x, gridDim.x);
    return (int64_t)tile * BLOCK + threadIdx.x;
}

// Per-block maxima of the two displacement arrays (the reference keeps the
// full arrays, particle.cu:1094-1100; only their maxima are ever read) and the
// rebuild request of Particle::checkFor{Cell,Verlet}ListUpdate
// (particle.cu:1141-1161): 3*max > threshold.  par < 0: no request.
__device__ __forceinline__ void disp_epilogue(const Dev& d, int tile, double nd, double cd,
                                              int method, double thr, double cthr, int par) {
    double bn = block_max(nd);
    double bc = block_max(cd);
    if (threadIdx.x == 0) {
        d.pmax_nd[tile] = bn;
        d.pmax_cd[tile] = bc;
        if (par >= 0) {
            bool need = false;
            if (method == DPM_NBR_CELL) need = (3.0 * bc > cthr) || (3.0 * bn > thr);
            else if (method == DPM_NBR_VERLET) need = 3.0 * bn > thr;
            if (need) d.ctl->flag[par] = 1;
            if (blockIdx.x == 0) d.ctl->flag[par ^ 1] = 0;
        }
    }
}

// NVE first half: kernelUpdateVelocities(dt/2) + kernelUpdatePositions(dt)
// (src/particles/base/kernels.cu:117-139, :48-81) fused.
__global__ __launch_bounds__(BLOCK) void k_nve_kick_drift(Dev d, double half_dt, double dt,
                                                          int method, double thr,
                                                          double cthr, int par) {
    if (d.ctl->halt) return;
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double2 v = d.vel[i], f = d.frc[i], x = d.pos[i];
        double dtm = half_dt / d.mass[i];
        v.x += f.x * dtm;
        v.y += f.y * dtm;
        x.x += v.x * dt;
        x.y += v.y * dt;
        d.vel[i] = v;
        d.pos[i] = x;
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = x.x - ln.x, dyn = x.y - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = x.x - lc.x, dyc = x.y - lc.y;
        cd = dxc * dxc + dyc * dyc;
    }
    disp_epilogue(d, tile, nd, cd, method, thr, cthr, par);
}

// Sum of harmonic pair forces/energies over particle i's ELL row:
// kernelCalcDiskForces, src/particles/disk/kernels.cu:11-57.
__device__ __forceinline__ void disk_row(const Dev& d, const PairParams& p, int64_t i, double& ax,
                                         double& ay, double& e) {
    double2 xi = d.pos[i];
    double ri = d.rad[i];
    int nn = min(d.nnum[i], d.cap);
    const int* col = d.nbr + i;
    ax = 0.0;
    ay = 0.0;
    e = 0.0;
    for (int k = 0; k < nn; ++k) {
        int j = col[(int64_t)k * d.N];
        double2 xj = d.pos[j];
        double rj = d.rad[j];
        double fx, fy;
        double ie = pair_law(xi.x, xi.y, ri, xj.x, xj.y, rj, p, fx, fy);
        ax += fx;
        ay += fy;
        e += ie;
    }
}

// NVE second half: zero + kernelCalcDiskForces + kernelUpdateVelocities(dt/2)
// fused (src/integrator/nve.cu:21-24).  Gated on the rebuild request of this
// step: when raised, the step halts here and the host rebuilds first.
__global__ __launch_bounds__(BLOCK) void k_nve_force_kick(Dev d, PairParams p, double half_dt,
                                                          int par, int cpos) {
    Ctl* c = d.ctl;
    if (c->halt) return;
    if (par >= 0 && c->flag[par]) {
        if (threadIdx.x == 0) {
            c->halt = 1;
            c->halt_pos = cpos;
        }
        return;
    }
    int tile;
    int64_t i = particle_of_thread(tile);
    if (i >= d.N) return;
    double ax, ay, e;
    disk_row(d, p, i, ax, ay, e);
    double fx = 0.0 + ax, fy = 0.0 + ay;
    d.frc[i] = make_double2(fx, fy);
    d.pe[i] = 0.0 + e;
    double dtm = half_dt / d.mass[i];
    double2 v = d.vel[i];
    v.x += fx * dtm;
    v.y += fy * dtm;
    d.vel[i] = v;
}

// Disk::calculateForces as called stand-alone: forces are accumulated (+=),
// exactly as the reference kernel does into caller-zeroed arrays.
__global__ __launch_bounds__(BLOCK) void k_disk_forces_acc(Dev d, PairParams p) {
    int tile;
    int64_t i = particle_of_thread(tile);
    if (i >= d.N) return;
    if (d.nnum[i] == 0) return;
    double ax, ay, e;
    disk_row(d, p, i, ax, ay, e);
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
    double dtm = dt / d.mass[i];
    v.x += f.x * dtm;
    v.y += f.y * dtm;
    d.vel[i] = v;
}

// kernelUpdatePositions, src/particles/base/kernels.cu:48-81
__global__ __launch_bounds__(BLOCK) void k_drift(Dev d, double dt) {
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double2 x = d.pos[i], v = d.vel[i];
        x.x += v.x * dt;
        x.y += v.y * dt;
        d.pos[i] = x;
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = x.x - ln.x, dyn = x.y - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = x.x - lc.x, dyc = x.y - lc.y;
        cd = dxc * dxc + dyc * dyc;
    }
    disp_epilogue(d, tile, nd, cd, 0, 0.0, 0.0, -1);
}

// kernelMixVelocitiesAndForces, src/particles/base/kernels.cu:91-112
__device__ __forceinline__ void mix_one(double2& v, double2 f, double alpha) {
    double fn = sqrt(f.x * f.x + f.y * f.y);
    double vn = sqrt(v.x * v.x + v.y * v.y);
    double mr = 0.0;
    if (fn > 1e-16) {
        mr = vn / fn * alpha;
    } else {
        v.x = 0.0;
        v.y = 0.0;
        mr = 0.0;
    }
    v.x = v.x * (1 - alpha) + f.x * mr;
    v.y = v.y * (1 - alpha) + f.y * mr;
}

__global__ void k_mix(Dev d, double alpha) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    mix_one(v, d.frc[i], alpha);
    d.vel[i] = v;
}

// kernelCalculateTranslationalKineticEnergy, src/particles/base/kernels.cu:163-180
__global__ void k_kinetic(Dev d, int rigid_rot) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    double v2 = v.x * v.x + v.y * v.y;
    if (rigid_rot) {
        double w = d.omega[i];
        d.ke[i] = 0.5 * d.mass[i] * v2 + 0.5 * d.moi[i] * w * w;
    } else {
        d.ke[i] = 0.5 * d.mass[i] * v2;
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
        case S_R2: return d.rad[i] * d.rad[i];
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
        case M_F2: { double2 f = d.frc[i]; return f.x * f.x + f.y * f.y; }
        case M_T2: { double t = d.torque[i]; return t * t; }
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
    double2 xi = d.pos[i], v = d.vel[i];
    double m = d.mass[i], ri = d.rad[i];
    double axx = m * v.x * v.x, axy = m * v.x * v.y, ayx = axy, ayy = m * v.y * v.y;
    int nn = min(d.nnum[i], d.cap);
    for (int k = 0; k < nn; ++k) {
        int j = d.nbr[(int64_t)k * d.N + i];
        double2 xj = d.pos[j];
        double xd = pbc_mul(xi.x, xj.x, p.Lx, p.invLx);
        double yd = pbc_mul(xi.y, xj.y, p.Ly, p.invLy);
        double fx, fy;
        pair_law(xi.x, xi.y, ri, xj.x, xj.y, d.rad[j], p, fx, fy);
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
    double2 xi = d.pos[i];
    double ri = d.rad[i];
    int c = 0;
    int nn = min(d.nnum[i], d.cap);
    for (int k = 0; k < nn; ++k) {
        int j = d.nbr[(int64_t)k * d.N + i];
        double2 xj = d.pos[j];
        double dx = pbc_mul(xi.x, xj.x, p.Lx, p.invLx);
        double dy = pbc_mul(xi.y, xj.y, p.Ly, p.invLy);
        double dist = sqrt(dx * dx + dy * dy);
        if (dist < ri + d.rad[j]) c++;
    }
    d.contacts[i] = c;
}

// positions.scale (Particle::scalePositionsFull, particle.cu:1070-1072)
__global__ void k_scale_positions(Dev d, double f) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 x = d.pos[i];
    if (d.vpos) {  // kernelScalePositions, rigid_bumpy/kernels.cu:1308-1333
        double nx = x.x * f, ny = x.y * f;
        int st = d.vstart[i], nv = d.nverts[i];
        for (int v = st; v < st + nv; ++v) {
            double2 q = d.vpos[v];
            q.x += (nx - x.x);
            q.y += (ny - x.y);
            d.vpos[v] = q;
        }
        d.pos[i] = make_double2(nx, ny);
        return;
    }
    x.x *= f;
    x.y *= f;
    d.pos[i] = x;
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
__global__ __launch_bounds__(BLOCK) void k_adam_update(Dev d, double alpha, double b1, double b2,
                                                       double omb1, double omb2, double eps) {
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
        double2 x = d.pos[i];
        x.x += ux;
        x.y += uy;
        d.pos[i] = x;
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = x.x - ln.x, dyn = x.y - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = x.x - lc.x, dyc = x.y - lc.y;
        cd = dxc * dxc + dyc * dyc;
        d.m1[i] = m1;
        d.m2[i] = m2;
    }
    disp_epilogue(d, tile, nd, cd, 0, 0.0, 0.0, -1);
}

__global__ void k_fill_int(int* a, int v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_double(double* a, double v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}

// ------------------------------------------------------------- neighbours
__device__ __forceinline__ int64_t cell_coord(double pos, double cs, int64_t n) {
    // getPBCCellIndex, include/particles/base/kernels.cuh:285-301
    int64_t idx = (int64_t)floor(pos / cs);
    return idx >= 0 ? idx % n : (idx % n + n) % n;
}

// kernelGetCellIndexForParticle (base/kernels.cu:309-330) + counting-sort histogram
__global__ void k_cell_bin(Dev d, double cs, int64_t ncd) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 x = d.pos[i];
    int64_t c = cell_coord(x.x, cs, ncd) + cell_coord(x.y, cs, ncd) * ncd;
    if (!(c >= 0 && c < ncd * ncd)) {  // non-finite position: keep the index in range, report
        c = 0;
        d.ctl->overflow = 1;
    }
    d.cell_tmp[i] = (int)c;
    atomicAdd(&d.cell_cnt[c], 1);
}

// Exclusive scan of the cell histogram into cell_start (one 1024-thread block).
// Replaces kernelGetFirstParticleIndexForCell's O(n_cells * N) scan
// (base/kernels.cu:332-438); empty cells get start == end instead of -1.
__global__ __launch_bounds__(1024) void k_cell_scan(Dev d, int64_t nc) {
    __shared__ int sums[1024];
    int t = threadIdx.x;
    int64_t per = (nc + 1023) / 1024;
    int64_t lo = t * per, hi = min(lo + per, nc);
    int loc = 0;
    for (int64_t c = lo; c < hi; ++c) loc += d.cell_cnt[c];
    sums[t] = loc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        int v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int base = sums[t] - loc;
    for (int64_t c = lo; c < hi; ++c) {
        d.cell_start[c] = base;
        d.cell_cur[c] = base;
        base += d.cell_cnt[c];
    }
    if (t == 1023) d.cell_start[nc] = (int)d.N;
}

__global__ void k_cell_scatter(Dev d) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    int c = d.cell_tmp[i];
    int p = atomicAdd(&d.cell_cur[c], 1);
    d.perm[p] = (int)i;
}

// Restore stability inside each cell: ascending previous index, which makes the
// order identical to thrust::sort_by_key's stable radix sort (particle.cu:1293).
__global__ void k_cell_sort(Dev d, int64_t nc) {
    int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
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

// kernelReorderParticleData, base/kernels.cu:508-543 (+ Adam moments, static ids)
__global__ void k_reorder_gather(Dev d, Spare sp) {
    int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= d.N) return;
    int o = d.perm[k];
    double2 x = d.pos[o];
    sp.pos[k] = x;
    sp.vel[k] = d.vel[o];
    sp.frc[k] = d.frc[o];
    sp.rad[k] = d.rad[o];
    sp.mass[k] = d.mass[o];
    sp.sidx[k] = d.sidx[o];
    d.cref[k] = x;
    d.cell_of[k] = d.cell_tmp[o];
    if (d.m1) {
        sp.m1[k] = d.m1[o];
        sp.m2[k] = d.m2[o];
    }
    if (d.vpos) {
        sp.angle[k] = d.angle[o];
        sp.omega[k] = d.omega[o];
        sp.torque[k] = d.torque[o];
        sp.moi[k] = d.moi[o];
        sp.nverts[k] = d.nverts[o];
        sp.old_to_new[o] = (int)k;
    }
}

__global__ void k_reorder_copyback(Dev d, Spare sp) {
    int64_t k = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k >= d.N) return;
    d.pos[k] = sp.pos[k];
    d.vel[k] = sp.vel[k];
    d.frc[k] = sp.frc[k];
    d.rad[k] = sp.rad[k];
    d.mass[k] = sp.mass[k];
    d.sidx[k] = sp.sidx[k];
    if (d.m1) {
        d.m1[k] = sp.m1[k];
        d.m2[k] = sp.m2[k];
    }
    if (d.vpos) {
        d.angle[k] = sp.angle[k];
        d.omega[k] = sp.omega[k];
        d.torque[k] = sp.torque[k];
        d.moi[k] = sp.moi[k];
        d.nverts[k] = sp.nverts[k];
    }
}

__device__ __forceinline__ bool within_div(double2 a, double2 b, double c2, double Lx, double Ly) {
    // isWithinCutoffSquared, include/particles/base/kernels.cuh:249-260
    double dx = pbc_div(a.x, b.x, Lx);
    double dy = pbc_div(a.y, b.y, Ly);
    return dx * dx + dy * dy < c2;
}

// kernelUpdateCellNeighborList, base/kernels.cu:440-505: 3x3 stencil in the
// reference's order (dy outer, dx inner, ascending sorted index), written to
// the column-major ELL.
__global__ __launch_bounds__(BLOCK) void k_build_cell(Dev d, double rc2, double Lx, double Ly,
                                                      int64_t ncd, int64_t cap) {
    int tile;
    int64_t i = particle_of_thread(tile);
    if (threadIdx.x == 0) d.pmax_nd[tile] = 0.0;
    if (i >= d.N) return;
    double2 xi = d.pos[i];
    int cid = d.cell_of[i];
    int64_t cx = cid % ncd, cy = cid / ncd;
    int cnt = 0;
    for (int64_t oy = -1; oy <= 1; ++oy) {
        int64_t yi = ((cy + oy) % ncd + ncd) % ncd;
        for (int64_t ox = -1; ox <= 1; ++ox) {
            int64_t xx = ((cx + ox) % ncd + ncd) % ncd;
            int64_t c = xx + yi * ncd;
            int s = d.cell_start[c], e = d.cell_start[c + 1];
            for (int j = s; j < e; ++j) {
                if (j == i) continue;
                if (within_div(xi, d.pos[j], rc2, Lx, Ly)) {
                    if (cnt < cap) d.nbr[(int64_t)cnt * d.N + i] = j;
                    cnt++;
                }
            }
        }
    }
    d.nnum[i] = cnt;
    d.lref[i] = xi;
    atomicMax(&d.ctl->max_nnum, cnt);
}

// kernelUpdateNeighborList, base/kernels.cu:256-293: all pairs, ascending j,
// j positions staged through LDS.
__global__ __launch_bounds__(BLOCK) void k_build_verlet(Dev d, double rc2, double Lx, double Ly,
                                                        int64_t cap) {
    __shared__ double2 tp[BLOCK];
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (threadIdx.x == 0) d.pmax_nd[blockIdx.x] = 0.0;
    double2 xi = (i < d.N) ? d.pos[i] : make_double2(0.0, 0.0);
    int cnt = 0;
    for (int64_t base = 0; base < d.N; base += BLOCK) {
        __syncthreads();
        if (base + threadIdx.x < d.N) tp[threadIdx.x] = d.pos[base + threadIdx.x];
        __syncthreads();
        int lim = (int)min((int64_t)BLOCK, d.N - base);
        if (i < d.N) {
            for (int t = 0; t < lim; ++t) {
                int64_t j = base + t;
                if (j == i) continue;
                if (within_div(xi, tp[t], rc2, Lx, Ly)) {
                    if (cnt < cap) d.nbr[(int64_t)cnt * d.N + i] = (int)j;
                    cnt++;
                }
            }
        }
    }
    if (i < d.N) {
        d.nnum[i] = cnt;
        d.lref[i] = xi;
        atomicMax(&d.ctl->max_nnum, cnt);
    }
}

// ------------------------------------------------------------- FIRE loop
// One FIRE iteration (src/integrator/fire.cu:13-46) is three launches:
//   k_fire_move   : [uphill back-step, v=0] + kick(dt/2) + mix + drift(dt)
//   k_fire_force  : zero + forces + kick(dt/2) + chunk partials of P, PE, |F|max
//   k_fire_ctl    : canonical sums, minimizeFire's stop test for this step and
//                   Fire::step's decision for the next (one wave)
__global__ __launch_bounds__(BLOCK) void k_fire_move(Dev d, int method, double thr, double cthr,
                                                     int par) {
    if (d.ctl->halt) return;
    const FireCtl* F = d.fire;
    if (F->done) return;
    double dt = F->dt, alpha = F->alpha;
    int uphill = F->uphill;
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double2 x = d.pos[i], v = d.vel[i], f = d.frc[i];
        if (uphill) {  // updatePositions(-dt/2) with the decreased dt; setVelocitiesToZero
            double bdt = -dt / 2.0;
            x.x += v.x * bdt;
            x.y += v.y * bdt;
            v.x = 0.0;
            v.y = 0.0;
        }
        double dtm = (dt / 2.0) / d.mass[i];
        v.x += f.x * dtm;
        v.y += f.y * dtm;
        mix_one(v, f, alpha);
        x.x += v.x * dt;
        x.y += v.y * dt;
        d.pos[i] = x;
        d.vel[i] = v;
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = x.x - ln.x, dyn = x.y - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = x.x - lc.x, dyc = x.y - lc.y;
        cd = dxc * dxc + dyc * dyc;
    }
    disp_epilogue(d, tile, nd, cd, method, thr, cthr, par);
}

// kick_mode: 0 = forces only (minimizeFire's initial evaluation), 1 = + closing kick
__global__ __launch_bounds__(BLOCK) void k_fire_force(Dev d, PairParams p, int par, int cpos,
                                                      int kick_mode) {
    Ctl* c = d.ctl;
    if (c->halt) return;
    const FireCtl* F = d.fire;
    if (F->done) return;
    if (par >= 0 && c->flag[par]) {
        if (threadIdx.x == 0) {
            c->halt = 1;
            c->halt_pos = cpos;
        }
        return;
    }
    double half = F->dt / 2.0;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double pxv = 0.0, pyv = 0.0, pe = 0.0, f2 = 0.0;
    if (i < d.N) {
        double ax, ay, e;
        disk_row(d, p, i, ax, ay, e);
        double fx = 0.0 + ax, fy = 0.0 + ay;
        d.frc[i] = make_double2(fx, fy);
        pe = 0.0 + e;
        d.pe[i] = pe;
        double2 v = d.vel[i];
        if (kick_mode) {
            double dtm = half / d.mass[i];
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

__global__ void k_fire_ctl(Dev d, int init) {
    if (d.ctl->halt) return;
    FireCtl* F = d.fire;
    if (F->done) return;
    int64_t nc = (d.N + CHUNK - 1) / CHUNK;
    int64_t nb = (d.N + BLOCK - 1) / BLOCK;
    const double* P = d.part;
    double px = wave_final_sum(P, nc);
    double py = wave_final_sum(P + nc, nc);
    double pe = wave_final_sum(P + 2 * nc, nc);
    double fm = 0.0;
    for (int64_t k = threadIdx.x; k < nb; k += WAVE) fm = fmax(fm, P[3 * nc + k]);
    fm = wave_max(fm);
    if (threadIdx.x != 0) return;
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
    // Fire::step decision for the next iteration (src/integrator/fire.cu:14-33)
    double power = px + py;
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
using namespace dpm;

namespace {

template <class T>
int dalloc(dpm_system* s, T** p, int64_t n) {
    size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(n, 1);
    void* q = nullptr;
    DPM_HIP(hipMalloc(&q, bytes));
    DPM_HIP(hipMemset(q, 0, bytes));
    s->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

void dfree(dpm_system* s, void* p) {
    if (!p) return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
    if (it != s->allocs.end()) s->allocs.erase(it);
    (void)hipFree(p);
}

int64_t pow2_at_least(int64_t m) { return (int64_t)std::pow(2.0, std::ceil(std::log2((double)m))); }

inline dim3 g1(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK)); }

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string("kernel launch: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

PairParams params_of(dpm_system* s) {
    if (s->kind == DPM_BUMPY) return pair_params(s, s->e_c, s->n_c);
    return pair_params(s, s->e_c, s->n_c);
}

}  // namespace

namespace dpm {

int launch_chunk_sum(dpm_system* s, int what, double* out_dev) {
    hipLaunchKernelGGL(k_chunk_sum, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, what, s->d.part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(WAVE), 0, s->stream, s->d.part, s->nchunk,
                       out_dev);
    return check_launch();
}

int sum_to_host(dpm_system* s, int what, double* out) {
    double* dv = &s->d.ctl->scal[0];
    if (launch_chunk_sum(s, what, dv)) return -1;
    DPM_HIP(hipMemcpyAsync(out, dv, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

int max_to_host(dpm_system* s, int what, double* out) {
    double* dv = &s->d.ctl->scal[1];
    const double* src = nullptr;
    int64_t n = s->nblk;
    if (what == M_ND) src = s->d.pmax_nd;
    else if (what == M_CD) src = s->d.pmax_cd;
    else {
        double* tmp = s->d.part;
        hipLaunchKernelGGL(k_block_max, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, what, tmp);
        src = tmp;
    }
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(BLOCK), 0, s->stream, src, n, dv);
    if (check_launch()) return -1;
    DPM_HIP(hipMemcpyAsync(out, dv, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

}  // namespace dpm

// defined in dpm_bumpy.hip
namespace dpm {
int bumpy_alloc(dpm_system* s);
int bumpy_forces(dpm_system* s, int accumulate);
int bumpy_vertex_build(dpm_system* s);
int bumpy_cell_reorder_vertices(dpm_system* s);
int bumpy_kick(dpm_system* s, double dt);
int bumpy_drift(dpm_system* s, double dt);
int bumpy_mix(dpm_system* s, double alpha);
int bumpy_stress(dpm_system* s);
int bumpy_contacts(dpm_system* s);
int bumpy_nve_run(dpm_system* s, double dt, int64_t n);
int bumpy_minimize_fire(dpm_system* s, const dpm_fire_params* p, dpm_fire_result* out);
int bumpy_fire_step(dpm_system* s, dpm_fire_state* st);
int bumpy_snapshot(dpm_system* s, bool save);
int bumpy_download(dpm_system* s, int field, void* data, int64_t count, const std::vector<int>& sidx);
int bumpy_upload(dpm_system* s, int field, const void* data, int64_t count, const std::vector<int>& sidx);
}  // namespace dpm

// =========================================================================
//                          neighbour management
// =========================================================================
namespace dpm {

static int ensure_ell(dpm_system* s, int64_t cap) {
    if (cap <= s->phys_cap) return 0;
    dfree(s, s->d.nbr);
    s->d.nbr = nullptr;
    if (dalloc(s, &s->d.nbr, s->N * cap)) return -1;
    s->phys_cap = cap;
    s->d.cap = (int)cap;
    s->gen++;
    return 0;
}

// Particle::updateVerletList / updateCellNeighborList with the reference's
// overflow regrowth (particle.cu:1119-1131, :1320-1331).
int neighbor_build(dpm_system* s) {
    for (int attempt = 0; attempt < 8; ++attempt) {
        DPM_HIP(hipMemsetAsync(&s->d.ctl->max_nnum, 0, sizeof(int), s->stream));
        if (s->method == DPM_NBR_CELL) {
            hipLaunchKernelGGL(k_build_cell, g1(s->N), dim3(BLOCK), 0, s->stream, s->d,
                               s->rc * s->rc, s->box[0], s->box[1], s->ncd, s->phys_cap);
        } else {
            hipLaunchKernelGGL(k_build_verlet, g1(s->N), dim3(BLOCK), 0, s->stream, s->d,
                               s->rc * s->rc, s->box[0], s->box[1], s->phys_cap);
        }
        if (check_launch()) return -1;
        int mx = 0;
        DPM_HIP(hipMemcpyAsync(&mx, &s->d.ctl->max_nnum, sizeof(int), hipMemcpyDeviceToHost,
                               s->stream));
        DPM_HIP(hipStreamSynchronize(s->stream));
        if (mx > s->max_alloc) s->max_alloc = pow2_at_least(mx);
        if (mx > s->phys_cap) {
            if (ensure_ell(s, std::max<int64_t>(s->max_alloc, pow2_at_least(mx)))) {
                s->nbr_ready = false;
                return -1;
            }
            continue;
        }
        s->n_rebuilds++;
        if (s->kind == DPM_BUMPY) return bumpy_vertex_build(s);
        return 0;
    }
    s->nbr_ready = false;
    set_error("neighbor_build: capacity did not converge");
    return -1;
}

// Particle::updateCellList + reorderParticleData (particle.cu:1292-1317)
int cell_rebuild(dpm_system* s) {
    const Dev& d = s->d;
    hipLaunchKernelGGL(k_fill_int, g1(s->ncells), dim3(BLOCK), 0, s->stream, d.cell_cnt, 0,
                       s->ncells);
    hipLaunchKernelGGL(k_cell_bin, g1(s->N), dim3(BLOCK), 0, s->stream, d, s->cell_size, s->ncd);
    hipLaunchKernelGGL(k_cell_scan, dim3(1), dim3(1024), 0, s->stream, d, s->ncells);
    hipLaunchKernelGGL(k_cell_scatter, g1(s->N), dim3(BLOCK), 0, s->stream, d);
    hipLaunchKernelGGL(k_cell_sort, g1(s->ncells), dim3(BLOCK), 0, s->stream, d, s->ncells);
    hipLaunchKernelGGL(k_reorder_gather, g1(s->N), dim3(BLOCK), 0, s->stream, d, s->sp);
    if (s->kind == DPM_BUMPY && bumpy_cell_reorder_vertices(s)) return -1;
    hipLaunchKernelGGL(k_reorder_copyback, g1(s->N), dim3(BLOCK), 0, s->stream, d, s->sp);
    hipLaunchKernelGGL(k_fill_double, g1(s->nblk), dim3(BLOCK), 0, s->stream, d.pmax_cd, 0.0,
                       s->nblk);
    s->n_cell_rebuilds++;
    return check_launch();
}

// Particle::checkForNeighborUpdate (particle.cu:1137-1161): returns 0 none,
// 1 neighbour rebuild, 2 cell + neighbour rebuild.
int rebuild_kind(dpm_system* s, int* kind) {
    *kind = 0;
    if (s->method != DPM_NBR_VERLET && s->method != DPM_NBR_CELL) return 0;
    double mn = 0.0, mc = 0.0;
    if (max_to_host(s, M_ND, &mn)) return -1;
    if (s->method == DPM_NBR_CELL) {
        if (max_to_host(s, M_CD, &mc)) return -1;
        if (3.0 * mc > s->cell_thr_sq) *kind = 2;
        else if (3.0 * mn > s->thr_sq) *kind = 1;
    } else if (3.0 * mn > s->thr_sq) {
        *kind = 1;
    }
    return 0;
}

int do_rebuild(dpm_system* s, int kind) {
    if (kind == 2 && cell_rebuild(s)) return -1;
    if (kind >= 1) return neighbor_build(s);
    return 0;
}

int check_and_rebuild(dpm_system* s, int* kind_out) {
    int kind = 0;
    if (rebuild_kind(s, &kind)) return -1;
    if (do_rebuild(s, kind)) return -1;
    if (kind_out) *kind_out = kind;
    return 0;
}

}  // namespace dpm

// =========================================================================
//                         timing (per-launch HIP events)
// =========================================================================
namespace {

hipEvent_t pool_event(dpm_system* s) {
    if (s->ev_next >= s->ev_pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_next++];
}

struct Timed {
    dpm_system* s;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(dpm_system* s_, int cls_) : s(s_), cls(cls_) {
        if (!s->timing) return;
        a = pool_event(s);
        b = pool_event(s);
        (void)hipEventRecord(a, s->stream);
    }
    ~Timed() {
        if (!s->timing) return;
        (void)hipEventRecord(b, s->stream);
        s->ev_pending.push_back({cls, {a, b}});
    }
};

void collect_timing(dpm_system* s) {
    if (s->ev_pending.empty()) return;
    (void)hipStreamSynchronize(s->stream);
    for (auto& pe : s->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            s->ktime_ms[pe.first] += ms;
            s->klaunch[pe.first] += 1;
        }
    }
    s->ev_pending.clear();
    s->ev_next = 0;
}

}  // namespace

// =========================================================================
//                               NVE loop
// =========================================================================
namespace dpm {

static void nve_step_launch(dpm_system* s, const PairParams& p, double dt, int par, int cpos) {
    {
        Timed t(s, DPM_K_DRIFT);
        hipLaunchKernelGGL(k_nve_kick_drift, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, 0.5 * dt,
                           dt, s->method, s->thr_sq, s->cell_thr_sq, par);
    }
    {
        Timed t(s, DPM_K_FORCE);
        hipLaunchKernelGGL(k_nve_force_kick, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, p,
                           0.5 * dt, par, cpos);
    }
}

static int ensure_graph(dpm_system* s, Graph& g, double dt, int steps,
                        void (*fill)(dpm_system*, double, int)) {
    if (g.exec && g.gen == s->gen && g.dt == dt && g.steps == steps) return 0;
    if (g.exec) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
        g.exec = nullptr;
    }
    DPM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    fill(s, dt, steps);
    DPM_HIP(hipStreamEndCapture(s->stream, &g.graph));
    DPM_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    g.gen = s->gen;
    g.dt = dt;
    g.steps = steps;
    return 0;
}

static void nve_fill(dpm_system* s, double dt, int steps) {
    PairParams p = params_of(s);
    for (int c = 0; c < steps; ++c) nve_step_launch(s, p, dt, c & 1, c);
}

static int clear_ctl(dpm_system* s) {
    DPM_HIP(hipMemsetAsync(s->d.ctl, 0, offsetof(Ctl, max_nnum), s->stream));
    return 0;
}

// Drives `n` steps of a chunked, halt-on-rebuild device loop.  `launch`
// enqueues k steps (graph or direct); `resume` finishes a halted step after the
// host-side rebuild; `finished` tells whether the device loop ended early
// (FIRE convergence).
struct LoopOps {
    int (*launch)(dpm_system*, double, int64_t);
    int (*resume)(dpm_system*, double);
};

static int run_chunked(dpm_system* s, double dt, int64_t n, const LoopOps& ops) {
    if (clear_ctl(s)) return -1;
    struct Inflight {
        int64_t k;
        int slot;
    };
    static thread_local hipEvent_t evs[2] = {nullptr, nullptr};
    if (!evs[0]) {
        DPM_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
        DPM_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
    }
    std::deque<Inflight> q;
    int64_t done = 0, launched = 0;
    while (done < n) {
        while (q.size() < 2 && launched < n) {
            int64_t k = std::min<int64_t>(s->chunk, n - launched);
            int slot = q.empty() ? 0 : (q.front().slot ^ 1);
            if (ops.launch(s, dt, k)) return -1;
            DPM_HIP(hipMemcpyAsync(&s->hctl[slot], s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                                   s->stream));
            DPM_HIP(hipEventRecord(evs[slot], s->stream));
            q.push_back({k, slot});
            launched += k;
        }
        Inflight f = q.front();
        q.pop_front();
        DPM_HIP(hipEventSynchronize(evs[f.slot]));
        if (s->timing) collect_timing(s);
        Ctl hc = s->hctl[f.slot];
        if (hc.halt) {
            DPM_HIP(hipStreamSynchronize(s->stream));
            done += hc.halt_pos;
            int kind = 0;
            if (rebuild_kind(s, &kind)) return -1;
            if (kind == 0) kind = 1;  // the flag was raised by the same test
            {
                Timed t(s, DPM_K_REBUILD);
                if (do_rebuild(s, kind)) return -1;
            }
            if (clear_ctl(s)) return -1;
            if (ops.resume(s, dt)) return -1;
            done += 1;
            launched = done;
            q.clear();
            Ctl after;
            DPM_HIP(hipMemcpyAsync(&after, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                                   s->stream));
            DPM_HIP(hipStreamSynchronize(s->stream));
            if (after.done) break;
            continue;
        }
        done += f.k;
        if (hc.done) break;
    }
    DPM_HIP(hipStreamSynchronize(s->stream));
    if (s->timing) collect_timing(s);
    return 0;
}

static int nve_launch(dpm_system* s, double dt, int64_t k) {
    if (s->use_graph && !s->timing && k == s->chunk && (k % 2) == 0) {
        if (ensure_graph(s, s->nve_graph, dt, (int)k, nve_fill)) return -1;
        DPM_HIP(hipGraphLaunch(s->nve_graph.exec, s->stream));
        return 0;
    }
    PairParams p = params_of(s);
    for (int64_t c = 0; c < k; ++c) nve_step_launch(s, p, dt, (int)(c & 1), (int)c);
    return check_launch();
}

static int nve_resume(dpm_system* s, double dt) {
    PairParams p = params_of(s);
    Timed t(s, DPM_K_FORCE);
    hipLaunchKernelGGL(k_nve_force_kick, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, p, 0.5 * dt,
                       -1, 0);
    return check_launch();
}

int nve_run(dpm_system* s, double dt, int64_t n) {
    if (s->kind == DPM_BUMPY) return bumpy_nve_run(s, dt, n);
    if (!s->nbr_ready) {
        set_error("nve_run: neighbour list not initialised");
        return -1;
    }
    LoopOps ops{nve_launch, nve_resume};
    return run_chunked(s, dt, n, ops);
}

// =========================================================================
//                               FIRE loop
// =========================================================================
static void fire_step_launch(dpm_system* s, const PairParams& p, int par, int cpos) {
    {
        Timed t(s, DPM_K_DRIFT);
        hipLaunchKernelGGL(k_fire_move, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, s->method,
                           s->thr_sq, s->cell_thr_sq, par);
    }
    {
        Timed t(s, DPM_K_FORCE);
        hipLaunchKernelGGL(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, s->stream, s->d,
                           p, par, cpos, 1);
    }
    hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, s->stream, s->d, 0);
}

static void fire_fill(dpm_system* s, double, int steps) {
    PairParams p = params_of(s);
    for (int c = 0; c < steps; ++c) fire_step_launch(s, p, c & 1, c);
}

static int fire_launch(dpm_system* s, double dt, int64_t k) {
    if (s->use_graph && !s->timing && k == s->chunk && (k % 2) == 0) {
        if (ensure_graph(s, s->fire_graph, 0.0, (int)k, fire_fill)) return -1;
        DPM_HIP(hipGraphLaunch(s->fire_graph.exec, s->stream));
        return 0;
    }
    PairParams p = params_of(s);
    for (int64_t c = 0; c < k; ++c) fire_step_launch(s, p, (int)(c & 1), (int)c);
    return check_launch();
}

static int fire_resume(dpm_system* s, double) {
    PairParams p = params_of(s);
    {
        Timed t(s, DPM_K_FORCE);
        hipLaunchKernelGGL(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, s->stream, s->d,
                           p, -1, 0, 1);
    }
    hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, s->stream, s->d, 0);
    return check_launch();
}

int minimize_fire(dpm_system* s, const dpm_fire_params* prm, dpm_fire_result* out) {
    if (s->kind == DPM_BUMPY) return bumpy_minimize_fire(s, prm, out);
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
    DPM_HIP(hipMemcpyAsync(s->d.fire, &f, sizeof(FireCtl), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_zero_velocities, g1(s->N), dim3(BLOCK), 0, s->stream, s->d);
    if (check_and_rebuild(s, nullptr)) return -1;
    if (clear_ctl(s)) return -1;
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_fire_force, dim3((unsigned)s->nblk), dim3(BLOCK), 0, s->stream, s->d, p,
                       -1, 0, 0);
    hipLaunchKernelGGL(k_fire_ctl, dim3(1), dim3(WAVE), 0, s->stream, s->d, 1);
    if (check_launch()) return -1;
    LoopOps ops{fire_launch, fire_resume};
    if (prm->max_steps > 0 && run_chunked(s, 0.0, prm->max_steps, ops)) return -1;
    hipLaunchKernelGGL(k_zero_velocities, g1(s->N), dim3(BLOCK), 0, s->stream, s->d);
    DPM_HIP(hipMemcpyAsync(&f, s->d.fire, sizeof(FireCtl), hipMemcpyDeviceToHost, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    if (out) {
        out->steps = std::min<int64_t>(f.step, prm->max_steps);
        out->stopped = f.stopped;
        out->converged = f.converged;
        out->avg_pe = f.avg_pe;
        out->dt = f.dt;
        out->alpha = f.alpha;
        double f2 = 0.0;
        if (max_to_host(s, M_F2, &f2)) return -1;
        out->fmax_sq = f2;
    }
    return check_launch();
}

// One Fire::step with host-held state (src/integrator/fire.cu:13-46), built from
// the single kernels; used by state-machine trace tests.
int fire_step_host(dpm_system* s, dpm_fire_state* st);

}  // namespace dpm

// =========================================================================
//                                 C ABI
// =========================================================================
using namespace dpm;

namespace {

// Host<->device copies ordered on the system stream (the stream is non-blocking,
// so the legacy null stream would not wait for kernels queued on it).
int scopy(dpm_system* s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    DPM_HIP(hipMemcpyAsync(dst, src, bytes, kind, s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}

int fetch_sidx(dpm_system* s, std::vector<int>& sidx) {
    sidx.resize(s->N);
    DPM_HIP(hipMemcpyAsync(sidx.data(), s->d.sidx, sizeof(int) * s->N, hipMemcpyDeviceToHost,
                           s->stream));
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
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

}  // namespace

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
    int ndev = dpm_device_count();
    if (ndev <= 0) {
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
    ok = ok && !dalloc(s, &d.pos, n) && !dalloc(s, &d.vel, n) && !dalloc(s, &d.frc, n) &&
         !dalloc(s, &d.lref, n) && !dalloc(s, &d.cref, n) && !dalloc(s, &d.rad, n) &&
         !dalloc(s, &d.mass, n) && !dalloc(s, &d.pe, n) && !dalloc(s, &d.ke, n) &&
         !dalloc(s, &d.sidx, n) && !dalloc(s, &d.nnum, n) && !dalloc(s, &d.cell_of, n) &&
         !dalloc(s, &d.cell_tmp, n) && !dalloc(s, &d.perm, n) && !dalloc(s, &d.pmax_nd, s->nblk) &&
         !dalloc(s, &d.pmax_cd, s->nblk) &&
         !dalloc(s, &d.part, 4 * s->nchunk + s->nblk + 64) && !dalloc(s, &d.ctl, 1) &&
         !dalloc(s, &d.fire, 1) && !dalloc(s, &s->sp.pos, n) && !dalloc(s, &s->sp.vel, n) &&
         !dalloc(s, &s->sp.frc, n) && !dalloc(s, &s->sp.rad, n) && !dalloc(s, &s->sp.mass, n) &&
         !dalloc(s, &s->sp.sidx, n);
    if (ok) {
        std::vector<int> id(n);
        for (int64_t k = 0; k < n; ++k) id[k] = (int)k;
        ok = hipMemcpy(d.sidx, id.data(), sizeof(int) * n, hipMemcpyHostToDevice) == hipSuccess;
    }
    ok = ok && hipHostMalloc((void**)&s->hctl, 2 * sizeof(Ctl), hipHostMallocDefault) == hipSuccess;
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
    for (auto* g : {&s->nve_graph, &s->fire_graph})
        if (g->exec) {
            (void)hipGraphExecDestroy(g->exec);
            (void)hipGraphDestroy(g->graph);
        }
    for (void* p : s->allocs) (void)hipFree(p);
    for (void* p : s->snap_allocs) (void)hipFree(p);
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
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
    return 0;
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

static int field_d2(dpm_system* s, int f, double2** p, int* comp) {
    switch (f) {
        case DPM_F_X: *p = s->d.pos; *comp = 0; return 1;
        case DPM_F_Y: *p = s->d.pos; *comp = 1; return 1;
        case DPM_F_VX: *p = s->d.vel; *comp = 0; return 1;
        case DPM_F_VY: *p = s->d.vel; *comp = 1; return 1;
        case DPM_F_FX: *p = s->d.frc; *comp = 0; return 1;
        case DPM_F_FY: *p = s->d.frc; *comp = 1; return 1;
    }
    return 0;
}

static double* field_d(dpm_system* s, int f) {
    switch (f) {
        case DPM_F_RAD: return s->d.rad;
        case DPM_F_MASS: return s->d.mass;
        case DPM_F_PE: return s->d.pe;
        case DPM_F_KE: return s->d.ke;
        case DPM_F_STRESS_XX: return s->d.sxx;
        case DPM_F_STRESS_XY: return s->d.sxy;
        case DPM_F_STRESS_YX: return s->d.syx;
        case DPM_F_STRESS_YY: return s->d.syy;
    }
    return nullptr;
}

int dpm_upload(dpm_system* s, int f, const void* data, int64_t count) {
    std::vector<int> sidx;
    if (fetch_sidx(s, sidx)) return -1;
    if (s->kind == DPM_BUMPY &&
        ((f >= DPM_F_ANGLE && f <= DPM_F_MOI) || f >= DPM_F_VERTEX_X))
        return bumpy_upload(s, f, data, count, sidx);
    if (count != s->N) {
        set_error("dpm_upload: count mismatch");
        return -1;
    }
    const double* src = (const double*)data;
    double2* p2;
    int comp;
    if (field_d2(s, f, &p2, &comp)) {
        std::vector<double2> h(s->N);
        if (scopy(s, h.data(), p2, sizeof(double2) * s->N, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->N; ++k) (comp ? h[k].y : h[k].x) = src[sidx[k]];
        if (scopy(s, p2, h.data(), sizeof(double2) * s->N, hipMemcpyHostToDevice)) return -1;
        return 0;
    }
    if (double* p = field_d(s, f)) {
        std::vector<double> h(s->N);
        for (int64_t k = 0; k < s->N; ++k) h[k] = src[sidx[k]];
        if (scopy(s, p, h.data(), sizeof(double) * s->N, hipMemcpyHostToDevice)) return -1;
        return 0;
    }
    set_error("dpm_upload: field not uploadable");
    return -1;
}

int dpm_download(dpm_system* s, int f, void* data, int64_t count) {
    std::vector<int> sidx;
    if (fetch_sidx(s, sidx)) return -1;
    if (s->kind == DPM_BUMPY &&
        ((f >= DPM_F_ANGLE && f <= DPM_F_MOI) || f >= DPM_F_VERTEX_X))
        return bumpy_download(s, f, data, count, sidx);
    if (count != s->N) {
        set_error("dpm_download: count mismatch");
        return -1;
    }
    double2* p2;
    int comp;
    if (field_d2(s, f, &p2, &comp)) {
        std::vector<double2> h(s->N);
        if (scopy(s, h.data(), p2, sizeof(double2) * s->N, hipMemcpyDeviceToHost)) return -1;
        double* dst = (double*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = comp ? h[k].y : h[k].x;
        return 0;
    }
    if (double* p = field_d(s, f)) {
        std::vector<double> h(s->N);
        if (scopy(s, h.data(), p, sizeof(double) * s->N, hipMemcpyDeviceToHost)) return -1;
        double* dst = (double*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    int* pi = nullptr;
    if (f == DPM_F_NUM_NEIGHBORS) pi = s->d.nnum;
    if (f == DPM_F_CONTACTS) pi = s->d.contacts;
    if (f == DPM_F_STATIC_INDEX) {
        int64_t* dst = (int64_t*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[k] = sidx[k];
        return 0;
    }
    if (pi) {
        std::vector<int> h(s->N);
        if (scopy(s, h.data(), pi, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
        int64_t* dst = (int64_t*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    set_error("dpm_download: unknown or unallocated field");
    return -1;
}

int dpm_neighbor_config(dpm_system* s, int method, double cutoff, double thr_sq, int64_t ncd,
                        double cell_thr_sq) {
    if (method != DPM_NBR_ALL && method != DPM_NBR_VERLET && method != DPM_NBR_CELL) {
        set_error("dpm_neighbor_config: invalid method");
        return -1;
    }
    s->method = method;
    s->max_alloc = 4;
    s->vmax_alloc = 4;
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
        if (s->ncells > s->ncells_alloc) {
            dfree(s, s->d.cell_start);
            dfree(s, s->d.cell_cnt);
            dfree(s, s->d.cell_cur);
            if (dalloc(s, &s->d.cell_start, s->ncells + 1) ||
                dalloc(s, &s->d.cell_cnt, s->ncells) || dalloc(s, &s->d.cell_cur, s->ncells))
                return -1;
            s->ncells_alloc = s->ncells;
        }
    }
    int64_t cap = 16;
    if (method == DPM_NBR_ALL) {
        // Particle::initAllToAllList (particle.cu:1186-1200)
        s->max_alloc = std::max<int64_t>(1, s->N - 1);
        cap = s->max_alloc;
        double diag = std::sqrt(s->box[0] * s->box[0] + s->box[1] * s->box[1]);
        s->rc = 10.0 * diag;
        s->thr_sq = s->rc * s->rc;
    }
    if (ensure_ell(s, cap)) return -1;
    s->gen++;
    return 0;
}

int dpm_neighbor_init(dpm_system* s) {
    if (s->method == DPM_NBR_NONE) {
        set_error("dpm_neighbor_init: no method configured");
        return -1;
    }
    // initVerletListVariables (particle.cu:1171-1178): references and displacements zeroed
    DPM_HIP(hipMemsetAsync(s->d.lref, 0, sizeof(double2) * s->N, s->stream));
    DPM_HIP(hipMemsetAsync(s->d.cref, 0, sizeof(double2) * s->N, s->stream));
    DPM_HIP(hipMemsetAsync(s->d.pmax_nd, 0, sizeof(double) * s->nblk, s->stream));
    DPM_HIP(hipMemsetAsync(s->d.pmax_cd, 0, sizeof(double) * s->nblk, s->stream));
    if (s->method == DPM_NBR_CELL && cell_rebuild(s)) return -1;
    if (neighbor_build(s)) return -1;
    s->nbr_ready = true;
    DPM_HIP(hipStreamSynchronize(s->stream));
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
    DPM_HIP(hipMemcpy(nb.data(), s->d.nbr, sizeof(int) * s->N * s->phys_cap,
                      hipMemcpyDeviceToHost));
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
    if (ma) *ma = s->max_alloc;
    if (tot) {
        std::vector<int> nn(s->N);
        if (scopy(s, nn.data(), s->d.nnum, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
        int64_t t = 0;
        for (int v : nn) t += v;
        *tot = t;
    }
    if (nr) *nr = s->n_rebuilds;
    if (ncr) *ncr = s->n_cell_rebuilds;
    return 0;
}

int dpm_zero_forces(dpm_system* s) {
    hipLaunchKernelGGL(k_zero_forces, g1(s->N), dim3(BLOCK), 0, s->stream, s->d);
    if (s->kind == DPM_BUMPY) return bumpy_forces(s, -1);
    return check_launch();
}

int dpm_compute_forces(dpm_system* s) {
    if (require_nbr(s, "dpm_compute_forces")) return -1;
    if (s->kind == DPM_BUMPY) return bumpy_forces(s, 1);
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_disk_forces_acc, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, p);
    return check_launch();
}

int dpm_kick(dpm_system* s, double dt) {
    if (s->kind == DPM_BUMPY) return bumpy_kick(s, dt);
    hipLaunchKernelGGL(k_kick, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, dt);
    return check_launch();
}

int dpm_drift(dpm_system* s, double dt) {
    if (s->kind == DPM_BUMPY) return bumpy_drift(s, dt);
    hipLaunchKernelGGL(k_drift, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, dt);
    return check_launch();
}

int dpm_mix(dpm_system* s, double alpha) {
    if (s->kind == DPM_BUMPY) return bumpy_mix(s, alpha);
    hipLaunchKernelGGL(k_mix, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, alpha);
    return check_launch();
}

int dpm_zero_velocities(dpm_system* s) {
    hipLaunchKernelGGL(k_zero_velocities, g1(s->N), dim3(BLOCK), 0, s->stream, s->d);
    return check_launch();
}

int dpm_nve_run(dpm_system* s, double dt, int64_t n) {
    if (n < 0) {
        set_error("dpm_nve_run: negative step count");
        return -1;
    }
    return nve_run(s, dt, n);
}

static int64_t ndof_of(dpm_system* s) {
    if (s->kind == DPM_BUMPY) return s->rotation ? 3 * s->N : 2 * s->N;
    return 2 * s->N;
}

int dpm_reduce(dpm_system* s, int what, double* out) {
    switch (what) {
        case DPM_RED_PE: return sum_to_host(s, S_PE, out);
        case DPM_RED_KE:
        case DPM_RED_TEMPERATURE: {
            hipLaunchKernelGGL(k_kinetic, g1(s->N), dim3(BLOCK), 0, s->stream, s->d,
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
            int w = what == DPM_RED_STRESS_XX ? S_SXX
                    : what == DPM_RED_STRESS_XY ? S_SXY
                    : what == DPM_RED_STRESS_YX ? S_SYX : S_SYY;
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
    if (s->kind == DPM_BUMPY) {
        set_error("dpm_adam_step: rigid bumpy Adam is not supported");
        return -1;
    }
    if (ensure_adam(s)) return -1;
    // Adam::minimize (src/integrator/adam.cu:11-18)
    if (dpm_zero_forces(s) || check_and_rebuild(s, nullptr) || dpm_compute_forces(s)) return -1;
    double omb1 = 1 - std::pow(b1, step + 1);
    double omb2 = 1 - std::pow(b2, step + 1);
    hipLaunchKernelGGL(k_adam_update, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, alpha, b1, b2,
                       omb1, omb2, eps);
    return check_launch();
}

int dpm_minimize_adam(dpm_system* s, double alpha, double b1, double b2, double eps, double pe_t,
                      double diff_t, int64_t max_steps, int64_t* steps_out) {
    if (require_nbr(s, "dpm_minimize_adam")) return -1;
    if (ensure_adam(s)) return -1;
    // minimizeAdam (src/routines/minimization.cu:3-33)
    DPM_HIP(hipMemsetAsync(s->d.m1, 0, sizeof(double2) * s->N, s->stream));
    DPM_HIP(hipMemsetAsync(s->d.m2, 0, sizeof(double2) * s->N, s->stream));
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
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_disk_stress, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, p,
                       s->box[0] * s->box[1]);
    return check_launch();
}

int dpm_count_contacts(dpm_system* s) {
    if (require_nbr(s, "dpm_count_contacts")) return -1;
    if (ensure_stress(s)) return -1;
    if (s->kind == DPM_BUMPY) return bumpy_contacts(s);
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_disk_contacts, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, p);
    return check_launch();
}

int dpm_scale_positions(dpm_system* s, double f) {
    hipLaunchKernelGGL(k_scale_positions, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, f);
    return check_launch();
}

// Particle::scaleBox (particle.cu:1057-1064)
int dpm_scale_box(dpm_system* s, double f) {
    if (dpm_scale_positions(s, f)) return -1;
    double side = std::pow(s->box[0] * s->box[1], 1.0 / 2);
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
    if (al(&s->snap.pos, n) || al(&s->snap.vel, n) || al(&s->snap.frc, n) ||
        al(&s->snap.rad, n) || al(&s->snap.mass, n) || al(&s->snap.sidx, n)) {
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
    DPM_HIP(cp(s->snap.pos, s->d.pos, sizeof(double2) * n));
    DPM_HIP(cp(s->snap.vel, s->d.vel, sizeof(double2) * n));
    DPM_HIP(cp(s->snap.frc, s->d.frc, sizeof(double2) * n));
    DPM_HIP(cp(s->snap.rad, s->d.rad, sizeof(double) * n));
    DPM_HIP(cp(s->snap.mass, s->d.mass, sizeof(double) * n));
    DPM_HIP(cp(s->snap.sidx, s->d.sidx, sizeof(int) * n));
    if (s->kind == DPM_BUMPY && bumpy_snapshot(s, true)) return -1;
    s->snap_box[0] = s->box[0];
    s->snap_box[1] = s->box[1];
    s->snap_valid = true;
    return 0;
}

// Particle::revertToLastState (particle.cu:977-991) + a full cell rebuild.
int dpm_snapshot_restore(dpm_system* s) {
    if (!s->snap_valid) {
        set_error("dpm_snapshot_restore: no snapshot");
        return -1;
    }
    auto cp = [&](void* dst, const void* src, size_t b) {
        return hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s->stream);
    };
    int64_t n = s->N;
    DPM_HIP(cp(s->d.pos, s->snap.pos, sizeof(double2) * n));
    DPM_HIP(cp(s->d.vel, s->snap.vel, sizeof(double2) * n));
    DPM_HIP(cp(s->d.frc, s->snap.frc, sizeof(double2) * n));
    DPM_HIP(cp(s->d.rad, s->snap.rad, sizeof(double) * n));
    DPM_HIP(cp(s->d.mass, s->snap.mass, sizeof(double) * n));
    DPM_HIP(cp(s->d.sidx, s->snap.sidx, sizeof(int) * n));
    if (s->kind == DPM_BUMPY && bumpy_snapshot(s, false)) return -1;
    if (dpm_set_box(s, s->snap_box[0], s->snap_box[1])) return -1;
    if (s->nbr_ready) {
        if (s->method == DPM_NBR_CELL && cell_rebuild(s)) return -1;
        if (neighbor_build(s)) return -1;
    }
    return 0;
}

// Particle::removeMeanVelocities (particle.cu:865-869)
int dpm_remove_mean_velocity(dpm_system* s) {
    double a, b;
    if (sum_to_host(s, S_VX, &a) || sum_to_host(s, S_VY, &b)) return -1;
    double mx = a / s->N, my = b / s->N;
    hipLaunchKernelGGL(k_add_velocity, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, mx, my, 0.0);
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
    hipLaunchKernelGGL(k_add_velocity, g1(s->N), dim3(BLOCK), 0, s->stream, s->d, 0.0, 0.0, f);
    return check_launch();
}

// Pressure-bisection jamming, scripts/runArthurJamWithFire.cpp:36-106 (see oracle jam()).
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
        case DPM_OPT_TIMING: s->timing = v != 0; return 0;
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
