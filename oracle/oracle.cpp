// oracle.cpp -- CPU restatement of dpCUDA-v2's particle hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never the thing
// measured or shipped.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it.
//
// Every routine below follows a reference function line by line; the
// reference file:line is cited in the comment above it.  Paths are relative
// to the reference repository root.  Build: g++ -O2 -fopenmp -ffp-contract=off
// (no FMA contraction, so the arithmetic is the plain IEEE sequence written).
//
// Deliberate, documented differences from the reference (each is a defect in
// the reference that scrambles state; the HIP library behaves the same way as
// this oracle):
//   * Adam moments are permuted with the particles on a cell reorder (the
//     reference swaps them with un-gathered temp buffers,
//     src/particles/base/particle.cu:1302-1303).
//   * restoring a snapshot forces a full cell rebuild (the reference rebuilds
//     only the neighbour list against a cell structure that may belong to a
//     different permutation, particle.cu:987-991).
//   * snapshots also hold moments of inertia and vertex masses (the reference
//     omits them, rigid_bumpy.cu:1144-1179, so a reorder in between leaves
//     them permuted differently from everything else).
//   * scalar sums use one fixed reduction tree (canonical_sum below) instead
//     of CUB's; the HIP kernels use the same tree, so GPU and oracle sums are
//     bit-identical.  The reference's CUB tree order is not reproducible here.
//   * sin/cos of the rigid-body rotations come from one shared implementation
//     (dpm_sincos, Cody-Waite reduction + Taylor polynomials, within 2.5 ulp of
//     the exact values) compiled into both the HIP kernels and this oracle; the
//     reference calls CUDA's sin/cos, which no CPU library reproduces bit for bit.
#include "oracle.h"
#include "../dpcuda-v2_amd/csrc/dpm_trig.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

// Loops over fewer elements than this run on one thread: every parallel loop here is
// element-wise (or a maximum), so the thread count never changes a result, and small
// systems (the batched and jamming tests) do not pay a parallel region per kernel.
constexpr long OMP_MIN_N = 2048;

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return -1; }

constexpr int CH = 64;  // chunk width of the canonical reduction tree

// Pairwise halving tree over 64 values: v[t] += v[t+s] for s = 32..1.
double tree64(double* v) {
    for (int s = CH / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) v[t] = v[t] + v[t + s];
    return v[0];
}

// Canonical sum: 64-wide chunks reduced by tree64; chunk partials P[c] are
// accumulated into 64 lanes (lane t sums P[t], P[t+64], ... in order, from
// 0.0), and the lanes reduced by tree64.  The HIP reductions realise exactly
// this order (wave64 shuffles + last-block final stage).
template <class F>
double canonical_sum_f(long n, F f) {
    long nc = (n + CH - 1) / CH;
    std::vector<double> P(nc);
#pragma omp parallel for schedule(static) if (nc >= OMP_MIN_N)
    for (long c = 0; c < nc; ++c) {
        double v[CH];
        for (int t = 0; t < CH; ++t) {
            long i = c * CH + t;
            v[t] = i < n ? f(i) : 0.0;
        }
        P[c] = tree64(v);
    }
    double A[CH];
    for (int t = 0; t < CH; ++t) {
        double s = 0.0;
        for (long k = t; k < nc; k += CH) s = s + P[k];
        A[t] = s;
    }
    return tree64(A);
}

template <class F>
double max_f(long n, F f) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) if (n >= OMP_MIN_N)
    for (long i = 0; i < n; ++i) m = std::max(m, f(i));
    return m;
}

// ------------------------------------------------------------------ helpers
// mod: include/particles/base/kernels.cuh:53-55
inline long pmod(long a, long b) { return (a % b + b) % b; }

// pbcDistance: include/particles/base/kernels.cuh:65-72 (round = half away from zero)
inline double pbc(double a, double b, double L) {
    double d = a - b;
    return d - L * std::round(d / L);
}

// overlap^(n_c-1); n_c == 2 gives exactly overlap (x^1 == x), else libm pow.
inline double powc(double ov, double p) { return p == 1.0 ? ov : std::pow(ov, p); }

// calcPointPointInteraction: include/particles/base/kernels.cuh:154-183
inline double pair_law(double px, double py, double r, double ox, double oy, double orad,
                       double Lx, double Ly, double e_c, double n_c, double& fx, double& fy) {
    double dx = pbc(px, ox, Lx);
    double dy = pbc(py, oy, Ly);
    double rs = r + orad;
    double d2 = dx * dx + dy * dy;
    if (d2 >= rs * rs) { fx = 0.0; fy = 0.0; return 0.0; }
    double d = std::sqrt(d2);
    double ov = 1.0 - (d / rs);
    double op = powc(ov, n_c - 1.0);
    double e = e_c * ov * op / (2.0 * n_c);
    double fm = e_c * op / (rs * d);
    fx = fm * dx;
    fy = fm * dy;
    return e;
}

// isWithinCutoffSquared: include/particles/base/kernels.cuh:249-260
inline bool within(double ax, double ay, double bx, double by, double c2, double Lx, double Ly) {
    double dx = pbc(ax, bx, Lx);
    double dy = pbc(ay, by, Ly);
    return dx * dx + dy * dy < c2;
}

// calcTorque: include/particles/base/kernels.cuh:237-241
inline double torque_of(double fx, double fy, double px, double py, double cx, double cy) {
    double dx = px - cx, dy = py - cy;
    return fy * dx - fx * dy;
}

template <class T>
void permute(std::vector<T>& a, const std::vector<long>& old_of_new) {
    if (a.empty()) return;
    std::vector<T> t(a.size());
    for (size_t i = 0; i < old_of_new.size(); ++i) t[i] = a[old_of_new[i]];
    a.swap(t);
}

long pow2_at_least(long m) {  // std::pow(2, std::ceil(std::log2(m))) as in particle.cu:1124
    return (long)std::pow(2.0, std::ceil(std::log2((double)m)));
}

}  // namespace

struct orc_system {
    int kind = ORC_DISK;
    long N = 0, Nv = 0;
    double box[2] = {1.0, 1.0};
    double e_c = 1.0, n_c = 2.0;
    // per particle (internal order)
    std::vector<double> x, y, vx, vy, fx, fy, r, m, pe, ke;
    std::vector<double> lnx, lny, lcx, lcy, nd2, cd2;
    std::vector<double> m1x, m1y, m2x, m2y, m1a, m2a;  // m1a, m2a: rigid angle moments
    std::vector<double> sxx, sxy, syx, syy;
    std::vector<long> contacts, vcontacts;
    // neighbour structures (reference layout: row-major, -1 padded, int64)
    std::vector<long> nbr, nnum;
    long max_alloc = 4;
    std::vector<long> cell_index, pidx, sidx, cell_start;
    int method = ORC_NBR_NONE;
    double rc = -1, thr_sq = -1, cell_thr_sq = -1, cell_size = -1;
    long ncd = -1, ncells = -1;
    long n_rebuilds = 0, n_cell_rebuilds = 0;
    // rigid bumpy
    bool rotation = true;
    double rv = 0.0, rc_v = -1, pcut_v = -1;
    std::vector<double> th, w, tq, moi, lpx, lpy, dlx, dly, dth;
    std::vector<long> nverts, vstart;
    std::vector<double> vxp, vyp, vm, vfx, vfy, vtq, vpe;
    std::vector<long> vpart, svidx;
    std::vector<long> vnbr, vnum;
    long vmax_alloc = 4;
    // pair analytics: pair_fields blocks of N * pair_row values (ids as int64 bits)
    std::vector<double> pairs;
    long pair_row = 0;
    int pair_fields = 0;
    // snapshot
    struct Snap {
        bool valid = false;
        std::vector<double> x, y, fx, fy, vx, vy, r, m, th, w, tq, moi, vxp, vyp, vm;
        std::vector<long> pidx, sidx, nverts, vstart, vpart, svidx;
        double box[2];
    } snap;

    long ndof() const { return kind == ORC_BUMPY ? (rotation ? 3 * N : 2 * N) : 2 * N; }
};

namespace {

void resize_particles(orc_system* s) {
    long N = s->N;
    for (auto* v : {&s->x, &s->y, &s->vx, &s->vy, &s->fx, &s->fy, &s->r, &s->m, &s->pe, &s->ke,
                    &s->lnx, &s->lny, &s->lcx, &s->lcy, &s->nd2, &s->cd2, &s->m1x, &s->m1y,
                    &s->m2x, &s->m2y, &s->m1a, &s->m2a, &s->sxx, &s->sxy, &s->syx, &s->syy, &s->th, &s->w,
                    &s->tq, &s->moi, &s->lpx, &s->lpy, &s->dlx, &s->dly, &s->dth})
        v->assign(N, 0.0);
    s->contacts.assign(N, 0);
    s->vcontacts.assign(N, 0);
    s->nnum.assign(N, 0);
    s->pidx.resize(N);
    s->sidx.resize(N);
    std::iota(s->pidx.begin(), s->pidx.end(), 0L);
    std::iota(s->sidx.begin(), s->sidx.end(), 0L);
    s->cell_index.assign(N, -1);
    s->nverts.assign(N, 0);
    s->vstart.assign(N, 0);
}

// ------------------------------------------------------------ neighbours
// kernelUpdateNeighborList: src/particles/base/kernels.cu:256-293
void verlet_kernel(orc_system* s) {
    long N = s->N;
    double c2 = s->rc * s->rc;
#pragma omp parallel for schedule(static) if (N >= OMP_MIN_N)
    for (long i = 0; i < N; ++i) {
        long added = 0;
        for (long j = 0; j < N; ++j) {
            if (i == j) continue;
            if (within(s->x[i], s->y[i], s->x[j], s->y[j], c2, s->box[0], s->box[1])) {
                if (added < s->max_alloc) s->nbr[i * s->max_alloc + added] = j;
                added++;
            }
        }
        s->lnx[i] = s->x[i];
        s->lny[i] = s->y[i];
        s->nd2[i] = 0.0;
        s->nnum[i] = added;
    }
}

// kernelUpdateCellNeighborList: src/particles/base/kernels.cu:440-505
void cell_nbr_kernel(orc_system* s) {
    long N = s->N, n = s->ncd;
    double c2 = s->rc * s->rc;
#pragma omp parallel for schedule(static) if (N >= OMP_MIN_N)
    for (long i = 0; i < N; ++i) {
        long added = 0;
        long cid = s->cell_index[i];
        long cx = cid % n, cy = cid / n;
        for (long oy = -1; oy <= 1; ++oy) {
            long yi = pmod(cy + oy, n);
            for (long ox = -1; ox <= 1; ++ox) {
                long xi = pmod(cx + ox, n);
                long nc = xi + yi * n;
                long st = s->cell_start[nc];
                if (st == -1) continue;
                long en = s->cell_start[nc + 1];
                while (en == -1) { nc++; en = s->cell_start[nc + 1]; }
                for (long j = st; j < en; ++j) {
                    if (j == i) continue;
                    if (within(s->x[i], s->y[i], s->x[j], s->y[j], c2, s->box[0], s->box[1])) {
                        if (added < s->max_alloc) s->nbr[i * s->max_alloc + added] = j;
                        added++;
                    }
                }
            }
        }
        s->nnum[i] = added;
        s->nd2[i] = 0.0;
        s->lnx[i] = s->x[i];
        s->lny[i] = s->y[i];
    }
}

// kernelUpdateVertexNeighborList: src/particles/rigid_bumpy/kernels.cu:870-922
void vertex_nbr_kernel(orc_system* s) {
    double c2 = s->rc_v * s->rc_v, p2 = s->pcut_v * s->pcut_v;
#pragma omp parallel for schedule(static) if (s->Nv >= OMP_MIN_N)
    for (long v = 0; v < s->Nv; ++v) {
        long added = 0;
        double px = s->vxp[v], py = s->vyp[v];
        long p = s->vpart[v];
        for (long k = 0; k < s->nnum[p]; ++k) {
            long q = s->nbr[p * s->max_alloc + k];
            if (q == -1 || q == p) continue;
            if (within(px, py, s->x[q], s->y[q], p2, s->box[0], s->box[1])) {
                long f0 = s->vstart[q], nv = s->nverts[q];
                for (long u = f0; u < f0 + nv; ++u) {
                    if (within(px, py, s->vxp[u], s->vyp[u], c2, s->box[0], s->box[1])) {
                        if (added < s->vmax_alloc) s->vnbr[v * s->vmax_alloc + added] = u;
                        added++;
                    }
                }
            }
        }
        s->vnum[v] = added;
    }
}

// RigidBumpy::updateVertexVerletList: src/particles/rigid_bumpy/rigid_bumpy.cu:1198-1212
void update_vertex_list(orc_system* s) {
    s->vnbr.assign(s->Nv * s->vmax_alloc, -1);
    vertex_nbr_kernel(s);
    long mx = -1;
    for (long v = 0; v < s->Nv; ++v) mx = std::max(mx, s->vnum[v]);
    if (mx > s->vmax_alloc) {
        s->vmax_alloc = pow2_at_least(mx);
        s->vnbr.assign(s->Nv * s->vmax_alloc, -1);
        vertex_nbr_kernel(s);
    }
}

// Particle::updateVerletList: src/particles/base/particle.cu:1119-1131
// (+ RigidBumpy::updateVerletList, rigid_bumpy.cu:1214-1217)
void update_verlet_list(orc_system* s) {
    s->nbr.assign(s->N * s->max_alloc, -1);
    verlet_kernel(s);
    long mx = -1;
    for (long i = 0; i < s->N; ++i) mx = std::max(mx, s->nnum[i]);
    if (mx > s->max_alloc) {
        s->max_alloc = pow2_at_least(mx);
        s->nbr.assign(s->N * s->max_alloc, -1);
        verlet_kernel(s);
    }
    s->n_rebuilds++;
    if (s->kind == ORC_BUMPY) update_vertex_list(s);
}

// Particle::updateCellNeighborList: particle.cu:1320-1331 (+ rigid_bumpy.cu:1251-1255)
void update_cell_neighbor_list(orc_system* s) {
    s->nbr.assign(s->N * s->max_alloc, -1);
    cell_nbr_kernel(s);
    long mx = -1;
    for (long i = 0; i < s->N; ++i) mx = std::max(mx, s->nnum[i]);
    if (mx > s->max_alloc) {
        s->max_alloc = pow2_at_least(mx);
        s->nbr.assign(s->N * s->max_alloc, -1);
        cell_nbr_kernel(s);
    }
    s->n_rebuilds++;
    if (s->kind == ORC_BUMPY) update_vertex_list(s);
}

// getPBCCellIndex: include/particles/base/kernels.cuh:285-301
long cell_coord(double pos, double cs, long n) {
    long idx = (long)std::floor(pos / cs);
    return idx >= 0 ? idx % n : (idx % n + n) % n;
}

// Particle::updateCellList + reorderParticleData: particle.cu:1292-1317;
// kernelGetCellIndexForParticle base/kernels.cu:309-330; kernelReorderParticleData
// base/kernels.cu:508-543; rigid: RigidBumpy::reorderParticleData rigid_bumpy.cu:1303-1376.
void update_cell_list(orc_system* s) {
    long N = s->N, n = s->ncd;
    for (long i = 0; i < N; ++i) {
        long cx = cell_coord(s->x[i], s->cell_size, n);
        long cy = cell_coord(s->y[i], s->cell_size, n);
        s->cell_index[i] = cx + cy * n;
        s->pidx[i] = i;
    }
    // thrust::sort_by_key on int64 keys is a stable radix sort
    std::vector<long> order(N);
    std::iota(order.begin(), order.end(), 0L);
    std::stable_sort(order.begin(), order.end(),
                     [&](long a, long b) { return s->cell_index[a] < s->cell_index[b]; });
    std::vector<long> ci(N), pi(N), si(N);
    for (long k = 0; k < N; ++k) {
        ci[k] = s->cell_index[order[k]];
        pi[k] = s->pidx[order[k]];
        si[k] = s->sidx[order[k]];
    }
    s->cell_index = ci;
    s->pidx = pi;
    s->sidx = si;
    const std::vector<long>& old = s->pidx;
    // last cell positions reset to the gathered positions, cell disp = 0
    for (long k = 0; k < N; ++k) {
        s->lcx[k] = s->x[old[k]];
        s->lcy[k] = s->y[old[k]];
        s->cd2[k] = 0.0;
    }
    for (auto* v : {&s->x, &s->y, &s->fx, &s->fy, &s->vx, &s->vy, &s->m, &s->r})
        permute(*v, old);
    // Adam moments follow their particles (documented deviation)
    for (auto* v : {&s->m1x, &s->m1y, &s->m2x, &s->m2y, &s->m1a, &s->m2a}) permute(*v, old);
    if (s->kind == ORC_BUMPY) {
        std::vector<long> old_to_new(N);
        for (long k = 0; k < N; ++k) old_to_new[old[k]] = k;
        std::vector<long> old_start = s->vstart;
        for (auto* v : {&s->w, &s->tq, &s->th, &s->moi}) permute(*v, old);
        permute(s->nverts, old);
        long acc = 0;
        for (long k = 0; k < N; ++k) { s->vstart[k] = acc; acc += s->nverts[k]; }
        std::vector<double> nvx(s->Nv), nvy(s->Nv), nvm(s->Nv);
        std::vector<long> nvp(s->Nv), nsv(s->Nv);
        // kernelReorderRigidBumpyVertexData: rigid_bumpy/kernels.cu:844-868
        for (long ov = 0; ov < s->Nv; ++ov) {
            long op = s->vpart[ov];
            long np = old_to_new[op];
            long rel = ov - old_start[op];
            long nv = s->vstart[np] + rel;
            nvp[nv] = np;
            nvx[nv] = s->vxp[ov];
            nvy[nv] = s->vyp[ov];
            nvm[nv] = s->vm[ov];
            nsv[nv] = s->svidx[ov];
        }
        s->vxp = nvx; s->vyp = nvy; s->vm = nvm; s->vpart = nvp; s->svidx = nsv;
    }
    // kernelGetFirstParticleIndexForCell: base/kernels.cu:332-348 (first i with cell == c, or -1)
    s->cell_start.assign(s->ncells + 1, -1);
    for (long k = N - 1; k >= 0; --k) s->cell_start[s->cell_index[k]] = k;
    s->cell_start[s->ncells] = N;
    s->n_cell_rebuilds++;
}

void update_neighbor_list(orc_system* s) {
    if (s->method == ORC_NBR_CELL) update_cell_neighbor_list(s);
    else update_verlet_list(s);
}

// Particle::checkForVerletListUpdate / checkForCellListUpdate / checkForAllToAllUpdate:
// particle.cu:1133-1161
int check_for_update(orc_system* s) {
    const double tol = 3.0;
    if (s->method == ORC_NBR_ALL || s->method == ORC_NBR_NONE) return 0;
    if (s->method == ORC_NBR_VERLET) {
        double mx = max_f(s->N, [&](long i) { return s->nd2[i]; });
        if (tol * mx > s->thr_sq) { update_verlet_list(s); return 1; }
        return 0;
    }
    double mc = max_f(s->N, [&](long i) { return s->cd2[i]; });
    if (tol * mc > s->cell_thr_sq) {
        update_cell_list(s);
        update_cell_neighbor_list(s);
        return 2;
    }
    double mn = max_f(s->N, [&](long i) { return s->nd2[i]; });
    if (tol * mn > s->thr_sq) { update_cell_neighbor_list(s); return 1; }
    return 0;
}

// ------------------------------------------------------------ dynamics
// kernelZeroForceAndPotentialEnergy base/kernels.cu:150-161 (+ rigid kernels.cu:234-252)
void zero_forces(orc_system* s) {
    std::fill(s->fx.begin(), s->fx.end(), 0.0);
    std::fill(s->fy.begin(), s->fy.end(), 0.0);
    std::fill(s->pe.begin(), s->pe.end(), 0.0);
    if (s->kind == ORC_BUMPY) {
        std::fill(s->tq.begin(), s->tq.end(), 0.0);
        std::fill(s->vfx.begin(), s->vfx.end(), 0.0);
        std::fill(s->vfy.begin(), s->vfy.end(), 0.0);
        std::fill(s->vtq.begin(), s->vtq.end(), 0.0);
        std::fill(s->vpe.begin(), s->vpe.end(), 0.0);
    }
}

// kernelCalcDiskForces: src/particles/disk/kernels.cu:11-57
void disk_forces(orc_system* s) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        long nn = s->nnum[i];
        if (nn == 0) continue;
        double ax = 0.0, ay = 0.0, e = 0.0;
        for (long k = 0; k < nn; ++k) {
            long j = s->nbr[i * s->max_alloc + k];
            if (j == -1 || j == i) continue;
            double fx, fy;
            double ie = pair_law(s->x[i], s->y[i], s->r[i], s->x[j], s->y[j], s->r[j], s->box[0],
                                 s->box[1], s->e_c, s->n_c, fx, fy);
            ax += fx;
            ay += fy;
            e += ie;
        }
        s->fx[i] += ax;
        s->fy[i] += ay;
        s->pe[i] += e;
    }
}

// kernelCalcRigidBumpyForces1 + kernelCalcRigidBumpyParticleForces1:
// src/particles/rigid_bumpy/kernels.cu:366-442
void bumpy_forces(orc_system* s) {
#pragma omp parallel for schedule(static) if (s->Nv >= OMP_MIN_N)
    for (long v = 0; v < s->Nv; ++v) {
        long nn = s->vnum[v];
        if (nn == 0) continue;
        double px = s->vxp[v], py = s->vyp[v];
        long p = s->vpart[v];
        double cx = s->x[p], cy = s->y[p];
        double ax = 0.0, ay = 0.0, e = 0.0;
        for (long k = 0; k < nn; ++k) {
            long u = s->vnbr[v * s->vmax_alloc + k];
            if (u == -1 || u == v) continue;
            double fx, fy;
            double ie = pair_law(px, py, s->rv, s->vxp[u], s->vyp[u], s->rv, s->box[0], s->box[1],
                                 s->e_c, s->n_c, fx, fy);
            ax += fx;
            ay += fy;
            e += ie;
        }
        double t = torque_of(ax, ay, px, py, cx, cy);
        s->vfx[v] += ax;
        s->vfy[v] += ay;
        s->vtq[v] += t;
        s->vpe[v] += e;
    }
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long p = 0; p < s->N; ++p) {
        double ax = 0.0, ay = 0.0, at = 0.0, e = 0.0;
        for (long k = 0; k < s->nverts[p]; ++k) {
            long v = s->vstart[p] + k;
            ax += s->vfx[v];
            ay += s->vfy[v];
            at += s->vtq[v];
            e += s->vpe[v];
        }
        s->fx[p] += ax;
        s->fy[p] += ay;
        s->tq[p] += at;
        s->pe[p] += e;
    }
}

void compute_forces(orc_system* s) {
    if (s->kind == ORC_BUMPY) bumpy_forces(s);
    else disk_forces(s);
}

// kernelUpdateVelocities base/kernels.cu:117-139; kernelUpdateRigidVelocities
// rigid_bumpy/kernels.cu:123-150
void kick(orc_system* s, double dt) {
    if (s->kind == ORC_BUMPY) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
        for (long i = 0; i < s->N; ++i) {
            double w = s->w[i];
            if (s->rotation) w += s->tq[i] * dt / s->moi[i];
            else w = 0.0;
            s->vx[i] += s->fx[i] * dt / s->m[i];
            s->vy[i] += s->fy[i] * dt / s->m[i];
            s->w[i] = w;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        double dtm = dt / s->m[i];
        s->vx[i] += s->fx[i] * dtm;
        s->vy[i] += s->fy[i] * dtm;
    }
}

// kernelTranslateAndRotateVertices1 rigid_bumpy/kernels.cu:152-201: each vertex turns by
// its particle's last angle change about the old centre and lands on the new one
void move_vertices(orc_system* s) {
#pragma omp parallel for schedule(static) if (s->Nv >= OMP_MIN_N)
    for (long v = 0; v < s->Nv; ++v) {
        long p = s->vpart[v];
        double rx = s->vxp[v] - s->lpx[p], ry = s->vyp[v] - s->lpy[p];
        double sn, c;
        dpm_sincos(s->dth[p], &sn, &c);
        double rxr = rx * c - ry * sn;
        double ryr = rx * sn + ry * c;
        s->vxp[v] = s->x[p] + rxr;
        s->vyp[v] = s->y[p] + ryr;
    }
}

// kernelUpdatePositions base/kernels.cu:48-81; rigid: kernelUpdateRigidPositions
// rigid_bumpy/kernels.cu:21-78 + kernelTranslateAndRotateVertices1 :152-201
void drift(orc_system* s, double dt) {
    if (s->kind == ORC_BUMPY) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
        for (long i = 0; i < s->N; ++i) {
            double ox = s->x[i], oy = s->y[i], a = s->th[i];
            s->lpx[i] = ox;
            s->lpy[i] = oy;
            double ddx = s->vx[i] * dt, ddy = s->vy[i] * dt, dda = s->w[i] * dt;
            s->dlx[i] = ddx;
            s->dly[i] = ddy;
            s->dth[i] = dda;
            double nx = ox + ddx, ny = oy + ddy;
            s->x[i] = nx;
            s->y[i] = ny;
            s->th[i] = a + dda;
            double dxn = nx - s->lnx[i], dyn = ny - s->lny[i];
            s->nd2[i] = dxn * dxn + dyn * dyn;
            double dxc = nx - s->lcx[i], dyc = ny - s->lcy[i];
            s->cd2[i] = dxc * dxc + dyc * dyc;
        }
        move_vertices(s);
        return;
    }
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        double px = s->x[i], py = s->y[i];
        px += s->vx[i] * dt;
        py += s->vy[i] * dt;
        s->x[i] = px;
        s->y[i] = py;
        double dxn = px - s->lnx[i], dyn = py - s->lny[i];
        s->nd2[i] = dxn * dxn + dyn * dyn;
        double dxc = px - s->lcx[i], dyc = py - s->lcy[i];
        s->cd2[i] = dxc * dxc + dyc * dyc;
    }
}

// kernelMixVelocitiesAndForces base/kernels.cu:91-112;
// kernelMixRigidVelocitiesAndForces rigid_bumpy/kernels.cu:89-121
void mix(orc_system* s, double alpha) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        double fx = s->fx[i], fy = s->fy[i], vx = s->vx[i], vy = s->vy[i];
        double fn = std::sqrt(fx * fx + fy * fy);
        double vn = std::sqrt(vx * vx + vy * vy);
        double mr = 0.0;
        if (fn > 1e-16) mr = vn / fn * alpha;
        else { vx = 0.0; vy = 0.0; mr = 0.0; }
        s->vx[i] = vx * (1 - alpha) + fx * mr;
        s->vy[i] = vy * (1 - alpha) + fy * mr;
        if (s->kind == ORC_BUMPY) {
            double w = s->w[i], t = s->tq[i];
            double tn = std::abs(t), wn = std::abs(w), tmr = 0.0;
            if (tn > 1e-16) tmr = wn / tn * alpha;
            else { w = 0.0; tmr = 0.0; }
            s->w[i] = w * (1 - alpha) + t * tmr;
        }
    }
}

void zero_velocities(orc_system* s) {
    std::fill(s->vx.begin(), s->vx.end(), 0.0);
    std::fill(s->vy.begin(), s->vy.end(), 0.0);
    if (s->kind == ORC_BUMPY) std::fill(s->w.begin(), s->w.end(), 0.0);
}

// kinetic energy: base/kernels.cu:163-180; rigid_bumpy/kernels.cu:254-269
void kinetic(orc_system* s) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        double v2 = s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i];
        if (s->kind == ORC_BUMPY && s->rotation)
            s->ke[i] = 0.5 * s->m[i] * v2 + 0.5 * s->moi[i] * s->w[i] * s->w[i];
        else
            s->ke[i] = 0.5 * s->m[i] * v2;
    }
}

double sum_pe(orc_system* s) { return canonical_sum_f(s->N, [&](long i) { return s->pe[i]; }); }

// Particle::getPowerFire particle.cu:993-1009 (+ RigidBumpy rigid_bumpy.cu:1181-1191)
double power_fire(orc_system* s) {
    double px = canonical_sum_f(s->N, [&](long i) { return s->fx[i] * s->vx[i]; });
    double py = canonical_sum_f(s->N, [&](long i) { return s->fy[i] * s->vy[i]; });
    double p = px + py;
    if (s->kind == ORC_BUMPY) p = p + canonical_sum_f(s->N, [&](long i) { return s->tq[i] * s->w[i]; });
    return p;
}

double fmax_sq(orc_system* s) {
    double m = max_f(s->N, [&](long i) { return s->fx[i] * s->fx[i] + s->fy[i] * s->fy[i]; });
    if (s->kind == ORC_BUMPY)
        m = std::max(m, max_f(s->N, [&](long i) { return s->tq[i] * s->tq[i]; }));
    return m;
}

// kernelCalcDiskStressTensor disk/kernels.cu:203-251;
// kernelCalcRigidBumpyStressTensor rigid_bumpy/kernels.cu:705-789
void stress(orc_system* s) {
    double area = s->box[0] * s->box[1];
    if (s->kind == ORC_DISK) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
        for (long i = 0; i < s->N; ++i) {
            double px = s->x[i], py = s->y[i], mm = s->m[i], rr = s->r[i], vx = s->vx[i],
                   vy = s->vy[i];
            double axx = mm * vx * vx, axy = mm * vx * vy, ayx = axy, ayy = mm * vy * vy;
            for (long k = 0; k < s->max_alloc; ++k) {
                long j = s->nbr[i * s->max_alloc + k];
                if (j == -1 || j == i) continue;
                double xd = pbc(px, s->x[j], s->box[0]);
                double yd = pbc(py, s->y[j], s->box[1]);
                double fx, fy;
                pair_law(px, py, rr, s->x[j], s->y[j], s->r[j], s->box[0], s->box[1], s->e_c,
                         s->n_c, fx, fy);
                axx += xd * fx / 2;
                axy += xd * fy / 2;
                ayx += yd * fx / 2;
                ayy += yd * fy / 2;
            }
            s->sxx[i] = axx / area;
            s->sxy[i] = axy / area;
            s->syx[i] = ayx / area;
            s->syy[i] = ayy / area;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long p = 0; p < s->N; ++p) {
        double axx = 0.0, axy = 0.0, ayx = 0.0, ayy = 0.0;
        double px = s->x[p], py = s->y[p], vx = s->vx[p], vy = s->vy[p], w = s->w[p];
        for (long k = 0; k < s->nverts[p]; ++k) {
            long v = s->vstart[p] + k;
            double vpx = s->vxp[v], vpy = s->vyp[v], vmass = s->vm[v];
            double rx = vpx - px, ry = vpy - py;
            double prad = std::sqrt(rx * rx + ry * ry);
            double vvx = vx, vvy = vy;
            if (w != 0.0) {
                double ang = std::atan2(ry, rx);
                vvx -= std::sin(ang) * w * prad;
                vvy += std::cos(ang) * w * prad;
            }
            axx += vmass * vvx * vvx;
            axy += vmass * vvx * vvy;
            ayx += vmass * vvy * vvx;
            ayy += vmass * vvy * vvy;
            for (long k2 = 0; k2 < s->vnum[v]; ++k2) {
                long u = s->vnbr[v * s->vmax_alloc + k2];
                long q = (u == -1) ? -1 : s->vpart[u];
                if (u == -1 || u == v || q == p) continue;
                double xd = pbc(vpx, s->vxp[u], s->box[0]);
                double yd = pbc(vpy, s->vyp[u], s->box[1]);
                double fx, fy;
                pair_law(vpx, vpy, s->rv, s->vxp[u], s->vyp[u], s->rv, s->box[0], s->box[1],
                         s->e_c, s->n_c, fx, fy);
                axx += xd * fx / 2;
                axy += xd * fy / 2;
                ayx += yd * fx / 2;
                ayy += yd * fy / 2;
            }
        }
        s->sxx[p] = axx / area;
        s->sxy[p] = axy / area;
        s->syx[p] = ayx / area;
        s->syy[p] = ayy / area;
    }
}

// Particle::getPressure particle.cu:1371-1376
double pressure(orc_system* s) {
    double a = canonical_sum_f(s->N, [&](long i) { return s->sxx[i]; });
    double b = canonical_sum_f(s->N, [&](long i) { return s->syy[i]; });
    return (a + b) / 2;
}

// kernelCountDiskContacts disk/kernels.cu:177-201;
// kernelCountRigidBumpyContacts rigid_bumpy/kernels.cu:924-997
void count_contacts(orc_system* s) {
    if (s->kind == ORC_DISK) {
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
        for (long i = 0; i < s->N; ++i) {
            long c = 0;
            for (long k = 0; k < s->nnum[i]; ++k) {
                long j = s->nbr[i * s->max_alloc + k];
                if (j == -1 || j == i) continue;
                double dx = pbc(s->x[i], s->x[j], s->box[0]);
                double dy = pbc(s->y[i], s->y[j], s->box[1]);
                double d = std::sqrt(dx * dx + dy * dy);
                if (d < s->r[i] + s->r[j]) c++;
            }
            s->contacts[i] = c;
        }
        return;
    }
    double vrs = 2 * s->rv;
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long p = 0; p < s->N; ++p) {
        long nc = 0, nvc = 0;
        for (long k = 0; k < s->nnum[p]; ++k) {
            long q = s->nbr[p * s->max_alloc + k];
            if (q == -1 || q == p) continue;
            bool is_contact = false;
            for (long a = 0; a < s->nverts[p]; ++a) {
                long v = s->vstart[p] + a;
                for (long k2 = 0; k2 < s->vnum[v]; ++k2) {
                    long u = s->vnbr[v * s->vmax_alloc + k2];
                    long uq = (u == -1) ? -1 : s->vpart[u];
                    if (u == -1 || u == v || uq != q) continue;
                    double dx = pbc(s->vxp[v], s->vxp[u], s->box[0]);
                    double dy = pbc(s->vyp[v], s->vyp[u], s->box[1]);
                    double d = std::sqrt(dx * dx + dy * dy);
                    if (d < vrs) {
                        if (!is_contact) nc++;
                        nvc++;
                        is_contact = true;
                    }
                }
                if (is_contact) break;
            }
        }
        s->contacts[p] = nc;
        s->vcontacts[p] = nvc;
    }
}

double particle_area(orc_system* s) {  // Disk::getParticleArea disk.cu:205-207 (pi * sum r^2)
    const double PI = 3.141592653589793238462643383279502884197;
    return canonical_sum_f(s->N, [&](long i) { return s->r[i] * s->r[i]; }) * PI;
}

// Particle::updateCellSize particle.cu:1236-1240
void update_cell_size(orc_system* s) {
    if (s->method == ORC_NBR_CELL) s->cell_size = s->box[0] / s->ncd;
}

// positions.scale (Particle::scalePositionsFull particle.cu:1070-1072);
// kernelScalePositions rigid_bumpy/kernels.cu:1308-1333
void scale_positions(orc_system* s, double f) {
    if (s->kind == ORC_BUMPY) {
        for (long p = 0; p < s->N; ++p) {
            double ox = s->x[p], oy = s->y[p];
            double nx = ox * f, ny = oy * f;
            s->x[p] = nx;
            s->y[p] = ny;
            for (long v = s->vstart[p]; v < s->vstart[p] + s->nverts[p]; ++v) {
                s->vxp[v] += (nx - ox);
                s->vyp[v] += (ny - oy);
            }
        }
        return;
    }
    for (long i = 0; i < s->N; ++i) {
        s->x[i] *= f;
        s->y[i] *= f;
    }
}

// Particle::scaleBox particle.cu:1057-1064 (pow(area, 1/2) taken as the correctly
// rounded square root)
void scale_box(orc_system* s, double f) {
    scale_positions(s, f);
    double side = std::sqrt(s->box[0] * s->box[1]);
    double ns = side * f;
    s->box[0] = ns;
    s->box[1] = ns;
    update_cell_size(s);
}

// Particle::setLastState / revertToLastState: particle.cu:956-991 (+ rigid_bumpy.cu:1144-1179)
void snapshot_save(orc_system* s) {
    auto& S = s->snap;
    S.valid = true;
    S.x = s->x; S.y = s->y; S.fx = s->fx; S.fy = s->fy; S.vx = s->vx; S.vy = s->vy;
    S.r = s->r; S.m = s->m; S.pidx = s->pidx; S.sidx = s->sidx;
    S.box[0] = s->box[0]; S.box[1] = s->box[1];
    if (s->kind == ORC_BUMPY) {
        S.th = s->th; S.w = s->w; S.tq = s->tq; S.moi = s->moi; S.vxp = s->vxp; S.vyp = s->vyp;
        S.vm = s->vm; S.nverts = s->nverts; S.vstart = s->vstart; S.vpart = s->vpart;
        S.svidx = s->svidx;
    }
}

void snapshot_restore(orc_system* s) {
    auto& S = s->snap;
    s->x = S.x; s->y = S.y; s->fx = S.fx; s->fy = S.fy; s->vx = S.vx; s->vy = S.vy;
    s->r = S.r; s->m = S.m; s->pidx = S.pidx; s->sidx = S.sidx;
    if (s->kind == ORC_BUMPY) {
        s->th = S.th; s->w = S.w; s->tq = S.tq; s->moi = S.moi; s->vxp = S.vxp; s->vyp = S.vyp;
        s->vm = S.vm; s->nverts = S.nverts; s->vstart = S.vstart; s->vpart = S.vpart;
        s->svidx = S.svidx;
    }
    s->box[0] = S.box[0];
    s->box[1] = S.box[1];
    update_cell_size(s);
    // full rebuild (documented deviation): cell structure + neighbour list
    if (s->method == ORC_NBR_CELL) {
        update_cell_list(s);
        update_cell_neighbor_list(s);
    } else if (s->method != ORC_NBR_NONE) {
        update_verlet_list(s);
    }
}

// ------------------------------------------------------------ FIRE
// Fire::Fire / Fire::step: src/integrator/fire.cu:3-46 (constants include/integrator/fire.h:25-31)
void fire_init(orc_fire_state* st, double alpha_init, double dt) {
    st->alpha = alpha_init;
    st->alpha_init = alpha_init;
    st->dt = dt;
    st->dt_max = 10.0 * dt;
    st->dt_min = 0.001 * dt;
    st->n_good = 0;
    st->n_bad = 0;
    st->stopped = 0;
    st->power = 0.0;
}

void fire_step(orc_system* s, orc_fire_state* st) {
    const double f_inc = 1.1, f_dec = 0.5, f_alpha = 0.99;
    const long N_min = 5, N_bad_max = 10;
    double power = power_fire(s);
    st->power = power;
    if (power > 0) {
        st->n_good++;
        st->n_bad = 0;
        if (st->n_good > N_min) {
            st->dt = std::min(st->dt * f_inc, st->dt_max);
            st->alpha *= f_alpha;
        }
    } else {
        st->n_good = 0;
        st->n_bad++;
        if (st->n_bad > N_bad_max) {
            st->stopped = 1;
            return;
        }
        st->dt = std::max(st->dt * f_dec, st->dt_min);
        st->alpha = st->alpha_init;
        drift(s, -st->dt / 2.0);
        zero_velocities(s);
    }
    kick(s, st->dt / 2.0);
    mix(s, st->alpha);
    drift(s, st->dt);
    zero_forces(s);
    check_for_update(s);
    compute_forces(s);
    kick(s, st->dt / 2.0);
}

// minimizeFire: src/routines/minimization.cu:75-112, plus an optional
// max-force criterion (fx^2+fy^2 [and tau^2] < fmax_target^2 for every particle).
void minimize_fire(orc_system* s, const orc_fire_params* p, orc_fire_result* out) {
    orc_fire_state st;
    fire_init(&st, p->alpha_init, p->dt);
    long step = 0;
    double dof = (double)s->N;
    double diff = 1e9;
    zero_velocities(s);
    zero_forces(s);
    check_for_update(s);
    compute_forces(s);
    double last = sum_pe(s) / dof;
    double avg = last, f2 = 0.0;
    int converged = 0;
    while (step < p->max_steps) {
        fire_step(s, &st);
        avg = sum_pe(s) / dof;
        diff = std::abs(avg / last - 1);
        last = avg;
        bool fok = false;
        if (p->fmax_target > 0) {
            f2 = fmax_sq(s);
            fok = f2 < p->fmax_target * p->fmax_target;
        }
        if (diff < p->avg_pe_diff_target || avg < p->avg_pe_target || st.stopped || fok) {
            converged = st.stopped ? 0 : 1;
            break;
        }
        step++;
    }
    zero_velocities(s);
    if (out) {
        out->steps = std::min(step + 1, p->max_steps);
        out->stopped = st.stopped;
        out->converged = converged;
        out->avg_pe = avg;
        out->fmax_sq = fmax_sq(s);
        out->dt = st.dt;
        out->alpha = st.alpha;
    }
}

// ------------------------------------------------------------ Adam
// Particle::updatePositionsAdam particle.cu:329-333 + kernelAdamStep base/kernels.cu:550-611
void adam_update(orc_system* s, long step, double alpha, double b1, double b2, double eps) {
    double omb1 = 1 - std::pow(b1, step + 1);
    double omb2 = 1 - std::pow(b2, step + 1);
    if (s->kind == ORC_BUMPY) {
        // RigidBumpy::updatePositionsAdam rigid_bumpy.cu:1473-1486 + kernelRigidBumpyAdamStep
        // rigid_bumpy/kernels.cu:1004-1093; without rotation the angle and its recorded
        // change stay as they are, and the vertex pass still turns by that change
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
        for (long i = 0; i < s->N; ++i) {
            double fx = s->fx[i], fy = s->fy[i], t = s->tq[i];
            double m1x = std::fma(b1, s->m1x[i], (b1 - 1) * fx);
            double m1y = std::fma(b1, s->m1y[i], (b1 - 1) * fy);
            double m1a = std::fma(b1, s->m1a[i], (b1 - 1) * t);
            double m2x = std::fma(b2, s->m2x[i], (1 - b2) * fx * fx);
            double m2y = std::fma(b2, s->m2y[i], (1 - b2) * fy * fy);
            double m2a = std::fma(b2, s->m2a[i], (1 - b2) * t * t);
            double ux = -alpha * (m1x / omb1) / (std::sqrt(m2x / omb2) + eps);
            double uy = -alpha * (m1y / omb1) / (std::sqrt(m2y / omb2) + eps);
            double ua = -alpha * (m1a / omb1) / (std::sqrt(m2a / omb2) + eps);
            double px = s->x[i], py = s->y[i];
            s->lpx[i] = px;
            s->lpy[i] = py;
            px += ux;
            py += uy;
            s->dlx[i] = ux;
            s->dly[i] = uy;
            if (s->rotation) {
                s->th[i] += ua;
                s->dth[i] = ua;
            }
            s->x[i] = px;
            s->y[i] = py;
            double dxn = px - s->lnx[i], dyn = py - s->lny[i];
            s->nd2[i] = dxn * dxn + dyn * dyn;
            double dxc = px - s->lcx[i], dyc = py - s->lcy[i];
            s->cd2[i] = dxc * dxc + dyc * dyc;
            s->m1x[i] = m1x; s->m1y[i] = m1y; s->m1a[i] = m1a;
            s->m2x[i] = m2x; s->m2y[i] = m2y; s->m2a[i] = m2a;
        }
        move_vertices(s);
        return;
    }
#pragma omp parallel for schedule(static) if (s->N >= OMP_MIN_N)
    for (long i = 0; i < s->N; ++i) {
        double fx = s->fx[i], fy = s->fy[i];
        double m1x = s->m1x[i], m1y = s->m1y[i], m2x = s->m2x[i], m2y = s->m2y[i];
        m1x = std::fma(b1, m1x, (b1 - 1) * fx);
        m1y = std::fma(b1, m1y, (b1 - 1) * fy);
        m2x = std::fma(b2, m2x, (1 - b2) * fx * fx);
        m2y = std::fma(b2, m2y, (1 - b2) * fy * fy);
        double mhx = m1x / omb1, mhy = m1y / omb1;
        double vhx = m2x / omb2, vhy = m2y / omb2;
        double ux = -alpha * mhx / (std::sqrt(vhx) + eps);
        double uy = -alpha * mhy / (std::sqrt(vhy) + eps);
        s->x[i] += ux;
        s->y[i] += uy;
        double px = s->x[i], py = s->y[i];
        double dxn = px - s->lnx[i], dyn = py - s->lny[i];
        s->nd2[i] = dxn * dxn + dyn * dyn;
        double dxc = px - s->lcx[i], dyc = py - s->lcy[i];
        s->cd2[i] = dxc * dxc + dyc * dyc;
        s->m1x[i] = m1x; s->m1y[i] = m1y; s->m2x[i] = m2x; s->m2y[i] = m2y;
    }
}

// Adam::minimize src/integrator/adam.cu:11-18
void adam_step(orc_system* s, long step, double alpha, double b1, double b2, double eps) {
    zero_forces(s);
    check_for_update(s);
    compute_forces(s);
    adam_update(s, step, alpha, b1, b2, eps);
}

// Particle::initAdamVariables particle.cu:319-322
void adam_reset(orc_system* s) {
    std::fill(s->m1x.begin(), s->m1x.end(), 0.0);
    std::fill(s->m1y.begin(), s->m1y.end(), 0.0);
    std::fill(s->m2x.begin(), s->m2x.end(), 0.0);
    std::fill(s->m2y.begin(), s->m2y.end(), 0.0);
    // RigidBumpy::initAdamVariables rigid_bumpy.cu:1461-1465
    std::fill(s->m1a.begin(), s->m1a.end(), 0.0);
    std::fill(s->m2a.begin(), s->m2a.end(), 0.0);
}

long minimize_adam(orc_system* s, double alpha, double b1, double b2, double eps, double pe_t,
                   double diff_t, long max_steps) {
    // minimizeAdam src/routines/minimization.cu:3-33
    adam_reset(s);
    long step = 0;
    double dof = (double)s->ndof();
    double last = 0.0;
    while (step < max_steps) {
        adam_step(s, step, alpha, b1, b2, eps);
        double avg = sum_pe(s) / dof;
        double diff = std::abs(avg / last - 1);
        last = avg;
        if (diff < diff_t || avg < pe_t) break;
        step++;
    }
    return step;
}

// ------------------------------------------------------------ jamming
// Pressure-bisection compress/minimize loop restated from
// scripts/runArthurJamWithFire.cpp:36-106 with the reference's file round trip
// replaced by an in-memory snapshot, minimizeFire given the max-force
// criterion, and the pressure window [P_t, P_t(1+window_rel)].
void jam(orc_system* s, const orc_jam_params* P, orc_jam_result* out) {
    double P_low = P->pressure_target, P_high = P->pressure_target * (1.0 + P->window_rel);
    long fire_total = 0;
    orc_fire_result fr;
    // decompress until unjammed: runArthurJamWithFire.cpp:37-45
    for (int it = 0; it < 1000; ++it) {
        minimize_fire(s, &P->fire, &fr);
        fire_total += fr.steps;
        stress(s);
        double pr = pressure(s);
        if (pr < P_low) break;
        scale_box(s, P->decompress_factor);
    }
    double L_low = -1, L_high = -1, L_last = -1;
    double rate = 1.0 - P->strain_step;
    long cstep = 0;
    int status = 1;
    double pr = 0.0;
    long ncont = 0;
    snapshot_save(s);
    while (cstep < P->max_compression_steps) {
        minimize_fire(s, &P->fire, &fr);
        fire_total += fr.steps;
        stress(s);
        pr = pressure(s);
        count_contacts(s);
        ncont = 0;
        for (long i = 0; i < s->N; ++i) ncont += s->contacts[i];
        double L = std::sqrt(s->box[0] * s->box[1]);
        double L_prior = L;
        if (pr < P_low) {
            L_low = L;
            snapshot_save(s);
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
                snapshot_restore(s);
                L_prior = L_last;
                L = (L_high + L_low) / 2;
            } else if ((double)ncont > P->min_contacts_per_particle * (double)s->N) {
                status = 0;
                break;
            }
        }
        if (L_low > 0 && L_high > 0 && std::abs(L_low / L_high - 1) < P->length_tol) {
            status = 2;
            break;
        }
        scale_box(s, L / L_prior);
        cstep++;
    }
    if (out) {
        out->compression_steps = cstep;
        out->fire_steps_total = fire_total;
        out->status = status;
        out->pressure = pr;
        out->packing_fraction = particle_area(s) / (s->box[0] * s->box[1]);
        out->pe = sum_pe(s);
        out->box_length = s->box[0];
        out->contacts = ncont;
    }
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

double orc_canonical_sum(const double* a, int64_t n) {
    return canonical_sum_f(n, [&](long i) { return a[i]; });
}

orc_system* orc_create(int kind, int64_t n, int64_t nv) {
    if (n <= 0 || (kind != ORC_DISK && kind != ORC_BUMPY)) { fail("bad create args"); return nullptr; }
    auto* s = new orc_system();
    s->kind = kind;
    s->N = n;
    s->Nv = kind == ORC_BUMPY ? nv : 0;
    resize_particles(s);
    if (kind == ORC_BUMPY) {
        for (auto* v : {&s->vxp, &s->vyp, &s->vm, &s->vfx, &s->vfy, &s->vtq, &s->vpe})
            v->assign(s->Nv, 0.0);
        s->vpart.assign(s->Nv, 0);
        s->svidx.resize(s->Nv);
        std::iota(s->svidx.begin(), s->svidx.end(), 0L);
        s->vnum.assign(s->Nv, 0);
    }
    return s;
}

void orc_destroy(orc_system* s) { delete s; }

int orc_set_box(orc_system* s, double lx, double ly) {
    s->box[0] = lx;
    s->box[1] = ly;
    update_cell_size(s);
    return 0;
}
int orc_get_box(orc_system* s, double* lx, double* ly) {
    *lx = s->box[0];
    *ly = s->box[1];
    return 0;
}
int orc_set_energy(orc_system* s, double e_c, double n_c) {
    s->e_c = e_c;
    s->n_c = n_c;
    return 0;
}

static std::vector<double>* dfield(orc_system* s, int f, bool& vertex) {
    vertex = false;
    switch (f) {
        case ORC_F_X: return &s->x;
        case ORC_F_Y: return &s->y;
        case ORC_F_VX: return &s->vx;
        case ORC_F_VY: return &s->vy;
        case ORC_F_FX: return &s->fx;
        case ORC_F_FY: return &s->fy;
        case ORC_F_RAD: return &s->r;
        case ORC_F_MASS: return &s->m;
        case ORC_F_PE: return &s->pe;
        case ORC_F_KE: return &s->ke;
        case ORC_F_ANGLE: return &s->th;
        case ORC_F_OMEGA: return &s->w;
        case ORC_F_TORQUE: return &s->tq;
        case ORC_F_MOI: return &s->moi;
        case ORC_F_STRESS_XX: return &s->sxx;
        case ORC_F_STRESS_XY: return &s->sxy;
        case ORC_F_STRESS_YX: return &s->syx;
        case ORC_F_STRESS_YY: return &s->syy;
        case ORC_F_VERTEX_X: vertex = true; return &s->vxp;
        case ORC_F_VERTEX_Y: vertex = true; return &s->vyp;
        case ORC_F_VERTEX_MASS: vertex = true; return &s->vm;
    }
    return nullptr;
}
static std::vector<long>* lfield(orc_system* s, int f) {
    switch (f) {
        case ORC_F_STATIC_INDEX: return &s->sidx;
        case ORC_F_NUM_NEIGHBORS: return &s->nnum;
        case ORC_F_CONTACTS: return &s->contacts;
        case ORC_F_NUM_VERTICES: return &s->nverts;
        case ORC_F_VERTEX_CONTACTS: return &s->vcontacts;
    }
    return nullptr;
}

// Values travel in static (original) order: static id sidx[k] lives at internal k.
int orc_upload(orc_system* s, int f, const void* data, int64_t count) {
    bool vert;
    if (auto* v = dfield(s, f, vert)) {
        long n = vert ? s->Nv : s->N;
        if (count != n) return fail("upload: count mismatch");
        const double* d = (const double*)data;
        if (vert) for (long k = 0; k < n; ++k) (*v)[k] = d[s->svidx[k]];
        else for (long k = 0; k < n; ++k) (*v)[k] = d[s->sidx[k]];
        return 0;
    }
    if (f == ORC_F_NUM_VERTICES) {
        if (count != s->N) return fail("upload: count mismatch");
        const int64_t* d = (const int64_t*)data;
        for (long k = 0; k < s->N; ++k) s->nverts[k] = d[s->sidx[k]];
        return 0;
    }
    return fail("upload: field not uploadable");
}

int orc_download(orc_system* s, int f, void* data, int64_t count) {
    bool vert;
    if (auto* v = dfield(s, f, vert)) {
        long n = vert ? s->Nv : s->N;
        if (count != n) return fail("download: count mismatch");
        double* d = (double*)data;
        if (vert) for (long k = 0; k < n; ++k) d[s->svidx[k]] = (*v)[k];
        else for (long k = 0; k < n; ++k) d[s->sidx[k]] = (*v)[k];
        return 0;
    }
    if (auto* v = lfield(s, f)) {
        if (count != s->N) return fail("download: count mismatch");
        int64_t* d = (int64_t*)data;
        if (f == ORC_F_STATIC_INDEX) for (long k = 0; k < s->N; ++k) d[k] = (*v)[k];
        else for (long k = 0; k < s->N; ++k) d[s->sidx[k]] = (*v)[k];
        return 0;
    }
    return fail("download: unknown field");
}

// Particle::setNeighborMethod/setNeighborSize/setCellSize (particle.cu:237-265, :1223-1272),
// with the numbers given directly as in scripts/runShortNVE.cpp:43-54.
int orc_neighbor_config(orc_system* s, int method, double cutoff, double thr_sq, int64_t ncd,
                        double cell_thr_sq) {
    s->method = method;
    s->max_alloc = 4;
    s->vmax_alloc = 4;
    s->rc = cutoff;
    s->thr_sq = thr_sq;
    std::iota(s->pidx.begin(), s->pidx.end(), 0L);
    std::iota(s->sidx.begin(), s->sidx.end(), 0L);
    if (s->kind == ORC_BUMPY) std::iota(s->svidx.begin(), s->svidx.end(), 0L);
    if (method == ORC_NBR_CELL) {
        if (ncd < 3) return fail("neighbor_config: need at least 3 cells per side");
        s->ncd = ncd;
        s->ncells = ncd * ncd;
        s->cell_thr_sq = cell_thr_sq;
        s->cell_size = s->box[0] / ncd;
    }
    if (method == ORC_NBR_ALL) {
        // Particle::initAllToAllList particle.cu:1186-1200
        s->max_alloc = s->N - 1 > 0 ? s->N - 1 : 1;
        double diag = std::sqrt(s->box[0] * s->box[0] + s->box[1] * s->box[1]);
        s->rc = 10.0 * diag;
        s->thr_sq = s->rc * s->rc;
    }
    return 0;
}

int orc_bumpy_neighbor_config(orc_system* s, double vertex_cutoff, double vertex_particle_cutoff) {
    s->rc_v = vertex_cutoff;
    s->pcut_v = vertex_particle_cutoff;
    return 0;
}

// Particle::initNeighborList -> initVerletList / initCellList / initAllToAllList
// (particle.cu:1163-1221)
int orc_neighbor_init(orc_system* s) {
    std::fill(s->lnx.begin(), s->lnx.end(), 0.0);
    std::fill(s->lny.begin(), s->lny.end(), 0.0);
    std::fill(s->nd2.begin(), s->nd2.end(), 0.0);
    std::fill(s->lcx.begin(), s->lcx.end(), 0.0);
    std::fill(s->lcy.begin(), s->lcy.end(), 0.0);
    std::fill(s->cd2.begin(), s->cd2.end(), 0.0);
    s->nbr.assign(s->N * s->max_alloc, -1);
    if (s->kind == ORC_BUMPY) s->vnbr.assign(s->Nv * s->vmax_alloc, -1);
    if (s->method == ORC_NBR_CELL) {
        s->cell_index.assign(s->N, -1);
        update_cell_list(s);
        update_cell_neighbor_list(s);
    } else if (s->method == ORC_NBR_VERLET || s->method == ORC_NBR_ALL) {
        update_verlet_list(s);
    } else {
        return fail("neighbor_init: no method configured");
    }
    return 0;
}

int orc_neighbor_check(orc_system* s, int* rebuilt) {
    int r = check_for_update(s);
    if (rebuilt) *rebuilt = r;
    return 0;
}

int orc_neighbor_pairs(orc_system* s, int64_t* oi, int64_t* oj, int64_t cap, int64_t* count) {
    int64_t c = 0;
    for (long i = 0; i < s->N; ++i)
        for (long k = 0; k < s->nnum[i] && k < s->max_alloc; ++k) {
            long j = s->nbr[i * s->max_alloc + k];
            if (j < 0) continue;
            if (c < cap) { oi[c] = s->sidx[i]; oj[c] = s->sidx[j]; }
            c++;
        }
    *count = c;
    return c > cap ? fail("neighbor_pairs: capacity too small") : 0;
}

int orc_vertex_pairs(orc_system* s, int64_t* oi, int64_t* oj, int64_t cap, int64_t* count) {
    int64_t c = 0;
    for (long v = 0; v < s->Nv; ++v)
        for (long k = 0; k < s->vnum[v] && k < s->vmax_alloc; ++k) {
            long u = s->vnbr[v * s->vmax_alloc + k];
            if (u < 0) continue;
            if (c < cap) { oi[c] = s->svidx[v]; oj[c] = s->svidx[u]; }
            c++;
        }
    *count = c;
    return c > cap ? fail("vertex_pairs: capacity too small") : 0;
}

// ------------------------------------------------------------ pair analytics
namespace {

void put_i64(std::vector<double>& out, long S, int f, long t, long long v) {
    std::memcpy(&out[(size_t)(f * S + t)], &v, sizeof v);
}

// the fill values of Disk::calculateForceDistancePairs (disk.cu:229-245) and
// RigidBumpy::calculateForceDistancePairs (rigid_bumpy.cu:1384-1415)
void pair_fill(orc_system* s, long S, int nf) {
    s->pairs.assign((size_t)(nf * S), 0.0);
    for (long t = 0; t < S; ++t) {
        for (int f : {ORC_P_POTENTIAL, ORC_P_DIST_X, ORC_P_DIST_Y, ORC_P_OVERLAP, ORC_P_RADSUM,
                      ORC_P_ANGLE})
            s->pairs[(size_t)(f * S + t)] = -1.0;
        put_i64(s->pairs, S, ORC_P_ID_I, t, -1);
        put_i64(s->pairs, S, ORC_P_ID_J, t, -1);
        if (nf > ORC_P_ANGLE_I) {
            s->pairs[(size_t)(ORC_P_ANGLE_I * S + t)] = -1.0;
            s->pairs[(size_t)(ORC_P_ANGLE_J * S + t)] = -1.0;
            s->pairs[(size_t)(ORC_P_FRICTION * S + t)] = -1.0;
            put_i64(s->pairs, S, ORC_P_VERTEX_CONTACTS, t, -1);
        }
    }
}

// kernelCalcDiskForceDistancePairs: src/particles/disk/kernels.cu:73-175
void disk_pairs(orc_system* s, long S) {
    const long M = s->max_alloc;
    const double e_c = s->e_c;
    std::vector<double>& P = s->pairs;
    for (long i = 0; i < s->N; ++i) {
        if (s->nnum[i] == 0) continue;
        for (long n = 0; n < M; ++n) {
            long j = s->nbr[(size_t)(i * M + n)];
            if (j == -1 || j == i) continue;
            long t = i * M + n;
            double fx, fy;
            double e = pair_law(s->x[i], s->y[i], s->r[i], s->x[j], s->y[j], s->r[j], s->box[0],
                                s->box[1], s->e_c, s->n_c, fx, fy);
            P[ORC_P_POTENTIAL * S + t] = e;
            P[ORC_P_FORCE_X * S + t] = fx;
            P[ORC_P_FORCE_Y * S + t] = fy;
            double dx = pbc(s->x[i], s->x[j], s->box[0]);
            double dy = pbc(s->y[i], s->y[j], s->box[1]);
            P[ORC_P_DIST_X * S + t] = dx;
            P[ORC_P_DIST_Y * S + t] = dy;
            put_i64(P, S, ORC_P_ID_I, t, s->sidx[i]);
            put_i64(P, S, ORC_P_ID_J, t, s->sidx[j]);
            double dist = std::sqrt(dx * dx + dy * dy);
            P[ORC_P_OVERLAP * S + t] = dist - (s->r[i] + s->r[j]);
            P[ORC_P_RADSUM * S + t] = s->r[i] + s->r[j];
            P[ORC_P_ANGLE * S + t] = std::atan2(dy, dx);
            if (dist < s->r[i] + s->r[j]) {
                double rs = s->r[i] + s->r[j];
                double t_ij = -e_c / rs * (1 - dist / rs);
                double c_ij = e_c / (rs * rs);
                double hij[4] = {0, 0, 0, 0}, hii[4] = {0, 0, 0, 0};
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        // unit derivatives of (dx, dy) by coordinate a of i and b of j
                        double ax = 1 * (a == 0), ay = 1 * (a == 1);
                        double bx = -1 * (b == 0), by = -1 * (b == 1);
                        double zx = 0, zy = 0;
                        double qa = dx * ax + dy * ay;
                        double qb = dx * bx + dy * by;
                        double dq = ax * bx + dx * zx + ay * by + dy * zy;
                        hij[a * 2 + b] += c_ij * (qa * qb) / (dist * dist) +
                                          t_ij * (dq / dist - (qa * qb) / (dist * dist * dist));
                        double cx = 1 * (b == 0), cy = 1 * (b == 1);
                        double wx = 0, wy = 0;
                        double qc = dx * cx + dy * cy;
                        double dqc = ax * cx + dx * wx + ay * cy + dy * wy;
                        hii[a * 2 + b] += c_ij * (qa * qc) / (dist * dist) +
                                          t_ij * (dqc / dist - (qa * qc) / (dist * dist * dist));
                    }
                for (int c = 0; c < 4; ++c) {
                    P[(ORC_P_H_XX + c) * S + t] += hij[c];
                    P[(ORC_P_HII_XX + c) * S + t] += hii[c];
                }
            }
        }
    }
}

// kernelCalcRigidBumpyForceDistancePairs: src/particles/rigid_bumpy/kernels.cu:482-702
void bumpy_pairs(orc_system* s, long S) {
    const long M = s->max_alloc;
    const double rv = s->rv, e_c = s->e_c;
    const double PI_R = 3.14159265358979323846;  // M_PI
    std::vector<double>& P = s->pairs;
    for (long i = 0; i < s->N; ++i) {
        long nn = s->nnum[i];
        for (long n = 0; n < nn; ++n) {
            long j = s->nbr[(size_t)(i * M + n)];
            if (j == -1 || j == i) continue;
            double ti = s->th[i], tj = s->th[j];
            double fx = 0.0, fy = 0.0, energy = 0.0, voverlap = 0.0;
            long count_i = 0;
            double hij[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, hii[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (long u = 0; u < s->nverts[i]; ++u) {
                long vi = s->vstart[i] + u;
                double phi_u = u * 2 * PI_R / s->nverts[i];
                double ex = s->vxp[vi] - s->x[i], ey = s->vyp[vi] - s->y[i];
                double Ri = std::sqrt(ex * ex + ey * ey);
                double si, ci;
                dpm_sincos(ti + phi_u, &si, &ci);
                bool contact = false;
                for (long m = 0; m < s->vnum[vi]; ++m) {
                    long vo = s->vnbr[(size_t)(vi * s->vmax_alloc + m)];
                    if (vo == -1 || vo == vi || s->vpart[vo] != j) continue;
                    long v = vo - s->vstart[j];
                    double phi_v = v * 2 * PI_R / s->nverts[j];
                    double gx = s->vxp[vo] - s->x[j], gy = s->vyp[vo] - s->y[j];
                    double Rj = std::sqrt(gx * gx + gy * gy);
                    double tfx, tfy;
                    double te = pair_law(s->vxp[vi], s->vyp[vi], rv, s->vxp[vo], s->vyp[vo], rv,
                                         s->box[0], s->box[1], e_c, s->n_c, tfx, tfy);
                    energy += te;
                    fx += tfx;
                    fy += tfy;
                    if (!(te > 0.0)) continue;
                    contact = true;
                    double dx = pbc(s->vxp[vi], s->vxp[vo], s->box[0]);
                    double dy = pbc(s->vyp[vi], s->vyp[vo], s->box[1]);
                    voverlap += (rv * 2) - std::sqrt(dx * dx + dy * dy);
                    double dist = std::sqrt(dx * dx + dy * dy);
                    double rs = rv * 2.0;
                    double t_ij = -e_c / rs * (1 - dist / rs);
                    double c_ij = e_c / (rs * rs);
                    double sj, cj;
                    dpm_sincos(tj + phi_v, &sj, &cj);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            // derivatives of the vertex separation by (x, y, theta) of i and j
                            double ax = (a == 0) - Ri * si * (a == 2);
                            double ay = (a == 1) + Ri * ci * (a == 2);
                            double bx = -1.0 * ((b == 0) - Rj * sj * (b == 2));
                            double by = -1.0 * ((b == 1) + Rj * cj * (b == 2));
                            double qa = dx * ax + dy * ay;
                            double qb = dx * bx + dy * by;
                            double dq = (ax * bx + ay * by) / dist;
                            hij[a * 3 + b] += c_ij * qa * qb / (dist * dist) +
                                              t_ij * (dq - qa * qb / (dist * dist * dist));
                            double cx = (b == 0) - Ri * si * (b == 2);
                            double cy = (b == 1) + Ri * ci * (b == 2);
                            double wx = -Ri * ci * (a == 2) * (b == 2);
                            double wy = -Ri * si * (a == 2) * (b == 2);
                            double qc = dx * cx + dy * cy;
                            double dqc = (ax * cx + dx * wx + ay * cy + dy * wy) / dist;
                            hii[a * 3 + b] += c_ij * qa * qc / (dist * dist) +
                                              t_ij * (dqc - qa * qc / (dist * dist * dist));
                        }
                }
                if (contact) count_i++;
            }
            long t = i * M + n;
            P[ORC_P_POTENTIAL * S + t] = energy;
            P[ORC_P_FORCE_X * S + t] = fx;
            P[ORC_P_FORCE_Y * S + t] = fy;
            double xd = pbc(s->x[i], s->x[j], s->box[0]);
            double yd = pbc(s->y[i], s->y[j], s->box[1]);
            P[ORC_P_DIST_X * S + t] = xd;
            P[ORC_P_DIST_Y * S + t] = yd;
            put_i64(P, S, ORC_P_ID_I, t, s->sidx[i]);
            put_i64(P, S, ORC_P_ID_J, t, s->sidx[j]);
            P[ORC_P_VERTEX_OVERLAP * S + t] = voverlap;
            double dist = std::sqrt(xd * xd + yd * yd);
            P[ORC_P_OVERLAP * S + t] = dist - (s->r[i] + s->r[j]);
            P[ORC_P_RADSUM * S + t] = s->r[i] + s->r[j];
            P[ORC_P_ANGLE * S + t] = std::atan2(yd, xd);
            P[ORC_P_ANGLE_I * S + t] = ti;
            P[ORC_P_ANGLE_J * S + t] = tj;
            put_i64(P, S, ORC_P_VERTEX_CONTACTS, t, count_i);
            double mu = 0.0;
            if (fx != 0.0 || fy != 0.0) {  // |tangential| / |normal| force
                double nx = fx * xd / dist, ny = fy * yd / dist;
                double nf = std::sqrt(nx * nx + ny * ny);
                double qx = fx - nx, qy = fy - ny;
                double tf = std::sqrt(qx * qx + qy * qy);
                mu = tf / nf;
            }
            P[ORC_P_FRICTION * S + t] = mu;
            const int hp[9] = {ORC_P_H_XX, ORC_P_H_XY, ORC_P_H_XT, ORC_P_H_YX, ORC_P_H_YY,
                               ORC_P_H_YT, ORC_P_H_TX, ORC_P_H_TY, ORC_P_H_TT};
            const int hs[9] = {ORC_P_HII_XX, ORC_P_HII_XY, ORC_P_HII_XT, ORC_P_HII_YX, ORC_P_HII_YY,
                               ORC_P_HII_YT, ORC_P_HII_TX, ORC_P_HII_TY, ORC_P_HII_TT};
            for (int c = 0; c < 9; ++c) {
                P[hp[c] * S + t] += hij[c];
                P[hs[c] * S + t] += hii[c];
            }
        }
    }
}

}  // namespace

int orc_pair_analytics(orc_system* s, int64_t* row_length) {
    if (s->method == ORC_NBR_NONE) return fail("orc_pair_analytics: neighbour list not initialised");
    const long row = std::max(1L, s->max_alloc);
    const long S = s->N * row;
    const int nf = s->kind == ORC_BUMPY ? ORC_P_NFIELDS : ORC_P_HII_YY + 1;
    pair_fill(s, S, nf);
    if (s->kind == ORC_BUMPY) bumpy_pairs(s, S);
    else disk_pairs(s, S);
    s->pair_row = row;
    s->pair_fields = nf;
    if (row_length) *row_length = row;
    return 0;
}

int orc_pair_download(orc_system* s, int field, void* out, int64_t count) {
    if (s->pair_row == 0) return fail("orc_pair_download: call orc_pair_analytics first");
    if (field < 0 || field >= s->pair_fields) return fail("orc_pair_download: unknown field");
    const long S = s->N * s->pair_row;
    if (count != S) return fail("orc_pair_download: count must be n_particles * row_length");
    std::memcpy(out, s->pairs.data() + (size_t)field * S, sizeof(double) * (size_t)S);
    return 0;
}

// RigidBumpy::calculateParticleArea: rigid_bumpy.cu:1018-1024 with
// kernelCalculateBumpyParticleAreaFull rigid_bumpy/kernels.cu:1226-1283, angleBetweenVectors
// and calcOverlapLenseArea include/particles/base/kernels.cuh:367-381
int orc_bumpy_area(orc_system* s, double* out, int64_t count) {
    if (s->kind != ORC_BUMPY) return fail("orc_bumpy_area: not a rigid bumpy system");
    if (count != s->N) return fail("orc_bumpy_area: count must be n_particles");
    const double PI_R = 3.14159265358979323846;
    const double rv = s->rv;
    auto seg = [](double R, double d) { return R * R * std::acos(d / R) - d * std::sqrt(R * R - d * d); };
    for (long i = 0; i < s->N; ++i) {
        double poly = 0.0, exposed = 0.0, va = PI_R * rv * rv;
        long f = s->vstart[i], nv = s->nverts[i];
        double cx = s->vxp[f], cy = s->vyp[f];
        double px = s->vxp[f + nv - 1], py = s->vyp[f + nv - 1];
        for (long k = 0; k < nv; ++k) {
            double own = 0.0;
            long nx_i = f + ((k + 1) % nv);
            double nx = s->vxp[nx_i], ny = s->vyp[nx_i];
            poly += cx * ny - nx * cy;
            double ms = (nx - cx) * (cy - py) - (ny - cy) * (cx - px);
            double mc = (nx - cx) * (cx - px) + (ny - cy) * (cy - py);
            double ang = std::atan2(ms, mc);
            double dx = nx - cx, dy = ny - cy;
            double r = std::sqrt(dx * dx + dy * dy);
            if (r < (2 * rv - 1e-10)) {
                double d1 = (r * r - rv * rv + rv * rv) / (2 * r);
                double d2 = r - d1;
                own -= (seg(rv, d1) + seg(rv, d2)) / 2.0;
            }
            exposed += va * (PI_R - ang) / (2.0 * PI_R) - own;
            px = cx;
            py = cy;
            cx = nx;
            cy = ny;
        }
        out[s->sidx[i]] = std::fabs(poly) * 0.5 + exposed;
    }
    return 0;
}

int orc_neighbor_stats(orc_system* s, int64_t* ma, int64_t* tot, int64_t* nr, int64_t* ncr) {
    if (ma) *ma = s->max_alloc;
    if (tot) { int64_t t = 0; for (long i = 0; i < s->N; ++i) t += s->nnum[i]; *tot = t; }
    if (nr) *nr = s->n_rebuilds;
    if (ncr) *ncr = s->n_cell_rebuilds;
    return 0;
}

int orc_zero_forces(orc_system* s) { zero_forces(s); return 0; }
int orc_compute_forces(orc_system* s) { compute_forces(s); return 0; }
int orc_kick(orc_system* s, double dt) { kick(s, dt); return 0; }
int orc_drift(orc_system* s, double dt) { drift(s, dt); return 0; }
int orc_mix(orc_system* s, double a) { mix(s, a); return 0; }
int orc_zero_velocities(orc_system* s) { zero_velocities(s); return 0; }

// NVE::step src/integrator/nve.cu:15-25, looped as scripts/runNVE.cpp:43-47
int orc_nve_run(orc_system* s, double dt, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        kick(s, 0.5 * dt);
        drift(s, dt);
        check_for_update(s);
        zero_forces(s);
        compute_forces(s);
        kick(s, 0.5 * dt);
    }
    return 0;
}

int orc_reduce(orc_system* s, int what, double* out) {
    switch (what) {
        case ORC_RED_PE: *out = sum_pe(s); return 0;
        case ORC_RED_KE:
            kinetic(s);
            *out = canonical_sum_f(s->N, [&](long i) { return s->ke[i]; });
            return 0;
        case ORC_RED_TEMPERATURE:  // Particle::calculateTemperature particle.cu:1338-1341
            kinetic(s);
            *out = canonical_sum_f(s->N, [&](long i) { return s->ke[i]; }) * 2.0 / s->ndof();
            return 0;
        case ORC_RED_POWER: *out = power_fire(s); return 0;
        case ORC_RED_FMAX_SQ: *out = fmax_sq(s); return 0;
        case ORC_RED_FORCE_BALANCE: {  // Particle::getForceBalance particle.cu:231-235
            double a = canonical_sum_f(s->N, [&](long i) { return s->fx[i]; });
            double b = canonical_sum_f(s->N, [&](long i) { return s->fy[i]; });
            *out = std::sqrt(a * a + b * b);
            return 0;
        }
        case ORC_RED_PRESSURE: stress(s); *out = pressure(s); return 0;
        case ORC_RED_STRESS_XX: stress(s); *out = canonical_sum_f(s->N, [&](long i) { return s->sxx[i]; }); return 0;
        case ORC_RED_STRESS_XY: stress(s); *out = canonical_sum_f(s->N, [&](long i) { return s->sxy[i]; }); return 0;
        case ORC_RED_STRESS_YX: stress(s); *out = canonical_sum_f(s->N, [&](long i) { return s->syx[i]; }); return 0;
        case ORC_RED_STRESS_YY: stress(s); *out = canonical_sum_f(s->N, [&](long i) { return s->syy[i]; }); return 0;
        case ORC_RED_CONTACTS: {
            count_contacts(s);
            long t = 0;
            for (long i = 0; i < s->N; ++i) t += s->contacts[i];
            *out = (double)t;
            return 0;
        }
        case ORC_RED_VERTEX_CONTACTS: {
            count_contacts(s);
            long t = 0;
            for (long i = 0; i < s->N; ++i) t += s->vcontacts[i];
            *out = (double)t;
            return 0;
        }
        case ORC_RED_PACKING_FRACTION: *out = particle_area(s) / (s->box[0] * s->box[1]); return 0;
        case ORC_RED_MAX_NEIGH_DISP_SQ: *out = max_f(s->N, [&](long i) { return s->nd2[i]; }); return 0;
        case ORC_RED_MAX_CELL_DISP_SQ: *out = max_f(s->N, [&](long i) { return s->cd2[i]; }); return 0;
    }
    return fail("reduce: unknown quantity");
}

int orc_minimize_fire(orc_system* s, const orc_fire_params* p, orc_fire_result* out) {
    minimize_fire(s, p, out);
    return 0;
}
int orc_fire_state_init(orc_fire_state* st, double a, double dt) { fire_init(st, a, dt); return 0; }
int orc_fire_step(orc_system* s, orc_fire_state* st) { fire_step(s, st); return 0; }

int orc_sincos(const double* x, double* s_out, double* c_out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) dpm_sincos(x[i], &s_out[i], &c_out[i]);
    return 0;
}

int orc_adam_reset(orc_system* s) {
    adam_reset(s);
    return 0;
}
int orc_adam_step(orc_system* s, int64_t step, double a, double b1, double b2, double eps) {
    adam_step(s, step, a, b1, b2, eps);
    return 0;
}
int orc_minimize_adam(orc_system* s, double a, double b1, double b2, double eps, double pe_t,
                      double diff_t, int64_t max_steps, int64_t* steps_out) {
    long st = minimize_adam(s, a, b1, b2, eps, pe_t, diff_t, max_steps);
    if (steps_out) *steps_out = st;
    return 0;
}

int orc_compute_stress(orc_system* s) { stress(s); return 0; }
int orc_count_contacts(orc_system* s) { count_contacts(s); return 0; }
int orc_scale_positions(orc_system* s, double f) { scale_positions(s, f); return 0; }
int orc_scale_box(orc_system* s, double f) { scale_box(s, f); return 0; }

// Particle::scaleToPackingFractionFull particle.cu:1047-1055
int orc_scale_to_packing_fraction(orc_system* s, double phi) {
    double ns = std::pow(particle_area(s) / phi, 1.0 / 2);
    double side = std::pow(s->box[0] * s->box[1], 1.0 / 2);
    scale_positions(s, ns / side);
    s->box[0] = ns;
    s->box[1] = ns;
    update_cell_size(s);
    return 0;
}
int orc_snapshot_save(orc_system* s) { snapshot_save(s); return 0; }
int orc_snapshot_restore(orc_system* s) {
    if (!s->snap.valid) return fail("snapshot_restore: no snapshot");
    snapshot_restore(s);
    return 0;
}

// Particle::removeMeanVelocities particle.cu:865-869
int orc_remove_mean_velocity(orc_system* s) {
    double mx = canonical_sum_f(s->N, [&](long i) { return s->vx[i]; }) / s->N;
    double my = canonical_sum_f(s->N, [&](long i) { return s->vy[i]; }) / s->N;
    for (long i = 0; i < s->N; ++i) { s->vx[i] -= mx; s->vy[i] -= my; }
    return 0;
}
// Particle::scaleVelocitiesToTemperature particle.cu:899-906 (velocities.scale: x *= f)
int orc_scale_to_temperature(orc_system* s, double T) {
    kinetic(s);
    double cur = canonical_sum_f(s->N, [&](long i) { return s->ke[i]; }) * 2.0 / s->ndof();
    if (cur <= 0.0) return fail("scale_to_temperature: zero temperature");
    double f = std::sqrt(T / cur);
    for (long i = 0; i < s->N; ++i) { s->vx[i] *= f; s->vy[i] *= f; }
    return 0;
}

int orc_jam(orc_system* s, const orc_jam_params* p, orc_jam_result* out) {
    if (s->method == ORC_NBR_NONE) return fail("jam: neighbours not configured");
    jam(s, p, out);
    return 0;
}

// Rigid bumpy geometry: RigidBumpy::setParticleStartIndex (rigid_bumpy.cu:476-478),
// kernelSetVertexParticleIndex (rigid_bumpy/kernels.cu:1099-1113),
// kernelInitializeVerticesOnParticles (:1137-1162), kernelGetVertexMasses (:1164-1169),
// kernelGetMomentsOfInertia (:1171-1189), RigidBumpy::setDegreesOfFreedom (rigid_bumpy.cu:618-624).
int orc_bumpy_setup(orc_system* s, const int64_t* nvp, double rv, int rotation) {
    if (s->kind != ORC_BUMPY) return fail("bumpy_setup: not a bumpy system");
    long tot = 0;
    for (long p = 0; p < s->N; ++p) tot += nvp[p];
    if (tot != s->Nv) return fail("bumpy_setup: vertex counts do not sum to n_vertices");
    s->rv = rv;
    s->rotation = rotation != 0;
    // counts arrive in static order; at set-up internal order == static order
    for (long p = 0; p < s->N; ++p) s->nverts[p] = nvp[s->sidx[p]];
    long acc = 0;
    for (long p = 0; p < s->N; ++p) { s->vstart[p] = acc; acc += s->nverts[p]; }
    const double PI = M_PI;
    for (long p = 0; p < s->N; ++p) {
        double a = s->th[p], rad = s->r[p];
        long nv = s->nverts[p], st = s->vstart[p];
        double inc = 2 * PI / nv;
        double inner = rad - rv;
        for (long i = 0; i < nv; ++i) {
            s->vpart[st + i] = p;
            double va = i * inc + a;
            double sv, cv;
            dpm_sincos(va, &sv, &cv);
            s->vxp[st + i] = s->x[p] + inner * cv;
            s->vyp[st + i] = s->y[p] + inner * sv;
        }
    }
    for (long v = 0; v < s->Nv; ++v) {
        long p = s->vpart[v];
        s->vm[v] = s->m[p] / static_cast<double>(s->nverts[p]);
    }
    for (long p = 0; p < s->N; ++p) {
        long f0 = s->vstart[p];
        double mass = s->vm[f0], I = 0.0;
        for (long i = 0; i < s->nverts[p]; ++i) {
            double rx = s->vxp[f0 + i] - s->x[p], ry = s->vyp[f0 + i] - s->y[p];
            I += mass * (rx * rx + ry * ry);
        }
        s->moi[p] = I;
    }
    return 0;
}

}  // extern "C"
