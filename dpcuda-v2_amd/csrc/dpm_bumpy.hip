// dpm_bumpy.hip -- rigid bumpy particles (rings of vertex disks) on gfx950.
//
// Layout in HBM (in addition to the per-particle arrays of dpm_core.hip):
//   vpos      : double2[Nv]  vertex positions, particle-contiguous (vstart[p] .. +nverts[p])
//   vmass     : double[Nv]
//   vpart     : int32[Nv]    owning particle of each vertex
//   vnbr      : int32 ELL column-major vnbr[k*Nv + v]; vnum[Nv]
//   vfrc,vtq,vpe : per-vertex force/torque/energy of the last evaluation
//   angle, omega, torque, moi : double[N]
#include <algorithm>
#include <cmath>
#include <vector>

#include "dpm_kernels.hpp"
#include "dpm_rigid.hpp"

namespace dpm {

// ------------------------------------------------------------------ kernels
// kernelCalcRigidBumpyForces1, src/particles/rigid_bumpy/kernels.cu:366-417: the force,
// torque and energy of vertex v from its vertex row.  mode 0: write (arrays known zero),
// 1: accumulate (+=)
__device__ __forceinline__ void vertex_force_one(const Dev& d, const PairParams& p, double rv,
                                                 int64_t v, int mode) {
    int nn = min(d.vnum[v], d.vcap);
    if (nn == 0) {
        if (mode == 0) {
            d.vfrc[v] = make_double2(0.0, 0.0);
            d.vtq[v] = 0.0;
            d.vpe[v] = 0.0;
        }
        return;
    }
    double2 xv = d.vpos[v];
    double4 cq = d.P0[d.vpart[v]];
    double2 c = make_double2(cq.x, cq.y);
    double ax = 0.0, ay = 0.0, e = 0.0;
    for (int k = 0; k < nn; ++k) {
        int u = d.vnbr[(int64_t)k * d.Nv + v];
        double2 xu = d.vpos[u];
        double fx, fy;
        double ie = pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
        ax += fx;
        ay += fy;
        e += ie;
    }
    // calcTorque, include/particles/base/kernels.cuh:237-241
    double dx = xv.x - c.x, dy = xv.y - c.y;
    double t = ay * dx - ax * dy;
    if (mode == 0) {
        d.vfrc[v] = make_double2(0.0 + ax, 0.0 + ay);
        d.vtq[v] = 0.0 + t;
        d.vpe[v] = 0.0 + e;
    } else {
        double2 f = d.vfrc[v];
        f.x += ax;
        f.y += ay;
        d.vfrc[v] = f;
        d.vtq[v] += t;
        d.vpe[v] += e;
    }
}

__global__ __launch_bounds__(BLOCK) void k_bumpy_vertex_forces(Dev d, PairParams p, double rv,
                                                               int mode) {
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= d.Nv) return;
    vertex_force_one(d, p, rv, v, mode);
}

// kernelCalcRigidBumpyParticleForces1, rigid_bumpy/kernels.cu:419-442 (sequential
// vertex order, as the reference).  mode 0 write, 1 accumulate.
__global__ void k_bumpy_particle_reduce(Dev d, int mode) {
    int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= d.N) return;
    double ax = 0.0, ay = 0.0, at = 0.0, e = 0.0;
    int st = d.vstart[p], nv = d.nverts[p];
    for (int k = 0; k < nv; ++k) {
        double2 f = d.vfrc[st + k];
        ax += f.x;
        ay += f.y;
        at += d.vtq[st + k];
        e += d.vpe[st + k];
    }
    if (mode == 0) {
        d.frc[p] = make_double2(0.0 + ax, 0.0 + ay);
        d.torque[p] = 0.0 + at;
        d.pe[p] = 0.0 + e;
    } else {
        double2 f = d.frc[p];
        f.x += ax;
        f.y += ay;
        d.frc[p] = f;
        d.torque[p] += at;
        d.pe[p] += e;
    }
}

__global__ void k_bumpy_zero_vertex(Dev d) {
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= d.Nv) return;
    d.vfrc[v] = make_double2(0.0, 0.0);
    d.vtq[v] = 0.0;
    d.vpe[v] = 0.0;
}

// kernelUpdateRigidVelocities, rigid_bumpy/kernels.cu:123-150
__global__ void k_bumpy_kick(Dev d, double dt, int rotation) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 f = d.frc[i], v = d.vel[i];
    double w = d.omega[i];
    if (rotation) w += d.torque[i] * dt / d.moi[i];
    else w = 0.0;
    double m = d.P0[i].w;
    v.x += f.x * dt / m;
    v.y += f.y * dt / m;
    d.vel[i] = v;
    d.omega[i] = w;
}

// kernelUpdateRigidPositions, rigid_bumpy/kernels.cu:21-78
__global__ __launch_bounds__(BLOCK) void k_bumpy_drift_com(Dev d, double dt) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double4 q = d.P0[i];
        double2 v = d.vel[i];
        double a = d.angle[i];
        d.lpos[i] = make_double2(q.x, q.y);
        double ddx = v.x * dt, ddy = v.y * dt, dda = d.omega[i] * dt;
        d.dth[i] = dda;
        double nx = q.x + ddx, ny = q.y + ddy;
        d.P0[i] = make_double4(nx, ny, q.z, q.w);
        d.angle[i] = a + dda;
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = nx - ln.x, dyn = ny - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = nx - lc.x, dyc = ny - lc.y;
        cd = dxc * dxc + dyc * dyc;
    }
    double bn = block_max(nd);
    double bc = block_max(cd);
    if (threadIdx.x == 0) {
        d.pmax_nd[blockIdx.x] = bn;
        d.pmax_cd[blockIdx.x] = bc;
    }
}

// kernelTranslateAndRotateVertices1, rigid_bumpy/kernels.cu:152-201
__global__ void k_bumpy_drift_vertices(Dev d) {
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= d.Nv) return;
    int p = d.vpart[v];
    double4 nq = d.P0[p];
    double2 o = d.lpos[p], n = make_double2(nq.x, nq.y), x = d.vpos[v];
    double dth = d.dth[p];
    double rx = x.x - o.x, ry = x.y - o.y;
    double sn, c;
    dpm_sincos(dth, &sn, &c);
    double rxr = rx * c - ry * sn;
    double ryr = rx * sn + ry * c;
    d.vpos[v] = make_double2(n.x + rxr, n.y + ryr);
}

// kernelRigidBumpyAdamStep, rigid_bumpy/kernels.cu:1004-1093: moments of the force and
// the torque, the bias-corrected step; the angle (and its change, which the vertex pass
// reads) moves only with rotation on, so without it the vertices turn by the last
// recorded change, as in the reference.  The vertices follow in k_bumpy_drift_vertices
// (kernelTranslateAndRotateVertices1, launched after it in rigid_bumpy.cu:1473-1486).
__global__ __launch_bounds__(BLOCK) void k_badam_update(Dev d, double alpha, double b1,
                                                        double b2, double omb1, double omb2,
                                                        double eps, int rotation) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double nd = 0.0, cd = 0.0;
    if (i < d.N) {
        double2 f = d.frc[i], m1 = d.m1[i], m2 = d.m2[i], ma = d.ma[i];
        double t = d.torque[i];
        m1.x = fma(b1, m1.x, (b1 - 1) * f.x);
        m1.y = fma(b1, m1.y, (b1 - 1) * f.y);
        ma.x = fma(b1, ma.x, (b1 - 1) * t);
        m2.x = fma(b2, m2.x, (1 - b2) * f.x * f.x);
        m2.y = fma(b2, m2.y, (1 - b2) * f.y * f.y);
        ma.y = fma(b2, ma.y, (1 - b2) * t * t);
        double mhx = m1.x / omb1, mhy = m1.y / omb1, mha = ma.x / omb1;
        double vhx = m2.x / omb2, vhy = m2.y / omb2, vha = ma.y / omb2;
        double ux = -alpha * mhx / (sqrt(vhx) + eps);
        double uy = -alpha * mhy / (sqrt(vhy) + eps);
        double ua = -alpha * mha / (sqrt(vha) + eps);
        double4 q = d.P0[i];
        d.lpos[i] = make_double2(q.x, q.y);
        double nx = q.x + ux, ny = q.y + uy;
        if (rotation) {
            d.angle[i] = d.angle[i] + ua;
            d.dth[i] = ua;
        }
        d.P0[i] = make_double4(nx, ny, q.z, q.w);
        double2 ln = d.lref[i], lc = d.cref[i];
        double dxn = nx - ln.x, dyn = ny - ln.y;
        nd = dxn * dxn + dyn * dyn;
        double dxc = nx - lc.x, dyc = ny - lc.y;
        cd = dxc * dxc + dyc * dyc;
        d.m1[i] = m1;
        d.m2[i] = m2;
        d.ma[i] = ma;
    }
    double bn = block_max(nd);
    double bc = block_max(cd);
    if (threadIdx.x == 0) {
        d.pmax_nd[blockIdx.x] = bn;
        d.pmax_cd[blockIdx.x] = bc;
    }
}

__global__ void k_bumpy_mix(Dev d, double alpha) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double2 v = d.vel[i];
    double w = d.omega[i];
    rigid_mix_one(v, d.frc[i], w, d.torque[i], alpha);
    d.vel[i] = v;
    d.omega[i] = w;
}

__device__ __forceinline__ bool within_v(double ax, double ay, double bx, double by,
                                         const NbrParams& nb, double c2) {
    double dx = pbc(ax, bx, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
    double dy = pbc(ay, by, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
    return dx * dx + dy * dy < c2;
}

// kernelUpdateVertexNeighborList, rigid_bumpy/kernels.cu:870-922
// gated: run only when the rebuild chain chose a list rebuild
// ring: the vertices sit on their rings (dpm_system::vring), so q's vertices are scanned
// only when v lies within reach of q's centre (vertex_ring_near); the rows are unchanged
__global__ __launch_bounds__(BLOCK) void k_bumpy_vertex_build(Dev d, NbrParams nb, double rv,
                                                              double rmax, int ring,
                                                              int gated) {
    if (gated && d.ctl->kind < 1) return;
    const double vc = sqrt(nb.vc2);
    // the centre's image serves q's vertices (pbc_known): needs the ring gate
    const bool shift = ring && fmin(nb.halfLx, nb.halfLy) > vertex_shift_limit(rmax, rv, vc);
    int mx = 0;
    for (int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x; v < d.Nv;
         v += (int64_t)gridDim.x * BLOCK) {
        double2 xv = d.vpos[v];
        int p = d.vpart[v];
        int nn = min(d.nnum[p], d.cap);
        int cnt = 0;
        for (int k = 0; k < nn; ++k) {
            int q = d.nbr[(int64_t)k * d.N + p];
            double4 cq = d.P0[q];
            const double kx = image_index(xv.x - cq.x, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
            const double ky = image_index(xv.y - cq.y, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
            const double dx = pbc_known(xv.x, cq.x, nb.Lx, kx, nb.fast);  // = pbc(xv, cq)
            const double dy = pbc_known(xv.y, cq.y, nb.Ly, ky, nb.fast);
            const double d2 = dx * dx + dy * dy;
            if (d2 < nb.vp2 && (!ring || vertex_ring_near(d2, cq.z, rv, vc))) {
                int f0 = d.vstart[q], nv = d.nverts[q];
                for (int u = f0; u < f0 + nv; ++u) {
                    double2 xu = d.vpos[u];
                    bool in;
                    if (shift) {
                        const double ux = pbc_known(xv.x, xu.x, nb.Lx, kx, nb.fast);
                        const double uy = pbc_known(xv.y, xu.y, nb.Ly, ky, nb.fast);
                        in = ux * ux + uy * uy < nb.vc2;
                    } else {
                        in = within_v(xv.x, xv.y, xu.x, xu.y, nb, nb.vc2);
                    }
                    if (in) {
                        if (cnt < d.vcap) d.vnbr[(int64_t)cnt * d.Nv + v] = u;
                        cnt++;
                    }
                }
            }
        }
        d.vnum[v] = cnt;
        mx = max(mx, cnt);
    }
    // row-length statistics: one atomic per block
    int m = block_max_int(mx);
    if (threadIdx.x == 0) {
        atomicMax(&d.ctl->max_vnum, m);
        atomicMax(&d.ctl->max_vnum_all, m);
    }
}

// The same rows for vertices on their rings (dpm_system::vring): a half-wave per particle,
// lanes over its vertices; q's vertices are tested against one gated vertex at a time by
// all 32 lanes (q's row entry and centre are the same address for the half), and the hits
// append to that vertex's row in ascending order (ballot ranks).  Grid-stride over the
// particles.
constexpr int VB_HALF = 32;
__global__ __launch_bounds__(BLOCK) void k_bumpy_vertex_rows(Dev d, NbrParams nb, double rv,
                                                             double rmax, int gated) {
    if (gated && d.ctl->kind < 1) return;
    const double vc = sqrt(nb.vc2);
    const bool shift = fmin(nb.halfLx, nb.halfLy) > vertex_shift_limit(rmax, rv, vc);
    const int lane = (int)(threadIdx.x % VB_HALF);
    const int hs = (int)(threadIdx.x % WAVE) / VB_HALF * VB_HALF;
    const unsigned lt = (1u << lane) - 1u;
    const int64_t halves = (int64_t)gridDim.x * (BLOCK / VB_HALF);
    int mx = 0;
    for (int64_t p = ((int64_t)blockIdx.x * BLOCK + threadIdx.x) / VB_HALF; p < d.N;
         p += halves) {
        const int st = d.vstart[p], nv = d.nverts[p];
        const int nn = min(d.nnum[p], d.cap);
        for (int a0 = 0; a0 < nv; a0 += VB_HALF) {
            const bool has = a0 + lane < nv;
            const double2 xv = has ? d.vpos[st + a0 + lane] : make_double2(0.0, 0.0);
            int cnt = 0;
            for (int k = 0; k < nn; ++k) {
                const int q = d.nbr[(int64_t)k * d.N + p];
                const double4 cq = d.P0[q];
                const double kx = image_index(xv.x - cq.x, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
                const double ky = image_index(xv.y - cq.y, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
                const double dx = pbc_known(xv.x, cq.x, nb.Lx, kx, nb.fast);  // = pbc(xv, cq)
                const double dy = pbc_known(xv.y, cq.y, nb.Ly, ky, nb.fast);
                const double d2 = dx * dx + dy * dy;
                const bool gate = has && d2 < nb.vp2 && vertex_ring_near(d2, cq.z, rv, vc);
                const unsigned gm = (unsigned)(__builtin_amdgcn_ballot_w64(gate) >> hs);
                if (gm == 0) continue;
                const int f0 = d.vstart[q], nq = d.nverts[q];
                for (int u0 = 0; u0 < nq; u0 += VB_HALF) {
                    const bool uin = u0 + lane < nq;
                    const int u = f0 + u0 + lane;
                    const double2 xu = uin ? d.vpos[u] : make_double2(0.0, 0.0);
                    for (unsigned m = gm; m != 0; m &= m - 1) {
                        const int g = __builtin_ctz(m);
                        const double gx = __shfl(xv.x, g, VB_HALF), gy = __shfl(xv.y, g, VB_HALF);
                        double ex, ey;
                        if (shift) {
                            ex = pbc_known(gx, xu.x, nb.Lx, __shfl(kx, g, VB_HALF), nb.fast);
                            ey = pbc_known(gy, xu.y, nb.Ly, __shfl(ky, g, VB_HALF), nb.fast);
                        } else {
                            ex = pbc(gx, xu.x, nb.Lx, nb.invLx, nb.halfLx, nb.fast);
                            ey = pbc(gy, xu.y, nb.Ly, nb.invLy, nb.halfLy, nb.fast);
                        }
                        const bool hit = uin && ex * ex + ey * ey < nb.vc2;
                        const unsigned hm = (unsigned)(__builtin_amdgcn_ballot_w64(hit) >> hs);
                        if (hm == 0) continue;
                        const int pos = __shfl(cnt, g, VB_HALF) + __builtin_popcount(hm & lt);
                        if (hit && pos < d.vcap)
                            d.vnbr[(int64_t)pos * d.Nv + st + a0 + g] = u;
                        if (lane == g) cnt += __builtin_popcount(hm);
                    }
                }
            }
            if (has) {
                d.vnum[st + a0 + lane] = cnt;
                mx = max(mx, cnt);
            }
        }
    }
    int m = block_max_int(mx);
    if (threadIdx.x == 0) {
        atomicMax(&d.ctl->max_vnum, m);
        atomicMax(&d.ctl->max_vnum_all, m);
    }
}

// exclusive scan of nverts (new order) into vstart for the vertex reorder
__global__ __launch_bounds__(1024) void k_scan_counts(const Ctl* ctl, int gated, const int* cnt,
                                                      int* start, int64_t n) {
    if (gated && ctl->kind != 2) return;
    __shared__ int sums[1024];
    int t = threadIdx.x;
    int64_t per = (n + 1023) / 1024;
    int64_t lo = t * per, hi = min(lo + per, n);
    int loc = 0;
    for (int64_t c = lo; c < hi; ++c) loc += cnt[c];
    sums[t] = loc;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int base = sums[t] - loc;
    for (int64_t c = lo; c < hi; ++c) {
        start[c] = base;
        base += cnt[c];
    }
}

// kernelReorderRigidBumpyVertexData, rigid_bumpy/kernels.cu:844-868
__global__ void k_bumpy_reorder_vertices(Dev d, Spare sp, int gated) {
    if (gated && d.ctl->kind != 2) return;
    int64_t ov = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (ov >= d.Nv) return;
    int op = d.vpart[ov];
    int np = sp.old_to_new[op];
    int rel = (int)ov - d.vstart[op];
    int nv = sp.vstart[np] + rel;
    sp.vpart[nv] = np;
    sp.vpos[nv] = d.vpos[ov];
    sp.vmass[nv] = d.vmass[ov];
    sp.svidx[nv] = d.svidx[ov];
}

__global__ void k_bumpy_vertex_copyback(Dev d, Spare sp, int gated) {
    if (gated && d.ctl->kind != 2) return;
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v < d.Nv) {
        d.vpart[v] = sp.vpart[v];
        d.vpos[v] = sp.vpos[v];
        d.vmass[v] = sp.vmass[v];
        d.svidx[v] = sp.svidx[v];
    }
    if (v < d.N) d.vstart[v] = sp.vstart[v];
}

// kernelCalcRigidBumpyStressTensor, rigid_bumpy/kernels.cu:705-789
__global__ void k_bumpy_stress(Dev d, PairParams p, double rv, double area) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double axx = 0.0, axy = 0.0, ayx = 0.0, ayy = 0.0;
    double4 cq = d.P0[i];
    double2 c = make_double2(cq.x, cq.y), vel = d.vel[i];
    double w = d.omega[i];
    int st = d.vstart[i], nv = d.nverts[i];
    for (int a = 0; a < nv; ++a) {
        int v = st + a;
        double2 xv = d.vpos[v];
        double vm = d.vmass[v];
        double rx = xv.x - c.x, ry = xv.y - c.y;
        double prad = sqrt(rx * rx + ry * ry);
        double vvx = vel.x, vvy = vel.y;
        if (w != 0.0) {
            double ang = atan2(ry, rx);
            vvx -= sin(ang) * w * prad;
            vvy += cos(ang) * w * prad;
        }
        axx += vm * vvx * vvx;
        axy += vm * vvx * vvy;
        ayx += vm * vvy * vvx;
        ayy += vm * vvy * vvy;
        int nn = min(d.vnum[v], d.vcap);
        for (int k = 0; k < nn; ++k) {
            int u = d.vnbr[(int64_t)k * d.Nv + v];
            if (d.vpart[u] == i) continue;
            double2 xu = d.vpos[u];
            double xd = pbc(xv.x, xu.x, p.Lx, p.invLx, p.halfLx, p.fast);
            double yd = pbc(xv.y, xu.y, p.Ly, p.invLy, p.halfLy, p.fast);
            double fx, fy;
            pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
            axx += xd * fx / 2;
            axy += xd * fy / 2;
            ayx += yd * fx / 2;
            ayy += yd * fy / 2;
        }
    }
    d.sxx[i] = axx / area;
    d.sxy[i] = axy / area;
    d.syx[i] = ayx / area;
    d.syy[i] = ayy / area;
}

// kernelCountRigidBumpyContacts, rigid_bumpy/kernels.cu:924-997
__global__ void k_bumpy_contacts(Dev d, PairParams p, double rv) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double vrs = 2 * rv;
    int nc = 0, nvc = 0;
    int nn = min(d.nnum[i], d.cap);
    int st = d.vstart[i], nv = d.nverts[i];
    for (int k = 0; k < nn; ++k) {
        int q = d.nbr[(int64_t)k * d.N + i];
        bool is_contact = false;
        for (int a = 0; a < nv; ++a) {
            int v = st + a;
            double2 xv = d.vpos[v];
            int vn = min(d.vnum[v], d.vcap);
            for (int k2 = 0; k2 < vn; ++k2) {
                int u = d.vnbr[(int64_t)k2 * d.Nv + v];
                if (d.vpart[u] != q) continue;
                double2 xu = d.vpos[u];
                double dx = pbc(xv.x, xu.x, p.Lx, p.invLx, p.halfLx, p.fast);
                double dy = pbc(xv.y, xu.y, p.Ly, p.invLy, p.halfLy, p.fast);
                double dist = sqrt(dx * dx + dy * dy);
                if (dist < vrs) {
                    if (!is_contact) nc++;
                    nvc++;
                    is_contact = true;
                }
            }
            if (is_contact) break;
        }
    }
    d.contacts[i] = nc;
    d.vcontacts[i] = nvc;
}

// kernelInitializeVerticesOnParticles (rigid_bumpy/kernels.cu:1137-1162),
// kernelGetVertexMasses (:1164-1169), kernelGetMomentsOfInertia (:1171-1189)
__global__ void k_bumpy_init_vertices(Dev d, double rv) {
    int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= d.N) return;
    double4 cq = d.P0[p];
    double a = d.angle[p], r = cq.z;
    int nv = d.nverts[p], st = d.vstart[p];
    double2 c = make_double2(cq.x, cq.y);
    for (int i = 0; i < nv; ++i) {
        d.vpos[st + i] = ring_vertex(c, r, a, nv, rv, i);
        d.vpart[st + i] = (int)p;
    }
    double vm = cq.w / static_cast<double>(nv);
    double I = 0.0;
    for (int i = 0; i < nv; ++i) {
        d.vmass[st + i] = vm;
        double rx = d.vpos[st + i].x - c.x, ry = d.vpos[st + i].y - c.y;
        I += vm * (rx * rx + ry * ry);
    }
    d.moi[p] = I;
}

// ------------------------------------------------------------------ FIRE loop
// One FIRE iteration of a rigid bumpy system (Fire::step, src/integrator/fire.cu:13-46,
// with RigidBumpy's updatePositions / updateVelocities / mixVelocitiesAndForces /
// calculateForces) as a chain of launches that a graph replays; each kernel reads the
// control block, so a halted or finished loop turns the rest of the chunk into no-ops:
//   k_bfire_move      particles: [uphill: drift(-dt/2), zero v and w] + kick(dt/2) + mix
//                     + drift(dt), and the rebuild requests of the new positions
//   k_bfire_vertices  vertices: the rotations of those moves, in the same order
//   k_bfire_decide    the neighbour check of the step: a list rebuild runs inside the
//                     chunk, a cell rebuild halts the loop until the chunk's end
//   k_build_cell / k_build_verlet, k_bumpy_vertex_build   (gated on that decision)
//   k_bfire_forces    an in-chunk rebuild that overflowed the capacity halts the loop;
//                     else vertex forces, torques, energies (kernelCalcRigidBumpyForces1)
//                     and their per-particle sums (kernelCalcRigidBumpyParticleForces1),
//                     one pass
//   k_bfire_partials  kick(dt/2), then the chunk partials of P = F.v + tau.w, PE and
//                     max(F^2, tau^2)
//   k_fire_ctl        stop test of the step and Fire::step's decision for the next
__device__ __forceinline__ bool bfire_skip(const Dev& d, int gate) {
    const Ctl* c = d.ctl;
    if (gate == G_NORMAL) return c->halt || d.fire->done;
    if (gate == G_RESUME) return !c->halt || !resumable(d) || c->halt_mode != H_FIRE;
    return false;
}

// The rigid moves of Fire::step for particle i (src/integrator/fire.cu:20-40 with the
// rigid overrides): [uphill back-step with the decreased dt, zero v and omega] + kick
// (dt/2) + mix + drift(dt).  Returns the old centre, the centre after the back-step, the
// two rotation angles and the displacements of the new centre from the list / cell
// references.
struct BMove {
    double2 old, mid, now;
    double dth0, dth;
};
__device__ __forceinline__ BMove bfire_move_one(const Dev& d, int64_t i, double dt, double alpha,
                                                int uphill, int rotation, double& nd,
                                                double& cd) {
    BMove m;
    double4 q = d.P0[i];
    double2 v = d.vel[i];
    double a = d.angle[i], w = d.omega[i];
    rigid_fire_move(q, v, a, w, d.frc[i], d.torque[i], d.moi[i], dt, alpha, uphill, rotation,
                    m.old, m.mid, m.dth0, m.dth);
    d.lpos[i] = m.old;
    d.mpos[i] = m.mid;
    if (uphill) {
        double s0, c0;
        dpm_sincos(m.dth0, &s0, &c0);
        d.rot0[i] = make_double2(s0, c0);
    }
    d.dth[i] = m.dth;
    double s1, c1;
    dpm_sincos(m.dth, &s1, &c1);
    d.rot1[i] = make_double2(s1, c1);
    m.now = make_double2(q.x, q.y);
    d.P0[i] = q;
    d.angle[i] = a;
    d.vel[i] = v;
    d.omega[i] = w;
    disp_of(d, i, q.x, q.y, nd, cd);
    return m;
}

__global__ __launch_bounds__(BLOCK) void k_bfire_move(Dev d, NbrParams nb, int pos,
                                                      int rotation) {
    Ctl* c = d.ctl;
    if (c->halt) return;
    const FireCtl* F = d.fire;
    if (F->done) return;
    long long e = c->epoch;
    int tile;
    int64_t i = particle_of_thread(tile);
    double nd = 0.0, cd = 0.0;
    if (i < d.N) bfire_move_one(d, i, F->dt, F->alpha, F->uphill, rotation, nd, cd);
    (void)tile;
    disp_flags(d, nb, nd, cd, pos & 1, stamp(e, pos));  // maxima: k_disp_partials at the end
}

__global__ void k_bfire_vertices(Dev d) {
    if (d.ctl->halt || d.fire->done) return;
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= d.Nv) return;
    int p = d.vpart[v];
    const int uphill = d.fire->uphill;
    // the rotations' sin/cos, formed once per particle by k_bfire_move
    const double2 r0 = uphill ? d.rot0[p] : make_double2(0.0, 1.0), r1 = d.rot1[p];
    double4 nq = d.P0[p];
    d.vpos[v] = bfire_vertex_one(d.vpos[v], d.lpos[p], d.mpos[p], make_double2(nq.x, nq.y), r0.x,
                                 r0.y, r1.x, r1.y, uphill);
}

// Particle::checkForNeighborUpdate of the step (particle.cu:1137-1161) from the
// requests k_bfire_move recorded: 1 = list (rebuilt in the chunk), 2 = cells + list
// (the loop halts; the chain at the chunk's end rebuilds and finishes the step).
__global__ void k_bfire_decide(Dev d, int pos) {
    Ctl* c = d.ctl;
    if (c->halt || d.fire->done) {
        c->kind = 0;
        return;
    }
    int need = pending_rebuild(c, c->epoch, pos);
    if (need >= 2) {
        c->halt = 1;
        c->halt_pos = pos;
        c->halt_mode = H_FIRE;
        c->halt_kind = 2;
        c->kind = 0;
        return;
    }
    c->kind = need;
    if (need == 1) {
        c->max_nnum = 0;
        c->max_vnum = 0;
        c->n_rebuilds += 1;
    }
}

// An in-chunk list rebuild whose rows did not fit: halt with the step unfinished; the
// host regrows the lists, rebuilds and finishes the step (fix_overflow).
__global__ void k_bfire_check(Dev d, int pos) {
    Ctl* c = d.ctl;
    if (c->kind < 1) return;
    if (c->max_nnum > d.cap || c->max_vnum > d.vcap) {
        c->overflow = 1;
        c->halt = 1;
        c->halt_pos = pos;
        c->halt_mode = H_FIRE;
        c->halt_kind = 1;
    }
    c->kind = 0;
}

__global__ __launch_bounds__(BLOCK) void k_bfire_vforce(Dev d, PairParams p, double rv,
                                                        int gate) {
    if (bfire_skip(d, gate)) return;
    int64_t v = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (v >= d.Nv) return;
    vertex_force_one(d, p, rv, v, 0);
}

// The force phase of a FIRE step in one pass over the vertices.  Block b owns the
// particles whose first vertex lies in [256 b, 256 b + 256) and all of their vertices
// (at most 256 + nv_max - 1): lanes compute vertex forces, torques and energies into
// LDS, then one lane per particle sums its vertices in vertex order (the order of
// kernelCalcRigidBumpyParticleForces1), applies the closing kick and stores the
// particle's force, torque and energy.  The per-vertex arrays are not written.
constexpr int BF_TILE = 256;
constexpr int BF_LDS_SMALL = 320;  // vertices per block (BF_TILE + nv_max - 1 must fit):
constexpr int BF_LDS = 1024;       // the small form keeps more blocks resident per CU

__device__ __forceinline__ int first_particle_from(const Dev& d, int64_t v) {
    if (v >= d.Nv) return (int)d.N;
    int p = d.vpart[v];
    return d.vstart[p] == v ? p : p + 1;
}

// threads per block: the small form gives every vertex of the tile its own lane (one
// pass over the vertices); the large form loops
template <int LDSV>
constexpr int bf_threads() { return BF_TILE; }

// Eight waves per SIMD (no spills): the closing kick runs in k_bfire_partials, and the
// vertex ranges of the per-particle sums are read before the vertex pass (two registers).
// (the general pair law and the large, LDS-bound form keep the compiler's own budget)
template <int LDSV, bool HOT>
constexpr int bf_waves() { return LDSV == BF_LDS_SMALL && HOT ? 8 : 1; }

template <int LDSV, bool HOT>
__global__ __launch_bounds__(bf_threads<LDSV>(), (bf_waves<LDSV, HOT>())) void k_bfire_forces(
    Dev d, PairParams p, double rv, int gate, int check_pos) {
    constexpr int NT = bf_threads<LDSV>();
    if (bfire_skip(d, gate)) return;
    if (check_pos >= 0) {  // k_bfire_check's test, made by every block (same words read)
        Ctl* c = d.ctl;
        if (c->kind >= 1 && (c->max_nnum > d.cap || c->max_vnum > d.vcap)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                c->overflow = 1;
                c->halt = 1;
                c->halt_pos = check_pos;
                c->halt_mode = H_FIRE;
                c->halt_kind = 1;
            }
            return;
        }
    }
    __shared__ double sv[4][LDSV];       // per vertex: fx, fy, torque, energy
    __shared__ double sres[4][BF_TILE];  // per particle: the sums of those
    // XCD-aware tile order: each XCD's L2 holds the vertex positions its tiles gather
    const int64_t b0 = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * BF_TILE;
    const int p0 = first_particle_from(d, b0);
    const int p1 = first_particle_from(d, b0 + BF_TILE);
    if (p0 >= p1) return;
    const int np = p1 - p0;
    const int64_t lo = d.vstart[p0];
    const int64_t hi = p1 < d.N ? (int64_t)d.vstart[p1] : d.Nv;
    // the per-particle phase below: np <= BF_TILE <= NT, one particle per lane; the
    // first of its work items reads its particle's vertex range now
    const bool own = (int)threadIdx.x < np;
    int st0 = 0, nv0 = 0;
    if ((int)threadIdx.x < 4 * np) {
        st0 = d.vstart[p0 + (threadIdx.x >> 2)];
        nv0 = d.nverts[p0 + (threadIdx.x >> 2)];
    }
    for (int64_t v = lo + threadIdx.x; v < hi; v += NT) {
        const double4 cq = d.P0[d.vpart[v]];
        const int nn = min(d.vnum[v], d.vcap);
        const double2 xv = d.vpos[v];
        double ax = 0.0, ay = 0.0, e = 0.0;
        if constexpr (HOT) {
            // wave-uniform row loop, two neighbours per batch with the next batch's
            // indices loaded ahead; every ELL slot holds a valid vertex index (zero-filled
            // at allocation, only written with indices), and a slot past the row end
            // adds +0.0, which leaves the sums unchanged
            const int* col = d.vnbr + v;
            const int64_t Nv = d.Nv;
            const int last = d.vcap - 1;
            const double4 qv = make_double4(xv.x, xv.y, rv, 0.0);
            int u0 = col[0], u1 = col[(int64_t)min(1, last) * Nv];
            for (int k = 0; __builtin_amdgcn_ballot_w64(k < nn) != 0; k += 2) {
                const double2 x0 = d.vpos[u0], x1 = d.vpos[u1];
                if (__builtin_amdgcn_ballot_w64(k + 2 < nn) != 0) {
                    u0 = col[(int64_t)min(k + 2, last) * Nv];
                    u1 = col[(int64_t)min(k + 3, last) * Nv];
                }
                double fx, fy, ie;
                ie = pair_hot<true>(qv, make_double4(x0.x, x0.y, rv, 0.0), k < nn, p, fx, fy);
                ax += fx;
                ay += fy;
                e += ie;
                ie = pair_hot<true>(qv, make_double4(x1.x, x1.y, rv, 0.0), k + 1 < nn, p, fx,
                                    fy);
                ax += fx;
                ay += fy;
                e += ie;
            }
        } else {
            for (int k = 0; k < nn; ++k) {
                int u = d.vnbr[(int64_t)k * d.Nv + v];
                double2 xu = d.vpos[u];
                double fx, fy, ie;
                ie = pair_law(xv.x, xv.y, rv, xu.x, xu.y, rv, p, fx, fy);
                ax += fx;
                ay += fy;
                e += ie;
            }
        }
        // calcTorque, include/particles/base/kernels.cuh:237-241
        double dx = xv.x - cq.x, dy = xv.y - cq.y;
        double t = ay * dx - ax * dy;
        int64_t l = v - lo;
        sv[0][l] = 0.0 + ax;
        sv[1][l] = 0.0 + ay;
        sv[2][l] = 0.0 + t;
        sv[3][l] = 0.0 + e;
    }
    __syncthreads();
    // four lanes per particle, one per quantity, each summing in vertex order
    for (int j = threadIdx.x; j < 4 * np; j += NT) {
        int q = j >> 2, comp = j & 3;
        int i = p0 + q;
        const bool first = j == (int)threadIdx.x;
        int64_t st = (first ? st0 : d.vstart[i]) - lo;
        int nv = first ? nv0 : d.nverts[i];
        const double* src = sv[comp] + st;
        double acc = 0.0;
#pragma unroll 8
        for (int k = 0; k < nv; ++k) acc += src[k];
        sres[comp][q] = 0.0 + acc;
    }
    __syncthreads();
    if (own) {
        const int q = threadIdx.x, i = p0 + q;
        d.frc[i] = make_double2(sres[0][q], sres[1][q]);
        d.torque[i] = sres[2][q];
        d.pe[i] = sres[3][q];
    }
}

// The closing kick of k_bfire_forces' particles (kick: 0 = forces only, minimizeFire's
// initial evaluation; 1 = + kick(dt/2)), then the chunk partials of P = F.v + tau.w, PE
// and max(F^2, tau^2) over the particles, in the canonical order (64-particle chunks)
// that k_fire_ctl finishes.  (The kick here rather than at the end of the force pass
// takes a dependent load off the force pass's tail.)
__global__ __launch_bounds__(BLOCK) void k_bfire_partials(Dev d, int gate, int kick,
                                                          int rotation) {
    if (bfire_skip(d, gate)) return;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double pxv = 0.0, pyv = 0.0, ptw = 0.0, pe = 0.0, f2 = 0.0;
    if (i < d.N) {
        double2 f = d.frc[i], v = d.vel[i];
        double t = d.torque[i], w = d.omega[i];
        pe = d.pe[i];
        if (kick) {
            const double hdt = d.fire->dt / 2.0;
            if (rotation) w += t * hdt / d.moi[i];
            else w = 0.0;
            const double mq = d.P0[i].w;
            v.x += f.x * hdt / mq;
            v.y += f.y * hdt / mq;
            d.vel[i] = v;
            d.omega[i] = w;
        }
        pxv = f.x * v.x;
        pyv = f.y * v.y;
        ptw = t * w;
        f2 = fmax(f.x * f.x + f.y * f.y, t * t);
    }
    int64_t nc = (d.N + CHUNK - 1) / CHUNK;
    int64_t nblk = (d.N + BLOCK - 1) / BLOCK;
    double* P = d.part;
    double s0 = wave_tree_sum(pxv);
    double s1 = wave_tree_sum(pyv);
    double s2 = wave_tree_sum(pe);
    double s3 = wave_tree_sum(ptw);
    int64_t ch = i / CHUNK;
    if ((threadIdx.x % WAVE) == 0 && ch < nc) {
        P[ch] = s0;
        P[nc + ch] = s1;
        P[2 * nc + ch] = s2;
        P[3 * nc + nblk + ch] = s3;
    }
    double b = block_max(f2);
    if (threadIdx.x == 0) P[3 * nc + blockIdx.x] = b;
}

// Two-pass form for particles with more vertices than the fused kernel's LDS holds.
// kick: 0 = forces only (minimizeFire's initial evaluation), 1 = + closing kick
__global__ __launch_bounds__(BLOCK) void k_bfire_reduce(Dev d, int kick, int gate,
                                                        int rotation) {
    if (bfire_skip(d, gate)) return;
    const FireCtl* F = d.fire;
    double hdt = F->dt / 2.0;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    double pxv = 0.0, pyv = 0.0, ptw = 0.0, pe = 0.0, f2 = 0.0;
    if (i < d.N) {
        double ax = 0.0, ay = 0.0, at = 0.0, e = 0.0;
        int st = d.vstart[i], nv = d.nverts[i];
        for (int k = 0; k < nv; ++k) {
            double2 f = d.vfrc[st + k];
            ax += f.x;
            ay += f.y;
            at += d.vtq[st + k];
            e += d.vpe[st + k];
        }
        double fx = 0.0 + ax, fy = 0.0 + ay, t = 0.0 + at;
        pe = 0.0 + e;
        d.frc[i] = make_double2(fx, fy);
        d.torque[i] = t;
        d.pe[i] = pe;
        double2 v = d.vel[i];
        double w = d.omega[i];
        if (kick) {
            if (rotation) w += t * hdt / d.moi[i];
            else w = 0.0;
            double m = d.P0[i].w;
            v.x += fx * hdt / m;
            v.y += fy * hdt / m;
            d.vel[i] = v;
            d.omega[i] = w;
        }
        pxv = fx * v.x;
        pyv = fy * v.y;
        ptw = t * w;
        f2 = fmax(fx * fx + fy * fy, t * t);
    }
    int64_t nc = (d.N + CHUNK - 1) / CHUNK;
    int64_t nblk = (d.N + BLOCK - 1) / BLOCK;
    double* P = d.part;
    double s0 = wave_tree_sum(pxv);
    double s1 = wave_tree_sum(pyv);
    double s2 = wave_tree_sum(pe);
    double s3 = wave_tree_sum(ptw);
    int64_t ch = i / CHUNK;
    if ((threadIdx.x % WAVE) == 0 && ch < nc) {
        P[ch] = s0;
        P[nc + ch] = s1;
        P[2 * nc + ch] = s2;
        P[3 * nc + nblk + ch] = s3;
    }
    double b = block_max(f2);
    if (threadIdx.x == 0) P[3 * nc + blockIdx.x] = b;
}

// ------------------------------------------------------------------ host side
int bumpy_alloc(dpm_system* s) {
    Dev& d = s->d;
    int64_t n = s->N, v = s->Nv;
    if (dalloc(s, &d.angle, n) || dalloc(s, &d.omega, n) || dalloc(s, &d.torque, n) ||
        dalloc(s, &d.moi, n) || dalloc(s, &d.vpos, v) || dalloc(s, &d.vmass, v) ||
        dalloc(s, &d.vpart, v) || dalloc(s, &d.vstart, n) || dalloc(s, &d.nverts, n) ||
        dalloc(s, &d.svidx, v) || dalloc(s, &d.vnum, v) || dalloc(s, &d.vfrc, v) ||
        dalloc(s, &d.vtq, v) || dalloc(s, &d.vpe, v) || dalloc(s, &d.lpos, n) ||
        dalloc(s, &d.dth, n) || dalloc(s, &d.mpos, n) || dalloc(s, &d.rot0, n) ||
        dalloc(s, &d.rot1, n))
        return -1;
    Spare& sp = s->sp;
    if (dalloc(s, &sp.angle, n) || dalloc(s, &sp.omega, n) || dalloc(s, &sp.torque, n) ||
        dalloc(s, &sp.moi, n) || dalloc(s, &sp.nverts, n) || dalloc(s, &sp.vstart, n) ||
        dalloc(s, &sp.old_to_new, n) || dalloc(s, &sp.vpos, v) || dalloc(s, &sp.vmass, v) ||
        dalloc(s, &sp.vpart, v) || dalloc(s, &sp.svidx, v))
        return -1;
    std::vector<int> id(v);
    for (int64_t k = 0; k < v; ++k) id[k] = (int)k;
    if (scopy(s, d.svidx, id.data(), sizeof(int) * v, hipMemcpyHostToDevice)) return -1;
    return bumpy_ensure_vcap(s, 16);
}

int bumpy_ensure_vcap(dpm_system* s, int64_t cap) {
    if (cap <= s->vphys_cap) return 0;
    dfree(s, s->d.vnbr);
    s->d.vnbr = nullptr;
    if (dalloc(s, &s->d.vnbr, s->Nv * cap)) return -1;
    s->vphys_cap = cap;
    s->d.vcap = (int)cap;
    s->gen++;
    return 0;
}

// RigidBumpy::updateVertexVerletList (rigid_bumpy.cu:1198-1212), enqueued after the
// particle list build; overflow is detected by the chain's end kernel.
void bumpy_enqueue_vertex_build(dpm_system* s, int gated) {
    // grid-stride: a gated launch that skips stays cheap
    if (s->vring) {
        dim3 g(std::min<unsigned>(pgrid(s->N * VB_HALF).x, 2048u));
        hipLaunchKernelGGL(k_bumpy_vertex_rows, g, dim3(BLOCK), 0, s->stream, s->d,
                           nbr_params(s), s->rv, s->rmax, gated);
        return;
    }
    dim3 g(std::min<unsigned>(pgrid(s->Nv).x, 2048u));
    hipLaunchKernelGGL(k_bumpy_vertex_build, g, dim3(BLOCK), 0, s->stream, s->d, nbr_params(s),
                       s->rv, s->rmax, (int)s->vring, gated);
}

// Vertex half of the cell reorder: new vertex starts from the permuted counts, then
// the vertex arrays follow their particles.
void bumpy_enqueue_vertex_reorder(dpm_system* s, int gated) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s->stream, s->d.ctl, gated,
                       s->sp.nverts, s->sp.vstart, s->N);
    hipLaunchKernelGGL(k_bumpy_reorder_vertices, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d,
                       s->sp, gated);
    hipLaunchKernelGGL(k_bumpy_vertex_copyback, pgrid(std::max(s->Nv, s->N)), dim3(BLOCK), 0,
                       s->stream, s->d, s->sp, gated);
}

// mode -1: zero the vertex arrays; 1: accumulate (Particle::calculateForces semantics)
int bumpy_forces(dpm_system* s, int mode) {
    if (mode < 0) {
        hipLaunchKernelGGL(k_bumpy_zero_vertex, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d);
        return check_launch();
    }
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_bumpy_vertex_forces, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d, p,
                       s->rv, 1);
    hipLaunchKernelGGL(k_bumpy_particle_reduce, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, 1);
    return check_launch();
}

int bumpy_kick(dpm_system* s, double dt) {
    hipLaunchKernelGGL(k_bumpy_kick, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, dt,
                       (int)s->rotation);
    return check_launch();
}

int bumpy_drift(dpm_system* s, double dt) {
    hipLaunchKernelGGL(k_bumpy_drift_com, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, dt);
    hipLaunchKernelGGL(k_bumpy_drift_vertices, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d);
    return check_launch();
}

// RigidBumpy::updatePositionsAdam, rigid_bumpy.cu:1473-1486
int bumpy_adam_update(dpm_system* s, double alpha, double b1, double b2, double omb1,
                      double omb2, double eps) {
    hipLaunchKernelGGL(k_badam_update, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, alpha, b1,
                       b2, omb1, omb2, eps, (int)s->rotation);
    hipLaunchKernelGGL(k_bumpy_drift_vertices, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d);
    return check_launch();
}

int bumpy_mix(dpm_system* s, double alpha) {
    hipLaunchKernelGGL(k_bumpy_mix, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, alpha);
    return check_launch();
}

int bumpy_stress(dpm_system* s) {
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_bumpy_stress, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, p, s->rv,
                       s->box[0] * s->box[1]);
    return check_launch();
}

// Whether every vertex lies at r - r_v from its particle's centre to a relative 1e-9
// (vertex_ring_near's assumption): true after dpm_bumpy_setup, checked again on the host
// after an upload of positions, radii or vertex positions.
int bumpy_check_ring(dpm_system* s) {
    std::vector<double4> P(s->N);
    std::vector<int> st(s->N), nv(s->N);
    std::vector<double2> vp(s->Nv);
    if (scopy(s, P.data(), s->d.P0, sizeof(double4) * s->N, hipMemcpyDeviceToHost) ||
        scopy(s, st.data(), s->d.vstart, sizeof(int) * s->N, hipMemcpyDeviceToHost) ||
        scopy(s, nv.data(), s->d.nverts, sizeof(int) * s->N, hipMemcpyDeviceToHost) ||
        scopy(s, vp.data(), s->d.vpos, sizeof(double2) * s->Nv, hipMemcpyDeviceToHost))
        return -1;
    auto img = [](double d, double L) { return d - L * std::round(d / L); };
    bool ok = true;
    for (int64_t p = 0; p < s->N && ok; ++p) {
        const double inner = P[p].z - s->rv;
        for (int k = 0; k < nv[p] && ok; ++k) {
            const double2 u = vp[st[p] + k];
            const double dx = img(u.x - P[p].x, s->box[0]), dy = img(u.y - P[p].y, s->box[1]);
            ok = inner > 0 && std::fabs(std::sqrt(dx * dx + dy * dy) - inner) <= 1e-9 * inner;
        }
    }
    s->vring = ok;
    s->gen++;  // captured graphs bake the flag
    return 0;
}

int bumpy_contacts(dpm_system* s) {
    PairParams p = params_of(s);
    hipLaunchKernelGGL(k_bumpy_contacts, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, p, s->rv);
    return check_launch();
}

// NVE::step for rigid bodies, one step at a time (src/integrator/nve.cu:15-25)
int bumpy_nve_run(dpm_system* s, double dt, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        if (bumpy_kick(s, 0.5 * dt) || bumpy_drift(s, dt) || check_and_rebuild(s, nullptr) ||
            dpm_zero_forces(s) || bumpy_forces(s, 1) || bumpy_kick(s, 0.5 * dt))
            return -1;
    }
    DPM_HIP(hipStreamSynchronize(s->stream));
    return 0;
}


// ---- device-resident FIRE (the loop driver is minimize_fire in dpm_core.hip) ----
void bumpy_enqueue_fire_forces(dpm_system* s, int kick, int gate, int check_pos) {
    PairParams p = params_of(s);
    if (BF_TILE + s->nv_max - 1 <= BF_LDS) {
        unsigned nb = (unsigned)((s->Nv + BF_TILE - 1) / BF_TILE);
        bool small = BF_TILE + s->nv_max - 1 <= BF_LDS_SMALL, hot = p.fast && p.fastdiv;
        auto k = small ? (hot ? k_bfire_forces<BF_LDS_SMALL, true>
                              : k_bfire_forces<BF_LDS_SMALL, false>)
                       : (hot ? k_bfire_forces<BF_LDS, true> : k_bfire_forces<BF_LDS, false>);
        hipLaunchKernelGGL(k, dim3(nb), dim3(small ? bf_threads<BF_LDS_SMALL>() : bf_threads<BF_LDS>()),
                           0, s->stream, s->d, p, s->rv, gate, check_pos);
        hipLaunchKernelGGL(k_bfire_partials, dim3((unsigned)s->nblk), dim3(BLOCK), 0, s->stream,
                           s->d, gate, kick, (int)s->rotation);
        return;
    }
    if (check_pos >= 0)
        hipLaunchKernelGGL(k_bfire_check, dim3(1), dim3(1), 0, s->stream, s->d, check_pos);
    hipLaunchKernelGGL(k_bfire_vforce, pgrid(s->Nv), dim3(BLOCK), 0, s->stream, s->d, p, s->rv,
                       gate);
    hipLaunchKernelGGL(k_bfire_reduce, dim3((unsigned)s->nblk), dim3(BLOCK), 0, s->stream, s->d,
                       kick, gate, (int)s->rotation);
}

void bumpy_enqueue_fire_step(dpm_system* s, int pos) {
    const Dev& d = s->d;
    NbrParams nb = nbr_params(s);
    hipStream_t st = s->stream;
    {
        Timed t(s, DPM_K_DRIFT);
        hipLaunchKernelGGL(k_bfire_move, pgrid(s->N), dim3(BLOCK), 0, st, d, nb, pos,
                           (int)s->rotation);
        hipLaunchKernelGGL(k_bfire_vertices, pgrid(s->Nv), dim3(BLOCK), 0, st, d);
    }
    if (s->method == DPM_NBR_CELL || s->method == DPM_NBR_VERLET) {
        Timed t(s, DPM_K_REBUILD);
        hipLaunchKernelGGL(k_bfire_decide, dim3(1), dim3(1), 0, st, d, pos);
        enqueue_list_build(s, 1);
        bumpy_enqueue_vertex_build(s, 1);
    }
    Timed t(s, DPM_K_FORCE);
    // the overflow test of an in-step rebuild (k_bfire_check) runs in the force kernel
    bumpy_enqueue_fire_forces(s, 1, G_NORMAL,
                              s->method == DPM_NBR_CELL || s->method == DPM_NBR_VERLET ? pos : -1);
}


int bumpy_snapshot(dpm_system* s, bool save) {
    auto cp = [&](void* a, void* b, size_t n) {
        return save ? hipMemcpyAsync(a, b, n, hipMemcpyDeviceToDevice, s->stream)
                    : hipMemcpyAsync(b, a, n, hipMemcpyDeviceToDevice, s->stream);
    };
    int64_t n = s->N, v = s->Nv;
    DPM_HIP(cp(s->snap.angle, s->d.angle, sizeof(double) * n));
    DPM_HIP(cp(s->snap.omega, s->d.omega, sizeof(double) * n));
    DPM_HIP(cp(s->snap.torque, s->d.torque, sizeof(double) * n));
    DPM_HIP(cp(s->snap.moi, s->d.moi, sizeof(double) * n));
    DPM_HIP(cp(s->snap.nverts, s->d.nverts, sizeof(int) * n));
    DPM_HIP(cp(s->snap.vstart, s->d.vstart, sizeof(int) * n));
    DPM_HIP(cp(s->snap.vpos, s->d.vpos, sizeof(double2) * v));
    DPM_HIP(cp(s->snap.vmass, s->d.vmass, sizeof(double) * v));
    DPM_HIP(cp(s->snap.vpart, s->d.vpart, sizeof(int) * v));
    DPM_HIP(cp(s->snap.svidx, s->d.svidx, sizeof(int) * v));
    return 0;
}

static int fetch_svidx(dpm_system* s, std::vector<int>& sv) {
    sv.resize(s->Nv);
    return scopy(s, sv.data(), s->d.svidx, sizeof(int) * s->Nv, hipMemcpyDeviceToHost);
}

static double* pfield(dpm_system* s, int f) {
    switch (f) {
        case DPM_F_ANGLE: return s->d.angle;
        case DPM_F_OMEGA: return s->d.omega;
        case DPM_F_TORQUE: return s->d.torque;
        case DPM_F_MOI: return s->d.moi;
    }
    return nullptr;
}

int bumpy_upload(dpm_system* s, int f, const void* data, int64_t count,
                 const std::vector<int>& sidx) {
    if (double* p = pfield(s, f)) {
        if (count != s->N) { set_error("dpm_upload: count mismatch"); return -1; }
        const double* src = (const double*)data;
        std::vector<double> h(s->N);
        for (int64_t k = 0; k < s->N; ++k) h[k] = src[sidx[k]];
        if (scopy(s, p, h.data(), sizeof(double) * s->N, hipMemcpyHostToDevice)) return -1;
        return 0;
    }
    if (f == DPM_F_VERTEX_X || f == DPM_F_VERTEX_Y || f == DPM_F_VERTEX_MASS) {
        if (count != s->Nv) { set_error("dpm_upload: count mismatch"); return -1; }
        std::vector<int> sv;
        if (fetch_svidx(s, sv)) return -1;
        const double* src = (const double*)data;
        if (f == DPM_F_VERTEX_MASS) {
            std::vector<double> h(s->Nv);
            for (int64_t k = 0; k < s->Nv; ++k) h[k] = src[sv[k]];
            if (scopy(s, s->d.vmass, h.data(), sizeof(double) * s->Nv, hipMemcpyHostToDevice)) return -1;
            return 0;
        }
        std::vector<double2> h(s->Nv);
        if (scopy(s, h.data(), s->d.vpos, sizeof(double2) * s->Nv, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->Nv; ++k) (f == DPM_F_VERTEX_X ? h[k].x : h[k].y) = src[sv[k]];
        if (scopy(s, s->d.vpos, h.data(), sizeof(double2) * s->Nv, hipMemcpyHostToDevice)) return -1;
        return 0;
    }
    set_error("dpm_upload: field not uploadable");
    return -1;
}

int bumpy_download(dpm_system* s, int f, void* data, int64_t count,
                   const std::vector<int>& sidx) {
    if (double* p = pfield(s, f)) {
        if (count != s->N) { set_error("dpm_download: count mismatch"); return -1; }
        std::vector<double> h(s->N);
        if (scopy(s, h.data(), p, sizeof(double) * s->N, hipMemcpyDeviceToHost)) return -1;
        double* dst = (double*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    if (f == DPM_F_NUM_VERTICES || f == DPM_F_VERTEX_CONTACTS) {
        if (count != s->N) { set_error("dpm_download: count mismatch"); return -1; }
        int* src = f == DPM_F_NUM_VERTICES ? s->d.nverts : s->d.vcontacts;
        if (!src) { set_error("dpm_download: field not computed"); return -1; }
        std::vector<int> h(s->N);
        if (scopy(s, h.data(), src, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
        int64_t* dst = (int64_t*)data;
        for (int64_t k = 0; k < s->N; ++k) dst[sidx[k]] = h[k];
        return 0;
    }
    if (f == DPM_F_VERTEX_X || f == DPM_F_VERTEX_Y || f == DPM_F_VERTEX_MASS) {
        if (count != s->Nv) { set_error("dpm_download: count mismatch"); return -1; }
        std::vector<int> sv;
        if (fetch_svidx(s, sv)) return -1;
        double* dst = (double*)data;
        if (f == DPM_F_VERTEX_MASS) {
            std::vector<double> h(s->Nv);
            if (scopy(s, h.data(), s->d.vmass, sizeof(double) * s->Nv, hipMemcpyDeviceToHost)) return -1;
            for (int64_t k = 0; k < s->Nv; ++k) dst[sv[k]] = h[k];
            return 0;
        }
        std::vector<double2> h(s->Nv);
        if (scopy(s, h.data(), s->d.vpos, sizeof(double2) * s->Nv, hipMemcpyDeviceToHost)) return -1;
        for (int64_t k = 0; k < s->Nv; ++k) dst[sv[k]] = f == DPM_F_VERTEX_X ? h[k].x : h[k].y;
        return 0;
    }
    set_error("dpm_download: unknown field");
    return -1;
}

}  // namespace dpm

using namespace dpm;

extern "C" {

int dpm_bumpy_neighbor_config(dpm_system* s, double vc, double vpc) {
    if (s->kind != DPM_BUMPY) {
        set_error("dpm_bumpy_neighbor_config: not a bumpy system");
        return -1;
    }
    s->rc_v = vc;
    s->pcut_v = vpc;
    s->gen++;
    return 0;
}

int dpm_vertex_pairs(dpm_system* s, int64_t* oi, int64_t* oj, int64_t cap, int64_t* count) {
    if (s->kind != DPM_BUMPY || !s->nbr_ready) {
        set_error("dpm_vertex_pairs: no vertex list");
        return -1;
    }
    std::vector<int> sv, nn(s->Nv), nb((size_t)(s->Nv * s->vphys_cap));
    if (fetch_svidx(s, sv)) return -1;
    if (scopy(s, nn.data(), s->d.vnum, sizeof(int) * s->Nv, hipMemcpyDeviceToHost)) return -1;
    if (scopy(s, nb.data(), s->d.vnbr, sizeof(int) * s->Nv * s->vphys_cap,
              hipMemcpyDeviceToHost))
        return -1;
    int64_t c = 0;
    for (int64_t v = 0; v < s->Nv; ++v)
        for (int64_t k = 0; k < nn[v] && k < s->vphys_cap; ++k) {
            int u = nb[k * s->Nv + v];
            if (c < cap) {
                oi[c] = sv[v];
                oj[c] = sv[u];
            }
            c++;
        }
    *count = c;
    if (c > cap) {
        set_error("dpm_vertex_pairs: capacity too small");
        return -1;
    }
    return 0;
}

int dpm_bumpy_setup(dpm_system* s, const int64_t* nvp, double rv, int rotation) {
    if (s->kind != DPM_BUMPY) {
        set_error("dpm_bumpy_setup: not a bumpy system");
        return -1;
    }
    std::vector<int> sidx(s->N);
    if (scopy(s, sidx.data(), s->d.sidx, sizeof(int) * s->N, hipMemcpyDeviceToHost)) return -1;
    std::vector<int> nv(s->N), st(s->N);
    int64_t acc = 0;
    for (int64_t k = 0; k < s->N; ++k) {
        nv[k] = (int)nvp[sidx[k]];
        if (nv[k] <= 0) {
            set_error("dpm_bumpy_setup: every particle needs at least one vertex");
            return -1;
        }
        st[k] = (int)acc;
        acc += nv[k];
    }
    if (acc != s->Nv) {
        set_error("dpm_bumpy_setup: vertex counts do not sum to n_vertices");
        return -1;
    }
    s->nv_max = *std::max_element(nv.begin(), nv.end());
    if (scopy(s, s->d.nverts, nv.data(), sizeof(int) * s->N, hipMemcpyHostToDevice)) return -1;
    if (scopy(s, s->d.vstart, st.data(), sizeof(int) * s->N, hipMemcpyHostToDevice)) return -1;
    std::vector<int> id(s->Nv);
    for (int64_t k = 0; k < s->Nv; ++k) id[k] = (int)k;
    if (scopy(s, s->d.svidx, id.data(), sizeof(int) * s->Nv, hipMemcpyHostToDevice)) return -1;
    s->rv = rv;
    s->rotation = rotation != 0;
    s->gen++;
    hipLaunchKernelGGL(k_bumpy_init_vertices, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, rv);
    if (check_launch()) return -1;
    DPM_HIP(hipStreamSynchronize(s->stream));
    s->vring = rv > 0;  // rings of radius r - r_v (radii above r_v: checked below)
    if (refresh_rmax(s)) return -1;
    return bumpy_check_ring(s);
}

}  // extern "C"
