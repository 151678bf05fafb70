// dpm_pairs.hip -- per-pair analytics of the neighbour list: pair energies, forces,
// separations, static ids, overlaps, separation angles and the pair / self Hessian
// blocks, plus (rigid bumpy) orientations, vertex-contact counts, vertex overlaps and
// friction coefficients.
//
// Layout (the reference's): field f of slot k of row i sits at out[f*S + i*row + k],
// S = N*row, row = the reference's neighbour capacity (max_neighbors_allocated), i in
// the library's internal (cell) order.  One thread per slot: consecutive threads write
// consecutive slots of a row, so every field store is coalesced.  This is analysis
// output read once per logged state, not a per-step kernel.
//
// Arithmetic follows the reference kernels operation by operation (the same products,
// quotients and summation order), so every field except the separation angle (atan2,
// the device library's) is bit-identical to the CPU restatement; sin and cos of the
// vertex angles come from dpm_sincos as everywhere else on the rigid path.
#include <hip/hip_runtime.h>

#include "dpm_kernels.hpp"
#include "dpm_trig.hpp"

namespace dpm {
namespace {

constexpr int NF_DISK = 18;
constexpr double PI_D = 3.14159265358979323846;

__device__ __forceinline__ double& fld(double* out, int64_t S, int f, int64_t t) {
    return out[(int64_t)f * S + t];
}
__device__ __forceinline__ void put_id(double* out, int64_t S, int f, int64_t t, long long v) {
    reinterpret_cast<long long*>(out)[(int64_t)f * S + t] = v;
}

// fill values of Disk/RigidBumpy::calculateForceDistancePairs (disk.cu:229-245,
// rigid_bumpy.cu:1384-1415)
__device__ void fill_slot(double* out, int64_t S, int64_t t, int bumpy) {
    fld(out, S, DPM_P_POTENTIAL, t) = -1.0;
    fld(out, S, DPM_P_FORCE_X, t) = 0.0;
    fld(out, S, DPM_P_FORCE_Y, t) = 0.0;
    fld(out, S, DPM_P_DIST_X, t) = -1.0;
    fld(out, S, DPM_P_DIST_Y, t) = -1.0;
    put_id(out, S, DPM_P_ID_I, t, -1);
    put_id(out, S, DPM_P_ID_J, t, -1);
    fld(out, S, DPM_P_OVERLAP, t) = -1.0;
    fld(out, S, DPM_P_RADSUM, t) = -1.0;
    fld(out, S, DPM_P_ANGLE, t) = -1.0;
    for (int f = DPM_P_H_XX; f <= DPM_P_HII_YY; ++f) fld(out, S, f, t) = 0.0;
    if (!bumpy) return;
    fld(out, S, DPM_P_ANGLE_I, t) = -1.0;
    fld(out, S, DPM_P_ANGLE_J, t) = -1.0;
    put_id(out, S, DPM_P_VERTEX_CONTACTS, t, -1);
    fld(out, S, DPM_P_FRICTION, t) = -1.0;
    fld(out, S, DPM_P_VERTEX_OVERLAP, t) = 0.0;
    for (int f = DPM_P_H_XT; f <= DPM_P_HII_TT; ++f) fld(out, S, f, t) = 0.0;
}

// slot t -> (row i, entry k); rows are a power of two long except for all-pairs lists
__device__ __forceinline__ int64_t row_of(int64_t t, int64_t row, int shift) {
    return shift >= 0 ? (t >> shift) : t / row;
}

// kernelCalcDiskForceDistancePairs, disk/kernels.cu:73-175.  Every field of the slot is
// formed in registers (the reference's fill value where the slot is unused) and stored
// once.
__global__ __launch_bounds__(BLOCK) void k_pairs_disk(Dev d, PairParams p, double* out,
                                                      int64_t row, int shift, int64_t S) {
    int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= S) return;
    int64_t i = row_of(t, row, shift);
    int k = (int)(t - i * row);
    int nn = min(d.nnum[i], d.cap);
    int j = k < nn ? d.nbr[(int64_t)k * d.N + i] : -1;
    double e = -1.0, fx = 0.0, fy = 0.0, dx = -1.0, dy = -1.0, ovl = -1.0, rsum = -1.0,
           ang = -1.0;
    long long id_i = -1, id_j = -1;
    double hij[4] = {0.0, 0.0, 0.0, 0.0}, hii[4] = {0.0, 0.0, 0.0, 0.0};
    if (j >= 0 && j != i) {
        double4 qi = d.P0[i], qj = d.P0[j];
        dx = pbc_div(qi.x, qj.x, p.Lx);
        dy = pbc_div(qi.y, qj.y, p.Ly);
        e = pair_core(dx, dy, qi.z, qj.z, p, fx, fy);
        id_i = d.sidx[i];
        id_j = d.sidx[j];
        double dist = sqrt(dx * dx + dy * dy);
        ovl = dist - (qi.z + qj.z);
        rsum = qi.z + qj.z;
        ang = atan2(dy, dx);
        if (dist < qi.z + qj.z) {
            double radsum = qi.z + qj.z;
            double t_ij = -p.e_c / radsum * (1 - dist / radsum);
            double c_ij = p.e_c / (radsum * radsum);
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    double d_ia_dx = 1 * (a == 0);
                    double d_ia_dy = 1 * (a == 1);
                    double d_jb_dx = -1 * (b == 0);
                    double d_jb_dy = -1 * (b == 1);
                    double d_ia_d_jb_dx = 0;
                    double d_ia_d_jb_dy = 0;
                    double q_ia = dx * d_ia_dx + dy * d_ia_dy;
                    double q_jb = dx * d_jb_dx + dy * d_jb_dy;
                    double d_ia_q_jb = d_ia_dx * d_jb_dx + dx * d_ia_d_jb_dx + d_ia_dy * d_jb_dy +
                                       dy * d_ia_d_jb_dy;
                    hij[a * 2 + b] += c_ij * (q_ia * q_jb) / (dist * dist) +
                                      t_ij * (d_ia_q_jb / dist - (q_ia * q_jb) / (dist * dist * dist));
                    double d_ib_dx = 1 * (b == 0);
                    double d_ib_dy = 1 * (b == 1);
                    double d_ia_d_ib_dx = 0;
                    double d_ia_d_ib_dy = 0;
                    double q_ib = dx * d_ib_dx + dy * d_ib_dy;
                    double d_ia_q_ib = d_ia_dx * d_ib_dx + dx * d_ia_d_ib_dx + d_ia_dy * d_ib_dy +
                                       dy * d_ia_d_ib_dy;
                    hii[a * 2 + b] += c_ij * (q_ia * q_ib) / (dist * dist) +
                                      t_ij * (d_ia_q_ib / dist - (q_ia * q_ib) / (dist * dist * dist));
                }
            for (int c = 0; c < 4; ++c) {
                hij[c] = 0.0 + hij[c];
                hii[c] = 0.0 + hii[c];
            }
        }
    }
    fld(out, S, DPM_P_POTENTIAL, t) = e;
    fld(out, S, DPM_P_FORCE_X, t) = fx;
    fld(out, S, DPM_P_FORCE_Y, t) = fy;
    fld(out, S, DPM_P_DIST_X, t) = dx;
    fld(out, S, DPM_P_DIST_Y, t) = dy;
    put_id(out, S, DPM_P_ID_I, t, id_i);
    put_id(out, S, DPM_P_ID_J, t, id_j);
    fld(out, S, DPM_P_OVERLAP, t) = ovl;
    fld(out, S, DPM_P_RADSUM, t) = rsum;
    fld(out, S, DPM_P_ANGLE, t) = ang;
    for (int c = 0; c < 4; ++c) {
        fld(out, S, DPM_P_H_XX + c, t) = hij[c];
        fld(out, S, DPM_P_HII_XX + c, t) = hii[c];
    }
}

// kernelCalcRigidBumpyForceDistancePairs, rigid_bumpy/kernels.cu:482-702
__global__ __launch_bounds__(BLOCK) void k_pairs_bumpy(Dev d, PairParams p, double rv,
                                                       double* out, int64_t row, int shift,
                                                       int64_t S) {
    int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= S) return;
    int64_t i = row_of(t, row, shift);
    int k = (int)(t - i * row);
    int nn = min(d.nnum[i], d.cap);
    int j = k < nn ? d.nbr[(int64_t)k * d.N + i] : -1;
    fill_slot(out, S, t, 1);
    if (j < 0 || j == i) return;
    double4 qi = d.P0[i], qj = d.P0[j];
    double theta_i = d.angle[i], theta_j = d.angle[j];
    double force_x = 0.0, force_y = 0.0, interaction_energy = 0.0, vertex_overlap = 0.0;
    long long vertex_count_i = 0;
    double hij[9], hii[9];
    for (int c = 0; c < 9; ++c) hij[c] = hii[c] = 0.0;
    const int nvi = d.nverts[i], nvj = d.nverts[j];
    const int sti = d.vstart[i], stj = d.vstart[j];
    for (int u = 0; u < nvi; ++u) {
        int64_t vid = sti + u;
        double2 xv = d.vpos[vid];
        double phi_u = (double)((long long)u * 2) * PI_D / (double)nvi;
        double rdx = xv.x - qi.x, rdy = xv.y - qi.y;
        double R_i = sqrt(rdx * rdx + rdy * rdy);
        double si, ci;
        dpm_sincos(theta_i + phi_u, &si, &ci);
        bool is_contact = false;
        int vn = min(d.vnum[vid], d.vcap);
        for (int n = 0; n < vn; ++n) {
            int ov = d.vnbr[(int64_t)n * d.Nv + vid];
            if (ov < 0 || ov == vid || d.vpart[ov] != j) continue;
            double2 xo = d.vpos[ov];
            long long v = (long long)ov - stj;
            double phi_v = (double)(v * 2) * PI_D / (double)nvj;
            double odx = xo.x - qj.x, ody = xo.y - qj.y;
            double R_j = sqrt(odx * odx + ody * ody);
            double fx, fy;
            double e = pair_law(xv.x, xv.y, rv, xo.x, xo.y, rv, p, fx, fy);
            interaction_energy += e;
            force_x += fx;
            force_y += fy;
            if (e > 0.0) {
                is_contact = true;
                double dx = pbc_div(xv.x, xo.x, p.Lx);
                double dy = pbc_div(xv.y, xo.y, p.Ly);
                vertex_overlap += (rv * 2) - sqrt(dx * dx + dy * dy);
                double dist = sqrt(dx * dx + dy * dy);
                double radsum = rv * 2.0;
                double t_ij = -p.e_c / radsum * (1 - dist / radsum);
                double c_ij = p.e_c / (radsum * radsum);
                double sj, cj;
                dpm_sincos(theta_j + phi_v, &sj, &cj);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        double d_ia_dx = (a == 0) - R_i * si * (a == 2);
                        double d_ia_dy = (a == 1) + R_i * ci * (a == 2);
                        double d_jb_dx = -1.0 * ((b == 0) - R_j * sj * (b == 2));
                        double d_jb_dy = -1.0 * ((b == 1) + R_j * cj * (b == 2));
                        double q_ia = dx * d_ia_dx + dy * d_ia_dy;
                        double q_jb = dx * d_jb_dx + dy * d_jb_dy;
                        double d_ia_q_jb = (d_ia_dx * d_jb_dx + d_ia_dy * d_jb_dy) / dist;
                        hij[a * 3 + b] += c_ij * q_ia * q_jb / (dist * dist) +
                                          t_ij * (d_ia_q_jb - q_ia * q_jb / (dist * dist * dist));
                        double d_ib_dx = (b == 0) - R_i * si * (b == 2);
                        double d_ib_dy = (b == 1) + R_i * ci * (b == 2);
                        double d_ia_ib_dx = -R_i * ci * (a == 2) * (b == 2);
                        double d_ia_ib_dy = -R_i * si * (a == 2) * (b == 2);
                        double q_ib = dx * d_ib_dx + dy * d_ib_dy;
                        double d_ia_q_ib = (d_ia_dx * d_ib_dx + dx * d_ia_ib_dx + d_ia_dy * d_ib_dy +
                                            dy * d_ia_ib_dy) /
                                           dist;
                        hii[a * 3 + b] += c_ij * q_ia * q_ib / (dist * dist) +
                                          t_ij * (d_ia_q_ib - q_ia * q_ib / (dist * dist * dist));
                    }
            }
        }
        if (is_contact) vertex_count_i++;
    }
    fld(out, S, DPM_P_POTENTIAL, t) = interaction_energy;
    fld(out, S, DPM_P_FORCE_X, t) = force_x;
    fld(out, S, DPM_P_FORCE_Y, t) = force_y;
    double x_dist = pbc_div(qi.x, qj.x, p.Lx);
    double y_dist = pbc_div(qi.y, qj.y, p.Ly);
    fld(out, S, DPM_P_DIST_X, t) = x_dist;
    fld(out, S, DPM_P_DIST_Y, t) = y_dist;
    put_id(out, S, DPM_P_ID_I, t, d.sidx[i]);
    put_id(out, S, DPM_P_ID_J, t, d.sidx[j]);
    fld(out, S, DPM_P_VERTEX_OVERLAP, t) = vertex_overlap;
    double dist = sqrt(x_dist * x_dist + y_dist * y_dist);
    fld(out, S, DPM_P_OVERLAP, t) = dist - (qi.z + qj.z);
    fld(out, S, DPM_P_RADSUM, t) = qi.z + qj.z;
    fld(out, S, DPM_P_ANGLE, t) = atan2(y_dist, x_dist);
    fld(out, S, DPM_P_ANGLE_I, t) = theta_i;
    fld(out, S, DPM_P_ANGLE_J, t) = theta_j;
    put_id(out, S, DPM_P_VERTEX_CONTACTS, t, vertex_count_i);
    double mu = 0.0;
    if (force_x != 0.0 || force_y != 0.0) {
        double nfx = force_x * x_dist / dist;
        double nfy = force_y * y_dist / dist;
        double nf = sqrt(nfx * nfx + nfy * nfy);
        double tfx = force_x - nfx;
        double tfy = force_y - nfy;
        double tf = sqrt(tfx * tfx + tfy * tfy);
        mu = tf / nf;
    }
    fld(out, S, DPM_P_FRICTION, t) = mu;
    // hess[a*3+b], a, b in (x, y, theta)
    const int hp[9] = {DPM_P_H_XX, DPM_P_H_XY, DPM_P_H_XT, DPM_P_H_YX, DPM_P_H_YY,
                       DPM_P_H_YT, DPM_P_H_TX, DPM_P_H_TY, DPM_P_H_TT};
    const int hs[9] = {DPM_P_HII_XX, DPM_P_HII_XY, DPM_P_HII_XT, DPM_P_HII_YX, DPM_P_HII_YY,
                       DPM_P_HII_YT, DPM_P_HII_TX, DPM_P_HII_TY, DPM_P_HII_TT};
    for (int c = 0; c < 9; ++c) {
        fld(out, S, hp[c], t) = 0.0 + hij[c];
        fld(out, S, hs[c], t) = 0.0 + hii[c];
    }
}

// angleBetweenVectors and calcOverlapLenseArea, include/particles/base/kernels.cuh:367-381
__device__ __forceinline__ double turn_angle(double nx, double ny, double cx, double cy, double px,
                                             double py) {
    double ms = (nx - cx) * (cy - py) - (ny - cy) * (cx - px);
    double mc = (nx - cx) * (cx - px) + (ny - cy) * (cy - py);
    return atan2(ms, mc);
}
__device__ __forceinline__ double segment_area(double R, double dd) {
    return R * R * acos(dd / R) - dd * sqrt(R * R - dd * dd);
}
__device__ __forceinline__ double lens_area(double r_ij, double ri, double rj) {
    double d1 = (r_ij * r_ij - ri * ri + rj * rj) / (2 * r_ij);
    double d2 = r_ij - d1;
    return segment_area(ri, d1) + segment_area(rj, d2);
}

// kernelCalculateBumpyParticleAreaFull, rigid_bumpy/kernels.cu:1226-1283: shoelace area of
// the vertex ring plus, per vertex, the part of its disk outside the polygon (the
// exterior-angle share of the disk less half of each lens it shares with the next vertex)
__global__ __launch_bounds__(BLOCK) void k_bumpy_area(Dev d, double rv, double* out) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= d.N) return;
    double poly = 0.0, exposed = 0.0;
    double vertex_area = PI_D * rv * rv;
    int64_t first = d.vstart[i];
    int64_t nv = d.nverts[i];
    double2 cur = d.vpos[first], prev = d.vpos[first + nv - 1];
    for (int64_t k = 0; k < nv; ++k) {
        double own = 0.0;
        double2 nxt = d.vpos[first + ((k + 1) % nv)];
        poly += cur.x * nxt.y - nxt.x * cur.y;
        double ang = turn_angle(nxt.x, nxt.y, cur.x, cur.y, prev.x, prev.y);
        double dx = nxt.x - cur.x, dy = nxt.y - cur.y;
        double r_ij = sqrt(dx * dx + dy * dy);
        if (r_ij < (2 * rv - 1e-10)) own -= lens_area(r_ij, rv, rv) / 2.0;
        exposed += vertex_area * (PI_D - ang) / (2.0 * PI_D) - own;
        prev = cur;
        cur = nxt;
    }
    out[i] = fabs(poly) * 0.5 + exposed;
}

}  // namespace
}  // namespace dpm

using namespace dpm;

extern "C" {

int dpm_pair_analytics(dpm_system* s, int64_t* row_length) {
    if (!s->nbr_ready || !s->d.nbr) {
        set_error("dpm_pair_analytics: neighbour list not initialised (call dpm_neighbor_init)");
        return -1;
    }
    int64_t row = 0;
    if (dpm_neighbor_stats(s, &row, nullptr, nullptr, nullptr)) return -1;
    if (row < 1) row = 1;
    const int nf = s->kind == DPM_BUMPY ? DPM_P_NFIELDS : NF_DISK;
    const int64_t S = s->N * row;
    if ((int64_t)nf * S > s->pair_alloc) {
        if (s->pairs) dfree(s, s->pairs);
        s->pairs = nullptr;
        s->pair_alloc = 0;
        if (dalloc(s, &s->pairs, (int64_t)nf * S)) return -1;
        s->pair_alloc = (int64_t)nf * S;
    }
    PairParams p = params_of(s);
    p.fast = 0;  // the reference's minimum image for every slot
    dim3 grid((unsigned)std::max<int64_t>(1, (S + BLOCK - 1) / BLOCK));
    int shift = -1;
    if ((row & (row - 1)) == 0)
        for (shift = 0; ((int64_t)1 << shift) < row; ++shift) {
        }
    if (s->kind == DPM_BUMPY)
        hipLaunchKernelGGL(k_pairs_bumpy, grid, dim3(BLOCK), 0, s->stream, s->d, p, s->rv, s->pairs,
                           row, shift, S);
    else
        hipLaunchKernelGGL(k_pairs_disk, grid, dim3(BLOCK), 0, s->stream, s->d, p, s->pairs, row,
                           shift, S);
    if (check_launch()) return -1;
    DPM_HIP(hipStreamSynchronize(s->stream));
    s->pair_row = row;
    s->pair_fields = nf;
    if (row_length) *row_length = row;
    return 0;
}

int dpm_bumpy_area(dpm_system* s, double* out, int64_t count) {
    if (s->kind != DPM_BUMPY || !s->d.vstart || s->nv_max == 0) {
        set_error("dpm_bumpy_area: not a rigid bumpy system with vertices set up");
        return -1;
    }
    if (count != s->N) {
        set_error("dpm_bumpy_area: count must be n_particles");
        return -1;
    }
    double* a = nullptr;
    if (dalloc(s, &a, s->N)) return -1;
    hipLaunchKernelGGL(k_bumpy_area, pgrid(s->N), dim3(BLOCK), 0, s->stream, s->d, s->rv, a);
    int rc = check_launch();
    std::vector<double> h((size_t)s->N);
    std::vector<int> sidx((size_t)s->N);
    if (!rc) rc = scopy(s, h.data(), a, sizeof(double) * (size_t)s->N, hipMemcpyDeviceToHost);
    if (!rc) rc = scopy(s, sidx.data(), s->d.sidx, sizeof(int) * (size_t)s->N, hipMemcpyDeviceToHost);
    dfree(s, a);
    if (rc) return -1;
    for (int64_t k = 0; k < s->N; ++k) out[sidx[(size_t)k]] = h[(size_t)k];
    return 0;
}

int dpm_pair_download(dpm_system* s, int field, void* out, int64_t count) {
    if (s->pair_row == 0) {
        set_error("dpm_pair_download: call dpm_pair_analytics first");
        return -1;
    }
    if (field < 0 || field >= s->pair_fields) {
        set_error(field >= 0 && field < DPM_P_NFIELDS
                      ? "dpm_pair_download: field exists for rigid bumpy systems only"
                      : "dpm_pair_download: unknown field");
        return -1;
    }
    const int64_t S = s->N * s->pair_row;
    if (count != S) {
        set_error("dpm_pair_download: count must be n_particles * row_length");
        return -1;
    }
    return scopy(s, out, s->pairs + (int64_t)field * S, sizeof(double) * (size_t)S,
                 hipMemcpyDeviceToHost);
}

}  // extern "C"
