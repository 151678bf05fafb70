// dpm_kernels.hpp -- launch helpers shared between the library's .hip sources.
#pragma once
#include <mutex>
#include <vector>

#include "dpm_internal.hpp"

namespace dpm {

inline PairParams pair_params(const dpm_system* s, double e_c, double n_c) {
    PairParams p{};
    p.Lx = s->box[0];
    p.Ly = s->box[1];
    p.invLx = 1.0 / s->box[0];
    p.invLy = 1.0 / s->box[1];
    p.halfLx = 0.5 * s->box[0];
    p.halfLy = 0.5 * s->box[1];
    p.e_c = e_c;
    p.n_c = n_c;
    p.ncm1 = n_c - 1.0;
    p.two_nc = 2.0 * n_c;
    // x/(2 n_c) == x*(1/(2 n_c)) exactly when 2 n_c is a power of two
    int ex;
    double m = frexp(p.two_nc, &ex);
    p.inv2nc = (m == 0.5) ? 1.0 / p.two_nc : 0.0;
    p.fast = 0;
    return p;
}

inline dim3 pgrid(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, (n + BLOCK - 1) / BLOCK));
}

// which per-particle quantity a chunk-sum kernel reduces
enum SumWhat {
    S_PE = 0, S_KE, S_FXVX, S_FYVY, S_FX, S_FY, S_SXX, S_SYY, S_SXY, S_SYX, S_VX, S_VY, S_R2,
    S_TQW, S_COUNT
};
enum MaxWhat { M_ND = 0, M_CD, M_F2, M_T2 };

// How a loop kernel decides whether to run: G_NORMAL inside a chunk (skip after a
// halt, halt on a matching rebuild request), G_RESUME in the rebuild chain (run only
// to finish a halted step), G_ALWAYS unconditionally.
enum Gate { G_NORMAL = 0, G_RESUME = 1, G_ALWAYS = 2 };

// ---- control-block helpers shared by the loop kernels ----
constexpr long long STAMP_STRIDE = 8192;  // > any chunk length
__device__ __forceinline__ long long stamp(long long epoch, int pos) {
    return epoch * STAMP_STRIDE + pos + 1;
}

// The end of a launch unit: copies the control block into slot (pub_seq mod 4) of the
// pinned host mirror and advances pub_seq; the host reads the slot once the event behind
// the unit has completed.  Called by every thread of one block, after the block's last
// write to the control block.
__device__ __forceinline__ void publish_ctl(const Dev& d) {
    Ctl* c = d.ctl;
    const int k = c->pub_seq;
    constexpr int NW = (int)(sizeof(Ctl) / sizeof(int));
    int* dst = (int*)(d.hpub + (k & 3));
    const int* src = (const int*)c;
    for (int w = threadIdx.x; w < NW; w += blockDim.x) dst[w] = src[w];
    __threadfence_system();
    __syncthreads();  // every lane has read the sequence number and copied its words
    if (threadIdx.x == 0) c->pub_seq = k + 1;
}

__device__ __forceinline__ bool resumable(const Dev& d) {
    const Ctl* c = d.ctl;
    return c->max_nnum <= d.cap && (d.vnbr == nullptr || c->max_vnum <= d.vcap);
}


__device__ __forceinline__ int64_t particle_of_thread(int& tile) {
    tile = xcd_tile(blockIdx.x, gridDim.x);
    return (int64_t)tile * BLOCK + threadIdx.x;
}

// Per-block maxima of the two displacement arrays (the reference keeps the full
// arrays, particle.cu:1094-1100; only their maxima are ever read) and the rebuild
// request of Particle::checkFor{Cell,Verlet}ListUpdate (particle.cu:1141-1161):
// 3*max > threshold.  wr < 0: no request.
__device__ __forceinline__ void disp_epilogue(const Dev& d, const NbrParams& nb, int tile,
                                              double nd, double cd, int wr, long long st) {
    double bn = nd, bc = cd;
    block_max2(bn, bc);
    if (threadIdx.x == 0) {
        d.pmax_nd[tile] = bn;
        d.pmax_cd[tile] = bc;
        if (wr >= 0 && (nb.method == DPM_NBR_CELL || nb.method == DPM_NBR_VERLET)) {
            if (nb.method == DPM_NBR_CELL && 3.0 * bc > nb.cthr) d.ctl->cflag[wr] = st;
            if (3.0 * bn > nb.thr) d.ctl->flag[wr] = st;
        }
    }
}

// The rebuild decision of disp_epilogue without the block maxima: 3 max d^2 > thr holds
// exactly when 3 d^2 > thr holds for some particle (3x is monotone under rounding), so a
// wave ballot per test replaces the block reduction.  For the steps inside a chunk, whose
// maxima nothing reads; the partials are written at the end of the chunk.
__device__ __forceinline__ void disp_flags(const Dev& d, const NbrParams& nb, double nd,
                                           double cd, int wr, long long st) {
    if (nb.method != DPM_NBR_CELL && nb.method != DPM_NBR_VERLET) return;
    const bool cf = nb.method == DPM_NBR_CELL && 3.0 * cd > nb.cthr;
    const bool lf = 3.0 * nd > nb.thr;
    const uint64_t bc = __builtin_amdgcn_ballot_w64(cf);
    const uint64_t bl = __builtin_amdgcn_ballot_w64(lf);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        if (bc) d.ctl->cflag[wr] = st;
        if (bl) d.ctl->flag[wr] = st;
    }
}

__device__ __forceinline__ void disp_of(const Dev& d, int64_t i, double x, double y, double& nd,
                                        double& cd) {
    double2 ln = d.lref[i], lc = d.cref[i];
    double dxn = x - ln.x, dyn = y - ln.y;
    nd = dxn * dxn + dyn * dyn;
    double dxc = x - lc.x, dyc = y - lc.y;
    cd = dxc * dxc + dyc * dyc;
}

// The rebuild a step at (epoch, pos) must perform before its forces:
// 0 none, 1 list, 2 cells + list (the cell test is the reference's first,
// particle.cu:1141-1161), 3 the list of the last in-kernel rebuild again, from its
// stored build positions, with a larger capacity.  ovf holds the stamp of the step
// whose in-kernel rebuild overflowed; that step itself ignores it, so every block of
// a launch reaches the same decision.
__device__ __forceinline__ bool ovf_pending(const Ctl* c, long long st) {
    return c->ovf != 0 && c->ovf != st;
}

__device__ __forceinline__ int pending_rebuild(const Ctl* c, long long epoch, int pos) {
    long long st = stamp(epoch, pos);
    if (c->cflag[pos & 1] == st) return 2;
    bool list = c->flag[pos & 1] == st;
    if (ovf_pending(c, st)) return list ? 1 : 3;
    return list ? 1 : 0;
}

// kernelMixVelocitiesAndForces (src/particles/base/kernels.cu:91-112): per-particle FIRE
// mixing with the 1e-16 force-norm guard
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

// ---- host helpers (dpm_core.hip) ----
std::mutex& capture_mutex();
int dalloc_bytes(dpm_system* s, void** p, size_t bytes);
template <class T>
int dalloc(dpm_system* s, T** p, int64_t n) {
    void* q = nullptr;
    if (dalloc_bytes(s, &q, sizeof(T) * (size_t)std::max<int64_t>(n, 1))) return -1;
    *p = (T*)q;
    return 0;
}
void dfree(dpm_system* s, void* p);
int64_t pow2_at_least(int64_t m);
int check_launch();
int scopy(dpm_system* s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
PairParams params_of(const dpm_system* s);
NbrParams nbr_params(const dpm_system* s);
int launch_chunk_sum(dpm_system* s, int what, double* out_dev);
int sum_to_host(dpm_system* s, int what, double* out);
int max_to_host(dpm_system* s, int what, double* out);
int ensure_ell(dpm_system* s, int64_t cap);
int device_rebuild(dpm_system* s, int mode, int* kind_out);
int check_and_rebuild(dpm_system* s, int* kind_out);
int refresh_rmax(dpm_system* s);
void enqueue_list_build(dpm_system* s, int gated);
void reset_ctl(dpm_system* s, int config);
int ensure_cells(dpm_system* s, int64_t nc);
int update_verlet_grid(dpm_system* s);
int bumpy_check_ring(dpm_system* s);
void collect_timing(dpm_system* s);

// Brackets one launch with HIP events on the system stream when timing is on.
struct Timed {
    dpm_system* s;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(dpm_system* s_, int cls_);
    ~Timed();
};

// ---- rigid bumpy (dpm_bumpy.hip) ----
int bumpy_alloc(dpm_system* s);
int bumpy_forces(dpm_system* s, int mode);
void bumpy_enqueue_vertex_build(dpm_system* s, int gated);
void bumpy_enqueue_vertex_reorder(dpm_system* s, int gated);
int bumpy_ensure_vcap(dpm_system* s, int64_t cap);
int bumpy_kick(dpm_system* s, double dt);
int bumpy_drift(dpm_system* s, double dt);
int bumpy_adam_update(dpm_system* s, double alpha, double b1, double b2, double omb1,
                      double omb2, double eps);
int bumpy_mix(dpm_system* s, double alpha);
int bumpy_stress(dpm_system* s);
int bumpy_contacts(dpm_system* s);
int bumpy_nve_run(dpm_system* s, double dt, int64_t n);
void bumpy_enqueue_fire_forces(dpm_system* s, int kick, int gate, int check_pos = -1);
void bumpy_enqueue_fire_step(dpm_system* s, int pos);
int bumpy_snapshot(dpm_system* s, bool save);
int bumpy_download(dpm_system* s, int field, void* data, int64_t count,
                   const std::vector<int>& sidx);
int bumpy_upload(dpm_system* s, int field, const void* data, int64_t count,
                 const std::vector<int>& sidx);

}  // namespace dpm
