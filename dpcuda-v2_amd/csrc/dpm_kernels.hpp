// dpm_kernels.hpp -- launch helpers shared between the library's .hip sources.
#pragma once
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

// ---- host helpers (dpm_core.hip) ----
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
void reset_ctl(dpm_system* s, int config);
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
int bumpy_mix(dpm_system* s, double alpha);
int bumpy_stress(dpm_system* s);
int bumpy_contacts(dpm_system* s);
int bumpy_nve_run(dpm_system* s, double dt, int64_t n);
int bumpy_minimize_fire(dpm_system* s, const dpm_fire_params* p, dpm_fire_result* out);
int bumpy_snapshot(dpm_system* s, bool save);
int bumpy_download(dpm_system* s, int field, void* data, int64_t count,
                   const std::vector<int>& sidx);
int bumpy_upload(dpm_system* s, int field, const void* data, int64_t count,
                 const std::vector<int>& sidx);

}  // namespace dpm
