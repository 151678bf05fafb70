// dpm_kernels.hpp -- launch helpers shared between the library's .hip sources.
#pragma once
#include "dpm_internal.hpp"

namespace dpm {

inline PairParams pair_params(const dpm_system* s, double e_c, double n_c) {
    PairParams p;
    p.Lx = s->box[0];
    p.Ly = s->box[1];
    p.invLx = 1.0 / s->box[0];
    p.invLy = 1.0 / s->box[1];
    p.e_c = e_c;
    p.n_c = n_c;
    p.ncm1 = n_c - 1.0;
    p.two_nc = 2.0 * n_c;
    // x/(2 n_c) == x*(1/(2 n_c)) exactly when 2 n_c is a power of two
    int ex;
    double m = frexp(p.two_nc, &ex);
    p.inv2nc = (m == 0.5) ? 1.0 / p.two_nc : 0.0;
    return p;
}

inline dim3 pgrid(int64_t n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// which per-particle quantity a chunk-sum kernel reduces
enum SumWhat {
    S_PE = 0, S_KE, S_FXVX, S_FYVY, S_FX, S_FY, S_SXX, S_SYY, S_SXY, S_SYX, S_VX, S_VY, S_R2,
    S_TQW, S_COUNT
};
enum MaxWhat { M_ND = 0, M_CD, M_F2, M_T2 };

// host-side launchers (dpm_core.hip)
int launch_chunk_sum(dpm_system* s, int what, double* out_dev);
int sum_to_host(dpm_system* s, int what, double* out);
int max_to_host(dpm_system* s, int what, double* out);

}  // namespace dpm
