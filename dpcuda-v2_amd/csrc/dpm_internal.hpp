// dpm_internal.hpp -- host-side system layout shared by the library's sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dpm.h"
#include "dpm_device.hpp"

namespace dpm {

void set_error(const std::string& m);

#define DPM_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::dpm::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return -1;                                                         \
        }                                                                      \
    } while (0)

// Device control block of one system.  Written by kernels with plain stores
// (never atomics in the hot loop); read by the host through a pinned copy.
struct Ctl {
    int flag[2];     // rebuild request raised by the drift kernel of step parity p
    int halt;        // set by the force kernel when a rebuild must happen first
    int halt_pos;    // chunk position of the halted step
    int done;        // the device-resident loop has finished (FIRE convergence)
    int max_nnum;    // max neighbours of the last build (atomicMax)
    int max_vnum;    // max vertex neighbours of the last build
    int overflow;    // a build row did not fit the physical capacity
    double scal[16]; // reduction results
};

// Device control block of the device-resident FIRE loop.
struct FireCtl {
    double alpha, alpha_init, dt, dt_max, dt_min;
    double avg_pe_target, avg_pe_diff_target, fmax_target_sq, dof;
    double last_avg_pe, avg_pe, power, fmax_sq;
    long long n_good, n_bad, step, max_steps;
    int stopped, done, converged, uphill;
    int started;     // minimisation has begun (first step taken)
    int pad[3];
};

// Device pointers of one system (passed by value to kernels).
struct Dev {
    int64_t N, Nv;
    double2 *pos, *vel, *frc, *lref, *cref;
    double *rad, *mass, *pe, *ke;
    int *sidx;                  // static id of internal slot k
    // neighbour list, ELL column-major: nbr[k*N + i]
    int *nbr, *nnum;
    int cap;                    // physical ELL slots per particle
    int *cell_of;               // cell of each particle at the last cell rebuild
    int *cell_start, *cell_cnt, *cell_cur, *cell_tmp, *perm;
    double *pmax_nd, *pmax_cd;  // per-block max displacement partials
    double *part;               // reduction partials scratch (chunk partials)
    Ctl* ctl;
    FireCtl* fire;
    // Adam moments (allocated on first use)
    double2 *m1, *m2;
    // stress / contacts (allocated on first use)
    double *sxx, *sxy, *syx, *syy;
    int *contacts, *vcontacts;
    // rigid bumpy
    double *angle, *omega, *torque, *moi;
    double2 *vpos;
    double *vmass;
    int *vpart, *vstart, *nverts, *svidx;
    int *vnbr, *vnum;           // vertex ELL: vnbr[k*Nv + v]
    int vcap;                   // physical vertex ELL slots
    double2 *vfrc;              // per-vertex force (two-kernel path scratch)
    double *vtq, *vpe;
    double2 *lpos;              // COM before the last drift (bumpy rotate)
    double2 *dpos;              // COM displacement of the last drift
    double *dth;                // angle change of the last drift
};

// Reorder scratch: one spare copy of every permuted array.
struct Spare {
    double2 *pos, *vel, *frc, *m1, *m2;
    double *rad, *mass, *angle, *omega, *torque, *moi, *vmass;
    double2 *vpos;
    int *sidx, *nverts, *vstart, *vpart, *svidx, *old_to_new;
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t gen = ~0ull;
    int steps = 0;
    double dt = 0.0;
};

}  // namespace dpm

struct dpm_system {
    int kind = DPM_DISK;
    int64_t N = 0, Nv = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double box[2] = {1.0, 1.0};
    double e_c = 1.0, n_c = 2.0;
    dpm::Dev d{};
    dpm::Spare sp{};
    std::vector<void*> allocs;  // everything hipMalloc'd for this system
    int64_t nblk = 0;           // ceil(N/256)
    int64_t nchunk = 0;         // ceil(N/64)
    // neighbours
    int method = DPM_NBR_NONE;
    double rc = -1, thr_sq = -1, cell_thr_sq = -1, cell_size = -1;
    int64_t ncd = -1, ncells = -1;
    int64_t max_alloc = 4;      // reference-semantics logical capacity
    int64_t phys_cap = 0;       // allocated ELL slots per particle
    int64_t vmax_alloc = 4, vphys_cap = 0;
    int64_t n_rebuilds = 0, n_cell_rebuilds = 0;
    int64_t ncells_alloc = 0;
    bool nbr_ready = false;
    // rigid bumpy
    bool rotation = true;
    double rv = 0.0, rc_v = -1, pcut_v = -1;
    // snapshot (device resident)
    bool snap_valid = false;
    std::vector<void*> snap_allocs;
    dpm::Spare snap{};
    double snap_box[2] = {0, 0};
    // execution options
    bool use_graph = true;
    int64_t chunk = 64;
    bool timing = false;
    uint64_t gen = 0;           // bumped whenever baked kernel arguments change
    dpm::Graph nve_graph, fire_graph;
    dpm::Ctl* hctl = nullptr;   // pinned host mirrors (2 slots)
    dpm::FireCtl* hfire = nullptr;
    double ktime_ms[3] = {0, 0, 0};
    int64_t klaunch[3] = {0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    size_t ev_next = 0;
};
