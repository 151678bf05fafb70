// dpm_internal.hpp -- host-side system layout shared by the library's sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
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

// What a halted step still has to do once the rebuild chain has run.
enum HaltMode { H_NVE_FUSED = 0, H_NVE_FORCE = 1, H_FIRE = 2 };

// Device control block of one system.  Hot kernels only ever write it with plain
// stores (no atomics, no fences); the kernel that ends a launch unit copies it into a
// pinned host mirror, where the host reads it.
//
// Rebuild requests are "stamped": the kernel that moves particles writes, into
// flag[slot] (list rebuild) and cflag[slot] (cell + list rebuild), the stamp of the
// step that must act on it (epoch*8192 + position + 1).  A reader acts only on an
// exact stamp match, so stale requests never need resets.
struct Ctl {
    long long flag[2];
    long long cflag[2];
    long long epoch;          // bumped by the chain at the end of every chunk
    long long ovf;            // stamp of a step whose in-kernel list rebuild overflowed
    int halt;                 // a step stopped because it needs the rebuild chain
    int halt_pos;             // chunk position of the halted step
    int halt_mode;            // HaltMode of the halted step
    int halt_kind;            // rebuild it needs: 1 list, 2 cells + list, 3 list from lref
    int progress;             // steps completed so far in the running chunk
    int last_progress;        // steps completed in the previous chunk
    int done;                 // device-resident FIRE loop finished
    int kind;                 // rebuild chosen by the chain: 0 none, 1 list, 2 cells+list
    int overflow;             // a row exceeded the physical capacity: host must regrow
    int max_nnum, max_nnum_all;  // last build / running max over all builds
    int max_vnum, max_vnum_all;
    int n_rebuilds, n_cell_rebuilds;
    int bad_pos;              // a non-finite position was binned
    int closed;               // the closing force of an NVE run has been applied
    int pub_seq;              // launch units published to the host mirror since the reset
    double scal[16];          // small reduction results
};

// Device control block of the device-resident FIRE loop.
struct FireCtl {
    double alpha, alpha_init, dt, dt_max, dt_min;
    double avg_pe_target, avg_pe_diff_target, fmax_target_sq, dof;
    double last_avg_pe, avg_pe, power, fmax_sq;
    long long n_good, n_bad, step, max_steps;
    int stopped, done, converged, uphill;
    int pad[4];
};

// Device pointers of one system (passed by value to kernels).
struct Dev {
    int64_t N, Nv;
    // particle state: P0/P1 hold {x, y, radius, mass}; P0 is current outside fused
    // NVE chunks, which ping-pong the positions between the two.
    double4 *P0, *P1;
    double2 *vel, *frc, *lref, *cref;
    double *pe, *ke;
    int* sidx;                  // static id of internal slot k
    int *nbr, *nnum;            // ELL column-major: nbr[k*N + i]
    int cap;                    // physical ELL slots per particle
    int *cell_of;               // cell of each particle at the last cell rebuild
    int *cell_start, *cell_cnt, *cell_cur, *cell_tmp, *perm, *scan_bsum;
    double *pmax_nd, *pmax_cd;  // per-block max displacement partials
    double* part;               // canonical-sum chunk partials
    Ctl* ctl;
    Ctl* hpub;                  // device address of the pinned host mirror (4 slots)
    FireCtl* fire;
    double2 *m1, *m2;           // Adam moments (allocated on first use)
    double2* ma;                // rigid Adam: angle moments {first, second}
    double *sxx, *sxy, *syx, *syy;
    int *contacts, *vcontacts;
    // rigid bumpy
    double *angle, *omega, *torque, *moi;
    double2* vpos;
    double* vmass;
    int *vpart, *vstart, *nverts, *svidx;
    int *vnbr, *vnum;           // vertex ELL: vnbr[k*Nv + v]
    int vcap;
    double2* vfrc;
    double *vtq, *vpe;
    double2* lpos;              // COM before the last drift (vertex rotation)
    double* dth;                // angle change of the last drift
    double2* mpos;              // FIRE: COM after the uphill back-step (= lpos otherwise)
    double2 *rot0, *rot1;       // FIRE: (sin, cos) of the back-step's and the step's
                                // angle changes, formed once per particle
};

// Reorder scratch: one spare copy of every permuted array.
struct Spare {
    double4* P;
    double2 *vel, *frc, *m1, *m2, *ma;
    double *angle, *omega, *torque, *moi, *vmass;
    double2* vpos;
    int *sidx, *nverts, *vstart, *vpart, *svidx, *old_to_new;
};

// Numbers the neighbour kernels need (by value).
struct NbrParams {
    int method, ncd, fast;
    int pad;
    long long ncells;
    double rc2, thr, cthr, cs, Lx, Ly, invLx, invLy, halfLx, halfLy;
    double vc2, vp2;            // bumpy vertex lists
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t gen = ~0ull;
    double dt = 0.0;
    int64_t chunk = 0;  // the set marker: the chunk length its graphs were captured for
};

}  // namespace dpm

struct dpm_system {
    int kind = DPM_DISK;
    int64_t N = 0, Nv = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double box[2] = {1.0, 1.0};
    double e_c = 1.0, n_c = 2.0;
    double rmax = 0.0;          // largest radius (decides the fast minimum image)
    double rmin = 0.0;          // smallest radius (decides the short sqrt/division path)
    dpm::Dev d{};
    dpm::Spare sp{};
    std::vector<void*> allocs;
    int64_t nblk = 0;           // ceil(N/256)
    int64_t nchunk = 0;         // ceil(N/64)
    // neighbours
    int method = DPM_NBR_NONE;
    double rc = -1, thr_sq = -1, cell_thr_sq = -1, cell_size = -1;
    int64_t ncd = -1, ncells = -1;
    int64_t vg_ncd = 0;         // Verlet method: cells per side of the grid its rows are
                                // built from (0: the all-pairs scan)
    int64_t phys_cap = 0;
    int64_t vphys_cap = 0;
    int64_t nv_max = 0;         // most vertices of any particle
    bool vring = false;         // every vertex at r - r_v from its centre (bumpy_check_ring)
    int64_t ncells_alloc = 0;
    bool nbr_ready = false;
    // rigid bumpy
    bool rotation = true;
    double rv = 0.0, rc_v = -1, pcut_v = -1;
    // pair analytics (dpm_pairs.hip): pair_fields blocks of N * pair_row values
    double* pairs = nullptr;
    int64_t pair_alloc = 0, pair_row = 0;
    int pair_fields = 0;
    // snapshot (device resident)
    bool snap_valid = false;
    std::vector<void*> snap_allocs;
    dpm::Spare snap{};
    double snap_box[2] = {0, 0};
    // execution options
    bool use_graph = true;
    int64_t chunk = 32;
    int timing = 0;             // 0 off, 1 per launch, 2 per chunk
    uint64_t gen = 0;           // bumped whenever baked kernel arguments change
    std::map<int64_t, dpm::Graph> nve_graphs, fire_graphs;
    dpm::Ctl* hctl = nullptr;   // pinned host mirror of the control block (4 slots),
                                // written by the kernel that ends each launch unit
    double ktime_ms[3] = {0, 0, 0};
    int64_t klaunch[3] = {0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    std::vector<hipEvent_t> ev_free;
    hipEvent_t poll_ev[4] = {nullptr, nullptr, nullptr, nullptr};
};
