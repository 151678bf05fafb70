/*
 * oracle.h -- C ABI of the CPU oracle for the dpCUDA-v2 particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++17 (+OpenMP)
 * restatement of the reference's algorithms, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * It is never linked into, loaded by or called from the product library.
 *
 * The entry points mirror include/dpm.h one for one (orc_ <-> dpm_), with the
 * same argument meanings, so a test can drive both through the same sequence
 * and compare.  Enum values are identical to dpm.h's.
 */
#ifndef DPM_ORACLE_H
#define DPM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;

/* kinds, methods, fields: same numbering as dpm.h */
enum { ORC_DISK = 0, ORC_BUMPY = 1 };
enum { ORC_NBR_NONE = 0, ORC_NBR_ALL = 1, ORC_NBR_VERLET = 2, ORC_NBR_CELL = 3 };
enum {
    ORC_F_X = 0, ORC_F_Y = 1, ORC_F_VX = 2, ORC_F_VY = 3, ORC_F_FX = 4, ORC_F_FY = 5,
    ORC_F_RAD = 6, ORC_F_MASS = 7, ORC_F_PE = 8, ORC_F_KE = 9,
    ORC_F_ANGLE = 10, ORC_F_OMEGA = 11, ORC_F_TORQUE = 12, ORC_F_MOI = 13,
    ORC_F_STATIC_INDEX = 14,   /* int64 */
    ORC_F_NUM_NEIGHBORS = 15,  /* int64 */
    ORC_F_CONTACTS = 16,       /* int64 */
    ORC_F_STRESS_XX = 17, ORC_F_STRESS_XY = 18, ORC_F_STRESS_YX = 19, ORC_F_STRESS_YY = 20,
    ORC_F_VERTEX_X = 21, ORC_F_VERTEX_Y = 22, ORC_F_VERTEX_MASS = 23,
    ORC_F_NUM_VERTICES = 24,   /* int64, per particle */
    ORC_F_VERTEX_CONTACTS = 25 /* int64, per particle */
};

orc_system* orc_create(int kind, int64_t n_particles, int64_t n_vertices);
void orc_destroy(orc_system* s);
const char* orc_last_error(void);

int orc_set_box(orc_system* s, double lx, double ly);
int orc_get_box(orc_system* s, double* lx, double* ly);
int orc_set_energy(orc_system* s, double e_c, double n_c);
int orc_upload(orc_system* s, int field, const void* data, int64_t count);
int orc_download(orc_system* s, int field, void* data, int64_t count);

int orc_neighbor_config(orc_system* s, int method, double cutoff, double disp_thresh_sq,
                        int64_t n_cells_dim, double cell_disp_thresh_sq);
int orc_neighbor_init(orc_system* s);
int orc_neighbor_check(orc_system* s, int* rebuilt);
int orc_neighbor_pairs(orc_system* s, int64_t* out_i, int64_t* out_j, int64_t capacity, int64_t* count);
int orc_bumpy_neighbor_config(orc_system* s, double vertex_cutoff, double vertex_particle_cutoff);
int orc_vertex_pairs(orc_system* s, int64_t* out_i, int64_t* out_j, int64_t capacity, int64_t* count);
int orc_neighbor_stats(orc_system* s, int64_t* max_alloc, int64_t* total_entries, int64_t* n_rebuilds,
                       int64_t* n_cell_rebuilds);

int orc_zero_forces(orc_system* s);
int orc_compute_forces(orc_system* s);
int orc_kick(orc_system* s, double dt);
int orc_drift(orc_system* s, double dt);
int orc_mix(orc_system* s, double alpha);
int orc_zero_velocities(orc_system* s);
int orc_nve_run(orc_system* s, double dt, int64_t n_steps);

enum { ORC_RED_PE = 0, ORC_RED_KE = 1, ORC_RED_POWER = 2, ORC_RED_FMAX_SQ = 3,
       ORC_RED_FORCE_BALANCE = 4, ORC_RED_PRESSURE = 5, ORC_RED_CONTACTS = 6,
       ORC_RED_TEMPERATURE = 7, ORC_RED_PACKING_FRACTION = 8, ORC_RED_MAX_NEIGH_DISP_SQ = 9,
       ORC_RED_MAX_CELL_DISP_SQ = 10, ORC_RED_STRESS_XX = 11, ORC_RED_STRESS_XY = 12,
       ORC_RED_STRESS_YX = 13, ORC_RED_STRESS_YY = 14, ORC_RED_VERTEX_CONTACTS = 15 };
int orc_reduce(orc_system* s, int what, double* out);

typedef struct {
    double alpha_init, dt;
    double avg_pe_target, avg_pe_diff_target;
    double fmax_target;   /* <= 0 disables the max-force criterion */
    int64_t max_steps;
} orc_fire_params;
typedef struct {
    int64_t steps;        /* number of FIRE steps executed */
    int32_t stopped, converged;
    double avg_pe, fmax_sq, dt, alpha;
} orc_fire_result;
int orc_minimize_fire(orc_system* s, const orc_fire_params* p, orc_fire_result* out);
/* one FIRE iteration with explicit state, for state-machine trace tests */
typedef struct {
    double alpha, alpha_init, dt, dt_max, dt_min;
    int64_t n_good, n_bad;
    int32_t stopped;
    double power;          /* power seen at the start of the step */
} orc_fire_state;
int orc_fire_state_init(orc_fire_state* st, double alpha_init, double dt);
int orc_fire_step(orc_system* s, orc_fire_state* st);

int orc_adam_reset(orc_system* s);
/* the sin/cos pair used for rigid-body rotations (shared with the HIP kernels) */
int orc_sincos(const double* x, double* s_out, double* c_out, int64_t n);
int orc_adam_step(orc_system* s, int64_t step, double alpha, double beta1, double beta2, double eps);
int orc_minimize_adam(orc_system* s, double alpha, double beta1, double beta2, double eps,
                      double avg_pe_target, double avg_pe_diff_target, int64_t max_steps,
                      int64_t* steps_out);

int orc_compute_stress(orc_system* s);
int orc_count_contacts(orc_system* s);
int orc_scale_positions(orc_system* s, double factor);
int orc_scale_box(orc_system* s, double factor);
int orc_scale_to_packing_fraction(orc_system* s, double phi);
int orc_snapshot_save(orc_system* s);
int orc_snapshot_restore(orc_system* s);
int orc_remove_mean_velocity(orc_system* s);
int orc_scale_to_temperature(orc_system* s, double temperature);

typedef struct {
    double pressure_target;     /* jam when P in [P_t, P_t*(1+window_rel)] */
    double window_rel;
    double strain_step;         /* L <- L*(1-strain_step) while under-compressed */
    double decompress_factor;   /* initial L <- L*f while P > P_t*(1+window) */
    double length_tol;          /* give up when |L_low/L_high - 1| < length_tol */
    double min_contacts_per_particle;  /* extra stop criterion (reference: Z>2N); 0 disables */
    int64_t max_compression_steps;
    orc_fire_params fire;
} orc_jam_params;
typedef struct {
    int64_t compression_steps, fire_steps_total;
    int32_t status;             /* 0 jammed, 1 max steps, 2 length tolerance */
    double pressure, packing_fraction, pe, box_length;
    int64_t contacts;
} orc_jam_result;
int orc_jam(orc_system* s, const orc_jam_params* p, orc_jam_result* out);

/* rigid bumpy geometry set-up (vertex ring + masses + moments of inertia) */
int orc_bumpy_setup(orc_system* s, const int64_t* n_vertices_per_particle, double vertex_radius,
                    int rotation);

/* pair analytics (Disk/RigidBumpy::calculateForceDistancePairs): same fields, layout and
 * fill values as dpm_pair_analytics / dpm_pair_download */
enum {
    ORC_P_POTENTIAL = 0, ORC_P_FORCE_X = 1, ORC_P_FORCE_Y = 2, ORC_P_DIST_X = 3,
    ORC_P_DIST_Y = 4, ORC_P_ID_I = 5, ORC_P_ID_J = 6, ORC_P_OVERLAP = 7, ORC_P_RADSUM = 8,
    ORC_P_ANGLE = 9,
    ORC_P_H_XX = 10, ORC_P_H_XY = 11, ORC_P_H_YX = 12, ORC_P_H_YY = 13,
    ORC_P_HII_XX = 14, ORC_P_HII_XY = 15, ORC_P_HII_YX = 16, ORC_P_HII_YY = 17,
    ORC_P_ANGLE_I = 18, ORC_P_ANGLE_J = 19, ORC_P_VERTEX_CONTACTS = 20, ORC_P_FRICTION = 21,
    ORC_P_VERTEX_OVERLAP = 22,
    ORC_P_H_XT = 23, ORC_P_H_YT = 24, ORC_P_H_TX = 25, ORC_P_H_TY = 26, ORC_P_H_TT = 27,
    ORC_P_HII_XT = 28, ORC_P_HII_YT = 29, ORC_P_HII_TX = 30, ORC_P_HII_TY = 31,
    ORC_P_HII_TT = 32,
    ORC_P_NFIELDS = 33
};
int orc_pair_analytics(orc_system* s, int64_t* row_length);
int orc_pair_download(orc_system* s, int field, void* out, int64_t count);
int orc_bumpy_area(orc_system* s, double* out, int64_t count);

/* canonical fixed-order reduction used by both sides (exposed for tests) */
double orc_canonical_sum(const double* a, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
