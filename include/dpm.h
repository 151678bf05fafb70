/*
 * dpm.h -- C ABI of the MI355X-native dpCUDA-v2 particle hot path.
 *
 * One opaque handle per particle system (disk or rigid bumpy).  All state is
 * device resident (HBM); every call enqueues on the handle's own HIP stream.
 * Status returns: 0 = OK, negative = error (text from dpm_last_error()).
 * No exceptions cross this boundary and no torch types appear in it.
 *
 * The reference has no FFI: its boundary is the C++ virtual interface of
 * class Particle (include/particles/base/particle.h:28-717) and its
 * subclasses Disk (include/particles/disk/disk.h:18-80) and RigidBumpy
 * (include/particles/rigid_bumpy/rigid_bumpy.h:19-294), driven by the
 * integrators (src/integrator/) and routines (src/routines/).  Each
 * entry point below names the reference member(s) it replaces.  The C++
 * classes in dpcuda-v2_amd/host/ re-expose that interface on top of this ABI.
 *
 * Per-particle values cross the ABI in static (original upload) order; the
 * cell-list permutation stays internal, as the reference's
 * static_particle_index does (particle.h:136).
 */
#ifndef DPM_H
#define DPM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpm_system dpm_system;

enum { DPM_DISK = 0, DPM_BUMPY = 1 };
enum { DPM_NBR_NONE = 0, DPM_NBR_ALL = 1, DPM_NBR_VERLET = 2, DPM_NBR_CELL = 3 };
enum {
    DPM_F_X = 0, DPM_F_Y = 1, DPM_F_VX = 2, DPM_F_VY = 3, DPM_F_FX = 4, DPM_F_FY = 5,
    DPM_F_RAD = 6, DPM_F_MASS = 7, DPM_F_PE = 8, DPM_F_KE = 9,
    DPM_F_ANGLE = 10, DPM_F_OMEGA = 11, DPM_F_TORQUE = 12, DPM_F_MOI = 13,
    DPM_F_STATIC_INDEX = 14,   /* int64 */
    DPM_F_NUM_NEIGHBORS = 15,  /* int64 */
    DPM_F_CONTACTS = 16,       /* int64 */
    DPM_F_STRESS_XX = 17, DPM_F_STRESS_XY = 18, DPM_F_STRESS_YX = 19, DPM_F_STRESS_YY = 20,
    DPM_F_VERTEX_X = 21, DPM_F_VERTEX_Y = 22, DPM_F_VERTEX_MASS = 23,
    DPM_F_NUM_VERTICES = 24,   /* int64, per particle */
    DPM_F_VERTEX_CONTACTS = 25 /* int64, per particle */
};

/* ---- lifecycle ------------------------------------------------------------
 * Replaces Particle::setParticleCounts/initDynamicVariables
 * (src/particles/base/particle.cu:335-341, :394-403) and the RigidBumpy
 * vertex-array set-up (src/particles/rigid_bumpy/rigid_bumpy.cu:138-303). */
dpm_system* dpm_create(int kind, int64_t n_particles, int64_t n_vertices);
void dpm_destroy(dpm_system* s);
const char* dpm_last_error(void);
int dpm_device_count(void);

/* ---- parameters -------------------------------------------------------------
 * Particle::setBoxSize/getBoxSize (particle.cu:694-709),
 * Particle::setEnergyScale/setExponent (particle.cu:729-839). */
int dpm_set_box(dpm_system* s, double lx, double ly);
int dpm_get_box(dpm_system* s, double* lx, double* ly);
int dpm_set_energy(dpm_system* s, double e_c, double n_c);

/* ---- data transfer (host <-> HBM, static order) -----------------------------
 * Replaces the Data1D/Data2D setData/getData round trips
 * (src/data/data_1d.cu, src/data/data_2d.cu) used by Particle::tryLoadArrayData
 * (particle.cu:102-149) and Particle::getArrayData (particle.cu:485-692). */
int dpm_upload(dpm_system* s, int field, const void* data, int64_t count);
int dpm_download(dpm_system* s, int field, void* data, int64_t count);

/* ---- neighbours ----------------------------------------------------------------
 * dpm_neighbor_config: Particle::setNeighborMethod/setNeighborSize/setCellSize
 *   (particle.cu:237-265, :1223-1272), taking the final numbers as
 *   scripts/runShortNVE.cpp:43-54 assigns them.
 * dpm_neighbor_init:   Particle::initNeighborList (particle.cu:1163-1221).
 * dpm_neighbor_check:  Particle::checkForNeighborUpdate (particle.cu:1137-1161).
 * dpm_neighbor_pairs:  the (i,j) pairs of the list in static ids (parity tests). */
int dpm_neighbor_config(dpm_system* s, int method, double cutoff, double disp_thresh_sq,
                        int64_t n_cells_dim, double cell_disp_thresh_sq);
int dpm_bumpy_neighbor_config(dpm_system* s, double vertex_cutoff, double vertex_particle_cutoff);
int dpm_neighbor_init(dpm_system* s);
int dpm_neighbor_check(dpm_system* s, int* rebuilt);
int dpm_neighbor_pairs(dpm_system* s, int64_t* out_i, int64_t* out_j, int64_t capacity,
                       int64_t* count);
int dpm_vertex_pairs(dpm_system* s, int64_t* out_i, int64_t* out_j, int64_t capacity,
                     int64_t* count);
int dpm_neighbor_stats(dpm_system* s, int64_t* max_alloc, int64_t* total_entries,
                       int64_t* n_rebuilds, int64_t* n_cell_rebuilds);

/* ---- forces and integration ------------------------------------------------------
 * zero/compute: Particle::zeroForceAndPotentialEnergy (particle.cu:1334-1336),
 *   Disk::calculateForces (disk.cu:218-221), RigidBumpy::calculateForces
 *   (rigid_bumpy.cu:1088-1105).
 * kick/drift/mix: Particle::updateVelocities/updatePositions (particle.cu:1086-1092),
 *   Particle::mixVelocitiesAndForces (particle.cu:895-897), rigid overrides
 *   (rigid_bumpy.cu:1107-1128, :1271-1275); setVelocitiesToZero (particle.cu:871-873).
 * dpm_nve_run: n_steps of NVE::step (src/integrator/nve.cu:15-25) as looped by
 *   scripts/runNVE.cpp:43-47, run device-resident. */
int dpm_zero_forces(dpm_system* s);
int dpm_compute_forces(dpm_system* s);
int dpm_kick(dpm_system* s, double dt);
int dpm_drift(dpm_system* s, double dt);
int dpm_mix(dpm_system* s, double alpha);
int dpm_zero_velocities(dpm_system* s);
int dpm_nve_run(dpm_system* s, double dt, int64_t n_steps);

/* ---- reductions ----------------------------------------------------------------------
 * Particle::totalPotentialEnergy/totalKineticEnergy/calculateTemperature/getPowerFire/
 * getForceBalance/getPressure/getContactCount/getPackingFraction/
 * getMaxSquared{Neighbor,Cell}Displacement (particle.cu:231-235, :993-1015,
 * :1074-1100, :1338-1341, :1367-1385). */
enum { DPM_RED_PE = 0, DPM_RED_KE = 1, DPM_RED_POWER = 2, DPM_RED_FMAX_SQ = 3,
       DPM_RED_FORCE_BALANCE = 4, DPM_RED_PRESSURE = 5, DPM_RED_CONTACTS = 6,
       DPM_RED_TEMPERATURE = 7, DPM_RED_PACKING_FRACTION = 8, DPM_RED_MAX_NEIGH_DISP_SQ = 9,
       DPM_RED_MAX_CELL_DISP_SQ = 10, DPM_RED_STRESS_XX = 11, DPM_RED_STRESS_XY = 12,
       DPM_RED_STRESS_YX = 13, DPM_RED_STRESS_YY = 14, DPM_RED_VERTEX_CONTACTS = 15 };
int dpm_reduce(dpm_system* s, int what, double* out);

/* ---- FIRE ---------------------------------------------------------------------------
 * dpm_minimize_fire: minimizeFire (src/routines/minimization.cu:75-148) with
 *   Fire::step (src/integrator/fire.cu:13-46) run as a device-resident state
 *   machine; fmax_target > 0 adds the max-force stop criterion.
 * dpm_fire_step: one Fire::step with the state held by the caller. */
typedef struct {
    double alpha_init, dt;
    double avg_pe_target, avg_pe_diff_target;
    double fmax_target;
    int64_t max_steps;
} dpm_fire_params;
typedef struct {
    int64_t steps;
    int32_t stopped, converged;
    double avg_pe, fmax_sq, dt, alpha;
} dpm_fire_result;
typedef struct {
    double alpha, alpha_init, dt, dt_max, dt_min;
    int64_t n_good, n_bad;
    int32_t stopped;
    double power;
} dpm_fire_state;
int dpm_minimize_fire(dpm_system* s, const dpm_fire_params* p, dpm_fire_result* out);
int dpm_fire_state_init(dpm_fire_state* st, double alpha_init, double dt);
int dpm_fire_step(dpm_system* s, dpm_fire_state* st);

/* ---- Adam -------------------------------------------------------------------------------
 * Adam::minimize (src/integrator/adam.cu:11-18) + Particle::updatePositionsAdam
 * (particle.cu:329-333); rigid bumpy: RigidBumpy::updatePositionsAdam (rigid_bumpy.cu:
 * 1473-1486, moments of force and torque, then the vertex rotation); minimizeAdam
 * (src/routines/minimization.cu:3-33); dpm_adam_reset: Particle::initAdamVariables
 * (particle.cu:319-322) and RigidBumpy::initAdamVariables (rigid_bumpy.cu:1461-1465),
 * zero moments. */
int dpm_adam_reset(dpm_system* s);
int dpm_adam_step(dpm_system* s, int64_t step, double alpha, double beta1, double beta2,
                  double eps);
int dpm_minimize_adam(dpm_system* s, double alpha, double beta1, double beta2, double eps,
                      double avg_pe_target, double avg_pe_diff_target, int64_t max_steps,
                      int64_t* steps_out);

/* ---- analytics, box and state -------------------------------------------------------------
 * calculateStressTensor (disk.cu:281-285, rigid_bumpy.cu:1513-1519), countContacts
 * (disk.cu:250-255, rigid_bumpy.cu:1505-1511), scalePositionsFull/scaleBox/
 * scaleToPackingFractionFull (particle.cu:1047-1072, rigid_bumpy.cu:1044-1048),
 * setLastState/revertToLastState (particle.cu:956-991, rigid_bumpy.cu:1144-1179),
 * removeMeanVelocities/scaleVelocitiesToTemperature (particle.cu:865-906). */
int dpm_compute_stress(dpm_system* s);
int dpm_count_contacts(dpm_system* s);
int dpm_scale_positions(dpm_system* s, double factor);
int dpm_scale_box(dpm_system* s, double factor);
int dpm_scale_to_packing_fraction(dpm_system* s, double phi);
int dpm_snapshot_save(dpm_system* s);
int dpm_snapshot_restore(dpm_system* s);
int dpm_remove_mean_velocity(dpm_system* s);
int dpm_scale_to_temperature(dpm_system* s, double temperature);

/* ---- jamming ----------------------------------------------------------------------------
 * The compress/minimise pressure bisection of scripts/runArthurJamWithFire.cpp:36-106
 * (minimizeFire + getPressure + countContacts + scaleBox), with the reference's
 * file round trip replaced by a device snapshot. */
typedef struct {
    double pressure_target, window_rel, strain_step, decompress_factor, length_tol;
    double min_contacts_per_particle;
    int64_t max_compression_steps;
    dpm_fire_params fire;
} dpm_jam_params;
typedef struct {
    int64_t compression_steps, fire_steps_total;
    int32_t status;
    double pressure, packing_fraction, pe, box_length;
    int64_t contacts;
} dpm_jam_result;
int dpm_jam(dpm_system* s, const dpm_jam_params* p, dpm_jam_result* out);

/* ---- pair analytics -------------------------------------------------------------------
 * Disk::calculateForceDistancePairs (disk.cu:228-248, kernelCalcDiskForceDistancePairs
 * disk/kernels.cu:73-175) and RigidBumpy::calculateForceDistancePairs
 * (rigid_bumpy.cu:1383-1470, kernelCalcRigidBumpyForceDistancePairs
 * rigid_bumpy/kernels.cu:482-702): one entry per neighbour-list slot, laid out as the
 * reference lays them out: row i = the i-th particle in the library's current (cell)
 * order, row length = the max_alloc of dpm_neighbor_stats, entry k = the k-th neighbour
 * of that row; unused entries hold the reference's fill values.  dpm_pair_analytics
 * computes every field for the current state and returns the row length;
 * dpm_pair_download copies one field (n_particles * row_length values: int64 for the id
 * and vertex-count fields, double otherwise).  The *_T* Hessian entries, the angles, the
 * vertex-contact counts, the friction coefficient and the vertex overlaps exist for
 * rigid bumpy systems only. */
enum {
    DPM_P_POTENTIAL = 0, DPM_P_FORCE_X = 1, DPM_P_FORCE_Y = 2, DPM_P_DIST_X = 3,
    DPM_P_DIST_Y = 4, DPM_P_ID_I = 5, DPM_P_ID_J = 6, DPM_P_OVERLAP = 7, DPM_P_RADSUM = 8,
    DPM_P_ANGLE = 9,
    DPM_P_H_XX = 10, DPM_P_H_XY = 11, DPM_P_H_YX = 12, DPM_P_H_YY = 13,
    DPM_P_HII_XX = 14, DPM_P_HII_XY = 15, DPM_P_HII_YX = 16, DPM_P_HII_YY = 17,
    DPM_P_ANGLE_I = 18, DPM_P_ANGLE_J = 19, DPM_P_VERTEX_CONTACTS = 20, DPM_P_FRICTION = 21,
    DPM_P_VERTEX_OVERLAP = 22,
    DPM_P_H_XT = 23, DPM_P_H_YT = 24, DPM_P_H_TX = 25, DPM_P_H_TY = 26, DPM_P_H_TT = 27,
    DPM_P_HII_XT = 28, DPM_P_HII_YT = 29, DPM_P_HII_TX = 30, DPM_P_HII_TY = 31,
    DPM_P_HII_TT = 32,
    DPM_P_NFIELDS = 33
};
int dpm_pair_analytics(dpm_system* s, int64_t* row_length);
int dpm_pair_download(dpm_system* s, int field, void* out, int64_t count);
/* RigidBumpy::calculateParticleArea (rigid_bumpy.cu:1018-1024 with
 * kernelCalculateBumpyParticleAreaFull, rigid_bumpy/kernels.cu:1226-1283): the area of each
 * particle's vertex ring (polygon plus the exposed parts of the vertex disks), static order */
int dpm_bumpy_area(dpm_system* s, double* out, int64_t count);

/* ---- rigid bumpy set-up --------------------------------------------------------------------
 * RigidBumpy::setParticleStartIndex/setVertexParticleIndex/setVertexPositions/setMass/
 * setDegreesOfFreedom (rigid_bumpy.cu:476-478, :562-568, :783-787, :1068-1081, :618-624). */
int dpm_bumpy_setup(dpm_system* s, const int64_t* n_vertices_per_particle, double vertex_radius,
                    int rotation);

/* ---- batched jamming engine (no reference counterpart: the reference jams one system per
 * process; batching is a TODO there, todo.txt:43) ---------------------------------------------
 * Many independent systems of uneven N -- disks, and with dpm_batch_create_mixed rigid bumpy
 * systems beside them -- in one device arena.  dpm_batch_jam runs dpm_jam (the Verlet-list
 * form) on every system at once: one launch sequence per step advances each system's own
 * minimizeFire / controller state, and finished systems drop out.  Each disk system's
 * result is bit-identical to dpm_jam on a dpm_system holding the same particles with the
 * Verlet method, cutoff and threshold given here.
 * dpm_batch_config: pair law (e_c, n_c), Verlet cutoff and rebuild threshold (3 max d^2 >
 * disp_thresh_sq), common to all systems.  dpm_batch_jam: out[] has n_systems entries;
 * status -1 marks a system still running after max_batched_steps; steps_done returns the
 * batched steps executed.  dpm_batch_set_option takes DPM_OPT_GRAPH, DPM_OPT_CHUNK
 * (batched steps per graph launch) and DPM_OPT_TIMING. */
typedef struct dpm_batch dpm_batch;
dpm_batch* dpm_batch_create(int64_t n_systems, const int64_t* n_particles);
void dpm_batch_destroy(dpm_batch* b);
int dpm_batch_set_system(dpm_batch* b, int64_t system, const double* x, const double* y,
                         const double* radius, const double* mass, double box_side);
int dpm_batch_get_system(dpm_batch* b, int64_t system, double* x, double* y, double* box_side);
/* Rigid bumpy systems in the arena: n_vertices[s] > 0 makes system s a rigid bumpy system of
 * that many vertices (0: a disk system), set with dpm_batch_set_rigid_system (the geometry of
 * dpm_bumpy_setup: vertex rings from the angles, vertex masses, moments of inertia;
 * RigidBumpy::initializeVerticesOnParticles, rigid_bumpy.cu:476-478, :562-568).  A rigid
 * system jams with its own contact energy e_c (n_c of dpm_batch_config) and the neighbour
 * rules of RigidBumpy::setNeighborSize (rigid_bumpy.cu:1278-1300) with the standard
 * multipliers: Verlet cutoff 1.5 d_max, rebuild threshold (0.2 r_v)^2, vertex cutoff
 * 1.5 (2 r_v), vertex-particle cutoff d_max.  Its result is bit-identical to dpm_jam on a
 * dpm_system set up with dpm_bumpy_setup and those cutoffs with the Verlet method.
 * dpm_batch_get_rigid_system returns the centres, angles, vertex positions (n_vertices
 * entries, particle order) and box side; any output may be null. */
dpm_batch* dpm_batch_create_mixed(int64_t n_systems, const int64_t* n_particles,
                                  const int64_t* n_vertices);
int dpm_batch_set_rigid_system(dpm_batch* b, int64_t system, const double* x, const double* y,
                               const double* radius, const double* mass, const double* angle,
                               const int64_t* vertices_per_particle, double vertex_radius,
                               double e_c, int32_t rotation, double box_side);
int dpm_batch_get_rigid_system(dpm_batch* b, int64_t system, double* x, double* y,
                               double* angle, double* vertex_x, double* vertex_y,
                               double* box_side);
int dpm_batch_config(dpm_batch* b, double e_c, double n_c, double cutoff, double disp_thresh_sq);
int dpm_batch_set_option(dpm_batch* b, int option, int64_t value);
int dpm_batch_jam(dpm_batch* b, const dpm_jam_params* p, int64_t max_batched_steps,
                  dpm_jam_result* out, int64_t* steps_done);
/* minimizeFire (src/routines/minimization.cu:75-112) of every system at once, from the state
 * as set: each system's result equals dpm_minimize_fire on that system alone (Verlet method,
 * this batch's cutoff and threshold) followed by dpm_reduce of the pressure, contacts,
 * packing fraction and energy.  out[s]: fire_steps_total = the FIRE steps, status 0
 * converged / 1 not converged (-1 still running after max_batched_steps),
 * compression_steps 0, box_length unchanged. */
int dpm_batch_minimize(dpm_batch* b, const dpm_fire_params* p, int64_t max_batched_steps,
                       dpm_jam_result* out, int64_t* steps_done);
/* the batched step (1-based) in which each system finished its last dpm_batch_jam, -1 if
 * it has not */
int dpm_batch_finish_steps(dpm_batch* b, int64_t* steps_out);
/* Measurement: with dpm_batch_set_option(b, DPM_OPT_TIMING, 1) every batched step is
 * launched eagerly with HIP events around its force pass (the disk pass k_b_force, or the
 * rigid pass k_rb_force in a batch of rigid systems only); dpm_batch_kernel_time returns
 * the summed milliseconds and the number of timed passes since the option was set.
 * dpm_batch_row_entries: the stored neighbour-row entries of every system's particles. */
int dpm_batch_kernel_time(dpm_batch* b, double* force_ms, int64_t* launches);
int dpm_batch_row_entries(dpm_batch* b, int64_t* total);
/* The stored vertex-row entries of every rigid system's vertices (0 without vertex slots);
 * no counterpart in the reference (the batched engine has none), used for the force pass's
 * algorithmic bytes. */
int dpm_batch_vertex_row_entries(dpm_batch* b, int64_t* total);
/* Autotuner: times trial_steps batched steps of dpm_batch_jam for every pair of a Verlet
 * cutoff (rebuild threshold: half its skin) and a number of steps per graph launch,
 * restoring the configured state after each trial, and keeps the fastest pair.  Neither
 * choice changes a result bit. */
int dpm_batch_autotune(dpm_batch* b, const dpm_jam_params* p, int64_t trial_steps,
                       const double* cutoffs, int n_cutoffs, const int64_t* chunks,
                       int n_chunks, double* best_cutoff, int64_t* best_chunk,
                       double* best_ms_per_step);
/* Launch-shape autotuner: with the current cutoff and steps per launch, times trial_steps
 * batched steps for each force-pass workgroup size (64, 128, 256), keeps the fastest, then
 * for each cell width of the cell-grid rows (100, 125, 150 % of the cutoff), restoring the
 * configured state after each trial; the choices stay set (DPM_OPT_FORCE_BLOCK,
 * DPM_OPT_CELL_PCT). */
int dpm_batch_autotune_shape(dpm_batch* b, const dpm_jam_params* p, int64_t trial_steps,
                             int64_t* best_block, int64_t* best_cell_pct,
                             double* best_ms_per_step);

/* ---- serving loop (no reference counterpart) ---------------------------------------------
 * Continuous batching of jamming and minimisation requests.  dpm_server_submit queues a
 * request (arrays are copied); dpm_server_progress admits queued requests while the device
 * budget allows -- each joins the running batch of the batched engine, whose arena holds
 * arena_slots particle slots, vertex_slots vertex slots (rigid bumpy requests; 0: the
 * server takes none) and max_systems systems -- advances the batch by batched_steps and
 * collects finished requests.
 * Disk requests share the server's pair law (e_c = 1, n_c = 2, Verlet cutoff 1.7, rebuild at
 * 3 max d^2 > 3 (0.15)^2); a rigid bumpy request brings its own e_c and follows the rigid
 * rules of dpm_batch_set_rigid_system (rotation on).  A DPM_REQ_JAM request's result
 * equals dpm_jam on the request's system alone (Verlet method); a DPM_REQ_MINIMIZE request
 * runs only minimizeFire with jam.fire and its result is that of dpm_batch_minimize (status
 * 0 converged, 1 not).
 * dpm_server_result: state 0 queued, 1 running, 2 done (out, latency from submission to
 * completion, final positions and box side are filled), 3 failed.  dpm_server_release
 * forgets a done or failed request (its result memory is freed).  Entry points of one
 * server may be called from several threads; dpm_server_progress calls must not overlap. */
enum { DPM_REQ_JAM = 0, DPM_REQ_MINIMIZE = 1 };
typedef struct dpm_server dpm_server;
typedef struct {
    int32_t kind;             /* DPM_DISK or DPM_BUMPY */
    int32_t mode;             /* DPM_REQ_JAM or DPM_REQ_MINIMIZE */
    int64_t n_particles, n_vertices;
    const double *x, *y, *radius, *mass;
    const double* angle;                  /* rigid bumpy: orientation per particle */
    const int64_t* vertices_per_particle; /* rigid bumpy */
    double vertex_radius;                 /* rigid bumpy */
    double box_side, e_c;
    dpm_jam_params jam;
} dpm_request;
dpm_server* dpm_server_create(int64_t max_systems, int64_t arena_slots, int64_t budget_bytes,
                              int64_t vertex_slots);
void dpm_server_destroy(dpm_server* s);
int64_t dpm_server_submit(dpm_server* s, const dpm_request* r);
int dpm_server_progress(dpm_server* s, int64_t batched_steps, int64_t* finished_total,
                        int64_t* in_flight, int64_t* queued);
int dpm_server_result(dpm_server* s, int64_t id, int32_t* state, dpm_jam_result* out,
                      double* latency_s, double* x, double* y, double* box_side);
int dpm_server_release(dpm_server* s, int64_t id);

/* ---- allocation-free forces (no counterpart in the reference) ------------------------------
 * For callers that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  Disk::calculateForces (disk.cu:218-221)
 * and totalPotentialEnergy (particle.cu:1078-1080) of n disks held in the caller's device
 * arrays x, y, radius (n doubles each): fx, fy, pe (n doubles each) receive the forces and
 * per-particle energies and pe_total (one double, device memory) the total -- the values a
 * disk system with the Verlet method and this cutoff returns from dpm_neighbor_init +
 * dpm_zero_forces + dpm_compute_forces + dpm_reduce(DPM_RED_PE), bit for bit.  Kernels are
 * enqueued on `stream` (a hipStream_t; null for the null stream) and the call returns
 * without waiting.  The working memory is a 20 MiB array of the library, resident on each
 * device it runs on until the process ends; the call allocates nothing.  Limits: n <=
 * 1,048,576; a box under three cutoffs a side takes n <= 4,096.  Calls on one device share
 * that array: calls that overlap in time on one device (other streams, other threads)
 * race and must be serialised by the caller. */
int dpm_disk_forces_noalloc(const double* x, const double* y, const double* radius, int64_t n,
                            double lx, double ly, double e_c, double n_c, double cutoff,
                            double* fx, double* fy, double* pe, double* pe_total, void* stream);

/* ---- execution options and measurement (no reference counterpart) -------------------------
 * DPM_OPT_GRAPH: capture the NVE/FIRE chunk in a hipGraph (default 1).
 * DPM_OPT_CHUNK: steps per host poll of the device status (default 32).
 * DPM_OPT_TIMING: 0 off (default); 1 HIP events around every launch of the hot kernels
 *   (disables graph capture); 2 HIP events around every graph-launched chunk of the NVE and
 *   FIRE loops (chunk time is charged to DPM_K_FORCE, recovery chains to DPM_K_REBUILD).
 *   dpm_kernel_time returns the summed milliseconds and the count per kernel class. */
enum { DPM_OPT_GRAPH = 0, DPM_OPT_CHUNK = 1, DPM_OPT_TIMING = 2 };
/* batched engine only: the force pass's workgroup size (64, 128 or 256 slots; default 256)
 * and the cell-grid build's cell width in percent of the cutoff (100-400; default 100).
 * Neither changes a result bit. */
enum { DPM_OPT_FORCE_BLOCK = 3, DPM_OPT_CELL_PCT = 4 };
enum { DPM_K_FORCE = 0, DPM_K_DRIFT = 1, DPM_K_REBUILD = 2 };
int dpm_set_option(dpm_system* s, int option, int64_t value);
int dpm_kernel_time(dpm_system* s, int kernel_class, double* total_ms, int64_t* launches);
int dpm_reset_kernel_time(dpm_system* s);
int dpm_synchronize(dpm_system* s);

#ifdef __cplusplus
}
#endif
#endif
