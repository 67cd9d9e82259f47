/*
 * dpmd.h -- C ABI of the MI355X-native (gfx950, HIP) batched 2D granular-MD hot path.
 *
 * This is the drop-in boundary for the reference's particle / integrator / routine API
 * (mmccraw98/dpmd).  Every entry point below replaces one reference call; the citation
 * on each names the reference interface (file:line) it stands in for.  The C++ facade in
 * include/dpmd/{device_fields,particles,integrators,routines}.hpp re-expose these under the
 * reference's own class and method names.
 *
 * Conventions
 *   - All entry points return 0 on success and a nonzero dpmd_status otherwise; the
 *     message is available from dpmd_last_error(ctx) (or dpmd_last_error(NULL) for
 *     failures that have no context yet).  No C++ exception crosses this boundary.
 *   - Floating point is IEEE FP64 throughout.
 *   - Per-system parameter arrays (dt, alpha, scale, flags, ...) are DEVICE pointers of
 *     length n_systems, exactly as the reference passes df::DeviceField1D<double>.
 *   - Particle arrays live in "current" order: cell rebuilds reorder particles so that
 *     each cell's particles are contiguous (reference: disk.cu:1102-1125).  The field
 *     DPMD_F_STATIC_INDEX maps canonical (input) index -> current index, as the
 *     reference's static_index does (base_particle_kernels.cu:172-180).
 *   - Every call is stream-ordered on the stream given at creation (NULL = the default
 *     stream).  Calls on one context must come from one host thread at a time.
 */
#ifndef DPMD_DPMD_H
#define DPMD_DPMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpmd_ctx dpmd_ctx;

enum dpmd_status {
    DPMD_OK = 0,
    DPMD_ERR_ARGUMENT = 1,   /* bad argument or precondition (std::invalid_argument)    */
    DPMD_ERR_STATE = 2,      /* call out of order (std::runtime_error in the reference) */
    DPMD_ERR_HIP = 3,        /* HIP runtime error                                        */
    DPMD_ERR_OVERFLOW = 4    /* neighbor list capacity exceeded                          */
};

enum dpmd_kind { DPMD_DISK = 0, DPMD_RIGID_BUMPY = 1 };

/* reference: enum class md::NeighborMethod (base_particle.hpp:24) */
enum dpmd_neighbor_method { DPMD_NEIGHBOR_NAIVE = 0, DPMD_NEIGHBOR_CELL = 1 };

/* PBC = compute_forces(), WALL = compute_wall_forces() (base_particle.hpp:380-383);
 * PER_SYSTEM picks per system from DPMD_F_BOUNDARY (0 = periodic, 1 = walls). */
enum dpmd_force_mode { DPMD_FORCE_PBC = 0, DPMD_FORCE_WALL = 1, DPMD_FORCE_PER_SYSTEM = 2 };

enum dpmd_order { DPMD_ORDER_CURRENT = 0, DPMD_ORDER_CANONICAL = 1 };

/* FIRE stopping predicate: REFERENCE = fire_convergence_kernel (fire.cu:59-78);
 * MAX_FORCE = every particle's |F_i| below a tolerance (per system); FORCE_OR_PE = MAX_FORCE
 * or the reference's unjammed test <pe> < avg_pe_target (per system). */
enum dpmd_fire_predicate { DPMD_FIRE_PRED_REFERENCE = 0, DPMD_FIRE_PRED_MAX_FORCE = 1, DPMD_FIRE_PRED_FORCE_OR_PE = 2 };

/* Fields (reference: public DeviceField members, base_particle.hpp:31-103,
 * base_point_particle.hpp:86-89, base_poly_particle.hpp:104-120, rigid_bumpy.cuh:31-45).
 * Type and length are given per field: P = n_particles, V = n_vertices, S = n_systems. */
enum dpmd_field {
    /* double, P */
    DPMD_F_POS_X = 0, DPMD_F_POS_Y, DPMD_F_VEL_X, DPMD_F_VEL_Y, DPMD_F_FORCE_X, DPMD_F_FORCE_Y,
    DPMD_F_PE, DPMD_F_KE, DPMD_F_RAD, DPMD_F_MASS, DPMD_F_AREA, DPMD_F_LAST_POS_X, DPMD_F_LAST_POS_Y,
    DPMD_F_ANGLE, DPMD_F_ANGULAR_VEL, DPMD_F_TORQUE, DPMD_F_MOMENT_INERTIA,
    DPMD_F_STRESS_XX, DPMD_F_STRESS_XY, DPMD_F_STRESS_YX, DPMD_F_STRESS_YY,
    DPMD_F_OVERLAP, DPMD_F_CONTACT_COUNT,   /* overlaps: total overlap, contact count (V for rigid) */
    /* double, S */
    DPMD_F_BOX_SIZE_X = 32, DPMD_F_BOX_SIZE_Y, DPMD_F_BOX_INV_X, DPMD_F_BOX_INV_Y,
    DPMD_F_E_INTERACTION, DPMD_F_NEIGHBOR_CUTOFF, DPMD_F_THRESH2,
    DPMD_F_PE_TOTAL, DPMD_F_KE_TOTAL, DPMD_F_FPOWER_TOTAL, DPMD_F_PRESSURE, DPMD_F_TEMPERATURE,
    DPMD_F_PACKING_FRACTION, DPMD_F_MAX_FORCE2,
    DPMD_F_STRESS_TOTAL_XX, DPMD_F_STRESS_TOTAL_XY, DPMD_F_STRESS_TOTAL_YX, DPMD_F_STRESS_TOTAL_YY,
    /* int, S (+1 where noted) */
    DPMD_F_SYSTEM_OFFSET = 64 /* S+1 */, DPMD_F_SYSTEM_SIZE, DPMD_F_CELL_DIM_X, DPMD_F_CELL_DIM_Y,
    DPMD_F_CELL_SYSTEM_START /* S+1 */, DPMD_F_BOUNDARY, DPMD_F_N_DOF, DPMD_F_REBUILD_FLAG,
    DPMD_F_FIRE_STEPS /* steps that did work per system in the last dpmd_fire_minimize /
                         dpmd_jam_compress: a system stops costing work one step after it stops */,
    DPMD_F_N_CONTACTS_TOTAL,
    /* int, P (or V for rigid bodies: the neighbor / cell objects are vertices) */
    DPMD_F_SYSTEM_ID = 80, DPMD_F_STATIC_INDEX, DPMD_F_CELL_ID, DPMD_F_NEIGHBOR_COUNT, DPMD_F_CONTACTS,
    /* rigid bumpy (base_poly_particle.hpp:104-120) */
    DPMD_F_VERTEX_POS_X = 96, DPMD_F_VERTEX_POS_Y, DPMD_F_VERTEX_FORCE_X, DPMD_F_VERTEX_FORCE_Y,
    DPMD_F_VERTEX_PE, DPMD_F_VERTEX_RAD, DPMD_F_VERTEX_MASS,              /* double, V */
    DPMD_F_VERTEX_PARTICLE_ID = 112, DPMD_F_PARTICLE_OFFSET /* P+1 */,
    DPMD_F_N_VERTICES_PER_PARTICLE, DPMD_F_VERTEX_SYSTEM_ID,
    DPMD_F_VERTEX_SYSTEM_OFFSET /* S+1 */                                 /* int */
};

/* ---- context lifecycle ------------------------------------------------------------ */
/* Replaces constructing md::disk::Disk / md::rigid_bumpy::RigidBumpy (disk.cuh:22,
 * rigid_bumpy.cuh:24).  hip_stream is a hipStream_t (NULL = default stream). */
int dpmd_ctx_create(dpmd_ctx** out, int kind, void* hip_stream);
int dpmd_ctx_destroy(dpmd_ctx* ctx);
const char* dpmd_last_error(const dpmd_ctx* ctx);

/* allocate_systems / allocate_particles / allocate_vertices (base_particle.hpp:315-338) */
int dpmd_allocate_systems(dpmd_ctx* ctx, int n_systems);
int dpmd_allocate_particles(dpmd_ctx* ctx, int n_particles);
int dpmd_allocate_vertices(dpmd_ctx* ctx, int n_vertices);
/* set_neighbor_method (base_particle.hpp:260) */
int dpmd_set_neighbor_method(dpmd_ctx* ctx, int method);
/* get_class_name / get_neighbor_method (base_particle.hpp:108-117, :743): the context's
 * dpmd_kind and dpmd_neighbor_method */
int dpmd_ctx_info(const dpmd_ctx* ctx, int* kind, int* neighbor_method);
/* n_particles / n_systems / n_vertices / n_cells (base_particle.hpp:291-312) */
int dpmd_sizes(const dpmd_ctx* ctx, int* n_systems, int* n_particles, int* n_vertices, int* n_cells);

/* ---- fields ----------------------------------------------------------------------- */
/* Device pointer and element count of a context-owned field (DeviceField::ptr(),
 * device_fields.cuh:303-463). */
int dpmd_field(dpmd_ctx* ctx, int field, void** device_ptr, int64_t* count);
/* from_host (device_fields.cuh:426-429): count must equal the field length. */
int dpmd_set_field(dpmd_ctx* ctx, int field, const void* host, int64_t count);
/* to_host, in current or canonical order (canonical applies to P- and V-length fields). */
int dpmd_get_field(dpmd_ctx* ctx, int field, void* host, int64_t count, int order);

/* sync_box / sync_system / sync_cells / sync_class_constants (base_particle.hpp:346-368):
 * validates the inputs and derives box inverses, system ids, cell layout. */
int dpmd_sync(dpmd_ctx* ctx);

/* ---- neighbors -------------------------------------------------------------------- */
int dpmd_init_neighbors(dpmd_ctx* ctx);                        /* base_particle.hpp:265-272 */
int dpmd_update_neighbors(dpmd_ctx* ctx);                      /* base_particle.hpp:275-280 */
int dpmd_check_neighbors(dpmd_ctx* ctx, int force_update);     /* base_particle.hpp:283-288 */
/* n_neighbors (base_particle.hpp:298-310) as *total, plus the fixed per-object capacity used
 * here and the largest count in the current lists.  Any pointer may be NULL. */
int dpmd_neighbor_info(dpmd_ctx* ctx, int* capacity, int* max_count, int64_t* total);
/* neighbor_start / neighbor_ids in the reference's CSR layout, current order. */
int dpmd_get_neighbors_csr(dpmd_ctx* ctx, int* start, int* ids, int64_t ids_len);
/* Neighbor-list capacity per object (default 32 for disks, 48 for vertices). */
int dpmd_set_neighbor_capacity(dpmd_ctx* ctx, int capacity);

/* ---- particle primitives (base_particle.hpp:380-420) -------------------------------- */
int dpmd_compute_forces(dpmd_ctx* ctx, int mode);
int dpmd_update_positions(dpmd_ctx* ctx, const double* scale, double scale2);
int dpmd_update_velocities(dpmd_ctx* ctx, const double* scale, double scale2);
int dpmd_scale_velocities(dpmd_ctx* ctx, const double* scale);
int dpmd_scale_positions(dpmd_ctx* ctx, const double* scale);
int dpmd_mix_velocities_and_forces(dpmd_ctx* ctx, const double* alpha);
int dpmd_zero_velocities(dpmd_ctx* ctx);
int dpmd_compute_ke(dpmd_ctx* ctx);
int dpmd_compute_ke_total(dpmd_ctx* ctx);            /* base_particle.hpp:601-606 */
int dpmd_compute_pe_total(dpmd_ctx* ctx);            /* base_particle.hpp:595-598 */
int dpmd_compute_fpower_total(dpmd_ctx* ctx);        /* disk.cu:941-960            */
int dpmd_compute_temperature(dpmd_ctx* ctx);         /* base_particle.hpp:609-615  */
int dpmd_compute_packing_fraction(dpmd_ctx* ctx);    /* base_particle.hpp:627-647  */
int dpmd_compute_stress_tensor(dpmd_ctx* ctx);       /* disk.cu:702-769            */
int dpmd_compute_pressure(dpmd_ctx* ctx);            /* base_particle.hpp:682-696  */
int dpmd_compute_max_force(dpmd_ctx* ctx);           /* new: max_i |F_i|^2 per system */
/* set_average_velocity (disk.cu:771-811) with per-system targets (device, S each). */
int dpmd_set_average_velocity(dpmd_ctx* ctx, const double* avg_x, const double* avg_y);
/* set_temperature (base_particle.hpp:449-462) */
int dpmd_set_temperature(dpmd_ctx* ctx, const double* temperature_target, int remove_average_velocity);
/* increment_packing_fraction (base_particle.hpp:505-529) */
int dpmd_increment_packing_fraction(dpmd_ctx* ctx, const double* increment);
/* save_state / restore_state (base_particle.hpp:708-736) */
int dpmd_save_state(dpmd_ctx* ctx, const int* flag, int true_val);
int dpmd_restore_state(dpmd_ctx* ctx, const int* flag, int true_val);
/* ---- contact observables ------------------------------------------------------------ */
/* compute_contacts (base_particle.hpp:649-651): disks (disk.cu:434-471): per particle the
 * number of contacts over the Verlet list (DPMD_F_CONTACTS), with the summed overlap and the
 * contact count as a double as compute_overlaps gives them; rigid bodies
 * (rigid_bumpy.cu:759-825,1598-1618): per particle the number of particle neighbors touched
 * by at least one vertex.  Minimum image in every box. */
int dpmd_compute_contacts(dpmd_ctx* ctx);
/* compute_overlaps (base_particle.hpp:698-699): DPMD_F_OVERLAP / DPMD_F_CONTACT_COUNT per
 * particle for disks (disk.cu:663-700) and per vertex, current order, for rigid bodies
 * (rigid_bumpy.cu:929-966). */
int dpmd_compute_overlaps(dpmd_ctx* ctx);
/* compute_n_contacts_total (base_particle.hpp:653-659): contacts, then their per-system sum
 * (DPMD_F_N_CONTACTS_TOTAL). */
int dpmd_compute_n_contacts_total(dpmd_ctx* ctx);
/* compute_pair_dist / compute_pair_forces / compute_hessian (base_particle.hpp:662-702;
 * disk.cu:473-661,976-1060) in one pass.  Outputs are DEVICE arrays of n entries, n = the
 * number of neighbor entries (dpmd_neighbor_info's total), in the CSR order of
 * dpmd_get_neighbors_csr; any group may be NULL.  pair_i, pair_j: canonical particle ids of
 * entry k; dist: |x_i - x_j| (minimum image); fx, fy: the contact force on i from j (0 when
 * not in contact); hessian: NULL or a HOST array of 8 device pointers, the reference's
 * hessian_xx/xy/yx/yy (N_neighbors, 2) columns in the order H_ix_ix, H_ix_jx, H_ix_iy, H_ix_jy,
 * H_iy_ix, H_iy_jx, H_iy_iy, H_iy_jy (0 when not in contact): entry k holds contact k's share
 * of the diagonal block of i and the off-diagonal block (i, j). */
int dpmd_compute_pair_observables(dpmd_ctx* ctx, int* pair_i, int* pair_j, double* dist, double* fx, double* fy,
                                  double* const* hessian, int64_t n);
/* build_particle_neighbors (base_poly_particle.hpp:299-375), rigid bodies: for each particle
 * (canonical order) the particles whose vertices are in its vertices' lists, ascending, no
 * repeats.  start: host [N+1]; ids: host [ids_len >= total]; either may be NULL;
 * total: number of entries. */
int dpmd_particle_neighbors(dpmd_ctx* ctx, int* start, int* ids, int64_t ids_len, int64_t* total);
/* compute_contacts / compute_pair_dist / compute_friction_coeff / compute_pair_forces for rigid
 * bodies (rigid_bumpy.cu:759-928,1261-1327,1598-1676) per particle-neighbor entry (the order of
 * dpmd_particle_neighbors).  DEVICE arrays of n entries, any group NULL: pair_i, pair_j
 * (particle ids); dist: centre distance (minimum image); fx, fy: summed contact force on i's
 * vertices from j's; vertex_contacts_i / _j: the reference's pair_vertex_contacts (i's
 * vertices touching j, j's vertices touching i); friction: |F_t| / |F_n| of that force about
 * the centre line (0 when F_n = 0).  Also fills DPMD_F_CONTACTS. */
/* compute_hessian for rigid bodies (rigid_bumpy.cu:1129-1259,1691-1730) per particle-neighbor
 * entry k = (i, j) (the order of dpmd_particle_neighbors): hessian is a HOST array of 18
 * DEVICE pointers of n entries each; column 2m holds entry k's contribution to the diagonal
 * block H_(i a),(i b) and column 2m + 1 the off-diagonal H_(i a),(j b), m = 3a + b with a, b
 * over (x, y, theta): the reference's hessian_xx, _xy, _xt, _yx, _yy, _yt, _tx, _ty, _tt
 * (N_pairs, 2) fields in that order.  Walls contribute nothing, as in the reference. */
int dpmd_compute_rigid_hessian(dpmd_ctx* ctx, double* const* hessian, int64_t n);
int dpmd_compute_rigid_pair_observables(dpmd_ctx* ctx, int* pair_i, int* pair_j, double* dist, double* fx,
                                        double* fy, int* vertex_contacts_i, int* vertex_contacts_j,
                                        double* friction, int64_t n);
/* compute_damping_forces (base_particle.hpp:386; disk.cu:249-263, rigid_bumpy.cu:310-327):
 * F_i += -v_i * scale[s] (rigid bodies also torque_i += -omega_i * scale[s]). */
int dpmd_compute_damping_forces(dpmd_ctx* ctx, const double* scale);

/* ---- stateless Philox4x32-10 (device_fields.cuh:219-298) ------------------------------ */
int dpmd_rand_uniform_2d(dpmd_ctx* ctx, int field_x, int field_y, uint64_t seed, uint64_t call,
                         double lo_x, double hi_x, double lo_y, double hi_y);
int dpmd_rand_uniform_1d(dpmd_ctx* ctx, int field, uint64_t seed, uint64_t call, double lo, double hi);
int dpmd_rand_normal_2d(dpmd_ctx* ctx, int field_x, int field_y, uint64_t seed, uint64_t call,
                        double mean_x, double sigma_x, double mean_y, double sigma_y);

/* ---- integrators (include/integrators/{velocity_verlet,fire}.cuh) ---------------------- */
/* VelocityVerlet[Wall](P, dt).init() / .step() (velocity_verlet.cuh:17-47). */
int dpmd_vv_init(dpmd_ctx* ctx, const double* dt, int mode);
int dpmd_vv_step(dpmd_ctx* ctx, int n_steps);
/* DampedVelocityVerlet[Wall](P, dt, damping).init() (damped_velocity_verlet.cuh:16-56): dt and
 * damping are device arrays of length S.  init computes the forces without the drag; each
 * later step is the velocity-Verlet step with compute_damping_forces(damping) applied after
 * the forces, at the half-step velocity, before the closing half kick.  Steps are taken with
 * dpmd_vv_step / dpmd_vv_run; dpmd_vv_init switches back to plain velocity Verlet. */
int dpmd_damped_vv_init(dpmd_ctx* ctx, const double* dt, const double* damping, int mode);
/* Rescale-thermostat NVT with affine compression: the step loop of the reference's
 * compression drivers (tests/nvt_rescale_disk_pbc_compress.cu:41-68: vv.step(), then
 * increment_packing_fraction every compression_frequency steps and set_temperature every
 * temperature_frequency steps; tests/damped_dynamics_compress_to_packing_fraction.cu:41-52,
 * the same with set_temperature every step) as one routine over the integrator set up by
 * dpmd_vv_init or dpmd_damped_vv_init.  For global step k = step0 + i, i in [0, n_steps):
 * one step; then increment_packing_fraction(phi_increment) if compress_every > 0 and
 * k % compress_every == 0; then set_temperature(temperature_target, remove_average_velocity)
 * if temperature_every > 0 and k % temperature_every == 0.  phi_increment and
 * temperature_target are device arrays of length S (NULL when the event is off).  Steps
 * between events run device-driven (graph batches when use_graph). */
int dpmd_nvt_rescale_run(dpmd_ctx* ctx, int n_steps, int step0, const double* phi_increment, int compress_every,
                         const double* temperature_target, int temperature_every, int remove_average_velocity,
                         int use_graph);
/* FIRE[Wall](P, dt).init() / .step() / .converged() / .get_dt() (fire.cuh:44-109). */
int dpmd_fire_init(dpmd_ctx* ctx, const double* dt, int mode);
int dpmd_fire_step(dpmd_ctx* ctx);
int dpmd_fire_converged(dpmd_ctx* ctx, double avg_pe_diff_target, double avg_pe_target, int* all_converged);
int dpmd_fire_converged_max_force(dpmd_ctx* ctx, double force_tol, int* all_converged);
int dpmd_fire_get_dt(dpmd_ctx* ctx, double* dt_out);
/* FIRE scalar state per system (host copies, length S each; any pointer may be NULL). */
int dpmd_fire_state(dpmd_ctx* ctx, double* dt, double* alpha, int* n_good, int* n_bad);

/* ---- fused, device-driven runs ------------------------------------------------------ */
/* minimize_fire[_wall] (minimizers.cuh:10-121) run on the device: every step's stopping
 * test is evaluated per system on the GPU, stopped systems are masked, and the host only
 * polls every poll_interval steps.  With use_graph the step is captured once into a
 * hipGraph and replayed.  predicate: dpmd_fire_predicate; (a, b) = (avg_pe_diff_target,
 * avg_pe_target) for REFERENCE, (force_tol, unused) for MAX_FORCE. */
int dpmd_fire_minimize(dpmd_ctx* ctx, const double* dt, int max_steps, int predicate, double a,
                       double b, int mode, int use_graph, int poll_interval, int* steps_taken,
                       int* all_converged);
/* Pressure-target jamming by compression, per system.  New driver with no single reference
 * function: the reference's jammer bisects on <pe> (jam_binary_search_pbc/_wall,
 * jammers.cuh:60-256); this one is built from the same primitives, minimize_fire
 * (minimizers.cuh:10-31) and increment_packing_fraction (base_particle.hpp:505-529).
 * Each round: dpmd_fire_minimize over the systems still compressing (the others get dt 0,
 * so they are masked and cost no work); P_s from dpmd_compute_pressure; a system with
 * P_s >= p_target is jammed; the others grow phi_s by dphi and go on, up to
 * max_compressions times.  jam_state (host, length S, may be NULL):
 *   1 jammed; 0 never run (dt[s] == 0) or still below p_target after max_compressions;
 *  -1 the initial minimisation (before any compression) did not stop within fire_max_steps
 *     (dropped, as jammers.cu:7-17 drops systems whose initial minimisation failed).  After a
 *     compression the state a minimisation reached decides, whether it stopped or not, as
 *     the reference's jammer decides after each minimisation (jammers.cuh:100-110).
 * n_compressions (host, S, may be NULL): compressions applied per system.  work (host, may
 * be NULL): sum over rounds and systems of N_s x (FIRE steps that did work).
 * DPMD_F_FIRE_STEPS then holds each system's FIRE steps summed over all rounds. */
int dpmd_jam_compress(dpmd_ctx* ctx, const double* dt, double dphi, double p_target, int max_compressions,
                      int fire_max_steps, int predicate, double a, double b, int mode, int use_graph,
                      int* jam_state, int* n_compressions, int64_t* work);
/* jam_binary_search_pbc / jam_binary_search_wall (jammers.cuh:60-256, jammers.cu:7-92) for
 * disk batches; mode DPMD_FORCE_PBC / _WALL picks the variant (DPMD_FORCE_PER_SYSTEM uses the
 * per-system boundary flag with the PBC variant's steps).  Per system: minimise with
 * minimize_fire_copy_dt (the reference's stopping test and argument order), drop systems
 * whose first minimisation did not stop, then bisect phi between the last unjammed state
 * (<pe> <= avg_pe_target, saved) and the first jammed one (reverted to the saved state),
 * growing phi by phi_increment until a jammed state is seen, until |phi_high / phi_low - 1| <
 * phi_tolerance or max_compression_steps rounds.  The box is resized and the lists rebuilt
 * before the positions are scaled, as in the reference.  dt: device [S].  action (host [S],
 * may be NULL): 0 done or dropped, -1 last state unjammed, 1 last state jammed (reverted).
 * phi (host [S], may be NULL): the bisection's packing fraction.  compression_steps (host,
 * may be NULL): rounds run.  Ends with the forces, pe totals and packing fractions computed. */
int dpmd_jam_binary_search(dpmd_ctx* ctx, const double* dt, int max_compression_steps, int max_minimization_steps,
                           double avg_pe_target, double avg_pe_diff_target, double phi_increment,
                           double phi_tolerance, int mode, int use_graph, int* action, double* phi,
                           int* compression_steps);
/* ---- adjoints (disks) --------------------------------------------------------------------
 * New; the reference has no derivatives of its dynamics.  Its Hessian observable
 * (compute_hessian, src/particles/disk.cu:565-661) fixes the pair block used here:
 * H_ij = k n n^T - (k delta / d)(I - n n^T) for each contact (delta = r_i + r_j - d > 0), plus
 * k on the diagonal per wall contact.  U = sum k delta^2 / 2 (+ walls); F = -dU/dx;
 * P = (sum m v^2 + sum_pairs k delta d) / (2 Lx Ly) (compute_pressure, T14 definition).
 * Box derivatives are taken at fixed reduced coordinates x / L (walls move with the box).
 * Particle vectors are canonical-order DEVICE arrays of length N; per-system arrays are
 * device arrays of length S.  mode: dpmd_force_mode. */
/* y = H w. */
int dpmd_adjoint_hvp(dpmd_ctx* ctx, int mode, const double* wx, const double* wy, double* yx, double* yy);
/* y = H w, the same bits as dpmd_adjoint_hvp, calling no allocator (no hipMalloc / hipFree,
 * no stream synchronisation: it only enqueues on the context's stream, so it can be called
 * under HIP graph capture and in tight solver loops).  No counterpart in the reference: for
 * callers that cannot allocate.  Its current-order work vectors live in a library-owned
 * device array of 4 x 2^20 doubles (32 MiB per GPU, held until the process ends); systems of
 * more than 2^20 particles in total are rejected with DPMD_ERR_ARGUMENT.  Calls that overlap
 * on one GPU share that array and race. */
int dpmd_adjoint_hvp_noalloc(dpmd_ctx* ctx, int mode, const double* wx, const double* wy, double* yx, double* yy);
/* Vector-Jacobian product of (U_s, P_s, F) at the current state: given cotangents gU[S],
 * gP[S], gF[N] (each may be NULL = 0), returns d/dx, d/dr and d/dL of
 * sum_s gU_s U_s + gP_s P_s + sum_i gF_i . F_i (any output may be NULL). */
int dpmd_adjoint_vjp(dpmd_ctx* ctx, int mode, const double* gU, const double* gP, const double* gfx,
                     const double* gfy, double* dx, double* dy, double* dr, double* dLx, double* dLy);
/* Solves H x = b per system by MINRES to |b - Hx| <= tol |b| over the particles with a
 * contact (pair or wall): b and x are projected off H's null directions (the coordinates of
 * particles without contacts, and uniform translations in periodic systems).  rel_resid: max
 * over systems of the true relative residual. */
int dpmd_adjoint_solve(dpmd_ctx* ctx, int mode, const double* bx, const double* by, double* xx, double* xy,
                       double tol, int max_iter, int* iters, double* rel_resid);
/* Implicit derivative of each system's pressure through its energy minimum (call at a
 * FIRE / Newton minimum): dP_s/dr_i and dP_s/dL_s = dP/dq + lambda^T dF/dq with
 * H lambda = dP/dx solved to tol as dpmd_adjoint_solve does.  dr: [N]; dLx, dLy: [S]; any
 * may be NULL. */
int dpmd_adjoint_pressure(dpmd_ctx* ctx, int mode, double tol, int max_iter, double* dr, double* dLx,
                          double* dLy, int* iters, double* rel_resid);
/* Levenberg-Marquardt Newton iterations x <- x + (H + 1e-10 k I)^-1 F from near a minimum
 * until max|F| <= ftol (polishes FIRE minima to rounding level).  max_force: final max|F_i|. */
int dpmd_newton_minimize(dpmd_ctx* ctx, int mode, int max_newton, double ftol, double cg_tol, int cg_max_iter,
                         double* max_force);
/* n_steps of velocity Verlet (as dpmd_vv_run) keeping canonical (x, y, vx, vy) checkpoints
 * every `every` steps, then reverse mode: cotangents of the final x, y, vx, vy (NULL = 0)
 * -> cotangents of the initial ones.  Each segment is replayed once from its checkpoint;
 * one Hessian-vector product per reverse step.  The context is left at the final state. */
int dpmd_vv_run_checkpointed(dpmd_ctx* ctx, int n_steps, int every, int use_graph);
int dpmd_vv_backprop(dpmd_ctx* ctx, const double* gx, const double* gy, const double* gvx, const double* gvy,
                     double* gx0, double* gy0, double* gvx0, double* gvy0, int use_graph);
/* n_steps of VelocityVerlet after dpmd_vv_init, device-driven (graph when use_graph). */
int dpmd_vv_run(dpmd_ctx* ctx, int n_steps, int use_graph);

/* ---- memory helpers so C++ callers need no HIP headers ------------------------------- */
int dpmd_malloc(void** device_ptr, int64_t bytes);
int dpmd_free(void* device_ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int dpmd_memcpy(void* dst, const void* src, int64_t bytes, int kind);
int dpmd_fill_f64(void* device_ptr, int64_t n, double value);
int dpmd_fill_i32(void* device_ptr, int64_t n, int32_t value);
int dpmd_synchronize(dpmd_ctx* ctx);
/* ---- instrumentation (no reference counterpart: the reference times with std::chrono,
 * tests/test_disk_system_cell_parallelism.cu:92-102) -------------------------------------
 * dpmd_profile(ctx, 1) starts recording a HIP event pair around every force-kernel launch
 * issued outside graph capture; dpmd_profile(ctx, 2) records one pair around the consecutive
 * fused velocity-Verlet launches of every directly launched batch that no rebuild halted;
 * dpmd_profile(ctx, 3) records one pair around the consecutive fused FIRE steps (fused and
 * settle launches; runs of one disk system, DPMD_OPT_FUSED_FIRE) of every directly launched
 * batch that no rebuild halted; dpmd_profile(ctx, 4) records one pair around the rebuild of
 * cells and lists that finishes every halted step batch (the in-step rebuild, whichever
 * engine runs it); 0 stops.  dpmd_profile_read returns the summed duration and the number of
 * launches (kind 4: rebuilds) covered. */
int dpmd_profile(dpmd_ctx* ctx, int enable);
int dpmd_profile_read(dpmd_ctx* ctx, double* total_ms, int* count);
/* The recorded event pairs one by one (ms, in order, up to capacity); *count = the number of
 * pairs.  Clears the record like dpmd_profile_read. */
int dpmd_profile_durations(dpmd_ctx* ctx, float* ms, int capacity, int* count);
/* A plain device copy kernel of n doubles (dst <- src), timed over `repeats` launches on the
 * default stream: the measured copy ceiling, and the calibration kernel for HBM counters. */
int dpmd_stream_copy(double* dst, const double* src, int64_t n, int repeats, double* ms_per_copy);

/* Execution options (no reference counterpart).
 *   DPMD_OPT_COOP_REBUILD, DPMD_OPT_COOP_IN_GRAPH: accept only 0 (a former cooperative
 *     single-launch rebuild was removed; setting 1 returns DPMD_ERR_ARGUMENT).
 *   DPMD_OPT_FUSED_VV: 1 (default) = device-driven velocity-Verlet runs of disks compute a
 *     step's forces and the next step's kick and drift in one launch (positions ping-pong
 *     between two buffers); 0 = one drift and one force launch per step.  Same results.
 *   DPMD_OPT_BLOCK: workgroup size of the disk step kernels (velocity-Verlet drift, force and
 *     fused launches, FIRE drift): 64, 128, 256 (default) or 512.  Same results.
 *   DPMD_OPT_MAX_BATCH: the largest number of steps captured into one hipGraph batch of a
 *     device-driven run, a power of two from 1 to 128 (default 128).  Same results.
 *   DPMD_OPT_PERSISTENT_FIRE: engine of device-driven FIRE runs of disk batches whose systems
 *     all have at most 16384 particles.  1 (default): persistent phases, one workgroup per
 *     system for DPMD_OPT_PERSIST_PHASE steps with no per-step launches, handing the systems
 *     still running to per-step launches once those are the cheaper engine by a cost model
 *     (few or large systems left); 2: persistent phases only; 0: graph-captured batches of
 *     per-step launches only.  Same results.
 *   DPMD_OPT_PERSIST_BLOCK: workgroup size of the persistent FIRE kernel: 256, 512 (default)
 *     or 1024.  Same results.
 *   DPMD_OPT_PERSIST_PHASE: steps per persistent phase (default 2000).  Same results.
 *   DPMD_OPT_DEVICE_REBUILD: how step batches of device-driven runs meet a rebuild.  1: the
 *     batch halts and the host finishes the step (cheapest when rebuilds are rare); 2: every
 *     step runs the rebuild chain on the device, gated by the rebuild flags, and batches never
 *     halt; 0 (default): 1 while rebuilds are rarer than one per 2 steps, 2 above (back below
 *     0.35 per step); for FIRE runs of one disk system with DPMD_OPT_FUSED_FIRE, whose halting
 *     batches take the fused step, the switch is at 0.8 (back below 0.6).  Same results.
 *   DPMD_OPT_PERSIST_LDS: 1 (default) runs persistent FIRE launches whose systems all have at
 *     most 1536 particles with each system's particle state held in LDS (one 512-thread
 *     workgroup per system); 0 keeps it in global memory.  Same results.
 *   DPMD_OPT_REBUILD_ENGINE: how cells and lists are rebuilt.  1: one workgroup per system,
 *     whose flagged systems run their whole rebuild in LDS (systems of up to about 8000
 *     objects); 2: the device-wide chain of launches over all objects with the tiled
 *     neighbor build; 0 (default): 1 when every system fits, else 2.  Same results.
 *   DPMD_OPT_JAM_RETRIES: dpmd_jam_compress runs a round whose minimisation hit the FIRE step
 *     cap again (FIRE restarted from where it stopped, no compression) up to this many times
 *     per system before the system is dropped (initial minimisation) or decided on the state
 *     reached (later rounds); default 0.
 *   DPMD_OPT_FUSED_FIRE: 1 (default) = device-driven FIRE runs of one disk system compute a
 *     step's forces, kick and power partials and the next step's drift in one launch, the
 *     drift taken with the scalars that F.v > 0 gives and redone by the fold's launch when
 *     the power shows otherwise (positions and velocities ping-pong between two buffer
 *     pairs); 0 = drift, force and fold launches per step.  Same results. */
enum dpmd_option {
    DPMD_OPT_COOP_REBUILD = 0, DPMD_OPT_COOP_IN_GRAPH = 1, DPMD_OPT_FUSED_VV = 2, DPMD_OPT_BLOCK = 3,
    DPMD_OPT_MAX_BATCH = 4, DPMD_OPT_PERSISTENT_FIRE = 5, DPMD_OPT_PERSIST_BLOCK = 6, DPMD_OPT_PERSIST_PHASE = 7,
    DPMD_OPT_DEVICE_REBUILD = 8, DPMD_OPT_PERSIST_LDS = 9, DPMD_OPT_REBUILD_ENGINE = 10, DPMD_OPT_JAM_RETRIES = 11,
    DPMD_OPT_FUSED_FIRE = 12
};
int dpmd_set_option(dpmd_ctx* ctx, int key, int value);
int dpmd_get_option(dpmd_ctx* ctx, int key, int* value);

/* ---- autotuner (no reference counterpart; the reference's cell size, skin and launch
 * shapes are fixed by the caller) -------------------------------------------------------
 * A configuration: for every system s the neighbor cutoff (Verlet skin) is
 * skin * 2 * r_min,s and the rebuild threshold is (cutoff / 2)^2; the cell grid is the
 * largest whose cells are at least cell_factor * (2 r_max,s + cutoff) wide; block, max_batch
 * and fused_vv are the DPMD_OPT_BLOCK / _MAX_BATCH / _FUSED_VV options.  None of these
 * changes the physics: the neighbor lists always hold every pair within the interaction
 * range, so trajectories differ only by the summation order of forces. */
typedef struct dpmd_tune_config {
    double skin;
    double cell_factor;
    int block;
    int max_batch;
    int fused_vv;
} dpmd_tune_config;
typedef struct dpmd_tune_trial {
    dpmd_tune_config config;
    double rate;      /* particle-steps per second */
    double seconds;   /* timed seconds of the trial */
} dpmd_tune_trial;
enum dpmd_tune_integrator { DPMD_TUNE_FIRE = 0, DPMD_TUNE_NVE = 1 };
/* The cell grid of one box for a configuration: nx = floor(Lx / (cell_factor * (2 r_max +
 * cutoff))), likewise ny.  Fails (DPMD_ERR_ARGUMENT) unless both are >= 3: the periodic 3x3
 * stencil visits a cell twice on a smaller grid, and cells narrower than 2 r_max + cutoff
 * miss pairs (reference precondition, base_point_particle_kernels.cu:72-94, unchecked
 * there).  Needs no context and no GPU. */
int dpmd_cell_grid(double Lx, double Ly, double r_max, double cutoff, double cell_factor, int* nx, int* ny);
/* The context's configuration (skin: largest cutoff / (2 r_min) over systems; cell_factor:
 * smallest cell side / (2 r_max + cutoff)). */
int dpmd_tune_get(dpmd_ctx* ctx, dpmd_tune_config* config);
/* Applies a configuration (disks, cell lists, after init_neighbors) and rebuilds. */
int dpmd_tune_apply(dpmd_ctx* ctx, const dpmd_tune_config* config);
/* The workload signature that keys the cache: kind, systems, particles, size range, boxes,
 * packing fraction, size ratio, integrator, force mode, device and library version. */
int dpmd_tune_signature(dpmd_ctx* ctx, int integrator, int mode, char* buf, int capacity);
/* Picks the configuration with the most particle-steps per second for `integrator` (FIRE
 * without stopping, or velocity Verlet) with per-system dt (device, S; 0 = masked) and force
 * mode, each trial timing `steps` device-driven, graph-captured steps from the current
 * state after a short warm-up, one knob at a time (skin, cell_factor, block, fused_vv for
 * NVE, max_batch).  Positions and velocities are restored after every trial, and the best
 * configuration is applied at the end; integrator state must be initialised again
 * afterwards.  With cache_path (JSON, may be NULL) a stored entry for this signature is
 * applied without trials (*n_trials = 0), and a new result is stored.  trials (host, may be
 * NULL): up to trial_capacity records of the trials run. */
int dpmd_autotune(dpmd_ctx* ctx, int integrator, const double* dt, int mode, int steps, const char* cache_path,
                  dpmd_tune_config* best, double* best_rate, dpmd_tune_trial* trials, int trial_capacity,
                  int* n_trials);
/* The JSON cache on its own (no context, no GPU): look up / store the entry for a signature. */
int dpmd_tune_cache_load(const char* path, const char* signature, dpmd_tune_config* config, double* rate,
                         int* found);
int dpmd_tune_cache_store(const char* path, const char* signature, const dpmd_tune_config* config, double rate);

/* Library version string. */
const char* dpmd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DPMD_DPMD_H */
