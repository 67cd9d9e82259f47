// Context layout and launcher declarations shared by the library's translation units.
#pragma once
#include "common.hpp"
#include "../../include/dpmd/dpmd.h"

namespace dpmd {

// Object view for the cell / neighbor pipeline.  Objects are disks, or the vertices of
// rigid bodies (then pid gives the owning particle, whose other vertices are excluded;
// reference: base_poly_particle_kernels.cu:78-195).
struct ObjArgs {
    const double* x; const double* y; const double* rad;
    const int* sid;      // object -> system
    const int* off;      // per-system object offsets (S+1)
    const int* pid;      // object -> particle (nullptr for disks)
    int n;
};

struct CellArgs {
    const int* cdx; const int* cdy; const int* cstart;        // S, S, S+1
    const double* Lx; const double* Ly; const double* iLx; const double* iLy;
    const double* cutoff;
    int* cell_id; int* cell_count; int* cell_start; int* cell_wptr; const int* cell_sys;
    int C;
    unsigned* flag;
};

// Arrays permuted by a rebuild (gathered by the new order, then copied back).
struct PermArgs {
    double* d[12]; int nd;
    int* i[4]; int ni;
    double* tmp_d; int* tmp_i;
};

struct NbrArgs {
    int* nn; int* nbr; int K;      // ELL, column-major: nbr[k * n + i]
    int* overflow;
};

// Run-control words (int array "run" of the context):
//   [1] max steps  [2] systems still running  [3] step active (device-conditional path)
//   [4] halted     [5] index of the halted step in its batch
//   [6] index of the step in its batch at which the last system stopped
//   [7] "some rebuild flag is set" (raised by the drift kernels, cleared by the rebuild)
//   [8] base slot for recorded positions (reverse-mode replays)
//   [9] second "some rebuild flag is set" word: fused velocity-Verlet launches raise the
//       flags of the positions they write into the word of the next step's parity, so a
//       launch never reads a word that its own workgroups may be setting
//  [10] gate of a device-side rebuild inside a step batch (a copy of [7], so clearing [7]
//       cannot race with the rebuild's own gate reads)
//  [11] number of device-side rebuilds run (batches that rebuild on the device)
//  [12] fused FIRE batches: k + 1 when step k's guessed next-step scalars were wrong, -(k + 1)
//       when the system stopped at step k, else 0 (written by the step's settle launch, read
//       by the next launch's gate)
//  [13] fused FIRE batches: rebuild word of positions from a redone (wrongly guessed) drift
enum { RUN_MAX = 1, RUN_RUNNING = 2, RUN_ACTIVE = 3, RUN_HALT = 4, RUN_HALT_AT = 5, RUN_CONV_AT = 6, RUN_ANY = 7,
       RUN_REC = 8, RUN_ANY2 = 9, RUN_REBUILDS = 11, RUN_MIS = 12, RUN_FIXW = 13, RUN_WORDS = 14 };

// Launch gate.  Plain form (run == nullptr): the kernel works iff *active != 0.  Batch form
// (lean fused runs): a step works while nothing halted and some system runs; the force
// kernel halts the batch instead of computing with a list that needs a rebuild.  Every
// workgroup of a launch reads the same words, so a whole launch takes one path.
struct Gate {
    const int* active;
    int* run;
    int k;
    int any = RUN_ANY;   // the run word a batch force kernel checks for rebuild flags
};

__device__ __forceinline__ bool gate_open(const Gate& g) {
    if (!g.run) return *g.active != 0;
    return g.run[RUN_HALT] == 0 && g.run[RUN_RUNNING] > 0;
}

__device__ __forceinline__ bool gate_force(const Gate& g) {
    if (!g.run) return *g.active != 0;
    if (g.run[RUN_HALT] != 0 || g.run[RUN_RUNNING] <= 0) return false;
    if (g.run[g.any] != 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { g.run[RUN_HALT_AT] = g.k; g.run[RUN_HALT] = 1; }
        return false;
    }
    return true;
}

// Raise rebuild flag s (and the any-flag word) for every lane with f set: one atomic per
// wave when the whole wave belongs to one system, else one per flagged lane.  Every lane of
// the wave must call (invalid lanes pass f = false, s = -1).
__device__ __forceinline__ void raise_flags(unsigned* flag, int* any, int s, bool f) {
    const unsigned long long m = __ballot(f);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int s0 = __shfl(s, 0, 64), s63 = __shfl(s, 63, 64);
    if (s0 == s63 && s0 >= 0) {
        if (lane == __ffsll((unsigned long long)m) - 1) { atomicOr(&flag[s], 1u); atomicOr(any, 1); }
    } else if (f) {
        atomicOr(&flag[s], 1u);
        atomicOr(any, 1);
    }
}

}  // namespace dpmd

struct dpmd_ctx {
    int kind = DPMD_DISK;
    hipStream_t st = nullptr;
    std::string err;
    int S = 0, N = 0, Nv = 0;
    int method = DPMD_NEIGHBOR_CELL;
    bool synced = false, nbr_ready = false;
    int K = 32;                 // neighbor capacity per object
    int C = 0;                  // total cells

    // ---- per system
    dpmd::DVec<int> off, size, cdx, cdy, cstart, boundary, ndof, voff;
    dpmd::DVec<unsigned> flag;
    dpmd::DVec<double> Lx, Ly, iLx, iLy, e, cutoff, thresh2;
    dpmd::DVec<double> pe_tot, ke_tot, pw_tot, press, temp, phi, maxf2;
    dpmd::DVec<double> st_xx, st_xy, st_yx, st_yy;
    std::vector<int> h_off, h_voff;
    std::vector<int> layout_sig;

    // ---- per particle
    dpmd::DVec<double> x, y, vx, vy, fx, fy, pe, ke, rad, mass, area, lx, ly;
    dpmd::DVec<double> angle, avel, torque, inertia;
    dpmd::DVec<double> sxx, sxy, syx, syy;
    dpmd::DVec<int> sid, static_index;

    // ---- per vertex (rigid bumpy)
    dpmd::DVec<double> vpx, vpy, vfx, vfy, vpe, vrad, vmass;
    dpmd::DVec<int> vpid, poff, nvpp, vsid;

    // ---- cells and neighbors (over objects: particles for disks, vertices for rigid)
    dpmd::DVec<int> cell_id, cell_count, cell_start, cell_wptr, cell_sys, order, order_inv;
    dpmd::DVec<int> tmp_i; dpmd::DVec<double> tmp_d;
    dpmd::DVec<int> nn, nbr, overflow;
    dpmd::DVec<int> scan_part;
    dpmd::DVec<int> tiles;      // 4 ints per cell-row tile (system, row, first cell, end cell)
    int ntiles = 0;
    dpmd::DVec<int> sys_tile;   // S+1 tile offsets per system
    dpmd::DVec<int> sys_list, sys_list_n;   // tiles of the systems the per-system engine rebuilt, and their count
    int max_sys_obj = 0, max_sys_cells = 0;   // largest system (objects, cells)
    // rebuild engine: 0 per-system workgroups when every system fits, else the device-wide
    // chain; 1 per-system workgroups; 2 device-wide chain (DPMD_OPT_REBUILD_ENGINE)
    int rebuild_engine = 0;
    // dpmd_jam_compress: times a round that hit its FIRE step cap is run again before the
    // system is dropped (DPMD_OPT_JAM_RETRIES)
    int jam_retries = 0;

    // ---- system-aligned 256-wide chunks over particles (deterministic reductions)
    int nchunks = 0;
    dpmd::DVec<int> chunk_sys, chunk_beg, sys_chunk;   // sys_chunk: S+1
    dpmd::DVec<double> part;                            // 4 * nchunks

    // ---- FIRE (fire.cuh:116-147): committed and next-step ("shadow") scalars
    dpmd::DVec<double> f_dt, f_dtmax, f_dtmin, f_alpha, f_peprev;
    dpmd::DVec<int> f_ngood, f_nbad, f_stopped;
    // frozen[s] = 1: system s stopped and its velocities were zeroed by the following step,
    // so further FIRE steps leave its state bit-identical; fused runs skip its work.
    dpmd::DVec<int> f_frozen;
    dpmd::DVec<int> f_nsteps;   // fused steps that did work, per system
    bool mask_frozen = false;   // honoured only inside fused minimisation runs
    dpmd::DVec<double> s_dt, s_alpha, s_dtrev, s_vscale;
    dpmd::DVec<int> s_ngood, s_nbad;
    int fire_mode = 0; bool fire_ready = false;
    // ---- VV
    dpmd::DVec<double> vv_dt, vv_dth;
    int vv_mode = 0; bool vv_ready = false;
    // damped velocity Verlet: per-system drag coefficients, applied when vv_damped is set
    dpmd::DVec<double> vv_damp;
    bool vv_damped = false;
    // ---- contact observables (compute_contacts / _overlaps / pair observables)
    dpmd::DVec<int> contacts, ncontacts, pair_start, canon;
    dpmd::DVec<double> overlap, ccount;   // per particle (disks) or per vertex (rigid bodies)
    // rigid bodies: particle-level neighbor lists (build_particle_neighbors)
    dpmd::DVec<int> pn_start, pn_count, pn_ids, pn_vc, pn_flag;
    int64_t pn_total = 0;
    // ---- run control: [0] steps_done [1] max_steps [2] n_running [3] active
    dpmd::DVec<int> run, one;
    // ---- saved state (save_state/restore_state)
    dpmd::DVec<double> sv_x, sv_y, sv_rad, sv_mass, sv_Lx, sv_Ly, sv_angle, sv_vx, sv_vy;
    dpmd::DVec<int> sv_static;
    dpmd::DVec<double> sv_vpx, sv_vpy, sv_vrad, sv_vmass;   // rigid bodies: vertex state
    dpmd::DVec<int> sv_vpid, sv_vstatic;
    bool saved = false;
    // scratch per-system
    dpmd::DVec<double> scr_a, scr_b, scr_c;
    // graphs
    hipGraphExec_t g_fire = nullptr, g_vv = nullptr;
    // Step batches kept across the device-driven runs of one routine call (GraphScope in
    // abi.hip); key = everything the captured launches take from the context.
    bool gcache_on = false;
    std::vector<int64_t> gcache_key;
    hipGraphExec_t gcache[16] = {};   // [0, 8): batches that halt on rebuilds, [8, 16): device-side rebuilds
    // step batches rebuild on the device instead of halting: 0 by the measured rebuild rate,
    // 1 never, 2 always (DPMD_OPT_DEVICE_REBUILD)
    int dev_rebuild = 0;
    int g_fire_key[4] = {0, 0, 0, 0};
    double g_fire_keyd[2] = {0, 0};
    int g_vv_key = -1;
    // adjoints: scratch, CG vectors (current order), NVE checkpoints (canonical order)
    dpmd::DVec<double> adj_bx, adj_by, adj_sa, adj_sb, adj_sc, adj_sd;
    dpmd::DVec<double> adj_t0, adj_t1, adj_t2, adj_t3, adj_t4, adj_t5, adj_t6;
    dpmd::DVec<double> adj_u0, adj_u1, adj_u2, adj_u3, adj_u4;   // per-system scratch
    dpmd::DVec<double> cg_rx, cg_ry, cg_px, cg_py, cg_qx, cg_qy;
    dpmd::DVec<double> cg_rr, cg_rrn, cg_pq, cg_bb, cg_beta;
    dpmd::DVec<int> cg_active;
    dpmd::DVec<double> mr_vec, mr_s;   // MINRES work vectors (6 x 2N) and per-system scalars
    dpmd::DVec<double> adj_live, adj_live2;   // contact mask (and its double buffer)
    dpmd::DVec<int> adj_flag;
    dpmd::DVec<double> ck_state;     // n_ck x 4N: x, y, vx, vy (canonical) at every checkpoint
    dpmd::DVec<double> ck_seg;       // (every + 1) x 2N positions of one replayed segment
    dpmd::DVec<double> ck_cur;       // the same slots of a reverse window, in the window's current order
    dpmd::DVec<double> adj_cur;      // 8N: x-bar, v-bar and the two w buffers of a reverse window
    dpmd::DVec<int> ck_nn, ck_nbr;   // canonical-order contact list for a segment
    int ck_steps = 0, ck_every = 0, ck_count = 0;
    double* rec_seg = nullptr;       // when set, lean VV batches record positions per step
    // fused velocity Verlet (disks): second position buffer and the switch
    dpmd::DVec<double> xy0, xy1;   // interleaved (x, y) ping-pong positions of fused VV batches (2N each)
    bool fuse_vv = true;
    // fused FIRE batches (runs of one disk system): interleaved (vx, vy) ping-pong velocities
    // (2N each; positions use xy0 / xy1), the snapshot of the FIRE scalars a step started from
    // (dt, alpha, previous <pe>; N_good, N_bad), and the switch
    dpmd::DVec<double> vxy0, vxy1, f_snap;
    dpmd::DVec<int> i_snap;
    bool fuse_fire = true;
    // workgroup size of the disk step kernels (64, 128, 256, 512) and the largest number of
    // steps captured into one hipGraph batch (a power of two, 1 .. 128)
    int block = DPMD_BLOCK;
    int max_batch = 128;
    // persistent per-system FIRE (disk batches of systems up to kPersistMaxN particles):
    // the switch, its workgroup size, and the per-run work lists
    int persist_fire = 1;        // 0 off, 1 with hand-off to per-step launches, 2 persistent only
    int persist_block = 512;
    // FIRE step launches restricted to the systems still running (set while a persistent run
    // hands its last systems to the per-step launches; 0 = launches over every chunk / system)
    dpmd::DVec<int> act_chunk, act_sys;
    int n_act_chunk = 0, n_act_sys = 0;
    // the one system of fused FIRE batches in a context of several (the last system running
    // in such launches; -1: none) and its first chunk
    int fsys = -1, fsys_ch0 = 0;
    int persist_lds = 1;         // systems of up to 1536 particles run with their state in LDS
    int persist_phase = 2000;    // steps per persistent phase
    dpmd::DVec<int> persist_order, persist_todo, persist_stop;
    // kernel timing (force kernels; outside graph capture only)
    int prof = 0;   // 0 off, 1 force kernels, 2 fused velocity-Verlet kernels
    bool capturing = false;
    std::vector<hipEvent_t> prof_ev;   // (begin, end) pairs
    std::vector<int> prof_n;           // launches covered by each pair
    hipEvent_t pend_b = nullptr, pend_e = nullptr;   // pair of the batch in flight (kind 2)
    int pend_n = 0;

    int nv_max = 0;   // rigid bodies: most vertices of one particle (set by sync)
    int n_obj() const { return kind == DPMD_RIGID_BUMPY ? Nv : N; }
};

namespace dpmd {

// Error plumbing of the C ABI: every entry point runs its body under guard(), which maps the
// exception types to dpmd_status and keeps the message on the context (or globally when
// there is no context yet).
std::string& global_error();
template <class F>
int guard(dpmd_ctx* c, F&& f) {
    try {
        f();
        return DPMD_OK;
    } catch (const ArgError& e) {
        (c ? c->err : global_error()) = e.what();
        return DPMD_ERR_ARGUMENT;
    } catch (const StateError& e) {
        (c ? c->err : global_error()) = e.what();
        return DPMD_ERR_STATE;
    } catch (const OverflowError& e) {
        (c ? c->err : global_error()) = e.what();
        return DPMD_ERR_OVERFLOW;
    } catch (const std::exception& e) {
        (c ? c->err : global_error()) = e.what();
        return DPMD_ERR_HIP;
    }
}
// Drag coefficients of the velocity-Verlet force kernels: null for plain velocity Verlet.
inline const double* vv_damping(const dpmd_ctx* c) { return c->vv_damped ? c->vv_damp.p : nullptr; }
inline void need(bool ok, const std::string& msg) { if (!ok) throw ArgError(msg); }
inline void need_state(bool ok, const std::string& msg) { if (!ok) throw StateError(msg); }

// Host copy of a device vector (synchronous on the context's stream).
template <class T>
std::vector<T> dl(dpmd_ctx* c, const dpmd::DVec<T>& v) {
    std::vector<T> h(v.n);
    if (v.n) {
        HIP_OK(hipMemcpyAsync(h.data(), v.p, v.n * sizeof(T), hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
    }
    return h;
}

// fire_persist.hip: device-driven FIRE with one workgroup per system (same results as the
// lean batches of lean_run, for disk batches whose systems all have <= kPersistMaxN particles)
bool fire_persist_ok(dpmd_ctx* c);
int fire_persistent_run(dpmd_ctx* c, int max_steps, int pred, double a, double b, bool use_graph, int* conv);
// abi.hip: device-driven runs as graph-captured batches of per-step launches
int lean_run_batches(dpmd_ctx* c, int max_steps, bool fire, int pred, double a, double b, bool use_graph, int* conv);

// abi.hip: the autotuner's engine (tune.hip holds the search and the cache)
void tune_get(dpmd_ctx* c, dpmd_tune_config* t);
void tune_apply(dpmd_ctx* c, const dpmd_tune_config& t);
void tune_snapshot(dpmd_ctx* c, double* dst);                   // canonical x, y, vx, vy (4N)
void tune_restore(dpmd_ctx* c, const double* src, int mode);    // and recompute forces
// Puts every particle back at the current position static_index0[k] had (canonical k), so
// later rebuilds break in-cell ties as they would have without the trials in between.
void tune_restore_order(dpmd_ctx* c, const std::vector<int>& static_index0);
// Runs `steps` steps of FIRE (no stopping: max |F| < 0) or velocity Verlet from the current
// state through the device-driven path; returns particle-steps done (steps x running N).
int64_t tune_run(dpmd_ctx* c, int integrator, const double* dt, int mode, int steps);

// cells.hip
void build_chunks(dpmd_ctx* c);
// in_step: the rebuild runs between a step's drift and its force evaluation, so forces and
// energies are dead and are not permuted
void rebuild_neighbors(dpmd_ctx* c, const int* active, bool force_all, bool in_step = false);
// cell-row tiles of the tiled neighbor build and the per-system extents (after setup_cells)
void build_tiles(dpmd_ctx* c, const std::vector<int>& cdx, const std::vector<int>& cdy);
// inside a step batch: rebuild the flagged systems iff some drift raised a flag this step
void rebuild_on_device(dpmd_ctx* c);
void permute_particles(dpmd_ctx* c, const int* order_host);
void naive_neighbors(dpmd_ctx* c);
void check_overflow(dpmd_ctx* c);

// adjoint.hip (disks; particle vectors in current order unless named canonical)
void adj_to_current(dpmd_ctx* c, const double* canon, double* cur);
// `rows` consecutive N-vectors from canonical into current order, and back
void adj_rows_to_current(dpmd_ctx* c, const double* canon, double* cur, int rows);
void adj_rows_to_canonical(dpmd_ctx* c, const double* cur, double* canon, int rows);
void adj_to_canonical(dpmd_ctx* c, const double* cur, double* canon);
void adj_hvp_current(dpmd_ctx* c, int mode, const double* wx, const double* wy, double* yx, double* yy,
                     const double* live = nullptr, double shift = 0.0);
void adj_vjp_current(dpmd_ctx* c, int mode, const double* gU, const double* gP, const double* gfx, const double* gfy,
                     double* dx, double* dy, double* dr, double* dLx, double* dLy, const double* live = nullptr);
const double* adj_touching(dpmd_ctx* c, int mode);
double adj_touching_max_force(dpmd_ctx* c, const double* live);
int adj_cg_current(dpmd_ctx* c, int mode, const double* bx, const double* by, double* xx, double* xy, double tol,
                   int max_iter, double* rel_resid, bool on_touching = true, double shift = 0.0);
// positions after a step's drift into the recording segment (canonical order); xy: read them
// from an interleaved (x, y) buffer instead of the context's arrays
void adj_record_positions(dpmd_ctx* c, Gate g, const double* xy = nullptr);
void adj_list_canonical(dpmd_ctx* c, int* inv_scratch);
void adj_seg_disp2(dpmd_ctx* c, int base, int len, double* d2);
void adj_rev_vw(dpmd_ctx* c, const double* m, double* xb, double* vb, double* w);
void adj_rev_w(dpmd_ctx* c, const double* m, const double* vb, double* w);
void adj_rev_x(dpmd_ctx* c, double coef, const double* z, double* xb);
void adj_hvp_segment(dpmd_ctx* c, int mode, int slot, const double* rad_c, const double* w, double* z);
// one reverse step with positions from slot `slot` of the segment array `seg` (2N per slot)
void adj_rev_step(dpmd_ctx* c, int mode, int slot, const double* rad_c, const double* m, const double* w, double coef,
                  double* xb, double* vb, double* w_next, bool next, const double* seg);
void adj_add(dpmd_ctx* c, int64_t n, const double* a, const double* b, double* out);
void adj_axpy2(dpmd_ctx* c, double a, const double* xx, const double* xy, double* yx, double* yy);
void adj_norm2(dpmd_ctx* c, const double* ax, const double* ay, double* out);

// disk.hip
void disk_compute_forces(dpmd_ctx* c, int mode);
void disk_update_positions(dpmd_ctx* c, const double* scale, double scale2);
void disk_update_velocities(dpmd_ctx* c, const double* scale, double scale2);
void disk_scale_velocities(dpmd_ctx* c, const double* scale);
void disk_scale_positions(dpmd_ctx* c, const double* scale);
void disk_mix(dpmd_ctx* c, const double* alpha);
void disk_stress(dpmd_ctx* c);
void disk_fire_force(dpmd_ctx* c, Gate g, int kick, int do_conv, int pred, double a, double b);
void disk_fire_drift(dpmd_ctx* c, Gate g);
void disk_vv_drift(dpmd_ctx* c, Gate g);
// fused velocity-Verlet batches: positions ping-pong between interleaved (x, y) buffers
void disk_vv_drift_to(dpmd_ctx* c, Gate g, double* xo, int any_word);
void disk_vv_fused(dpmd_ctx* c, Gate g, const double* xs, double* xo, int any_word);
void disk_vv_force_packed(dpmd_ctx* c, Gate g, const double* xs);
void disk_unpack_positions(dpmd_ctx* c, const double* xs);
// fused FIRE batches (one system): first drift into (xo, vo); step k's forces, kick and
// partials plus step k + 1's guessed drift; the fold that checks the guess (and redoes the
// drift when it was wrong); the batch's last forces (writing the state back); unpacking a
// halted step's buffers
constexpr int kSettleGrid = 128;   // workgroups of the settle launch (each folds the partials)
// The system of fused FIRE batches: its index, first particle, size, first chunk and number of
// chunks (chunk c of it covers particles b0 + 256 c .. b0 + 256 c + 255).
struct FusedSys {
    int s, b0, n, ch0, nch;
};
FusedSys fused_sys(dpmd_ctx* c);
void disk_fire_drift_to(dpmd_ctx* c, Gate g, double* xo, double* vo, int any_word);
void disk_fire_fused(dpmd_ctx* c, Gate g, const double* xs, const double* vs, double* xo, double* vo, int any_word);
void disk_fire_settle(dpmd_ctx* c, Gate g, int pred, double pa, double pb, const double* xs, const double* vs,
                      double* xo, double* vo, int any_word);
void disk_fire_force_packed(dpmd_ctx* c, Gate g, const double* xs, const double* vs, int pred, double pa, double pb);
void disk_unpack_pv(dpmd_ctx* c, const double* xs, const double* vs);
void disk_vv_force(dpmd_ctx* c, Gate g);
void disk_damping(dpmd_ctx* c, const double* scale);

// observe.hip (disks)
void disk_contacts(dpmd_ctx* c);
void contacts_total(dpmd_ctx* c);
int64_t pair_setup(dpmd_ctx* c);
int64_t rigid_particle_neighbors(dpmd_ctx* c);
void rigid_pair_obs(dpmd_ctx* c, int* pi, int* pj, double* dist, double* fx, double* fy, int* vci, int* vcj,
                    double* friction);
void rigid_overlaps(dpmd_ctx* c);
void rigid_hessian(dpmd_ctx* c, double* const* h18);
void disk_pair_obs(dpmd_ctx* c, int* pi, int* pj, double* dist, double* fx, double* fy, double* const* h);
void disk_ke(dpmd_ctx* c);
void fire_recompute_update(dpmd_ctx* c);
void rigid_compute_forces(dpmd_ctx* c, int mode);
void rigid_update_positions(dpmd_ctx* c, const double* scale, double scale2);
void rigid_update_velocities(dpmd_ctx* c, const double* scale, double scale2);
void rigid_scale_velocities(dpmd_ctx* c, const double* scale);
void rigid_scale_positions(dpmd_ctx* c, const double* scale);
void rigid_mix(dpmd_ctx* c, const double* alpha);
void rigid_stress(dpmd_ctx* c);
void rigid_ke(dpmd_ctx* c);
void rigid_fire_drift(dpmd_ctx* c, Gate g);
void rigid_fire_force(dpmd_ctx* c, Gate g, int kick, int do_conv, int pred, double a, double b);
void rigid_vv_drift(dpmd_ctx* c, Gate g);
void rigid_vv_force(dpmd_ctx* c, Gate g);
// Forces of the step at the current vertex positions, its closing half kick and the next
// step's opening half kick and drift (one launch pair; rigid_vv_fusable must hold).
bool rigid_vv_fusable(dpmd_ctx* c);
void rigid_vv_fused(dpmd_ctx* c, Gate g);
void rigid_damping(dpmd_ctx* c, const double* scale);

// reduce.hip
enum RedKind { RED_PE = 0, RED_KE, RED_POWER, RED_VX, RED_VY, RED_PHI, RED_SXX, RED_SXY, RED_SYX, RED_SYY, RED_MAXF2,
               RED_ARR, RED_DOT, RED_MAXARR };
void sys_reduce(dpmd_ctx* c, int kind, double* out);
void sys_reduce_arr(dpmd_ctx* c, const double* a, double* out);
void sys_reduce_max(dpmd_ctx* c, const double* a, double* out);
void sys_reduce_dot(dpmd_ctx* c, const double* a, const double* b, const double* cc, const double* d, double* out);
void fire_apply_convergence(dpmd_ctx* c, int pred, double a, double b);
void fire_setup(dpmd_ctx* c, const double* dt);
void run_begin_step(dpmd_ctx* c);
// Clears the halt and the batch words; flag_one: also flags the system of fused FIRE batches
// (fused_sys) for the rebuild (their launches raise only run words).
void halt_reset(dpmd_ctx* c, bool flag_one);
void per_system_temperature(dpmd_ctx* c);
void per_system_pressure(dpmd_ctx* c);
void temperature_scale(dpmd_ctx* c, const double* target, double* scale_out);
void subtract_mean_velocity(dpmd_ctx* c, const double* cur_x, const double* cur_y, const double* want_x, const double* want_y);
void box_inv(dpmd_ctx* c);
void scale_box_increment(dpmd_ctx* c, const double* inc, double* scale_out);
void save_restore(dpmd_ctx* c, const int* flag, int tv, bool restore);

// rng.hip
void rand_uniform_2d(hipStream_t st, double* ox, double* oy, int n, uint64_t seed, uint64_t call,
                     double lox, double hix, double loy, double hiy);
void rand_uniform_1d(hipStream_t st, double* o, int n, uint64_t seed, uint64_t call, double lo, double hi);
void rand_normal_2d(hipStream_t st, double* ox, double* oy, int n, uint64_t seed, uint64_t call,
                    double mx, double sx, double my, double sy);
void fill_f64(hipStream_t st, double* p, int64_t n, double v);
void fill_i32(hipStream_t st, int* p, int64_t n, int v);
void copy_f64(hipStream_t st, double* dst, const double* src, int64_t n);

// Records a (begin, end) event pair around the enclosed launch when timing is on.
// Records a HIP event pair around one launch when force-kernel profiling (kind 1) is on and
// the stream is not capturing.  Fused velocity-Verlet profiling (kind 2) brackets the
// consecutive fused launches of a batch instead (abi.hip).
struct ProfScope {
    dpmd_ctx* c; hipEvent_t b = nullptr;
    explicit ProfScope(dpmd_ctx* cc, int kind = 1) : c(cc) {
        if (c->prof == kind && !c->capturing) { HIP_OK(hipEventCreate(&b)); HIP_OK(hipEventRecord(b, c->st)); }
    }
    ~ProfScope() {
        if (!b) return;
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) == hipSuccess && hipEventRecord(e, c->st) == hipSuccess) {
            c->prof_ev.push_back(b); c->prof_ev.push_back(e); c->prof_n.push_back(1);
        }
    }
};

}  // namespace dpmd
