// C ABI (include/dpmd/dpmd.h): context lifecycle, fields, validation, and the host-side
// orchestration of the particle / integrator / routine calls.
#include "internal.hpp"
#include <array>
#include <map>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

using namespace dpmd;

namespace {

struct FieldRef {
    void* p; int64_t n; int elem;   // elem: 8 = double, 4 = int
    int domain;                     // 0 particle, 1 vertex, 2 system, 3 other
};

template <class T>
FieldRef ref(dpmd::DVec<T>& v, int domain) { return FieldRef{(void*)v.p, v.n, (int)sizeof(T), domain}; }

FieldRef field_ref(dpmd_ctx* c, int f) {
    switch (f) {
        case DPMD_F_POS_X: return ref(c->x, 0);
        case DPMD_F_POS_Y: return ref(c->y, 0);
        case DPMD_F_VEL_X: return ref(c->vx, 0);
        case DPMD_F_VEL_Y: return ref(c->vy, 0);
        case DPMD_F_FORCE_X: return ref(c->fx, 0);
        case DPMD_F_FORCE_Y: return ref(c->fy, 0);
        case DPMD_F_PE: return ref(c->pe, 0);
        case DPMD_F_KE: return ref(c->ke, 0);
        case DPMD_F_RAD: return ref(c->rad, 0);
        case DPMD_F_MASS: return ref(c->mass, 0);
        case DPMD_F_AREA: return ref(c->area, 0);
        case DPMD_F_LAST_POS_X: return ref(c->lx, 0);
        case DPMD_F_LAST_POS_Y: return ref(c->ly, 0);
        case DPMD_F_ANGLE: return ref(c->angle, 0);
        case DPMD_F_ANGULAR_VEL: return ref(c->avel, 0);
        case DPMD_F_TORQUE: return ref(c->torque, 0);
        case DPMD_F_MOMENT_INERTIA: return ref(c->inertia, 0);
        case DPMD_F_STRESS_XX: return ref(c->sxx, 0);
        case DPMD_F_STRESS_XY: return ref(c->sxy, 0);
        case DPMD_F_STRESS_YX: return ref(c->syx, 0);
        case DPMD_F_STRESS_YY: return ref(c->syy, 0);
        case DPMD_F_OVERLAP: return ref(c->overlap, c->kind == DPMD_RIGID_BUMPY ? 1 : 0);
        case DPMD_F_CONTACT_COUNT: return ref(c->ccount, c->kind == DPMD_RIGID_BUMPY ? 1 : 0);
        case DPMD_F_N_CONTACTS_TOTAL: return ref(c->ncontacts, 2);
        case DPMD_F_CONTACTS: return ref(c->contacts, 0);
        case DPMD_F_BOX_SIZE_X: return ref(c->Lx, 2);
        case DPMD_F_BOX_SIZE_Y: return ref(c->Ly, 2);
        case DPMD_F_BOX_INV_X: return ref(c->iLx, 2);
        case DPMD_F_BOX_INV_Y: return ref(c->iLy, 2);
        case DPMD_F_E_INTERACTION: return ref(c->e, 2);
        case DPMD_F_NEIGHBOR_CUTOFF: return ref(c->cutoff, 2);
        case DPMD_F_THRESH2: return ref(c->thresh2, 2);
        case DPMD_F_PE_TOTAL: return ref(c->pe_tot, 2);
        case DPMD_F_KE_TOTAL: return ref(c->ke_tot, 2);
        case DPMD_F_FPOWER_TOTAL: return ref(c->pw_tot, 2);
        case DPMD_F_PRESSURE: return ref(c->press, 2);
        case DPMD_F_TEMPERATURE: return ref(c->temp, 2);
        case DPMD_F_PACKING_FRACTION: return ref(c->phi, 2);
        case DPMD_F_MAX_FORCE2: return ref(c->maxf2, 2);
        case DPMD_F_STRESS_TOTAL_XX: return ref(c->st_xx, 2);
        case DPMD_F_STRESS_TOTAL_XY: return ref(c->st_xy, 2);
        case DPMD_F_STRESS_TOTAL_YX: return ref(c->st_yx, 2);
        case DPMD_F_STRESS_TOTAL_YY: return ref(c->st_yy, 2);
        case DPMD_F_SYSTEM_OFFSET: return ref(c->off, 3);
        case DPMD_F_SYSTEM_SIZE: return ref(c->size, 2);
        case DPMD_F_CELL_DIM_X: return ref(c->cdx, 2);
        case DPMD_F_CELL_DIM_Y: return ref(c->cdy, 2);
        case DPMD_F_CELL_SYSTEM_START: return ref(c->cstart, 3);
        case DPMD_F_BOUNDARY: return ref(c->boundary, 2);
        case DPMD_F_N_DOF: return ref(c->ndof, 2);
        case DPMD_F_REBUILD_FLAG: return ref(c->flag, 2);
        case DPMD_F_FIRE_STEPS: return ref(c->f_nsteps, 2);
        case DPMD_F_SYSTEM_ID: return ref(c->sid, 0);
        case DPMD_F_STATIC_INDEX: return ref(c->static_index, c->kind == DPMD_RIGID_BUMPY ? 1 : 0);
        case DPMD_F_CELL_ID: return ref(c->cell_id, c->kind == DPMD_RIGID_BUMPY ? 1 : 0);
        case DPMD_F_NEIGHBOR_COUNT: return ref(c->nn, c->kind == DPMD_RIGID_BUMPY ? 1 : 0);
        case DPMD_F_VERTEX_POS_X: return ref(c->vpx, 1);
        case DPMD_F_VERTEX_POS_Y: return ref(c->vpy, 1);
        case DPMD_F_VERTEX_FORCE_X: return ref(c->vfx, 1);
        case DPMD_F_VERTEX_FORCE_Y: return ref(c->vfy, 1);
        case DPMD_F_VERTEX_PE: return ref(c->vpe, 1);
        case DPMD_F_VERTEX_RAD: return ref(c->vrad, 1);
        case DPMD_F_VERTEX_MASS: return ref(c->vmass, 1);
        case DPMD_F_VERTEX_PARTICLE_ID: return ref(c->vpid, 1);
        case DPMD_F_PARTICLE_OFFSET: return ref(c->poff, 3);
        case DPMD_F_N_VERTICES_PER_PARTICLE: return ref(c->nvpp, 0);
        case DPMD_F_VERTEX_SYSTEM_ID: return ref(c->vsid, 1);
        case DPMD_F_VERTEX_SYSTEM_OFFSET: return ref(c->voff, 3);
    }
    throw ArgError("unknown field id " + std::to_string(f));
}


const int* active_one(dpmd_ctx* c) { return c->one.p; }
Gate plain(dpmd_ctx* c) { return Gate{c->one.p, nullptr, 0}; }

bool rigid(dpmd_ctx* c) { return c->kind == DPMD_RIGID_BUMPY; }

// Cell layout (init_cell_sizes, base_particle.hpp:552-592).
// Grid rule (the oracle applies the same one): a zero dimension is derived as the most cells
// whose side covers 2*r_max + cutoff; after a box change (CELLS_BOX_CHANGED) a grid whose
// cells became smaller than that is re-derived too.  A user grid is kept at set-up, as the
// reference does (its own one-particle-per-cell test uses cells smaller than the cutoff).
enum { CELLS_SYNC = 0, CELLS_INIT = 1, CELLS_BOX_CHANGED = 2 };
void setup_cells(dpmd_ctx* c, int mode) {
    const int S = c->S;
    auto cdx = dl(c, c->cdx), cdy = dl(c, c->cdy);
    auto Lx = dl(c, c->Lx), Ly = dl(c, c->Ly), cut = dl(c, c->cutoff);
    auto rad = rigid(c) ? dl(c, c->vrad) : dl(c, c->rad);
    const std::vector<int>& off = rigid(c) ? c->h_voff : c->h_off;
    bool changed = false;
    for (int s = 0; s < S; ++s) {
        double rmax = 0.0;
        for (int i = off[s]; i < off[s + 1]; ++i) rmax = std::max(rmax, rad[i]);
        const double need_sz = 2.0 * rmax + cut[s];
        const bool too_small = Lx[s] / cdx[s] < need_sz || Ly[s] / cdy[s] < need_sz;
        if (cdx[s] <= 0 || cdy[s] <= 0 || (mode == CELLS_BOX_CHANGED && too_small)) {
            const int nx = (int)std::floor(Lx[s] / need_sz), ny = (int)std::floor(Ly[s] / need_sz);
            if (mode != CELLS_SYNC)
                need(nx >= 3 && ny >= 3, "system " + std::to_string(s) +
                                             ": the box is too small for a 3x3 grid of cells covering "
                                             "2*max radius + neighbor cutoff (use the naive method)");
            cdx[s] = std::max(3, nx); cdy[s] = std::max(3, ny);
            changed = true;
        }
        if (mode != CELLS_SYNC)
            need(cdx[s] >= 3 && cdy[s] >= 3,
                 "cell grid must be at least 3x3 per system (system " + std::to_string(s) +
                     "): the periodic 3x3 stencil would visit a cell twice");
    }
    if (changed) {
        c->cdx.upload(cdx.data(), S, c->st);
        c->cdy.upload(cdy.data(), S, c->st);
    }
    std::vector<int> cs(S + 1, 0);
    for (int s = 0; s < S; ++s) cs[s + 1] = cs[s] + std::max(0, cdx[s]) * std::max(0, cdy[s]);
    c->cstart.upload(cs.data(), S + 1, c->st);
    c->C = cs[S];
    std::vector<int> csys(std::max(1, c->C), 0);
    for (int s = 0; s < S; ++s) for (int k = cs[s]; k < cs[s + 1]; ++k) csys[k] = s;
    c->cell_sys.upload(csys.data(), csys.size(), c->st);
    c->cell_count.resize(std::max(1, c->C)); c->cell_count.zero(c->st);
    c->cell_start.resize(c->C + 1); c->cell_start.zero(c->st);
    c->cell_wptr.resize(std::max(1, c->C));
    c->scan_part.resize(std::max(1, (c->C + 2047) / 2048));
    build_tiles(c, cdx, cdy);
}

void alloc_obj_buffers(dpmd_ctx* c) {
    const int n = c->n_obj();
    c->cell_id.resize(std::max(1, n));
    c->order.resize(std::max(1, n));
    c->order_inv.resize(std::max(1, n));
    c->tmp_d.resize((int64_t)std::max(1, n) * 10);
    c->tmp_i.resize((int64_t)std::max(1, n) * 2);
    c->nn.resize(std::max(1, n)); c->nn.zero(c->st);
    c->nbr.resize((int64_t)std::max(1, n) * c->K);
}

void sync_impl(dpmd_ctx* c) {
    need(c->S > 0, "dpmd_sync: allocate_systems first");
    need(c->N > 0, "dpmd_sync: allocate_particles first");
    const int S = c->S, N = c->N;
    c->h_off = dl(c, c->off);
    need(c->h_off[0] == 0 && c->h_off[S] == N, "system_offset must start at 0 and end at n_particles");
    for (int s = 0; s < S; ++s)
        need(c->h_off[s + 1] > c->h_off[s], "every system needs at least one particle (system_offset must increase)");
    std::vector<int> sid(N), size(S), ndof(S);
    for (int s = 0; s < S; ++s) {
        size[s] = c->h_off[s + 1] - c->h_off[s];
        for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) sid[i] = s;
    }
    c->sid.upload(sid.data(), N, c->st);
    c->size.upload(size.data(), S, c->st);
    if (rigid(c)) {
        need(c->Nv > 0, "dpmd_sync: allocate_vertices first");
        auto nvpp = dl(c, c->nvpp);
        std::vector<int> poff(N + 1, 0), vpid(c->Nv), vsid(c->Nv);
        for (int i = 0; i < N; ++i) { need(nvpp[i] > 0, "every rigid particle needs at least one vertex"); poff[i + 1] = poff[i] + nvpp[i]; }
        c->nv_max = *std::max_element(nvpp.begin(), nvpp.end());
        need(poff[N] == c->Nv, "sum of n_vertices_per_particle must equal n_vertices");
        c->h_voff.assign(S + 1, 0);
        for (int s = 0; s < S; ++s) c->h_voff[s + 1] = poff[c->h_off[s + 1]];
        for (int i = 0; i < N; ++i)
            for (int v = poff[i]; v < poff[i + 1]; ++v) { vpid[v] = i; vsid[v] = sid[i]; }
        c->poff.upload(poff.data(), N + 1, c->st);
        c->vpid.upload(vpid.data(), c->Nv, c->st);
        c->vsid.upload(vsid.data(), c->Nv, c->st);
        c->voff.upload(c->h_voff.data(), S + 1, c->st);
        auto inertia = dl(c, c->inertia);
        for (int s = 0; s < S; ++s) {   // set_n_dof_kernel (rigid_bumpy.cu:743-757)
            int d = 0;
            for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) d += 2 + (inertia[i] > 0.0 ? 1 : 0);
            ndof[s] = d;
        }
    } else {
        for (int s = 0; s < S; ++s) ndof[s] = 2 * size[s];   // Disk::set_n_dof_impl (disk.cu:1250-1253)
    }
    c->ndof.upload(ndof.data(), S, c->st);
    box_inv(c);
    // A repeated sync with an unchanged layout (the reference calls sync_box / sync_system /
    // sync_cells / ... in sequence, and again after editing a field) keeps the neighbor
    // state; any layout change invalidates it.
    std::vector<int> sig = {S, N, c->Nv, c->method, c->K};
    sig.insert(sig.end(), c->h_off.begin(), c->h_off.end());
    auto cdx = dl(c, c->cdx), cdy = dl(c, c->cdy);
    sig.insert(sig.end(), cdx.begin(), cdx.end());
    sig.insert(sig.end(), cdy.begin(), cdy.end());
    if (c->synced && c->nbr_ready && sig == c->layout_sig) return;
    build_chunks(c);
    if (c->method == DPMD_NEIGHBOR_CELL) setup_cells(c, CELLS_SYNC);
    alloc_obj_buffers(c);
    c->layout_sig = sig;
    c->synced = true;
    c->nbr_ready = false;
}

void zero_scale(dpmd_ctx* c) { fill_f64(c->st, c->scr_c.p, c->S, 0.0); }

void compute_forces_impl(dpmd_ctx* c, int mode) {
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    if (rigid(c)) rigid_compute_forces(c, mode); else disk_compute_forces(c, mode);
}

void check_neighbors_impl(dpmd_ctx* c, bool force) {
    if (c->method != DPMD_NEIGHBOR_CELL) return;
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    rebuild_neighbors(c, active_one(c), force);
}

void init_neighbors_impl(dpmd_ctx* c) {
    need_state(c->synced, "call dpmd_sync before init_neighbors");
    const int n = c->n_obj();
    HIP_OK(hipMemsetAsync(c->overflow.p, 0, sizeof(int), c->st));
    if (c->method == DPMD_NEIGHBOR_NAIVE) {
        const std::vector<int>& off = rigid(c) ? c->h_voff : c->h_off;
        int mx = 1;
        for (int s = 0; s < c->S; ++s) mx = std::max(mx, off[s + 1] - off[s] - 1);
        c->K = mx;
        alloc_obj_buffers(c);
        naive_neighbors(c);
    } else {
        setup_cells(c, CELLS_INIT);
        alloc_obj_buffers(c);
        std::vector<int> iota(n);
        std::iota(iota.begin(), iota.end(), 0);
        c->static_index.upload(iota.data(), n, c->st);
        if (!rigid(c)) {
            HIP_OK(hipMemcpyAsync(c->lx.p, c->x.p, c->N * 8, hipMemcpyDeviceToDevice, c->st));
            HIP_OK(hipMemcpyAsync(c->ly.p, c->y.p, c->N * 8, hipMemcpyDeviceToDevice, c->st));
        }
        c->nbr_ready = true;
        rebuild_neighbors(c, active_one(c), true);
    }
    c->nbr_ready = true;
    check_overflow(c);
}

void update_positions_impl(dpmd_ctx* c, const double* sc, double s2) {
    if (rigid(c)) rigid_update_positions(c, sc, s2); else disk_update_positions(c, sc, s2);
}
void update_velocities_impl(dpmd_ctx* c, const double* sc, double s2) {
    if (rigid(c)) rigid_update_velocities(c, sc, s2); else disk_update_velocities(c, sc, s2);
}
void scale_velocities_impl(dpmd_ctx* c, const double* sc) {
    if (rigid(c)) rigid_scale_velocities(c, sc); else disk_scale_velocities(c, sc);
}
void scale_positions_impl(dpmd_ctx* c, const double* sc) {
    if (rigid(c)) rigid_scale_positions(c, sc); else disk_scale_positions(c, sc);
}

void set_average_velocity_impl(dpmd_ctx* c, const double* ax, const double* ay) {
    sys_reduce(c, RED_VX, c->scr_a.p);
    sys_reduce(c, RED_VY, c->scr_b.p);
    subtract_mean_velocity(c, c->scr_a.p, c->scr_b.p, ax, ay);
}

// increment_packing_fraction (base_particle.hpp:505-529): s = sqrt(phi / (phi + inc)) per
// system scales box and positions, then cells and lists are rebuilt.
void increment_packing_fraction_impl(dpmd_ctx* c, const double* inc) {
    sys_reduce(c, RED_PHI, c->phi.p);
    scale_box_increment(c, inc, c->scr_c.p);
    scale_positions_impl(c, c->scr_c.p);
    box_inv(c);
    if (c->method == DPMD_NEIGHBOR_CELL) {
        setup_cells(c, CELLS_BOX_CHANGED);
        alloc_obj_buffers(c);
        // the cell grid may have changed: every system is rebuilt below
    }
    check_neighbors_impl(c, true);
    sys_reduce(c, RED_PHI, c->phi.p);
}

void fire_init_impl(dpmd_ctx* c, const double* dt, int mode) {
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    const int S = c->S;
    std::vector<double> h(S);
    HIP_OK(hipMemcpyAsync(h.data(), dt, S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    std::vector<double> mx(S), mn(S), al(S, 0.1), pp(S, 1e9);
    std::vector<int> z(S, 0), stopped(S, 0);
    int running = 0;
    for (int s = 0; s < S; ++s) {          // fire.cuh:136-147 define_arrays
        mx[s] = h[s] * 10.0; mn[s] = h[s] * 0.001;
        stopped[s] = h[s] == 0.0 ? 1 : 0;
        running += stopped[s] ? 0 : 1;
    }
    c->f_dt.upload(h.data(), S, c->st); c->f_dtmax.upload(mx.data(), S, c->st); c->f_dtmin.upload(mn.data(), S, c->st);
    c->f_alpha.upload(al.data(), S, c->st); c->f_peprev.upload(pp.data(), S, c->st);
    c->f_ngood.upload(z.data(), S, c->st); c->f_nbad.upload(z.data(), S, c->st);
    c->f_stopped.upload(stopped.data(), S, c->st);
    c->f_frozen.upload(stopped.data(), S, c->st);   // init zeroes v: a stopped system is final
    c->f_nsteps.upload(z.data(), S, c->st);
    int run[RUN_WORDS] = {0, 0, running, 1, 0, -1, -1, 0, 0};
    c->run.upload(run, RUN_WORDS, c->st);
    c->fire_mode = mode;
    // init_impl (fire.cuh:90-96): v *= 0, then forces; the first update sees F.v = 0.
    zero_scale(c);
    scale_velocities_impl(c, c->scr_c.p);
    if (rigid(c)) rigid_fire_force(c, plain(c), 0, 0, 0, 0.0, 0.0);
    else disk_fire_force(c, plain(c), 0, 0, 0, 0.0, 0.0);
    c->fire_ready = true;
}

// One full step in the reference's order, every launch unconditional: drift, the
// device-conditional rebuild chain, forces (+ FIRE bookkeeping).
void fire_step_body(dpmd_ctx* c, int do_conv, int pred, double a, double b) {
    if (rigid(c)) rigid_fire_drift(c, plain(c)); else disk_fire_drift(c, plain(c));
    if (c->method == DPMD_NEIGHBOR_CELL) rebuild_neighbors(c, active_one(c), false, true);
    if (rigid(c)) rigid_fire_force(c, plain(c), 1, do_conv, pred, a, b);
    else disk_fire_force(c, plain(c), 1, do_conv, pred, a, b);
}

void vv_step_body(dpmd_ctx* c) {
    if (rigid(c)) rigid_vv_drift(c, plain(c)); else disk_vv_drift(c, plain(c));
    if (c->method == DPMD_NEIGHBOR_CELL) rebuild_neighbors(c, active_one(c), false, true);
    if (rigid(c)) rigid_vv_force(c, plain(c)); else disk_vv_force(c, plain(c));
}

// ---------------------------------------------------------------------------------------
// Lean fused runs.  A batch of m steps issues only drift + force (+ FIRE finalize) per
// step.  A step whose drift raised a rebuild flag is halted by its force kernel (the rest
// of the batch becomes no-ops); the host then finishes that step with the full chain
// (rebuild, forces, bookkeeping) and continues.  Every step therefore does exactly what
// the full step does, while steps without a rebuild cost two or three launches.  Batches
// are captured once per size into hipGraphs when use_graph is set.
// ---------------------------------------------------------------------------------------
struct RunCtl {
    int running, halt, halt_at, conv_at, rebuilds;
};

// Turns on frozen-system skipping for the duration of one fused FIRE run.
struct FrozenMask {
    dpmd_ctx* c;
    explicit FrozenMask(dpmd_ctx* ctx) : c(ctx) { c->mask_frozen = true; }
    ~FrozenMask() { c->mask_frozen = false; }
};

RunCtl read_run(dpmd_ctx* c) {
    int run[RUN_WORDS];
    HIP_OK(hipMemcpyAsync(run, c->run.p, sizeof(run), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    return RunCtl{run[RUN_RUNNING], run[RUN_HALT], run[RUN_HALT_AT], run[RUN_CONV_AT], run[RUN_REBUILDS]};
}

// Fused velocity-Verlet batches (disks, even batch sizes): drift of step 0 from the
// context's positions A into the interleaved buffer P0, then m - 1 launches that each finish
// step q (forces at x_q, closing half kick) and start step q + 1 (opening half kick, drift
// into the other buffer), then the forces and closing half kick of step m - 1, which also
// write x_{m-1} back to A.  x_q lives in P(q mod 2).  The rebuild flags of x_q go to run word
// W(q) (RUN_ANY for even q, RUN_ANY2 for odd q) and gate the launch that computes forces at
// x_q; a halted batch copies x_q back to A before the host finishes the step.
bool fused_vv_batch(dpmd_ctx* c, int m, bool fire) {
    return !fire && !rigid(c) && c->fuse_vv && m >= 2 && m % 2 == 0;
}
int any_word(int q) { return q % 2 == 0 ? RUN_ANY : RUN_ANY2; }

// Fused FIRE batches (one disk system, even batch sizes): the first drift from the context's
// arrays into the buffer pair (P0, V0); then per step q < m - 1 one launch for step q's
// forces, kick and partials plus step q + 1's drift with the scalars F.v > 0 gives (into the
// other pair), and one that folds the partials, commits the next step's scalars and redoes
// the drift when the guess was wrong (or writes the final state after a stop); then the
// forces of step m - 1, which also write the state back, and its fold.  Rebuild tests go to
// the parity words as in fused velocity-Verlet batches (a redone drift to RUN_FIXW); a halted
// batch unpacks the halted step's pair and flags the system before the rebuild.
bool fused_fire_batch(dpmd_ctx* c, int m, bool fire) {
    return fire && !rigid(c) && c->fuse_fire && (c->S == 1 || c->fsys >= 0) && m >= 2 && m % 2 == 0;
}

// dev: every step rebuilds the flagged systems on the device between its drift and its
// forces (device-gated rebuild chain), so the batch never halts; used when rebuilds are
// frequent (many systems, each rebuilding now and then).
void lean_batch(dpmd_ctx* c, int m, bool fire, int pred, double a, double b, bool dev) {
    if (!dev && fused_fire_batch(c, m, fire)) {
        double* P[2] = {c->xy0.p, c->xy1.p};
        double* V[2] = {c->vxy0.p, c->vxy1.p};
        disk_fire_drift_to(c, Gate{nullptr, c->run.p, 0}, P[0], V[0], any_word(0));
        const bool prof = c->prof == 3 && !c->capturing;
        if (prof) {   // one pair around the batch's fused steps; kept only if no rebuild halts it
            HIP_OK(hipEventCreate(&c->pend_b)); HIP_OK(hipEventCreate(&c->pend_e));
            HIP_OK(hipEventRecord(c->pend_b, c->st));
            c->pend_n = m - 1;
        }
        for (int q = 0; q + 1 < m; ++q) {
            Gate g{nullptr, c->run.p, q};
            g.any = any_word(q);
            const int nx = (q + 1) % 2, nw = any_word(q + 1);
            disk_fire_fused(c, g, P[q % 2], V[q % 2], P[nx], V[nx], nw);
            disk_fire_settle(c, g, pred, a, b, P[q % 2], V[q % 2], P[nx], V[nx], nw);
        }
        if (prof) HIP_OK(hipEventRecord(c->pend_e, c->st));
        Gate gf{nullptr, c->run.p, m - 1};
        gf.any = any_word(m - 1);
        disk_fire_force_packed(c, gf, P[(m - 1) % 2], V[(m - 1) % 2], pred, a, b);
        return;
    }
    if (!dev && fused_vv_batch(c, m, fire)) {
        double* P[2] = {c->xy0.p, c->xy1.p};
        disk_vv_drift_to(c, Gate{nullptr, c->run.p, 0}, P[0], any_word(0));
        // a recording replay (reverse mode) stores the positions after every step's drift
        if (c->rec_seg) adj_record_positions(c, Gate{nullptr, c->run.p, 0}, P[0]);
        const bool prof = c->prof == 2 && !c->capturing;
        if (prof) {   // one pair around the batch's fused launches; kept only if no rebuild halts it
            HIP_OK(hipEventCreate(&c->pend_b)); HIP_OK(hipEventCreate(&c->pend_e));
            HIP_OK(hipEventRecord(c->pend_b, c->st));
            c->pend_n = m - 1;
        }
        for (int q = 0; q + 1 < m; ++q) {
            Gate g{nullptr, c->run.p, q};
            g.any = any_word(q);
            disk_vv_fused(c, g, P[q % 2], P[(q + 1) % 2], any_word(q + 1));
            if (c->rec_seg) adj_record_positions(c, Gate{nullptr, c->run.p, q + 1}, P[(q + 1) % 2]);
        }
        if (prof) HIP_OK(hipEventRecord(c->pend_e, c->st));
        Gate gf{nullptr, c->run.p, m - 1};
        gf.any = any_word(m - 1);
        disk_vv_force_packed(c, gf, P[(m - 1) % 2]);
        return;
    }
    if (!fire && rigid(c) && c->fuse_vv && !c->rec_seg && rigid_vv_fusable(c)) {
        // rigid bodies: drift of step 0, then per step the (gated) rebuild, the vertex forces
        // and one launch that closes the step and opens the next (the last step only closes)
        rigid_vv_drift(c, Gate{nullptr, c->run.p, 0});
        for (int k = 0; k < m; ++k) {
            const Gate g = Gate{nullptr, c->run.p, k};
            if (dev) rebuild_on_device(c);
            if (k + 1 < m) rigid_vv_fused(c, g); else rigid_vv_force(c, g);
        }
        return;
    }
    for (int k = 0; k < m; ++k) {
        const Gate g = Gate{nullptr, c->run.p, k};
        if (fire) {
            if (rigid(c)) rigid_fire_drift(c, g); else disk_fire_drift(c, g);
            if (dev) rebuild_on_device(c);
            if (rigid(c)) rigid_fire_force(c, g, 1, 1, pred, a, b);
            else disk_fire_force(c, g, 1, 1, pred, a, b);
        } else {
            if (rigid(c)) rigid_vv_drift(c, g); else disk_vv_drift(c, g);
            if (c->rec_seg) adj_record_positions(c, g);
            if (dev) rebuild_on_device(c);
            if (rigid(c)) rigid_vv_force(c, g); else disk_vv_force(c, g);
        }
    }
}

hipGraphExec_t capture_batch(dpmd_ctx* c, int m, bool fire, int pred, double a, double b, bool dev) {
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    try {
        lean_batch(c, m, fire, pred, a, b, dev);
    } catch (...) {
        c->capturing = false;
        (void)hipStreamEndCapture(c->st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    c->capturing = false;
    HIP_OK(hipStreamEndCapture(c->st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw Error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ex;
}

}  // namespace

namespace {

// Kernel arguments a captured step batch bakes in: the run's parameters, the execution
// options and the device buffers its launches touch.  A cached batch is replayed only while
// all of them are unchanged.
std::vector<int64_t> batch_key(dpmd_ctx* c, bool fire, int pred, double a, double b) {
    auto P = [](const void* p) { return (int64_t)(intptr_t)p; };
    int64_t ab[2];
    std::memcpy(&ab[0], &a, 8);
    std::memcpy(&ab[1], &b, 8);
    return {fire ? 1 : 0, pred, ab[0], ab[1], c->block, c->fuse_vv ? 1 : 0, c->fuse_fire ? 1 : 0, c->vv_mode,
            c->fire_mode, P(c->vxy0.p), P(c->vxy1.p), P(c->f_snap.p), P(c->i_snap.p), P(c->f_alpha.p), P(c->f_ngood.p),
            P(c->f_nbad.p), P(c->f_dtmax.p), P(c->f_dtmin.p), P(c->f_stopped.p), P(c->s_dt.p), P(c->pw_tot.p),
            c->mask_frozen ? 1 : 0, c->N, c->Nv, c->S, c->K, c->method, c->nchunks, P(c->rec_seg), P(vv_damping(c)),
            P(c->x.p), P(c->y.p), P(c->xy0.p), P(c->xy1.p), P(c->vx.p), P(c->vy.p), P(c->fx.p), P(c->fy.p), P(c->pe.p),
            P(c->rad.p), P(c->mass.p), P(c->lx.p), P(c->ly.p), P(c->nn.p), P(c->nbr.p), P(c->run.p), P(c->flag.p),
            P(c->vv_dt.p), P(c->vv_dth.p), P(c->f_dt.p), P(c->f_frozen.p), P(c->part.p), P(c->chunk_sys.p),
            P(c->static_index.p), P(c->angle.p), P(c->avel.p), P(c->torque.p), P(c->vpx.p), P(c->vpy.p),
            P(c->vfx.p), P(c->vfy.p), P(c->vpe.p), P(c->sid.p), P(c->off.p), P(c->Lx.p), P(c->thresh2.p),
            // the device-side rebuild chain's grid, buffers and engine
            c->C, c->ntiles, c->rebuild_engine, P(c->cell_id.p), P(c->cell_count.p), P(c->cell_start.p),
            P(c->cell_wptr.p), P(c->cell_sys.p), P(c->cstart.p), P(c->cdx.p), P(c->cdy.p), P(c->scan_part.p),
            P(c->order.p), P(c->order_inv.p), P(c->tmp_d.p), P(c->tmp_i.p), P(c->tiles.p), P(c->overflow.p),
            P(c->Ly.p), P(c->iLx.p), P(c->iLy.p), P(c->cutoff.p), P(c->vsid.p), P(c->voff.p), P(c->vpid.p),
            P(c->poff.p), P(c->vrad.p), P(c->vmass.p), P(c->inertia.p), P(c->area.p), P(c->sys_tile.p),
            P(c->sys_list.p), P(c->sys_list_n.p), c->n_act_chunk, c->n_act_sys, P(c->act_chunk.p), P(c->act_sys.p),
            c->fsys, c->fsys_ch0};
}

void drop_cached_graphs(dpmd_ctx* c) {
    for (auto& g : c->gcache) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    c->gcache_key.clear();
}

// Keeps captured step batches across the device-driven runs of one routine call (segments of
// a checkpointed run, the replays of a reverse sweep, the chunks of a thermostatted run).
struct GraphScope {
    dpmd_ctx* c;
    bool outer;
    explicit GraphScope(dpmd_ctx* ctx) : c(ctx), outer(!ctx->gcache_on) { c->gcache_on = true; }
    ~GraphScope() {
        if (outer) {
            drop_cached_graphs(c);
            c->gcache_on = false;
        }
    }
};

}  // namespace

namespace {
// A FIRE run whose launches cover only the running systems (act_chunk / act_sys): drops the
// systems that are frozen now (once at most half of the set is left, or one system) and,
// when one system is left and the fused FIRE step applies, makes it the system of fused
// batches.  Returns whether the launches changed.
bool shrink_active(dpmd_ctx* c) {
    const int S = c->S;
    std::vector<int> frozen = dl(c, c->f_frozen), sys = dl(c, c->act_sys);
    sys.resize(c->n_act_sys);
    std::vector<int> keep;
    for (int s : sys) if (!frozen[s]) keep.push_back(s);
    // each change costs the capture of new graphs: shrink when the set has halved, or to the
    // last system (which then takes the fused step)
    if (keep.empty() || (keep.size() != 1 && 2 * (int)keep.size() > c->n_act_sys)) return false;
    std::vector<int> first(S + 1, 0), chunks;
    for (int s = 0; s < S; ++s) first[s + 1] = first[s] + (c->h_off[s + 1] - c->h_off[s] + DPMD_BLOCK - 1) / DPMD_BLOCK;
    for (int s : keep)
        for (int k = first[s]; k < first[s + 1]; ++k) chunks.push_back(k);
    c->act_chunk.upload(chunks.data(), (int64_t)chunks.size(), c->st);
    c->act_sys.upload(keep.data(), (int64_t)keep.size(), c->st);
    c->n_act_chunk = (int)chunks.size();
    c->n_act_sys = (int)keep.size();
    if (keep.size() == 1 && c->fuse_fire && !rigid(c)) {
        c->fsys = keep[0];
        c->fsys_ch0 = first[keep[0]];
    }
    return true;
}
}  // namespace

// Runs up to max_steps steps as graph-captured batches of per-step launches; returns the
// number taken and (FIRE) whether all systems stopped.
int dpmd::lean_run_batches(dpmd_ctx* c, int max_steps, bool fire, int pred, double a, double b, bool use_graph,
                           int* conv) {
    hipGraphExec_t local[16] = {};   // batch sizes 1, 2, 4, ..., 128; halting / device-rebuild
    hipGraphExec_t* graphs = local;
    const bool cached = use_graph && c->gcache_on;
    if (cached) {
        std::vector<int64_t> key = batch_key(c, fire, pred, a, b);
        if (key != c->gcache_key) {
            drop_cached_graphs(c);
            c->gcache_key = std::move(key);
        }
        graphs = c->gcache;
    }
    // interval: steps between the last two halts (0 = none seen yet); after a batch that ran to
    // its end the next one is sized to end where the next halt is expected, since the launches
    // of a halted batch's remaining steps still cost their dispatch
    int total = 0, m = 8, since_halt = 0, interval = 0;
    bool done = false;
    // Rebuilds per step, smoothed over batches.  A halted batch costs a host round trip and
    // restarts with small batches; above one rebuild per 2 steps (back below 0.35 per step)
    // the batches rebuild on the device instead.  A halt costs about 95 us and the gated
    // device chain about 21 us per step, but the smoothed rate lags FIRE's early burst of
    // rebuilds, so the switch sits above the 0.2 break-even.  FIRE of one disk system halts
    // in fused batches (one fused step about 40 us) but rebuilds on the device in
    // three-launch steps (about 60 us), so there a device batch pays about 41 us more per
    // step and the switch moves to 0.8 (back below 0.6): at 2^20 disks 2000 FIRE steps from
    // random positions took 0.126-0.127 s against 0.139 s with the general thresholds.
    double rate = 0.0;
    bool dev = false;
    const bool fused_halts = fused_fire_batch(c, 2, fire);
    const double dev_up = fused_halts ? 0.8 : 0.5, dev_down = fused_halts ? 0.6 : 0.35;
    int rebuilds = read_run(c).rebuilds;
    try {
        while (!done && total < max_steps) {
            int lg = 0;
            const bool cell = c->method == DPMD_NEIGHBOR_CELL;
            dev = cell && (c->dev_rebuild == 2 || (c->dev_rebuild == 0 && rate > (dev ? dev_down : dev_up)));
            if (dev) m = std::max(m, 8);
            while ((2 << lg) <= std::min(std::min(m, c->max_batch), max_steps - total) && lg < 7) ++lg;
            const int bm = 1 << lg;
            if (c->rec_seg)   // recorded positions go to slots total + 1 ... of the segment buffer
                HIP_OK(hipMemcpyAsync(c->run.p + RUN_REC, &total, sizeof(int), hipMemcpyHostToDevice, c->st));
            const bool fused = !dev && fused_vv_batch(c, bm, fire);
            const bool fused_fire = !dev && fused_fire_batch(c, bm, fire);
            const int slot = lg + (dev ? 8 : 0);
            if (use_graph) {
                if (!graphs[slot]) graphs[slot] = capture_batch(c, bm, fire, pred, a, b, dev);
                HIP_OK(hipGraphLaunch(graphs[slot], c->st));
            } else {
                lean_batch(c, bm, fire, pred, a, b, dev);
            }
            RunCtl r = read_run(c);
            if (fire && c->n_act_sys > 1 && c->mask_frozen && r.running < c->n_act_sys && !r.halt) {
                // systems of a restricted run have stopped: once frozen they leave the launches,
                // and the last one left runs the fused FIRE step
                if (shrink_active(c)) {
                    if (cached) {
                        drop_cached_graphs(c);
                        c->gcache_key = batch_key(c, fire, pred, a, b);
                    } else {
                        for (auto& g : local) if (g) { HIP_OK(hipGraphExecDestroy(g)); g = nullptr; }
                    }
                }
            }
            {
                const int dn = r.rebuilds - rebuilds;
                rebuilds = r.rebuilds;
                int steps_b = bm, events = dn;
                if (r.halt) { steps_b = r.halt_at + 1; events += 1; }
                else if (fire && r.running <= 0) steps_b = r.conv_at + 1;
                rate = 0.5 * rate + 0.5 * (double)events / (double)std::max(1, steps_b);
            }
            if (c->pend_b) {   // profiled fused batch: keep its event pair if it ran to the end
                if (!r.halt) {
                    c->prof_ev.push_back(c->pend_b); c->prof_ev.push_back(c->pend_e); c->prof_n.push_back(c->pend_n);
                } else {
                    (void)hipEventDestroy(c->pend_b); (void)hipEventDestroy(c->pend_e);
                }
                c->pend_b = c->pend_e = nullptr;
            }
            if (r.halt) {
                total += r.halt_at + 1;
                since_halt += r.halt_at;
                if (fused)   // the halted step's positions are in the interleaved buffer P(halt_at mod 2)
                    disk_unpack_positions(c, r.halt_at % 2 == 0 ? c->xy0.p : c->xy1.p);
                if (fused_fire)   // and for FIRE its velocities in V(halt_at mod 2)
                    disk_unpack_pv(c, r.halt_at % 2 == 0 ? c->xy0.p : c->xy1.p,
                                   r.halt_at % 2 == 0 ? c->vxy0.p : c->vxy1.p);
                // clear the halt and the batch words (the rebuild clears RUN_ANY); the launches of
                // a fused FIRE batch raise only the run words, so its system is flagged here
                halt_reset(c, fused_fire);
                // finish the halted step: rebuild, forces, bookkeeping
                hipEvent_t rb0 = nullptr, rb1 = nullptr;
                if (c->prof == 4 && !c->capturing) {
                    HIP_OK(hipEventCreate(&rb0)); HIP_OK(hipEventCreate(&rb1));
                    HIP_OK(hipEventRecord(rb0, c->st));
                }
                try {
                    rebuild_neighbors(c, active_one(c), false, true);
                } catch (...) {
                    if (rb0) { (void)hipEventDestroy(rb0); (void)hipEventDestroy(rb1); }
                    throw;
                }
                if (rb0) {
                    HIP_OK(hipEventRecord(rb1, c->st));
                    c->prof_ev.push_back(rb0); c->prof_ev.push_back(rb1); c->prof_n.push_back(1);
                }
                if (fire) {
                    if (rigid(c)) rigid_fire_force(c, plain(c), 1, 1, pred, a, b);
                    else disk_fire_force(c, plain(c), 1, 1, pred, a, b);
                } else {
                    if (rigid(c)) rigid_vv_force(c, plain(c)); else disk_vv_force(c, plain(c));
                }
                r = read_run(c);
                if (fire && r.running <= 0) done = true;
                m = std::max(1, std::min(128, since_halt));
                interval = since_halt + 1;
                since_halt = 0;
            } else if (fire && r.running <= 0) {
                total += r.conv_at + 1;
                done = true;
            } else {
                total += bm;
                since_halt += bm;
                // steps to the expected next halt; up to 7 steps past it are allowed (a halted
                // step's tail costs less than another host round trip)
                const int rem = interval - since_halt;
                m = (interval > 0 && rem > 0) ? std::min(rem + 7, 2 * bm) : std::min(128, 2 * bm);
            }
        }
    } catch (...) {
        if (cached) drop_cached_graphs(c);
        else for (auto g : local) if (g) (void)hipGraphExecDestroy(g);
        throw;
    }
    if (!cached)
        for (auto g : local) if (g) HIP_OK(hipGraphExecDestroy(g));
    if (conv) *conv = (fire && read_run(c).running <= 0) ? 1 : 0;
    return total;
}

namespace {

// Runs up to max_steps device-driven steps; returns the number taken and (FIRE) whether all
// systems stopped.  FIRE runs of disk batches go to the persistent per-system engine.
int lean_run(dpmd_ctx* c, int max_steps, bool fire, int pred, double a, double b, bool use_graph, int* conv) {
    if (fire && fire_persist_ok(c)) return fire_persistent_run(c, max_steps, pred, a, b, use_graph, conv);
    return lean_run_batches(c, max_steps, fire, pred, a, b, use_graph, conv);
}


// ---------------------------------------------------------------------------------------
// Adjoints (disks).  Particle vectors at the C ABI are canonical-order device arrays.
// ---------------------------------------------------------------------------------------
void need_disk_cells(dpmd_ctx* c) {
    need(!rigid(c), "adjoints are implemented for disks");
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    need(c->method == DPMD_NEIGHBOR_CELL, "adjoints need the cell neighbor method");
}

void vv_run_impl(dpmd_ctx* c, int n, bool use_graph) {
    if (n <= 0) return;
    int any = 0;
    HIP_OK(hipMemcpyAsync(&any, c->run.p + RUN_ANY, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    int run[RUN_WORDS] = {0, n, 1, 0, 0, -1, -1, any, 0};
    c->run.upload(run, RUN_WORDS, c->st);
    (void)lean_run(c, n, false, 0, 0.0, 0.0, use_graph, nullptr);
    check_overflow(c);
}

void set_temperature_impl(dpmd_ctx* c, const double* T, bool remove_avg) {   // base_particle.hpp:449-462
    per_system_temperature(c);
    temperature_scale(c, T, c->scr_c.p);
    if (remove_avg) set_average_velocity_impl(c, nullptr, nullptr);
    scale_velocities_impl(c, c->scr_c.p);
}

// Rescale-thermostat NVT with affine compression.  Global step k = step0 + i: one
// velocity-Verlet step, then increment_packing_fraction(inc) if compress_every > 0 and
// k % compress_every == 0, then set_temperature(T) if temperature_every > 0 and
// k % temperature_every == 0.  The steps between two such events run as one device-driven
// batch run.  (The compression touches only positions, box and lists and the thermostat only
// velocities, so the two commute exactly.)
void nvt_rescale_impl(dpmd_ctx* c, int n, int step0, const double* inc, int cf, const double* T, int tf, bool rm,
                      bool use_graph) {
    need_state(c->vv_ready, "nvt_rescale_run: call vv_init first");
    need(n >= 0 && step0 >= 0 && cf >= 0 && tf >= 0, "nvt_rescale_run: negative step counts");
    need(cf == 0 || inc != nullptr, "nvt_rescale_run: compress_every > 0 needs phi_increment");
    need(tf == 0 || T != nullptr, "nvt_rescale_run: temperature_every > 0 needs temperature_target");
    GraphScope graphs(c);
    auto compress_at = [&](int k) { return cf > 0 && k % cf == 0; };
    auto thermo_at = [&](int k) { return tf > 0 && k % tf == 0; };
    int i = 0;
    while (i < n) {
        int j = i;
        while (j < n && !compress_at(step0 + j) && !thermo_at(step0 + j)) ++j;
        const int len = (j < n ? j + 1 : n) - i;
        vv_run_impl(c, len, use_graph);
        i += len;
        if (j < n) {
            const int k = step0 + j;
            if (compress_at(k)) increment_packing_fraction_impl(c, inc);
            if (thermo_at(k)) set_temperature_impl(c, T, rm);
        }
    }
}

void snapshot(dpmd_ctx* c, double* dst) {   // canonical x, y, vx, vy
    const int N = c->N;
    adj_to_canonical(c, c->x.p, dst);
    adj_to_canonical(c, c->y.p, dst + N);
    adj_to_canonical(c, c->vx.p, dst + 2 * (size_t)N);
    adj_to_canonical(c, c->vy.p, dst + 3 * (size_t)N);
}

void restore_positions(dpmd_ctx* c, const double* src) {   // canonical x, y -> current; rebuild
    const int N = c->N;
    adj_to_current(c, src, c->x.p);
    adj_to_current(c, src + N, c->y.p);
    check_neighbors_impl(c, true);
}

void restore_snapshot(dpmd_ctx* c, const double* src) {
    const int N = c->N;
    adj_to_current(c, src + 2 * (size_t)N, c->vx.p);
    adj_to_current(c, src + 3 * (size_t)N, c->vy.p);
    restore_positions(c, src);
    compute_forces_impl(c, c->vv_mode);
}

struct RecordTo {
    dpmd_ctx* c;
    RecordTo(dpmd_ctx* ctx, double* seg) : c(ctx) { c->rec_seg = seg; }
    ~RecordTo() { c->rec_seg = nullptr; }
};

void vv_checkpointed_impl(dpmd_ctx* c, int n, int every, bool use_graph) {
    need_disk_cells(c);
    need_state(c->vv_ready, "vv_run_checkpointed: call vv_init first");
    need(!c->vv_damped, "vv_run_checkpointed: the reverse mode covers undamped velocity Verlet");
    need(n >= 0 && every > 0, "vv_run_checkpointed: n_steps >= 0 and every > 0 required");
    const int N = c->N;
    const int J = (n + every - 1) / every;
    GraphScope graphs(c);
    c->ck_state.resize((int64_t)(J + 1) * 4 * N);
    snapshot(c, c->ck_state.p);
    for (int j = 0; j < J; ++j) {
        const int len = std::min(every, n - j * every);
        vv_run_impl(c, len, use_graph);
        snapshot(c, c->ck_state.p + (size_t)(j + 1) * 4 * N);
    }
    c->ck_steps = n; c->ck_every = every; c->ck_count = J + 1;
}

// Reverse mode through the checkpointed run: cotangents of the final (x, y, vx, vy) ->
// cotangents of the initial ones.  Per reverse step one Hessian-vector product:
//   vbar += dt xbar;  z = H(x_k) (vbar / m);  xbar -= dt z  (dt / 2 at k = 0 and, before the
//   loop, at k = T with vbar = the final cotangent).
// Each segment is replayed from its checkpoint with positions recorded per step; the
// Hessian products use one canonical-order contact list per segment, built at the segment's
// first configuration with a cutoff covering every displacement in the segment.
void vv_backprop_impl(dpmd_ctx* c, const double* gx, const double* gy, const double* gvx, const double* gvy, double* ox,
                      double* oy, double* ovx, double* ovy, bool use_graph) {
    need_disk_cells(c);
    need(!c->vv_damped, "vv_backprop: the reverse mode covers undamped velocity Verlet");
    need_state(c->ck_count > 0, "vv_backprop: run vv_run_checkpointed first");
    GraphScope graphs(c);
    const int N = c->N, S = c->S, n = c->ck_steps, every = c->ck_every, J = c->ck_count - 1;
    const int mode = c->vv_mode;
    hipStream_t st = c->st;
    c->adj_t0.resize(2 * (int64_t)N); c->adj_t1.resize(2 * (int64_t)N); c->adj_t2.resize(2 * (int64_t)N);
    c->adj_bx.resize(N);
    c->adj_u0.resize(S);
    // x-bar, v-bar and w in canonical order between windows
    double *xb = c->adj_t0.p, *vb = c->adj_t1.p, *w_c = c->adj_t2.p;
    auto in = [&](const double* src, double* dst) {
        if (src) HIP_OK(hipMemcpyAsync(dst, src, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
        else fill_f64(st, dst, N, 0.0);
    };
    in(gx, xb); in(gy, xb + N); in(gvx, vb); in(gvy, vb + N);
    c->ck_seg.resize((int64_t)(every + 1) * 2 * N);
    c->ck_cur.resize((int64_t)(every + 1) * 2 * N);
    c->adj_cur.resize(8 * (int64_t)N);
    c->ck_nn.resize(N);
    c->ck_nbr.resize((int64_t)N * c->K);
    double *xc = c->adj_cur.p, *vc = xc + 2 * (size_t)N, *wA = xc + 4 * (size_t)N, *wB = xc + 6 * (size_t)N;
    std::vector<double> cut0(S), d2(S), cut(S), Lx(S), Ly(S), rmax(S, 0.0);
    HIP_OK(hipMemcpyAsync(cut0.data(), c->cutoff.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(Lx.data(), c->Lx.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(Ly.data(), c->Ly.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
    {
        const auto rad = dl(c, c->rad);
        for (int s = 0; s < S; ++s)
            for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) rmax[s] = std::max(rmax[s], rad[i]);
    }
    // captured reverse windows, destroyed once the stream has drained
    struct RevGraphs {
        hipStream_t st;
        std::map<std::array<int64_t, 6>, hipGraphExec_t> g;
        ~RevGraphs() {
            (void)hipStreamSynchronize(st);
            for (auto& kv : g) (void)hipGraphExecDestroy(kv.second);
        }
    } rev_graphs{st, {}};
    bool first = true;
    for (int j = J - 1; j >= 0; --j) {
        const int k0 = j * every, k1 = std::min(k0 + every, n), len = k1 - k0;
        const double* ck = c->ck_state.p + (size_t)j * 4 * N;
        restore_snapshot(c, ck);
        HIP_OK(hipMemcpyAsync(c->ck_seg.p, ck, (size_t)2 * N * 8, hipMemcpyDeviceToDevice, st));
        {
            RecordTo rec(c, c->ck_seg.p);
            vv_run_impl(c, len, use_graph);
        }
        // Reverse sweep over the segment in windows [a, hi]: one canonical contact list per
        // window, built at x_a with a cutoff covering every displacement in the window; a
        // window whose enlarged cutoff would leave fewer than 3 cells per side, or exceed the
        // cutoff by more than the largest radius (neighbor capacity), is halved.
        int hi = k1;
        while (hi > k0 || first) {
            int w = hi - k0, a = hi - w;
            bool enlarge = false;
            for (;;) {
                adj_seg_disp2(c, a - k0, hi - a, c->adj_bx.p);
                sys_reduce_max(c, c->adj_bx.p, c->adj_u0.p);
                HIP_OK(hipMemcpyAsync(d2.data(), c->adj_u0.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                bool fits = true;
                enlarge = false;
                for (int s = 0; s < S; ++s) {
                    const double span = 2.0 * std::sqrt(d2[s]);
                    cut[s] = cut0[s];
                    if (span >= cut0[s]) {
                        cut[s] = span * (1.0 + 1e-9) + 1e-12;
                        enlarge = true;
                        const double need_sz = 2.0 * rmax[s] + cut[s];
                        fits = fits && std::floor(Lx[s] / need_sz) >= 3 && std::floor(Ly[s] / need_sz) >= 3 &&
                               cut[s] <= cut0[s] + rmax[s];
                    }
                }
                if (fits || hi - a <= 1) break;
                w = std::max(1, (hi - a) / 2);
                a = hi - w;
            }
            if (enlarge) {
                c->cutoff.upload(cut.data(), S, st);
                setup_cells(c, CELLS_BOX_CHANGED);
            }
            restore_positions(c, c->ck_seg.p + (size_t)(a - k0) * 2 * N);
            check_overflow(c);
            // The window runs in the context's current order (the cells of x_a): its contact
            // list as built, and the recorded positions of slots a .. hi and the adjoint
            // vectors permuted into that order, so that a particle's neighbors lie near it in
            // memory.  Every per-particle sum runs over the same list in the same order as in
            // canonical order, so the results are the same bits.
            HIP_OK(hipMemcpyAsync(c->ck_nn.p, c->nn.p, (size_t)N * 4, hipMemcpyDeviceToDevice, st));
            HIP_OK(hipMemcpyAsync(c->ck_nbr.p, c->nbr.p, (size_t)N * c->K * 4, hipMemcpyDeviceToDevice, st));
            adj_rows_to_current(c, c->ck_seg.p + (size_t)(a - k0) * 2 * N, c->ck_cur.p + (size_t)(a - k0) * 2 * N,
                                2 * (hi - a + 1));
            adj_rows_to_current(c, xb, xc, 2);
            adj_rows_to_current(c, vb, vc, 2);
            adj_rows_to_current(c, w_c, wA, 2);
            if (enlarge) {
                c->cutoff.upload(cut0.data(), S, st);
                setup_cells(c, CELLS_BOX_CHANGED);
            }
            const double* rad_w = c->rad.p;
            const double* m_w = c->mass.p;
            double* wc = wA;
            double* wn = wB;
            // Each launch is one reverse step (Hessian product, xbar update) plus the next
            // step's vbar and w, which need only per-particle data; w ping-pongs between
            // two buffers because the product reads the neighbors' w.
            if (first) {
                adj_rev_w(c, m_w, vc, wc);
                adj_rev_step(c, mode, hi - k0, rad_w, m_w, wc, 0.5, xc, vc, wn, n > 0, c->ck_cur.p);
                std::swap(wc, wn);
                first = false;
            }
            auto steps = [&](double* w0, double* w1) {
                for (int k = hi - 1; k >= a; --k) {
                    adj_rev_step(c, mode, k - k0, rad_w, m_w, w0, k > 0 ? 1.0 : 0.5, xc, vc, w1, k > 0, c->ck_cur.p);
                    std::swap(w0, w1);
                }
            };
            if (use_graph && hi - a >= 4) {
                // Every argument of a window's launches is fixed by its first and last slot in
                // the segment, whether it ends at step 0 and the w buffer it starts from (the
                // other buffers are the context's), so windows of one shape (the full segments)
                // replay one captured graph instead of launching step by step.
                const std::array<int64_t, 6> key = {hi - k0, a - k0, a == 0 ? 1 : 0, (int64_t)(intptr_t)wc,
                                                    (int64_t)(intptr_t)c->ck_nn.p, (int64_t)(intptr_t)c->ck_cur.p};
                auto it = rev_graphs.g.find(key);
                if (it == rev_graphs.g.end()) {
                    hipGraph_t gr = nullptr;
                    HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                    c->capturing = true;
                    try {
                        steps(wc, wn);
                    } catch (...) {
                        c->capturing = false;
                        (void)hipStreamEndCapture(st, &gr);
                        if (gr) (void)hipGraphDestroy(gr);
                        throw;
                    }
                    c->capturing = false;
                    HIP_OK(hipStreamEndCapture(st, &gr));
                    hipGraphExec_t ex = nullptr;
                    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(gr);
                    if (e != hipSuccess) throw Error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                    it = rev_graphs.g.emplace(key, ex).first;
                }
                HIP_OK(hipGraphLaunch(it->second, st));
            } else {
                steps(wc, wn);
            }
            if ((hi - a) % 2) std::swap(wc, wn);
            // back to canonical order for the next window (another cell order)
            adj_rows_to_canonical(c, xc, xb, 2);
            adj_rows_to_canonical(c, vc, vb, 2);
            adj_rows_to_canonical(c, wc, w_c, 2);
            hi = a;
        }
    }
    auto out = [&](const double* src, double* dst) {
        if (dst) HIP_OK(hipMemcpyAsync(dst, src, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
    };
    out(xb, ox); out(xb + N, oy); out(vb, ovx); out(vb + N, ovy);
    restore_snapshot(c, c->ck_state.p + (size_t)J * 4 * N);   // leave the final state
    HIP_OK(hipStreamSynchronize(st));
}

// d P_s / d(radii, box) through the minimum: dP = dP/dq + lambda^T dF/dq with H lambda = dP/dx.
void adjoint_pressure_impl(dpmd_ctx* c, int mode, double tol, int max_iter, double* dr, double* dLx, double* dLy,
                           int* iters, double* resid) {
    need_disk_cells(c);
    const int N = c->N, S = c->S;
    hipStream_t st = c->st;
    c->adj_t0.resize(2 * (int64_t)N); c->adj_t1.resize(N); c->adj_t2.resize(2 * (int64_t)N); c->adj_t3.resize(N);
    for (auto* v : {&c->adj_u0, &c->adj_u1, &c->adj_u2, &c->adj_u3, &c->adj_u4}) v->resize(S);
    fill_f64(st, c->adj_u0.p, S, 1.0);
    double *b = c->adj_t0.p, *lam = c->adj_t2.p;
    adj_vjp_current(c, mode, nullptr, c->adj_u0.p, nullptr, nullptr, b, b + N, c->adj_t1.p, c->adj_u1.p, c->adj_u2.p);
    const int it = adj_cg_current(c, mode, b, b + N, lam, lam + N, tol, max_iter, resid);
    // lambda lives on the particles with contacts (the others are null directions of H)
    adj_vjp_current(c, mode, nullptr, nullptr, lam, lam + N, nullptr, nullptr, c->adj_t3.p, c->adj_u3.p, c->adj_u4.p,
                    c->adj_live.p);
    adj_add(c, N, c->adj_t1.p, c->adj_t3.p, c->adj_t3.p);
    if (dr) adj_to_canonical(c, c->adj_t3.p, dr);
    if (dLx) adj_add(c, S, c->adj_u1.p, c->adj_u3.p, dLx);
    if (dLy) adj_add(c, S, c->adj_u2.p, c->adj_u4.p, dLy);
    HIP_OK(hipStreamSynchronize(st));
    if (iters) *iters = it;
}

// Levenberg-Marquardt Newton iterations x <- x + (H + mu I)^-1 F (MINRES, every particle and
// contact, mu = kNewtonShift * k) from a configuration near a minimum, until max|F| <= ftol.
// The shift makes the solve definite (rattlers, particles touching at the FIRE tolerance)
// while real modes converge quadratically.  Steps are damped to a largest displacement of
// kNewtonMaxStep; a solve that does not converge is an error.  Returns the final max|F|.
constexpr double kNewtonMaxStep = 0.05;
constexpr double kNewtonShift = 1e-10;
double newton_impl(dpmd_ctx* c, int mode, int max_newton, double ftol, double cg_tol, int cg_max) {
    need_disk_cells(c);
    const int N = c->N, S = c->S;
    c->adj_t0.resize(2 * (int64_t)N);
    c->adj_bx.resize(N);
    c->adj_u0.resize(S);
    std::vector<double> mf(S);
    std::vector<double> ek(S);
    HIP_OK(hipMemcpyAsync(ek.data(), c->e.p, (size_t)S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    double kmax = 0.0;
    for (double v : ek) kmax = std::max(kmax, v);
    auto maxf = [&]() {
        compute_forces_impl(c, mode);
        sys_reduce(c, RED_MAXF2, c->maxf2.p);
        HIP_OK(hipMemcpyAsync(mf.data(), c->maxf2.p, (size_t)S * 8, hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        double m = 0.0;
        for (double v : mf) m = std::max(m, v);
        return std::sqrt(m);
    };
    double f = maxf();
    for (int it = 0; it < max_newton && f > ftol; ++it) {
        double res = 0.0;
        (void)adj_cg_current(c, mode, c->fx.p, c->fy.p, c->adj_t0.p, c->adj_t0.p + N, cg_tol, cg_max, &res, false,
                             kNewtonShift * kmax);
        if (!(res <= 1e-6))
            throw StateError("newton_minimize: the Hessian solve did not converge (relative residual " +
                             std::to_string(res) + "); run FIRE closer to the minimum first");
        adj_norm2(c, c->adj_t0.p, c->adj_t0.p + N, c->adj_bx.p);
        sys_reduce_max(c, c->adj_bx.p, c->adj_u0.p);
        HIP_OK(hipMemcpyAsync(mf.data(), c->adj_u0.p, (size_t)S * 8, hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        double step = 0.0;
        for (double v : mf) step = std::max(step, std::sqrt(v));
        const double scale = step > kNewtonMaxStep ? kNewtonMaxStep / step : 1.0;
        adj_axpy2(c, scale, c->adj_t0.p, c->adj_t0.p + N, c->x.p, c->y.p);
        check_neighbors_impl(c, true);
        f = maxf();
    }
    return f;
}

// save_state / restore_state (base_particle.hpp:708-736; disk.cu:348-432,1158-1225)
void save_state_impl(dpmd_ctx* c, const int* flag, int tv) {
    const int N = c->N, S = c->S;
    if (!c->saved) {
        dpmd::DVec<double>* dv[] = {&c->sv_x, &c->sv_y, &c->sv_rad, &c->sv_mass, &c->sv_angle};
        for (auto* v : dv) { v->resize(N); v->zero(c->st); }
        c->sv_static.resize(N); c->sv_static.zero(c->st);
        c->sv_Lx.resize(S); c->sv_Ly.resize(S);
        if (rigid(c)) {
            const int Nv = c->Nv;
            dpmd::DVec<double>* vv[] = {&c->sv_vpx, &c->sv_vpy, &c->sv_vrad, &c->sv_vmass};
            for (auto* v : vv) { v->resize(Nv); v->zero(c->st); }
            c->sv_vpid.resize(Nv); c->sv_vpid.zero(c->st);
            c->sv_vstatic.resize(Nv); c->sv_vstatic.zero(c->st);
        }
        c->saved = true;
    }
    save_restore(c, flag, tv, false);
}
void restore_state_impl(dpmd_ctx* c, const int* flag, int tv) {
    need_state(c->saved, "restore_state: nothing has been saved");
    save_restore(c, flag, tv, true);
    box_inv(c);
    check_neighbors_impl(c, true);
}

// minimize_fire[_wall]_copy_dt (minimizers.cuh:34-58,85-121): FIRE until every system stopped
// or max_steps, stopping test converged(avg_pe_target, avg_pe_diff_target) as the reference
// passes it (the arguments land swapped in converged(diff, target), fire.cuh:98); returns the
// final dt per system and zeroes the velocities.
std::vector<double> minimize_copy_dt(dpmd_ctx* c, const std::vector<double>& dt, int max_steps, double pe_target,
                                     double pe_diff_target, int mode, bool use_graph) {
    const int S = c->S;
    c->scr_a.upload(dt.data(), S, c->st);
    fire_init_impl(c, c->scr_a.p, mode);
    if (max_steps > 0) {
        FrozenMask mask(c);
        (void)lean_run(c, max_steps, true, DPMD_FIRE_PRED_REFERENCE, pe_target, pe_diff_target, use_graph, nullptr);
    }
    check_overflow(c);
    std::vector<double> out = dl(c, c->f_dt);
    zero_scale(c);
    scale_velocities_impl(c, c->scr_c.p);
    return out;
}

// jam_binary_search_pbc / _wall (jammers.cuh:60-256; kernels jammers.cu:7-92).  The per-system
// scalar kernels (remove_badly_initialized_systems, reinit_dt, jamming_update, scale_box) act
// on S values and run on the host here, in the same arithmetic; every particle operation runs
// on the device.  As in the reference, the box is changed and the lists rebuilt before the
// positions are scaled (the next minimisation's skin test catches the displacement).
// action: -1 unjammed (state saved), 1 jammed (reverted), 0 done; phi: the bisection's phi.
void jam_bisect_impl(dpmd_ctx* c, const double* dt, int max_comp, int max_min, double pe_t, double pe_diff_t,
                     double dphi, double phi_tol, int mode, bool use_graph, int* action_out, double* phi_out,
                     int* steps_out) {
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    need(dt != nullptr && max_comp >= 0 && max_min >= 0, "jam_binary_search: bad arguments");
    const int S = c->S;
    std::vector<double> dt_copy(S);
    HIP_OK(hipMemcpyAsync(dt_copy.data(), dt, S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    std::vector<int> action(S, -1);
    dpmd::DVec<int> d_action;
    d_action.resize(S);
    auto push_action = [&] { d_action.upload(action.data(), S, c->st); };
    if (mode != DPMD_FORCE_WALL) {   // jam_binary_search_pbc zeroes the velocities first (jammers.cuh:85)
        zero_scale(c);
        scale_velocities_impl(c, c->scr_c.p);
    }
    std::vector<double> dtf = minimize_copy_dt(c, dt_copy, max_min, pe_t, pe_diff_t, mode, use_graph);
    for (int s = 0; s < S; ++s) if (dtf[s] != 0.0) action[s] = 0;   // remove_badly_initialized_systems_kernel
    push_action();
    save_state_impl(c, d_action.p, -1);
    sys_reduce(c, RED_PHI, c->phi.p);
    std::vector<double> phi = dl(c, c->phi), phi_low = phi, phi_high(S, -1.0);
    std::vector<double> dtw(S), sf(S), area(S);
    int step = 0;
    while (step < max_comp) {
        for (int s = 0; s < S; ++s) dtw[s] = action[s] != 0 ? dt_copy[s] : 0.0;   // reinit_dt_kernel
        (void)minimize_copy_dt(c, dtw, max_min, pe_t, pe_diff_t, mode, use_graph);
        sys_reduce(c, RED_PE, c->pe_tot.p);
        std::vector<double> pe = dl(c, c->pe_tot);
        for (int s = 0; s < S; ++s) {   // jamming_update_kernel
            if (action[s] == 0) continue;
            action[s] = -2;
            const int dof = c->h_off[s + 1] - c->h_off[s];
            const double avg_pe = pe[s] / dof;
            if (avg_pe > pe_t) {
                phi_high[s] = phi[s];
                phi[s] = (phi_high[s] + phi_low[s]) / 2.0;
                action[s] = 1;
            } else {
                action[s] = -1;
                phi_low[s] = phi[s];
                if (phi_high[s] > 0) phi[s] = (phi_high[s] + phi_low[s]) / 2.0;
                else phi[s] += dphi;
            }
            if (std::abs(phi_high[s] / phi_low[s] - 1) < phi_tol && phi_high[s] > 0) action[s] = 0;
        }
        push_action();
        save_state_impl(c, d_action.p, -1);
        restore_state_impl(c, d_action.p, 1);
        sys_reduce_arr(c, c->area.p, c->scr_a.p);   // compute_particle_area_total
        std::vector<double> A(S);
        HIP_OK(hipMemcpyAsync(A.data(), c->scr_a.p, S * 8, hipMemcpyDeviceToHost, c->st));
        std::vector<double> Lx = dl(c, c->Lx), Ly = dl(c, c->Ly);
        for (int s = 0; s < S; ++s) {   // scale_box_kernel
            if (action[s] == 0 || phi[s] <= 0.0) { sf[s] = 1.0; continue; }
            const double L = sqrt(A[s] / phi[s]);
            sf[s] = L / Lx[s];
            Lx[s] = L; Ly[s] = L;
        }
        c->Lx.upload(Lx.data(), S, c->st);
        c->Ly.upload(Ly.data(), S, c->st);
        box_inv(c);
        if (c->method == DPMD_NEIGHBOR_CELL) {
            setup_cells(c, CELLS_BOX_CHANGED);
            alloc_obj_buffers(c);
        }
        check_neighbors_impl(c, true);
        sys_reduce(c, RED_PHI, c->phi.p);
        c->scr_b.upload(sf.data(), S, c->st);
        scale_positions_impl(c, c->scr_b.p);
        bool done = true;
        for (int s = 0; s < S; ++s) done = done && action[s] == 0;
        if (done) break;
        ++step;
    }
    compute_forces_impl(c, mode);
    sys_reduce(c, RED_PE, c->pe_tot.p);
    sys_reduce(c, RED_PHI, c->phi.p);
    HIP_OK(hipStreamSynchronize(c->st));
    for (int s = 0; s < S; ++s) {
        if (action_out) action_out[s] = action[s];
        if (phi_out) phi_out[s] = phi[s];
    }
    if (steps_out) *steps_out = step;
}

// damping == nullptr: velocity Verlet (velocity_verlet.cuh:16-32); otherwise damped velocity
// Verlet (damped_velocity_verlet.cuh:16-56), whose init computes the forces without the drag.
void vv_init_impl(dpmd_ctx* c, const double* dt, int mode, const double* damping = nullptr) {
    const int S = c->S;
    std::vector<double> h(S), hh(S);
    HIP_OK(hipMemcpyAsync(h.data(), dt, S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    for (int s = 0; s < S; ++s) hh[s] = h[s] * 0.5;   // dt_half = 0.5 * dt (velocity_verlet.cuh:28-29)
    c->vv_dt.upload(h.data(), S, c->st);
    c->vv_dth.upload(hh.data(), S, c->st);
    c->vv_damped = damping != nullptr;
    if (damping) {
        c->vv_damp.resize(S);
        HIP_OK(hipMemcpyAsync(c->vv_damp.p, damping, S * 8, hipMemcpyDeviceToDevice, c->st));
    }
    c->vv_mode = mode;
    compute_forces_impl(c, mode);
    c->vv_ready = true;
}

}  // namespace

std::string& dpmd::global_error() {
    static std::string e;
    return e;
}

// ---------------------------------------------------------------------------------------
// Autotuner engine.  A configuration sets, for every system s, the neighbor cutoff (the
// Verlet skin) to skin * 2 r_min,s with the rebuild threshold (cutoff / 2)^2, and a cell
// grid whose side is at least cell_factor * (2 r_max,s + cutoff) -- the two preconditions of
// the periodic 3x3 stencil (a grid of at least 3 x 3 cells, cells no smaller than the
// interaction range) are checked here for every system -- plus the execution options.
// ---------------------------------------------------------------------------------------
namespace dpmd {

void tune_get(dpmd_ctx* c, dpmd_tune_config* t) {
    need(!rigid(c) && c->method == DPMD_NEIGHBOR_CELL, "the autotuner handles disk systems with cell lists");
    auto rad = dl(c, c->rad);
    auto cut = dl(c, c->cutoff);
    auto Lx = dl(c, c->Lx), Ly = dl(c, c->Ly);
    auto cdx = dl(c, c->cdx), cdy = dl(c, c->cdy);
    double skin = 0.0, factor = 1e300;
    for (int s = 0; s < c->S; ++s) {
        double rmin = 1e300, rmax = 0.0;
        for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) { rmin = std::min(rmin, rad[i]); rmax = std::max(rmax, rad[i]); }
        skin = std::max(skin, cut[s] / (2.0 * rmin));
        const double need_sz = 2.0 * rmax + cut[s];
        factor = std::min(factor, std::min(Lx[s] / cdx[s], Ly[s] / cdy[s]) / need_sz);
    }
    t->skin = skin;
    t->cell_factor = factor;
    t->block = c->block;
    t->max_batch = c->max_batch;
    t->fused_vv = c->fuse_vv ? 1 : 0;
}

void tune_apply(dpmd_ctx* c, const dpmd_tune_config& t) {
    need(!rigid(c) && c->method == DPMD_NEIGHBOR_CELL, "the autotuner handles disk systems with cell lists");
    need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
    need(t.skin > 0.0 && std::isfinite(t.skin), "tune config: skin must be positive");
    need(t.cell_factor >= 1.0 && std::isfinite(t.cell_factor), "tune config: cell_factor must be >= 1");
    need(t.block == 64 || t.block == 128 || t.block == 256 || t.block == 512, "tune config: block must be 64, 128, 256 or 512");
    need(t.max_batch >= 1 && t.max_batch <= 128 && (t.max_batch & (t.max_batch - 1)) == 0,
         "tune config: max_batch must be a power of two from 1 to 128");
    const int S = c->S;
    auto rad = dl(c, c->rad);
    auto Lx = dl(c, c->Lx), Ly = dl(c, c->Ly);
    std::vector<double> cut(S), th2(S);
    std::vector<int> nx(S), ny(S);
    for (int s = 0; s < S; ++s) {
        double rmin = 1e300, rmax = 0.0;
        for (int i = c->h_off[s]; i < c->h_off[s + 1]; ++i) { rmin = std::min(rmin, rad[i]); rmax = std::max(rmax, rad[i]); }
        cut[s] = t.skin * 2.0 * rmin;
        th2[s] = (0.5 * cut[s]) * (0.5 * cut[s]);
        const int rc = dpmd_cell_grid(Lx[s], Ly[s], rmax, cut[s], t.cell_factor, &nx[s], &ny[s]);
        if (rc != DPMD_OK)
            throw ArgError("tune config: system " + std::to_string(s) + ": " + global_error());
    }
    c->cutoff.upload(cut.data(), S, c->st);
    c->thresh2.upload(th2.data(), S, c->st);
    c->cdx.upload(nx.data(), S, c->st);
    c->cdy.upload(ny.data(), S, c->st);
    setup_cells(c, CELLS_SYNC);
    alloc_obj_buffers(c);
    check_neighbors_impl(c, true);
    check_overflow(c);
    c->block = t.block;
    c->max_batch = t.max_batch;
    c->fuse_vv = t.fused_vv != 0;
}

void tune_snapshot(dpmd_ctx* c, double* dst) { snapshot(c, dst); }

void tune_restore_order(dpmd_ctx* c, const std::vector<int>& si0) {
    const int n = c->N;
    need((int)si0.size() == n, "tune_restore_order: wrong length");
    const auto si = dl(c, c->static_index);
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[si0[k]] = si[k];
    permute_particles(c, order.data());
}

void tune_restore(dpmd_ctx* c, const double* src, int mode) {
    const int N = c->N;
    adj_to_current(c, src + 2 * (size_t)N, c->vx.p);
    adj_to_current(c, src + 3 * (size_t)N, c->vy.p);
    restore_positions(c, src);
    compute_forces_impl(c, mode);
}

int64_t tune_run(dpmd_ctx* c, int integrator, const double* dt, int mode, int steps) {
    need(integrator == DPMD_TUNE_FIRE || integrator == DPMD_TUNE_NVE, "unknown tuning integrator");
    const int S = c->S;
    std::vector<double> h(S);
    HIP_OK(hipMemcpyAsync(h.data(), dt, S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    int64_t live = 0;
    for (int s = 0; s < S; ++s)
        if (h[s] != 0.0) live += c->h_off[s + 1] - c->h_off[s];
    need(live > 0, "tuning run: every dt is zero");
    if (integrator == DPMD_TUNE_FIRE) {
        fire_init_impl(c, dt, mode);
        FrozenMask mask(c);
        const int done = lean_run(c, steps, true, DPMD_FIRE_PRED_MAX_FORCE, 0.0, 0.0, true, nullptr);
        check_overflow(c);
        return live * done;
    }
    vv_init_impl(c, dt, mode);
    vv_run_impl(c, steps, true);
    return (int64_t)c->N * steps;
}

}  // namespace dpmd

extern "C" {

const char* dpmd_version(void) { return "dpmd_amd 0.1 (gfx950, fp64)"; }

const char* dpmd_last_error(const dpmd_ctx* c) { return c ? c->err.c_str() : global_error().c_str(); }

// ---- scratch of the entry point that allocates nothing ------------------------------------
// dpmd_adjoint_hvp_noalloc keeps the current-order input and output vectors (4 x N doubles)
// in this file-scope device array instead of context buffers, so it calls no allocator.
// Each GPU the library runs on holds its own copy (32 MiB) until the process ends; every
// call on a GPU shares it, so calls that overlap on one GPU race.
constexpr int64_t kNoAllocMaxN = 1 << 20;
static __device__ __align__(256) double scratch[4 * kNoAllocMaxN];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};   // device address of `scratch` per device

// The device address of `scratch` on the current device (the array's name is a host-side
// stand-in).  Resolved once per device: at context creation, else at the first call.
double* scratch_ptr() {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    need(dev >= 0 && dev < kMaxDevices, "scratch: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        need(p != nullptr, "scratch: hipGetSymbolAddress returned a null address");
        g_scratch_dev[dev] = (double*)p;
    }
    return g_scratch_dev[dev];
}
}  // namespace

int dpmd_ctx_create(dpmd_ctx** out, int kind, void* stream) {
    return guard(nullptr, [&] {
        need(out != nullptr, "dpmd_ctx_create: out is NULL");
        need(kind == DPMD_DISK || kind == DPMD_RIGID_BUMPY, "dpmd_ctx_create: unknown kind");
        dpmd_ctx* c = new dpmd_ctx();
        c->kind = kind;
        if (kind == DPMD_RIGID_BUMPY) c->K = 48;
        try {
            if (stream) c->st = (hipStream_t)stream;
            else HIP_OK(hipStreamCreate(&c->st));
            c->one.resize(1);
            int one = 1;
            c->one.upload(&one, 1, c->st);
            c->run.resize(RUN_WORDS); c->run.zero(c->st);
            c->f_snap.resize(4); c->f_snap.zero(c->st);
            c->i_snap.resize(4); c->i_snap.zero(c->st);
            c->overflow.resize(1); c->overflow.zero(c->st);
            (void)scratch_ptr();   // set-up of the allocation-free entry point on this device
        } catch (...) { delete c; throw; }
        c->err.clear();
        *out = c;
    });
}

int dpmd_ctx_destroy(dpmd_ctx* c) {
    if (!c) return DPMD_OK;
    (void)hipStreamSynchronize(c->st);
    if (c->g_fire) (void)hipGraphExecDestroy(c->g_fire);
    if (c->g_vv) (void)hipGraphExecDestroy(c->g_vv);
    // Free the large buffers first; every DVec member also frees itself when c is deleted.
    dpmd::DVec<double>* dv[] = {&c->Lx, &c->Ly, &c->iLx, &c->iLy, &c->e, &c->cutoff, &c->thresh2, &c->pe_tot,
                                &c->ke_tot, &c->pw_tot, &c->press, &c->temp, &c->phi, &c->maxf2, &c->st_xx,
                                &c->st_xy, &c->st_yx, &c->st_yy, &c->x, &c->y, &c->vx, &c->vy, &c->fx, &c->fy,
                                &c->pe, &c->ke, &c->rad, &c->mass, &c->area, &c->lx, &c->ly, &c->angle, &c->avel,
                                &c->torque, &c->inertia, &c->sxx, &c->sxy, &c->syx, &c->syy, &c->xy0, &c->xy1,
                                &c->vxy0, &c->vxy1, &c->f_snap, &c->vpx, &c->vpy,
                                &c->vfx, &c->vfy, &c->vpe, &c->vrad, &c->vmass, &c->tmp_d, &c->part, &c->f_dt,
                                &c->f_dtmax, &c->f_dtmin, &c->f_alpha, &c->f_peprev, &c->s_dt, &c->s_alpha,
                                &c->s_dtrev, &c->s_vscale, &c->vv_dt, &c->vv_dth, &c->sv_x, &c->sv_y, &c->sv_rad,
                                &c->sv_mass, &c->sv_Lx, &c->sv_Ly, &c->sv_angle, &c->sv_vx, &c->sv_vy, &c->scr_a,
                                &c->scr_b, &c->scr_c};
    for (auto* v : dv) v->release();
    dpmd::DVec<int>* iv[] = {&c->off, &c->size, &c->cdx, &c->cdy, &c->cstart, &c->boundary, &c->ndof, &c->voff,
                             &c->sid, &c->static_index, &c->vpid, &c->poff, &c->nvpp, &c->vsid, &c->cell_id,
                             &c->cell_count, &c->cell_start, &c->cell_wptr, &c->cell_sys, &c->order,
                             &c->order_inv, &c->tmp_i, &c->nn, &c->nbr, &c->overflow,
                             &c->scan_part, &c->chunk_sys, &c->chunk_beg, &c->sys_chunk, &c->f_ngood,
                             &c->f_nbad, &c->f_stopped, &c->f_frozen, &c->f_nsteps, &c->s_ngood, &c->s_nbad, &c->run, &c->one, &c->sv_static, &c->i_snap,
                             &c->persist_order, &c->persist_todo, &c->persist_stop};
    for (auto* v : iv) v->release();
    c->flag.release();
    delete c;
    return DPMD_OK;
}

int dpmd_allocate_systems(dpmd_ctx* c, int S) {
    return guard(c, [&] {
        need(S > 0, "allocate_systems: n_systems must be positive");
        c->S = S;
        dpmd::DVec<double>* dv[] = {&c->Lx, &c->Ly, &c->iLx, &c->iLy, &c->e, &c->cutoff, &c->thresh2, &c->pe_tot,
                                    &c->ke_tot, &c->pw_tot, &c->press, &c->temp, &c->phi, &c->maxf2, &c->st_xx,
                                    &c->st_xy, &c->st_yx, &c->st_yy, &c->scr_a, &c->scr_b, &c->scr_c,
                                    &c->f_dt, &c->f_dtmax, &c->f_dtmin, &c->f_alpha, &c->f_peprev, &c->s_dt,
                                    &c->s_alpha, &c->s_dtrev, &c->s_vscale, &c->vv_dt, &c->vv_dth};
        for (auto* v : dv) { v->resize(S); v->zero(c->st); }
        dpmd::DVec<int>* iv[] = {&c->size, &c->cdx, &c->cdy, &c->boundary, &c->ndof, &c->f_ngood, &c->f_nbad,
                                 &c->f_stopped, &c->f_frozen, &c->f_nsteps, &c->s_ngood, &c->s_nbad,
                                 &c->ncontacts};
        for (auto* v : iv) { v->resize(S); v->zero(c->st); }
        c->off.resize(S + 1); c->off.zero(c->st);
        c->cstart.resize(S + 1); c->cstart.zero(c->st);
        c->voff.resize(S + 1); c->voff.zero(c->st);
        c->flag.resize(S); c->flag.zero(c->st);
        c->synced = false; c->nbr_ready = false; c->fire_ready = false; c->vv_ready = false;
        c->saved = false;   // a saved state no longer matches the arrays
    });
}

int dpmd_allocate_particles(dpmd_ctx* c, int N) {
    return guard(c, [&] {
        need(N > 0, "allocate_particles: n_particles must be positive");
        c->N = N;
        dpmd::DVec<double>* dv[] = {&c->x, &c->y, &c->vx, &c->vy, &c->fx, &c->fy, &c->pe, &c->ke, &c->rad,
                                    &c->mass, &c->area, &c->lx, &c->ly, &c->angle, &c->avel, &c->torque,
                                    &c->inertia, &c->sxx, &c->sxy, &c->syx, &c->syy,
                                    &c->overlap, &c->ccount};
        for (auto* v : dv) { v->resize(N); v->zero(c->st); }
        c->xy0.resize(2 * (int64_t)N); c->xy0.zero(c->st);
        c->xy1.resize(2 * (int64_t)N); c->xy1.zero(c->st);
        c->vxy0.resize(2 * (int64_t)N); c->vxy0.zero(c->st);
        c->vxy1.resize(2 * (int64_t)N); c->vxy1.zero(c->st);
        c->sid.resize(N); c->sid.zero(c->st);
        c->contacts.resize(N); c->contacts.zero(c->st);
        c->nvpp.resize(N); c->nvpp.zero(c->st);
        c->poff.resize(N + 1); c->poff.zero(c->st);
        if (!rigid(c)) {
            std::vector<int> iota(N);
            std::iota(iota.begin(), iota.end(), 0);
            c->static_index.upload(iota.data(), N, c->st);
        }
        c->synced = false; c->nbr_ready = false; c->fire_ready = false; c->vv_ready = false;
        c->saved = false;   // a saved state no longer matches the arrays
    });
}

int dpmd_allocate_vertices(dpmd_ctx* c, int Nv) {
    return guard(c, [&] {
        need(rigid(c), "allocate_vertices: only rigid bumpy contexts have vertices");
        need(Nv > 0, "allocate_vertices: n_vertices must be positive");
        c->Nv = Nv;
        dpmd::DVec<double>* dv[] = {&c->vpx, &c->vpy, &c->vfx, &c->vfy, &c->vpe, &c->vrad, &c->vmass};
        for (auto* v : dv) { v->resize(Nv); v->zero(c->st); }
        c->vpid.resize(Nv); c->vpid.zero(c->st);
        c->vsid.resize(Nv); c->vsid.zero(c->st);
        std::vector<int> iota(Nv);
        std::iota(iota.begin(), iota.end(), 0);
        c->static_index.upload(iota.data(), Nv, c->st);
        c->synced = false; c->nbr_ready = false;
        c->saved = false;
    });
}

int dpmd_set_neighbor_method(dpmd_ctx* c, int m) {
    return guard(c, [&] {
        need(m == DPMD_NEIGHBOR_NAIVE || m == DPMD_NEIGHBOR_CELL, "unknown neighbor method");
        c->method = m;
        c->nbr_ready = false;
    });
}

int dpmd_set_neighbor_capacity(dpmd_ctx* c, int k) {
    return guard(c, [&] {
        need(k > 0 && k <= 4096, "neighbor capacity must be in [1, 4096]");
        c->K = k;
        if (c->synced) alloc_obj_buffers(c);
        c->nbr_ready = false;
    });
}

int dpmd_ctx_info(const dpmd_ctx* c, int* kind, int* method) {
    if (!c) return DPMD_ERR_ARGUMENT;
    if (kind) *kind = c->kind;
    if (method) *method = c->method;
    return DPMD_OK;
}

int dpmd_sizes(const dpmd_ctx* c, int* S, int* N, int* Nv, int* C) {
    if (!c) return DPMD_ERR_ARGUMENT;
    if (S) *S = c->S;
    if (N) *N = c->N;
    if (Nv) *Nv = c->Nv;
    if (C) *C = c->C;
    return DPMD_OK;
}

int dpmd_field(dpmd_ctx* c, int f, void** p, int64_t* n) {
    return guard(c, [&] {
        FieldRef r = field_ref(c, f);
        if (p) *p = r.p;
        if (n) *n = r.n;
    });
}

int dpmd_set_field(dpmd_ctx* c, int f, const void* host, int64_t count) {
    return guard(c, [&] {
        FieldRef r = field_ref(c, f);
        need(count == r.n, "set_field: count " + std::to_string(count) + " != field length " + std::to_string(r.n));
        if (count) HIP_OK(hipMemcpyAsync(r.p, host, count * r.elem, hipMemcpyHostToDevice, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        if (f == DPMD_F_SYSTEM_OFFSET || f == DPMD_F_CELL_DIM_X || f == DPMD_F_CELL_DIM_Y ||
            f == DPMD_F_N_VERTICES_PER_PARTICLE)
            c->synced = false;
    });
}

int dpmd_get_field(dpmd_ctx* c, int f, void* host, int64_t count, int order) {
    return guard(c, [&] {
        FieldRef r = field_ref(c, f);
        need(count == r.n, "get_field: count " + std::to_string(count) + " != field length " + std::to_string(r.n));
        HIP_OK(hipStreamSynchronize(c->st));
        if (count == 0) return;
        const bool reordered = (r.domain == 0 && !rigid(c)) || (r.domain == 1 && rigid(c));
        if (order == DPMD_ORDER_CANONICAL && reordered && c->method == DPMD_NEIGHBOR_CELL &&
            f != DPMD_F_STATIC_INDEX && f != DPMD_F_SYSTEM_ID && f != DPMD_F_VERTEX_SYSTEM_ID) {
            std::vector<char> cur(count * r.elem);
            HIP_OK(hipMemcpy(cur.data(), r.p, count * r.elem, hipMemcpyDeviceToHost));
            auto st = dl(c, c->static_index);
            char* out = (char*)host;
            for (int64_t k = 0; k < count; ++k) std::memcpy(out + k * r.elem, cur.data() + (int64_t)st[k] * r.elem, r.elem);
        } else {
            HIP_OK(hipMemcpy(host, r.p, count * r.elem, hipMemcpyDeviceToHost));
        }
    });
}

int dpmd_sync(dpmd_ctx* c) { return guard(c, [&] { sync_impl(c); }); }

int dpmd_init_neighbors(dpmd_ctx* c) { return guard(c, [&] { init_neighbors_impl(c); }); }

int dpmd_update_neighbors(dpmd_ctx* c) {
    return guard(c, [&] {
        if (c->method == DPMD_NEIGHBOR_NAIVE) { naive_neighbors(c); return; }
        check_neighbors_impl(c, true);
    });
}

int dpmd_check_neighbors(dpmd_ctx* c, int force) { return guard(c, [&] { check_neighbors_impl(c, force != 0); }); }

int dpmd_neighbor_info(dpmd_ctx* c, int* capacity, int* max_count, int64_t* total) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised");
        if (capacity) *capacity = c->K;
        if (max_count || total) {
            // A list longer than the capacity is an error at build time, so the stored counts
            // are the true ones.
            auto nn = dl(c, c->nn);
            int64_t t = 0;
            int mx = 0;
            for (int k = 0; k < c->n_obj(); ++k) { t += nn[k]; mx = std::max(mx, nn[k]); }
            if (max_count) *max_count = mx;
            if (total) *total = t;
        }
    });
}

int dpmd_get_neighbors_csr(dpmd_ctx* c, int* start, int* ids, int64_t ids_len) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised");
        check_overflow(c);
        const int n = c->n_obj();
        auto nn = dl(c, c->nn);
        auto nbr = dl(c, c->nbr);
        int64_t w = 0;
        for (int i = 0; i < n; ++i) {
            start[i] = (int)w;
            for (int k = 0; k < nn[i]; ++k) {
                need(w < ids_len, "get_neighbors_csr: ids buffer too small");
                ids[w++] = nbr[(size_t)k * n + i];
            }
        }
        start[n] = (int)w;
    });
}

int dpmd_compute_forces(dpmd_ctx* c, int mode) { return guard(c, [&] { compute_forces_impl(c, mode); }); }
int dpmd_update_positions(dpmd_ctx* c, const double* sc, double s2) {
    return guard(c, [&] { update_positions_impl(c, sc, s2); });
}
int dpmd_update_velocities(dpmd_ctx* c, const double* sc, double s2) {
    return guard(c, [&] { update_velocities_impl(c, sc, s2); });
}
int dpmd_scale_velocities(dpmd_ctx* c, const double* sc) { return guard(c, [&] { scale_velocities_impl(c, sc); }); }
int dpmd_scale_positions(dpmd_ctx* c, const double* sc) { return guard(c, [&] { scale_positions_impl(c, sc); }); }
int dpmd_mix_velocities_and_forces(dpmd_ctx* c, const double* a) {
    return guard(c, [&] { if (rigid(c)) rigid_mix(c, a); else disk_mix(c, a); });
}
int dpmd_zero_velocities(dpmd_ctx* c) {
    return guard(c, [&] {
        fill_f64(c->st, c->vx.p, c->N, 0.0);
        fill_f64(c->st, c->vy.p, c->N, 0.0);
        if (rigid(c)) fill_f64(c->st, c->avel.p, c->N, 0.0);
    });
}
int dpmd_compute_ke(dpmd_ctx* c) { return guard(c, [&] { if (rigid(c)) rigid_ke(c); else disk_ke(c); }); }
int dpmd_compute_ke_total(dpmd_ctx* c) { return guard(c, [&] { sys_reduce(c, RED_KE, c->ke_tot.p); }); }
int dpmd_compute_pe_total(dpmd_ctx* c) { return guard(c, [&] { sys_reduce(c, RED_PE, c->pe_tot.p); }); }
int dpmd_compute_fpower_total(dpmd_ctx* c) { return guard(c, [&] { sys_reduce(c, RED_POWER, c->pw_tot.p); }); }
int dpmd_compute_temperature(dpmd_ctx* c) { return guard(c, [&] { per_system_temperature(c); }); }
int dpmd_compute_packing_fraction(dpmd_ctx* c) { return guard(c, [&] { sys_reduce(c, RED_PHI, c->phi.p); }); }
int dpmd_compute_max_force(dpmd_ctx* c) { return guard(c, [&] { sys_reduce(c, RED_MAXF2, c->maxf2.p); }); }
int dpmd_compute_stress_tensor(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised");
        if (rigid(c)) rigid_stress(c); else disk_stress(c);
    });
}
int dpmd_compute_pressure(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised");
        if (rigid(c)) rigid_stress(c); else disk_stress(c);
        per_system_pressure(c);
    });
}
int dpmd_set_average_velocity(dpmd_ctx* c, const double* ax, const double* ay) {
    return guard(c, [&] { set_average_velocity_impl(c, ax, ay); });
}
int dpmd_set_temperature(dpmd_ctx* c, const double* T, int rm) {
    return guard(c, [&] { set_temperature_impl(c, T, rm != 0); });
}
int dpmd_increment_packing_fraction(dpmd_ctx* c, const double* inc) {
    return guard(c, [&] { increment_packing_fraction_impl(c, inc); });
}
int dpmd_save_state(dpmd_ctx* c, const int* flag, int tv) {
    return guard(c, [&] { save_state_impl(c, flag, tv); });
}
int dpmd_restore_state(dpmd_ctx* c, const int* flag, int tv) {
    return guard(c, [&] { restore_state_impl(c, flag, tv); });
}

int dpmd_rand_uniform_2d(dpmd_ctx* c, int fx, int fy, uint64_t seed, uint64_t call, double lox, double hix,
                         double loy, double hiy) {
    return guard(c, [&] {
        FieldRef a = field_ref(c, fx), b = field_ref(c, fy);
        need(a.elem == 8 && b.elem == 8 && a.n == b.n, "rand_uniform_2d: two equal-length double fields");
        rand_uniform_2d(c->st, (double*)a.p, (double*)b.p, (int)a.n, seed, call, lox, hix, loy, hiy);
    });
}
int dpmd_rand_uniform_1d(dpmd_ctx* c, int f, uint64_t seed, uint64_t call, double lo, double hi) {
    return guard(c, [&] {
        FieldRef a = field_ref(c, f);
        need(a.elem == 8, "rand_uniform_1d: double field");
        rand_uniform_1d(c->st, (double*)a.p, (int)a.n, seed, call, lo, hi);
    });
}
int dpmd_rand_normal_2d(dpmd_ctx* c, int fx, int fy, uint64_t seed, uint64_t call, double mx, double sx, double my,
                        double sy) {
    return guard(c, [&] {
        FieldRef a = field_ref(c, fx), b = field_ref(c, fy);
        need(a.elem == 8 && b.elem == 8 && a.n == b.n, "rand_normal_2d: two equal-length double fields");
        rand_normal_2d(c->st, (double*)a.p, (double*)b.p, (int)a.n, seed, call, mx, sx, my, sy);
    });
}

int dpmd_vv_init(dpmd_ctx* c, const double* dt, int mode) {
    return guard(c, [&] { vv_init_impl(c, dt, mode); });
}

int dpmd_nvt_rescale_run(dpmd_ctx* c, int n_steps, int step0, const double* phi_increment, int compress_every,
                         const double* temperature_target, int temperature_every, int remove_average_velocity,
                         int use_graph) {
    return guard(c, [&] {
        nvt_rescale_impl(c, n_steps, step0, phi_increment, compress_every, temperature_target, temperature_every,
                         remove_average_velocity != 0, use_graph != 0);
        check_overflow(c);
    });
}

int dpmd_compute_contacts(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        if (rigid(c)) {
            rigid_particle_neighbors(c);
            rigid_pair_obs(c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        } else {
            disk_contacts(c);
        }
    });
}
int dpmd_compute_n_contacts_total(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        if (rigid(c)) {
            rigid_particle_neighbors(c);
            rigid_pair_obs(c, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        } else {
            disk_contacts(c);
        }
        contacts_total(c);
    });
}
int dpmd_compute_overlaps(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        if (rigid(c)) rigid_overlaps(c); else disk_contacts(c);
    });
}

int dpmd_particle_neighbors(dpmd_ctx* c, int* start, int* ids, int64_t ids_len, int64_t* total) {
    return guard(c, [&] {
        need(rigid(c), "particle_neighbors: rigid bumpy contexts (disk lists are particle lists)");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        const int64_t tot = rigid_particle_neighbors(c);
        if (total) *total = tot;
        if (start) {
            HIP_OK(hipMemcpyAsync(start, c->pn_start.p, (size_t)(c->N + 1) * 4, hipMemcpyDeviceToHost, c->st));
        }
        if (ids) {
            need(ids_len >= tot, "particle_neighbors: ids buffer too small");
            if (tot) HIP_OK(hipMemcpyAsync(ids, c->pn_ids.p, (size_t)tot * 4, hipMemcpyDeviceToHost, c->st));
        }
        HIP_OK(hipStreamSynchronize(c->st));
    });
}

int dpmd_compute_rigid_hessian(dpmd_ctx* c, double* const* hessian, int64_t n) {
    return guard(c, [&] {
        need(rigid(c), "rigid hessian: rigid bumpy contexts");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        need(hessian != nullptr, "rigid hessian: 18 column pointers are required");
        for (int m = 0; m < 18; ++m) need(hessian[m] != nullptr, "rigid hessian: 18 column pointers are required");
        const int64_t tot = rigid_particle_neighbors(c);
        need(n == tot, "rigid hessian: n must equal the number of particle neighbor entries (" + std::to_string(tot) +
                           ")");
        if (tot > 0) rigid_hessian(c, hessian);
    });
}

int dpmd_compute_rigid_pair_observables(dpmd_ctx* c, int* pair_i, int* pair_j, double* dist, double* fx, double* fy,
                                        int* vertex_contacts_i, int* vertex_contacts_j, double* friction, int64_t n) {
    return guard(c, [&] {
        need(rigid(c), "rigid pair observables: rigid bumpy contexts");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        need((pair_i == nullptr) == (pair_j == nullptr), "rigid pair observables: pair_i and pair_j go together");
        need((fx == nullptr) == (fy == nullptr), "rigid pair observables: fx and fy go together");
        const int64_t tot = rigid_particle_neighbors(c);
        need(n == tot, "rigid pair observables: n must equal the number of particle neighbor entries (" +
                           std::to_string(tot) + ")");
        rigid_pair_obs(c, pair_i, pair_j, dist, fx, fy, vertex_contacts_i, vertex_contacts_j, friction);
    });
}

int dpmd_compute_pair_observables(dpmd_ctx* c, int* pair_i, int* pair_j, double* dist, double* fx, double* fy,
                                  double* const* hessian, int64_t n) {
    return guard(c, [&] {
        need(!rigid(c), "pair observables: disks (rigid bodies: dpmd_compute_rigid_pair_observables)");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        need((pair_i == nullptr) == (pair_j == nullptr), "pair observables: pair_i and pair_j go together");
        need((fx == nullptr) == (fy == nullptr), "pair observables: fx and fy go together");
        if (hessian)
            for (int m = 0; m < 8; ++m) need(hessian[m] != nullptr, "pair observables: 8 Hessian arrays or none");
        const int64_t tot = pair_setup(c);
        need(n == tot, "pair observables: n must equal the number of neighbor entries (" + std::to_string(tot) + ")");
        disk_pair_obs(c, pair_i, pair_j, dist, fx, fy, hessian);
    });
}

int dpmd_damped_vv_init(dpmd_ctx* c, const double* dt, const double* damping, int mode) {
    return guard(c, [&] {
        need(dt != nullptr && damping != nullptr, "damped_vv_init: dt and damping are required");
        vv_init_impl(c, dt, mode, damping);
    });
}

int dpmd_compute_damping_forces(dpmd_ctx* c, const double* scale) {
    return guard(c, [&] {
        need(scale != nullptr, "compute_damping_forces: scale is required");
        if (rigid(c)) rigid_damping(c, scale); else disk_damping(c, scale);
    });
}

int dpmd_vv_step(dpmd_ctx* c, int n) {
    return guard(c, [&] {
        need_state(c->vv_ready, "vv_step: call vv_init first");
        for (int k = 0; k < n; ++k) vv_step_body(c);
    });
}

int dpmd_vv_run(dpmd_ctx* c, int n, int use_graph) {
    return guard(c, [&] {
        need_state(c->vv_ready, "vv_run: call vv_init first");
        if (n <= 0) return;
        int any = 0;
        HIP_OK(hipMemcpyAsync(&any, c->run.p + RUN_ANY, sizeof(int), hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        int run[RUN_WORDS] = {0, n, 1, 0, 0, -1, -1, any, 0};
        c->run.upload(run, RUN_WORDS, c->st);
        (void)lean_run(c, n, false, 0, 0.0, 0.0, use_graph != 0, nullptr);
        check_overflow(c);
    });
}

int dpmd_adjoint_hvp_noalloc(dpmd_ctx* c, int mode, const double* wx, const double* wy, double* yx, double* yy) {
    return guard(c, [&] {
        need_disk_cells(c);
        need(wx && wy && yx && yy, "adjoint_hvp_noalloc: null vector");
        const int64_t N = c->N;
        need(N <= kNoAllocMaxN, "adjoint_hvp_noalloc: " + std::to_string(N) + " particles exceed the scratch maximum of " +
                                    std::to_string(kNoAllocMaxN) + " (use dpmd_adjoint_hvp)");
        double* sp = scratch_ptr();
        double* tx = sp;                     // w in current order
        double* ty = sp + kNoAllocMaxN;
        double* ux = sp + 2 * kNoAllocMaxN;  // H w in current order
        double* uy = sp + 3 * kNoAllocMaxN;
        adj_to_current(c, wx, tx); adj_to_current(c, wy, ty);
        adj_hvp_current(c, mode, tx, ty, ux, uy);
        adj_to_canonical(c, ux, yx); adj_to_canonical(c, uy, yy);
    });
}

int dpmd_adjoint_hvp(dpmd_ctx* c, int mode, const double* wx, const double* wy, double* yx, double* yy) {
    return guard(c, [&] {
        need_disk_cells(c);
        need(wx && wy && yx && yy, "adjoint_hvp: null vector");
        const int N = c->N;
        c->adj_t0.resize(2 * (int64_t)N); c->adj_t1.resize(2 * (int64_t)N);
        adj_to_current(c, wx, c->adj_t0.p); adj_to_current(c, wy, c->adj_t0.p + N);
        adj_hvp_current(c, mode, c->adj_t0.p, c->adj_t0.p + N, c->adj_t1.p, c->adj_t1.p + N);
        adj_to_canonical(c, c->adj_t1.p, yx); adj_to_canonical(c, c->adj_t1.p + N, yy);
        HIP_OK(hipStreamSynchronize(c->st));
    });
}

int dpmd_adjoint_vjp(dpmd_ctx* c, int mode, const double* gU, const double* gP, const double* gfx, const double* gfy,
                     double* dx, double* dy, double* dr, double* dLx, double* dLy) {
    return guard(c, [&] {
        need_disk_cells(c);
        need((gfx == nullptr) == (gfy == nullptr) && (dx == nullptr) == (dy == nullptr) &&
                 (dLx == nullptr) == (dLy == nullptr), "adjoint_vjp: x and y arrays come in pairs");
        const int N = c->N;
        c->adj_t0.resize(2 * (int64_t)N); c->adj_t1.resize(2 * (int64_t)N); c->adj_t2.resize(N);
        if (gfx) { adj_to_current(c, gfx, c->adj_t0.p); adj_to_current(c, gfy, c->adj_t0.p + N); }
        adj_vjp_current(c, mode, gU, gP, gfx ? c->adj_t0.p : nullptr, gfx ? c->adj_t0.p + N : nullptr,
                        dx ? c->adj_t1.p : nullptr, dx ? c->adj_t1.p + N : nullptr, dr ? c->adj_t2.p : nullptr, dLx,
                        dLy);
        if (dx) { adj_to_canonical(c, c->adj_t1.p, dx); adj_to_canonical(c, c->adj_t1.p + N, dy); }
        if (dr) adj_to_canonical(c, c->adj_t2.p, dr);
        HIP_OK(hipStreamSynchronize(c->st));
    });
}

int dpmd_adjoint_solve(dpmd_ctx* c, int mode, const double* bx, const double* by, double* xx, double* xy, double tol,
                       int max_iter, int* iters, double* rel_resid) {
    return guard(c, [&] {
        need_disk_cells(c);
        need(bx && by && xx && xy && tol > 0.0 && max_iter > 0, "adjoint_solve: bad arguments");
        const int N = c->N;
        c->adj_t0.resize(2 * (int64_t)N); c->adj_t1.resize(2 * (int64_t)N);
        adj_to_current(c, bx, c->adj_t0.p); adj_to_current(c, by, c->adj_t0.p + N);
        const int it = adj_cg_current(c, mode, c->adj_t0.p, c->adj_t0.p + N, c->adj_t1.p, c->adj_t1.p + N, tol,
                                      max_iter, rel_resid);
        adj_to_canonical(c, c->adj_t1.p, xx); adj_to_canonical(c, c->adj_t1.p + N, xy);
        HIP_OK(hipStreamSynchronize(c->st));
        if (iters) *iters = it;
    });
}

int dpmd_adjoint_pressure(dpmd_ctx* c, int mode, double tol, int max_iter, double* dr, double* dLx, double* dLy,
                          int* iters, double* rel_resid) {
    return guard(c, [&] {
        need(tol > 0.0 && max_iter > 0 && (dLx == nullptr) == (dLy == nullptr), "adjoint_pressure: bad arguments");
        adjoint_pressure_impl(c, mode, tol, max_iter, dr, dLx, dLy, iters, rel_resid);
    });
}

int dpmd_newton_minimize(dpmd_ctx* c, int mode, int max_newton, double ftol, double cg_tol, int cg_max_iter,
                         double* max_force) {
    return guard(c, [&] {
        need(max_newton >= 0 && cg_tol > 0.0 && cg_max_iter > 0, "newton_minimize: bad arguments");
        const double f = newton_impl(c, mode, max_newton, ftol, cg_tol, cg_max_iter);
        if (max_force) *max_force = f;
    });
}

int dpmd_vv_run_checkpointed(dpmd_ctx* c, int n_steps, int every, int use_graph) {
    return guard(c, [&] { vv_checkpointed_impl(c, n_steps, every, use_graph != 0); });
}

int dpmd_vv_backprop(dpmd_ctx* c, const double* gx, const double* gy, const double* gvx, const double* gvy, double* gx0,
                     double* gy0, double* gvx0, double* gvy0, int use_graph) {
    return guard(c, [&] {
        need((gx == nullptr) == (gy == nullptr) && (gvx == nullptr) == (gvy == nullptr),
             "vv_backprop: x and y arrays come in pairs");
        vv_backprop_impl(c, gx, gy, gvx, gvy, gx0, gy0, gvx0, gvy0, use_graph != 0);
    });
}

int dpmd_fire_init(dpmd_ctx* c, const double* dt, int mode) { return guard(c, [&] { fire_init_impl(c, dt, mode); }); }

int dpmd_fire_step(dpmd_ctx* c) {
    return guard(c, [&] {
        need_state(c->fire_ready, "fire_step: call fire_init first");
        // compute_fpower_total + fire_update_kernel (fire.cuh:62-72) from the live state
        sys_reduce(c, RED_POWER, c->pw_tot.p);
        fire_recompute_update(c);
        fire_step_body(c, 0, 0, 0.0, 0.0);
    });
}

int dpmd_fire_converged(dpmd_ctx* c, double diff_target, double pe_target, int* all) {
    return guard(c, [&] {
        need_state(c->fire_ready, "fire_converged: call fire_init first");
        fire_apply_convergence(c, DPMD_FIRE_PRED_REFERENCE, diff_target, pe_target);
        check_overflow(c);
        if (all) *all = read_run(c).running <= 0 ? 1 : 0;
    });
}

int dpmd_fire_converged_max_force(dpmd_ctx* c, double tol, int* all) {
    return guard(c, [&] {
        need_state(c->fire_ready, "fire_converged: call fire_init first");
        fire_apply_convergence(c, DPMD_FIRE_PRED_MAX_FORCE, tol, 0.0);
        check_overflow(c);
        if (all) *all = read_run(c).running <= 0 ? 1 : 0;
    });
}

int dpmd_fire_get_dt(dpmd_ctx* c, double* out) {
    return guard(c, [&] {
        HIP_OK(hipMemcpyAsync(out, c->f_dt.p, c->S * 8, hipMemcpyDeviceToDevice, c->st));
    });
}

int dpmd_fire_state(dpmd_ctx* c, double* dt, double* alpha, int* ng, int* nb) {
    return guard(c, [&] {
        if (dt) { auto h = dl(c, c->f_dt); std::copy(h.begin(), h.end(), dt); }
        if (alpha) { auto h = dl(c, c->f_alpha); std::copy(h.begin(), h.end(), alpha); }
        if (ng) { auto h = dl(c, c->f_ngood); std::copy(h.begin(), h.end(), ng); }
        if (nb) { auto h = dl(c, c->f_nbad); std::copy(h.begin(), h.end(), nb); }
    });
}

int dpmd_fire_minimize(dpmd_ctx* c, const double* dt, int max_steps, int pred, double a, double b, int mode,
                       int use_graph, int poll, int* steps_taken, int* all_conv) {
    return guard(c, [&] {
        need(pred >= DPMD_FIRE_PRED_REFERENCE && pred <= DPMD_FIRE_PRED_FORCE_OR_PE, "unknown FIRE predicate");
        fire_init_impl(c, dt, mode);
        (void)poll;   // batches adapt to the rebuild rate; the host reads state once per batch
        int conv = 0;
        const int running0 = read_run(c).running;
        int steps = 0;
        if (running0 > 0) {
            FrozenMask mask(c);
            steps = lean_run(c, max_steps, true, pred, a, b, use_graph != 0, &conv);
        }
        check_overflow(c);
        if (steps_taken) *steps_taken = steps;
        if (all_conv) *all_conv = running0 > 0 ? conv : 1;
    });
}

int dpmd_jam_compress(dpmd_ctx* c, const double* dt, double dphi, double p_target, int max_comp, int fire_max,
                      int pred, double a, double b, int mode, int use_graph, int* jam_state, int* n_comp,
                      int64_t* work) {
    return guard(c, [&] {
        need(pred >= DPMD_FIRE_PRED_REFERENCE && pred <= DPMD_FIRE_PRED_FORCE_OR_PE, "unknown FIRE predicate");
        need(dt != nullptr && max_comp >= 0 && fire_max > 0, "jam_compress: bad arguments");
        need_state(c->nbr_ready, "neighbors are not initialised (call init_neighbors)");
        const int S = c->S;
        std::vector<double> dt0(S), dts(S), inc(S), P(S);
        HIP_OK(hipMemcpyAsync(dt0.data(), dt, S * 8, hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        // state: 0 compressing, 1 jammed, 2 budget used up, 3 not run, -1 dropped
        std::vector<int> state(S), ncomp(S, 0), nsteps(S, 0), stopped(S), ns(S), retries(S, 0);
        for (int s = 0; s < S; ++s) state[s] = dt0[s] == 0.0 ? 3 : 0;
        int64_t total = 0;
        for (;;) {
            bool any = false;
            for (int s = 0; s < S; ++s) {
                dts[s] = state[s] == 0 ? dt0[s] : 0.0;
                any = any || dts[s] != 0.0;
            }
            if (!any) break;
            c->scr_a.upload(dts.data(), S, c->st);
            fire_init_impl(c, c->scr_a.p, mode);
            {
                FrozenMask mask(c);
                (void)lean_run(c, fire_max, true, pred, a, b, use_graph != 0, nullptr);
            }
            check_overflow(c);
            if (rigid(c)) rigid_stress(c); else disk_stress(c);
            per_system_pressure(c);
            HIP_OK(hipMemcpyAsync(stopped.data(), c->f_stopped.p, S * 4, hipMemcpyDeviceToHost, c->st));
            HIP_OK(hipMemcpyAsync(ns.data(), c->f_nsteps.p, S * 4, hipMemcpyDeviceToHost, c->st));
            HIP_OK(hipMemcpyAsync(P.data(), c->press.p, S * 8, hipMemcpyDeviceToHost, c->st));
            HIP_OK(hipStreamSynchronize(c->st));
            bool grow = false, retry = false;
            for (int s = 0; s < S; ++s) {
                inc[s] = 0.0;
                if (dts[s] == 0.0) continue;
                nsteps[s] += ns[s];
                total += (int64_t)(c->h_off[s + 1] - c->h_off[s]) * ns[s];
                if (!stopped[s] && retries[s] < c->jam_retries) {
                    // a round that hit the step cap is run again from where it stopped (FIRE
                    // restarted, no compression), up to jam_retries times
                    retries[s] += 1;
                    retry = true;
                }
                // the reference's rule (jammers.cuh:85-89, jammers.cu:7-17): a system whose
                // initial minimisation does not stop is dropped; after a compression the
                // state the minimisation reached decides, whether it stopped or not
                else if (!stopped[s] && ncomp[s] == 0) state[s] = -1;
                else if (P[s] >= p_target) state[s] = 1;
                else if (ncomp[s] >= max_comp) state[s] = 2;
                else { inc[s] = dphi; ncomp[s] += 1; grow = true; }
            }
            if (!grow && !retry) break;
            // every round boundary rebuilds every system (the compression does), so a retried
            // system starts its next minimisation from the same lists alone or in a batch
            c->scr_b.upload(inc.data(), S, c->st);
            increment_packing_fraction_impl(c, c->scr_b.p);
        }
        c->f_nsteps.upload(nsteps.data(), S, c->st);
        HIP_OK(hipStreamSynchronize(c->st));
        for (int s = 0; s < S; ++s) {
            if (jam_state) jam_state[s] = state[s] == 1 ? 1 : state[s] == -1 ? -1 : 0;
            if (n_comp) n_comp[s] = ncomp[s];
        }
        if (work) *work = total;
    });
}

int dpmd_jam_binary_search(dpmd_ctx* c, const double* dt, int max_compression_steps, int max_minimization_steps,
                           double avg_pe_target, double avg_pe_diff_target, double phi_increment,
                           double phi_tolerance, int mode, int use_graph, int* action, double* phi,
                           int* compression_steps) {
    return guard(c, [&] {
        jam_bisect_impl(c, dt, max_compression_steps, max_minimization_steps, avg_pe_target, avg_pe_diff_target,
                        phi_increment, phi_tolerance, mode, use_graph != 0, action, phi, compression_steps);
    });
}

int dpmd_malloc(void** p, int64_t bytes) {
    return guard(nullptr, [&] {
        need(p != nullptr && bytes >= 0, "dpmd_malloc: bad arguments");
        HIP_OK(hipMalloc(p, bytes > 0 ? bytes : 1));
    });
}
int dpmd_free(void* p) { return guard(nullptr, [&] { if (p) HIP_OK(hipFree(p)); }); }
int dpmd_memcpy(void* dst, const void* src, int64_t bytes, int kind) {
    return guard(nullptr, [&] {
        need(kind >= 0 && kind <= 2, "dpmd_memcpy: kind must be 0, 1 or 2");
        const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
        if (bytes > 0) HIP_OK(hipMemcpy(dst, src, bytes, k));
    });
}
int dpmd_fill_f64(void* p, int64_t n, double v) {
    return guard(nullptr, [&] { fill_f64(nullptr, (double*)p, n, v); HIP_OK(hipDeviceSynchronize()); });
}
int dpmd_fill_i32(void* p, int64_t n, int32_t v) {
    return guard(nullptr, [&] { fill_i32(nullptr, (int*)p, n, v); HIP_OK(hipDeviceSynchronize()); });
}
int dpmd_set_option(dpmd_ctx* c, int key, int value) {
    return guard(c, [&] {
        switch (key) {
            case DPMD_OPT_COOP_REBUILD:
            case DPMD_OPT_COOP_IN_GRAPH:
                // The cooperative single-launch rebuild was removed: it relied on a grid-wide
                // barrier whose co-residency a graph replay of the node does not guarantee.
                need(value == 0, "DPMD_OPT_COOP_REBUILD / DPMD_OPT_COOP_IN_GRAPH: the cooperative rebuild is not "
                                 "available (only 0 is accepted)");
                break;
            case DPMD_OPT_FUSED_VV: c->fuse_vv = value != 0; break;
            case DPMD_OPT_FUSED_FIRE: c->fuse_fire = value != 0; break;
            case DPMD_OPT_BLOCK:
                need(value == 64 || value == 128 || value == 256 || value == 512,
                     "DPMD_OPT_BLOCK must be 64, 128, 256 or 512");
                c->block = value;
                break;
            case DPMD_OPT_MAX_BATCH:
                need(value >= 1 && value <= 128 && (value & (value - 1)) == 0,
                     "DPMD_OPT_MAX_BATCH must be a power of two from 1 to 128");
                c->max_batch = value;
                break;
            case DPMD_OPT_PERSISTENT_FIRE:
                need(value >= 0 && value <= 2, "DPMD_OPT_PERSISTENT_FIRE must be 0, 1 or 2");
                c->persist_fire = value;
                break;
            case DPMD_OPT_PERSIST_PHASE:
                need(value >= 1, "DPMD_OPT_PERSIST_PHASE must be positive");
                c->persist_phase = value;
                break;
            case DPMD_OPT_PERSIST_BLOCK:
                need(value == 256 || value == 512 || value == 1024, "DPMD_OPT_PERSIST_BLOCK must be 256, 512 or 1024");
                c->persist_block = value;
                break;
            case DPMD_OPT_DEVICE_REBUILD:
                need(value >= 0 && value <= 2, "DPMD_OPT_DEVICE_REBUILD must be 0, 1 or 2");
                c->dev_rebuild = value;
                break;
            case DPMD_OPT_PERSIST_LDS:
                need(value == 0 || value == 1, "DPMD_OPT_PERSIST_LDS must be 0 or 1");
                c->persist_lds = value;
                break;
            case DPMD_OPT_REBUILD_ENGINE:
                need(value >= 0 && value <= 2, "DPMD_OPT_REBUILD_ENGINE must be 0, 1 or 2");
                c->rebuild_engine = value;
                break;
            case DPMD_OPT_JAM_RETRIES:
                need(value >= 0, "DPMD_OPT_JAM_RETRIES must be non-negative");
                c->jam_retries = value;
                break;
            default: throw ArgError("unknown option key");
        }
    });
}

int dpmd_get_option(dpmd_ctx* c, int key, int* value) {
    return guard(c, [&] {
        switch (key) {
            case DPMD_OPT_COOP_REBUILD:
            case DPMD_OPT_COOP_IN_GRAPH: *value = 0; break;
            case DPMD_OPT_FUSED_VV: *value = c->fuse_vv ? 1 : 0; break;
            case DPMD_OPT_FUSED_FIRE: *value = c->fuse_fire ? 1 : 0; break;
            case DPMD_OPT_BLOCK: *value = c->block; break;
            case DPMD_OPT_PERSISTENT_FIRE: *value = c->persist_fire; break;
            case DPMD_OPT_PERSIST_PHASE: *value = c->persist_phase; break;
            case DPMD_OPT_PERSIST_BLOCK: *value = c->persist_block; break;
            case DPMD_OPT_MAX_BATCH: *value = c->max_batch; break;
            case DPMD_OPT_DEVICE_REBUILD: *value = c->dev_rebuild; break;
            case DPMD_OPT_PERSIST_LDS: *value = c->persist_lds; break;
            case DPMD_OPT_REBUILD_ENGINE: *value = c->rebuild_engine; break;
            case DPMD_OPT_JAM_RETRIES: *value = c->jam_retries; break;
            default: throw ArgError("unknown option key");
        }
    });
}

int dpmd_profile(dpmd_ctx* c, int enable) {
    return guard(c, [&] {
        HIP_OK(hipStreamSynchronize(c->st));
        for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
        c->prof_ev.clear();
        c->prof_n.clear();
        need(enable >= 0 && enable <= 4, "dpmd_profile: enable must be 0 .. 4");
        c->prof = enable;
    });
}

int dpmd_profile_read(dpmd_ctx* c, double* total_ms, int* count) {
    return guard(c, [&] {
        HIP_OK(hipStreamSynchronize(c->st));
        double t = 0.0;
        int n = 0;
        for (size_t k = 0; k + 1 < c->prof_ev.size(); k += 2) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, c->prof_ev[k], c->prof_ev[k + 1]));
            t += ms;
            n += c->prof_n[k / 2];
        }
        if (total_ms) *total_ms = t;
        if (count) *count = n;
        for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
        c->prof_ev.clear();
        c->prof_n.clear();
    });
}

int dpmd_profile_durations(dpmd_ctx* c, float* ms, int capacity, int* count) {
    return guard(c, [&] {
        need(capacity >= 0 && (ms != nullptr || capacity == 0), "dpmd_profile_durations: bad buffer");
        HIP_OK(hipStreamSynchronize(c->st));
        const int n = (int)(c->prof_ev.size() / 2);
        for (int k = 0; k < n && k < capacity; ++k)
            HIP_OK(hipEventElapsedTime(&ms[k], c->prof_ev[2 * k], c->prof_ev[2 * k + 1]));
        if (count) *count = n;
        for (hipEvent_t e : c->prof_ev) (void)hipEventDestroy(e);
        c->prof_ev.clear();
        c->prof_n.clear();
    });
}

int dpmd_stream_copy(double* dst, const double* src, int64_t n, int repeats, double* ms_per_copy) {
    return guard(nullptr, [&] {
        need(n > 0 && repeats > 0, "dpmd_stream_copy: n and repeats must be positive");
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a)); HIP_OK(hipEventCreate(&b));
        copy_f64(nullptr, dst, src, n);
        HIP_OK(hipEventRecord(a, nullptr));
        for (int k = 0; k < repeats; ++k) copy_f64(nullptr, dst, src, n);
        HIP_OK(hipEventRecord(b, nullptr));
        HIP_OK(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, a, b));
        (void)hipEventDestroy(a); (void)hipEventDestroy(b);
        if (ms_per_copy) *ms_per_copy = ms / repeats;
    });
}

int dpmd_synchronize(dpmd_ctx* c) {
    return guard(c, [&] {
        HIP_OK(hipStreamSynchronize(c->st));
        if (c->nbr_ready) check_overflow(c);
    });
}

}  // extern "C"
