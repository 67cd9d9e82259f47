// md::disk::Disk and md::rigid_bumpy::RigidBumpy -- header-only C++ facade over the C ABI.
//
// Reference interface: the CRTP classes md::BaseParticle (include/particles/base_particle.hpp:
// 27-1088), md::BasePointParticle (base_point_particle.hpp:81-456), md::disk::Disk
// (include/particles/disk.cuh:22), md::BasePolyParticle (base_poly_particle.hpp:99-620) and
// md::rigid_bumpy::RigidBumpy (include/particles/rigid_bumpy.cuh:24).  Public field names and
// method names and their call order are the reference's; each field is a view of the
// context's device array, each method one C ABI call.  Errors surface as
// std::runtime_error / std::invalid_argument, as in the reference.
//
// Both particle types are declared once here, so they can be used in one translation unit
// (the reference declares CellSortMethod twice and cannot, base_point_particle.hpp:78 and
// base_poly_particle.hpp:96).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "device_fields.hpp"
#include "dpmd.h"

namespace md {

enum class NeighborMethod : int { Naive = DPMD_NEIGHBOR_NAIVE, Cell = DPMD_NEIGHBOR_CELL };

template <class Derived>
class ParticleBase {
protected:
    dpmd_ctx* ctx_;   // declared first: every field view below is initialised from it
    uint64_t seed_ = 0;

public:
    using F1 = df::DeviceField1D<double>;
    using F2 = df::DeviceField2D<double>;
    using I1 = df::DeviceField1D<int>;

    explicit ParticleBase(int kind, void* hip_stream = nullptr) : ctx_(make(kind, hip_stream)) {}
    ParticleBase(const ParticleBase&) = delete;
    ParticleBase& operator=(const ParticleBase&) = delete;
    ~ParticleBase() { dpmd_ctx_destroy(ctx_); }

    dpmd_ctx* ctx() const { return ctx_; }

    // ---- fields (base_particle.hpp:31-103) -------------------------------------------
    F2 pos{ctx_, DPMD_F_POS_X, DPMD_F_POS_Y};
    F2 vel{ctx_, DPMD_F_VEL_X, DPMD_F_VEL_Y};
    F2 force{ctx_, DPMD_F_FORCE_X, DPMD_F_FORCE_Y};
    F1 pe{ctx_, DPMD_F_PE};
    F1 ke{ctx_, DPMD_F_KE};
    F1 area{ctx_, DPMD_F_AREA};
    F1 mass{ctx_, DPMD_F_MASS};
    F1 e_interaction{ctx_, DPMD_F_E_INTERACTION};
    F2 last_pos{ctx_, DPMD_F_LAST_POS_X, DPMD_F_LAST_POS_Y};
    I1 neighbor_count{ctx_, DPMD_F_NEIGHBOR_COUNT};
    I1 cell_id{ctx_, DPMD_F_CELL_ID};
    I1 static_index{ctx_, DPMD_F_STATIC_INDEX};
    df::DeviceField2D<int> cell_dim{ctx_, DPMD_F_CELL_DIM_X, DPMD_F_CELL_DIM_Y};
    I1 cell_system_start{ctx_, DPMD_F_CELL_SYSTEM_START};
    F1 neighbor_cutoff{ctx_, DPMD_F_NEIGHBOR_CUTOFF};
    F1 thresh2{ctx_, DPMD_F_THRESH2};
    I1 system_id{ctx_, DPMD_F_SYSTEM_ID};
    I1 system_size{ctx_, DPMD_F_SYSTEM_SIZE};
    I1 system_offset{ctx_, DPMD_F_SYSTEM_OFFSET};
    I1 boundary{ctx_, DPMD_F_BOUNDARY};   // addition: per-system 0 periodic / 1 walls
    F2 box_size{ctx_, DPMD_F_BOX_SIZE_X, DPMD_F_BOX_SIZE_Y};
    F2 box_inv{ctx_, DPMD_F_BOX_INV_X, DPMD_F_BOX_INV_Y};
    F1 packing_fraction{ctx_, DPMD_F_PACKING_FRACTION};
    F1 pressure{ctx_, DPMD_F_PRESSURE};
    F1 temperature{ctx_, DPMD_F_TEMPERATURE};
    F1 pe_total{ctx_, DPMD_F_PE_TOTAL};
    F1 ke_total{ctx_, DPMD_F_KE_TOTAL};
    F1 fpower_total{ctx_, DPMD_F_FPOWER_TOTAL};
    I1 n_dof{ctx_, DPMD_F_N_DOF};
    F2 stress_tensor_x{ctx_, DPMD_F_STRESS_XX, DPMD_F_STRESS_XY};
    F2 stress_tensor_y{ctx_, DPMD_F_STRESS_YX, DPMD_F_STRESS_YY};
    F2 stress_tensor_total_x{ctx_, DPMD_F_STRESS_TOTAL_XX, DPMD_F_STRESS_TOTAL_XY};
    F2 stress_tensor_total_y{ctx_, DPMD_F_STRESS_TOTAL_YX, DPMD_F_STRESS_TOTAL_YY};

    NeighborMethod neighbor_method = NeighborMethod::Naive;
    NeighborMethod get_neighbor_method() const { return neighbor_method; }

    // ---- set-up (base_particle.hpp:260-370) --------------------------------------------
    void set_neighbor_method(NeighborMethod m) {
        neighbor_method = m;
        ok(dpmd_set_neighbor_method(ctx_, (int)m));
    }
    void allocate_systems(int S) { ok(dpmd_allocate_systems(ctx_, S)); }
    void allocate_particles(int N) { ok(dpmd_allocate_particles(ctx_, N)); }
    void allocate_vertices(int Nv) { ok(dpmd_allocate_vertices(ctx_, Nv)); }
    // The five reference sync calls all map to dpmd_sync (idempotent).
    void sync_box() { ok(dpmd_sync(ctx_)); }
    void sync_system() { ok(dpmd_sync(ctx_)); }
    void sync_neighbors() { ok(dpmd_sync(ctx_)); }
    void sync_cells() { ok(dpmd_sync(ctx_)); }
    void sync_class_constants() { ok(dpmd_sync(ctx_)); }

    int n_particles() const { int n; sizes(nullptr, &n, nullptr, nullptr); return n; }
    int n_systems() const { int s; sizes(&s, nullptr, nullptr, nullptr); return s; }
    int n_vertices() const { int v; sizes(nullptr, nullptr, &v, nullptr); return v; }
    int n_cells() const { int c; sizes(nullptr, nullptr, nullptr, &c); return c; }
    int n_neighbors() const {
        int64_t t = 0;
        ok(dpmd_neighbor_info(ctx_, nullptr, nullptr, &t));
        return (int)t;
    }

    // ---- neighbors (base_particle.hpp:265-288) -------------------------------------------
    void init_neighbors() { ok(dpmd_init_neighbors(ctx_)); }
    void update_neighbors() { ok(dpmd_update_neighbors(ctx_)); }
    void check_neighbors(bool force_update = false) { ok(dpmd_check_neighbors(ctx_, force_update ? 1 : 0)); }
    // neighbor_start / neighbor_ids (CSR, current order)
    void neighbor_lists(std::vector<int>& start, std::vector<int>& ids) const {
        const int n = (Derived::kKind == DPMD_RIGID_BUMPY) ? n_vertices() : n_particles();
        start.assign(n + 1, 0);
        ids.assign(std::max(1, n_neighbors()), 0);
        ok(dpmd_get_neighbors_csr(ctx_, start.data(), ids.data(), (int64_t)ids.size()));
        ids.resize(start[n]);
    }

    // ---- primitives (base_particle.hpp:380-420) ------------------------------------------
    void compute_forces() { ok(dpmd_compute_forces(ctx_, DPMD_FORCE_PBC)); }
    void compute_wall_forces() { ok(dpmd_compute_forces(ctx_, DPMD_FORCE_WALL)); }
    void compute_forces_per_system() { ok(dpmd_compute_forces(ctx_, DPMD_FORCE_PER_SYSTEM)); }
    void update_positions(F1 scale, double scale2 = 1.0) { ok(dpmd_update_positions(ctx_, scale.ptr(), scale2)); }
    void update_velocities(F1 scale, double scale2 = 1.0) { ok(dpmd_update_velocities(ctx_, scale.ptr(), scale2)); }
    void mix_velocities_and_forces(F1 alpha) { ok(dpmd_mix_velocities_and_forces(ctx_, alpha.ptr())); }
    void compute_damping_forces(F1 scale) { ok(dpmd_compute_damping_forces(ctx_, scale.ptr())); }
    void scale_velocities(F1 scale) { ok(dpmd_scale_velocities(ctx_, scale.ptr())); }
    void scale_velocities(double s) { scale_velocities(per_system(s)); }
    void scale_positions(F1 scale) { ok(dpmd_scale_positions(ctx_, scale.ptr())); }
    void zero_velocities() { ok(dpmd_zero_velocities(ctx_)); }
    void set_average_velocity(double vx = 0.0, double vy = 0.0) {
        F1 ax = per_system(vx), ay = per_system(vy);
        ok(dpmd_set_average_velocity(ctx_, ax.ptr(), ay.ptr()));
    }
    void set_temperature(F1 target, bool remove_average_velocity = true) {
        ok(dpmd_set_temperature(ctx_, target.ptr(), remove_average_velocity ? 1 : 0));
    }
    void set_temperature(double t, bool remove_average_velocity = true) {
        set_temperature(per_system(t), remove_average_velocity);
    }
    void increment_packing_fraction(F1 inc) { ok(dpmd_increment_packing_fraction(ctx_, inc.ptr())); }
    void increment_packing_fraction(double inc) { increment_packing_fraction(per_system(inc)); }
    void compute_ke() { ok(dpmd_compute_ke(ctx_)); }
    void compute_ke_total() { ok(dpmd_compute_ke_total(ctx_)); }
    void compute_pe_total() { ok(dpmd_compute_pe_total(ctx_)); }
    void compute_temperature() { ok(dpmd_compute_temperature(ctx_)); }
    void compute_packing_fraction() { ok(dpmd_compute_packing_fraction(ctx_)); }
    void compute_stress_tensor() { ok(dpmd_compute_stress_tensor(ctx_)); }
    void compute_pressure() { ok(dpmd_compute_pressure(ctx_)); }
    void compute_fpower_total() { ok(dpmd_compute_fpower_total(ctx_)); }
    void save_state(df::DeviceField1D<int> flag, int true_val) { ok(dpmd_save_state(ctx_, flag.ptr(), true_val)); }
    void save_state() { save_state(flags(1), 1); }
    void restore_state(df::DeviceField1D<int> flag, int true_val) { ok(dpmd_restore_state(ctx_, flag.ptr(), true_val)); }
    void restore_state() { restore_state(flags(1), 1); }

    // stateless Philox draws into pos (device_fields.cuh:219-298, stateless_rand_uniform)
    void set_rng_seed(uint64_t seed) { seed_ = seed; }
    void stateless_rand_uniform_pos(double lox, double hix, double loy, double hiy, uint64_t call) {
        ok(dpmd_rand_uniform_2d(ctx_, DPMD_F_POS_X, DPMD_F_POS_Y, seed_, call, lox, hix, loy, hiy));
    }
    void stateless_rand_normal_vel(double mx, double sx, double my, double sy, uint64_t call) {
        ok(dpmd_rand_normal_2d(ctx_, DPMD_F_VEL_X, DPMD_F_VEL_Y, seed_, call, mx, sx, my, sy));
    }
    void synchronize() { ok(dpmd_synchronize(ctx_)); }

    F1 per_system(double v) const {
        F1 f(n_systems());
        f.fill(v);
        return f;
    }

protected:
    void ok(int st) const { df::detail::check(st, ctx_); }
    void sizes(int* s, int* n, int* v, int* c) const { ok(dpmd_sizes(ctx_, s, n, v, c)); }
    df::DeviceField1D<int> flags(int v) const {
        df::DeviceField1D<int> f(n_systems());
        f.fill(v);
        return f;
    }
    static dpmd_ctx* make(int kind, void* stream) {
        dpmd_ctx* c = nullptr;
        df::detail::check(dpmd_ctx_create(&c, kind, stream));
        return c;
    }
};

namespace disk {
// md::disk::Disk (include/particles/disk.cuh:22; base_point_particle.hpp:86-89)
class Disk : public ParticleBase<Disk> {
public:
    static constexpr int kKind = DPMD_DISK;
    explicit Disk(void* hip_stream = nullptr) : ParticleBase<Disk>(kKind, hip_stream) {}
    F1 rad{ctx_, DPMD_F_RAD};

    // ---- contact observables (base_particle.hpp:39-44,85,101-103,649-702; disk.cu:434-700) --
    I1 contacts{ctx_, DPMD_F_CONTACTS};
    F2 overlaps{ctx_, DPMD_F_OVERLAP, DPMD_F_CONTACT_COUNT};
    I1 n_contacts_total{ctx_, DPMD_F_N_CONTACTS_TOTAL};
    // per neighbor entry (CSR order of neighbor_lists), sized on each compute call
    df::DeviceField2D<int> pair_ids;
    F1 pair_dist;
    F2 pair_forces;
    F2 hessian_xx, hessian_xy, hessian_yx, hessian_yy;

    void compute_contacts() { ok(dpmd_compute_contacts(ctx_)); }
    void compute_overlaps() { ok(dpmd_compute_overlaps(ctx_)); }
    void compute_n_contacts_total() { ok(dpmd_compute_n_contacts_total(ctx_)); }
    void compute_pair_dist() {
        const int n = size_pairs();
        pair_dist.resize(n);
        ok(dpmd_compute_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), pair_dist.ptr(), nullptr, nullptr,
                                         nullptr, n));
    }
    void compute_pair_forces() {
        const int n = size_pairs();
        pair_forces.resize(n);
        ok(dpmd_compute_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), nullptr, pair_forces.xptr(),
                                         pair_forces.yptr(), nullptr, n));
    }
    void compute_hessian() {
        const int n = size_pairs();
        F2* h[] = {&hessian_xx, &hessian_xy, &hessian_yx, &hessian_yy};
        double* cols[8];
        for (int m = 0; m < 4; ++m) {
            h[m]->resize(n);
            cols[2 * m] = h[m]->xptr();
            cols[2 * m + 1] = h[m]->yptr();
        }
        ok(dpmd_compute_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), nullptr, nullptr, nullptr, cols, n));
    }

private:
    int size_pairs() {
        const int n = n_neighbors();
        pair_ids.resize(n);
        return n;
    }
};
}  // namespace disk

namespace rigid_bumpy {
// md::rigid_bumpy::RigidBumpy (include/particles/rigid_bumpy.cuh:24-45;
// base_poly_particle.hpp:104-120)
class RigidBumpy : public ParticleBase<RigidBumpy> {
public:
    static constexpr int kKind = DPMD_RIGID_BUMPY;
    explicit RigidBumpy(void* hip_stream = nullptr) : ParticleBase<RigidBumpy>(kKind, hip_stream) {}
    F1 angle{ctx_, DPMD_F_ANGLE};
    F1 angular_vel{ctx_, DPMD_F_ANGULAR_VEL};
    F1 torque{ctx_, DPMD_F_TORQUE};
    F1 moment_inertia{ctx_, DPMD_F_MOMENT_INERTIA};
    F2 vertex_pos{ctx_, DPMD_F_VERTEX_POS_X, DPMD_F_VERTEX_POS_Y};
    F2 vertex_force{ctx_, DPMD_F_VERTEX_FORCE_X, DPMD_F_VERTEX_FORCE_Y};
    F1 vertex_pe{ctx_, DPMD_F_VERTEX_PE};
    F1 vertex_rad{ctx_, DPMD_F_VERTEX_RAD};
    F1 vertex_mass{ctx_, DPMD_F_VERTEX_MASS};
    I1 vertex_particle_id{ctx_, DPMD_F_VERTEX_PARTICLE_ID};
    I1 particle_offset{ctx_, DPMD_F_PARTICLE_OFFSET};
    I1 n_vertices_per_particle{ctx_, DPMD_F_N_VERTICES_PER_PARTICLE};
    I1 vertex_system_id{ctx_, DPMD_F_VERTEX_SYSTEM_ID};
    I1 vertex_system_offset{ctx_, DPMD_F_VERTEX_SYSTEM_OFFSET};

    // ---- contact observables (base_poly_particle.hpp:299-375; rigid_bumpy.cu:759-966,
    // 1261-1327,1598-1690).  Per particle-neighbor entry fields are sized on each call.
    I1 contacts{ctx_, DPMD_F_CONTACTS};
    F2 overlaps{ctx_, DPMD_F_OVERLAP, DPMD_F_CONTACT_COUNT};   // per vertex (current order)
    I1 n_contacts_total{ctx_, DPMD_F_N_CONTACTS_TOTAL};
    std::vector<int> particle_neighbor_start, particle_neighbor_ids;   // host copies
    df::DeviceField2D<int> pair_ids, pair_vertex_contacts;
    F1 pair_dist, friction_coeff;
    F2 pair_forces;

    void build_particle_neighbors() {
        int64_t tot = 0;
        ok(dpmd_particle_neighbors(ctx_, nullptr, nullptr, 0, &tot));
        particle_neighbor_start.assign(n_particles() + 1, 0);
        particle_neighbor_ids.assign(std::max<int64_t>(1, tot), 0);
        ok(dpmd_particle_neighbors(ctx_, particle_neighbor_start.data(), particle_neighbor_ids.data(),
                                   (int64_t)particle_neighbor_ids.size(), &tot));
        particle_neighbor_ids.resize(tot);
    }
    void compute_contacts() {
        build_particle_neighbors();
        const int n = size_pairs();
        pair_vertex_contacts.resize(n);
        ok(dpmd_compute_rigid_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), nullptr, nullptr, nullptr,
                                               pair_vertex_contacts.xptr(), pair_vertex_contacts.yptr(), nullptr, n));
    }
    void compute_n_contacts_total() { ok(dpmd_compute_n_contacts_total(ctx_)); }
    void compute_overlaps() { ok(dpmd_compute_overlaps(ctx_)); }
    void compute_pair_dist() {
        const int n = size_pairs();
        pair_dist.resize(n);
        ok(dpmd_compute_rigid_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), pair_dist.ptr(), nullptr,
                                               nullptr, nullptr, nullptr, nullptr, n));
    }
    void compute_pair_forces() {
        const int n = size_pairs();
        pair_forces.resize(n);
        ok(dpmd_compute_rigid_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), nullptr, pair_forces.xptr(),
                                               pair_forces.yptr(), nullptr, nullptr, nullptr, n));
    }
    // hessian_ab (N_pairs, 2): .x = entry's share of H_(i a),(i b), .y = H_(i a),(j b)
    F2 hessian_xx, hessian_xy, hessian_xt, hessian_yx, hessian_yy, hessian_yt, hessian_tx, hessian_ty, hessian_tt;
    void compute_hessian() {
        const int n = size_pairs();
        F2* h[] = {&hessian_xx, &hessian_xy, &hessian_xt, &hessian_yx, &hessian_yy,
                   &hessian_yt, &hessian_tx, &hessian_ty, &hessian_tt};
        double* cols[18];
        for (int m = 0; m < 9; ++m) {
            h[m]->resize(n);
            cols[2 * m] = h[m]->xptr();
            cols[2 * m + 1] = h[m]->yptr();
        }
        ok(dpmd_compute_rigid_hessian(ctx_, cols, n));
    }
    void compute_friction_coeff() {
        const int n = size_pairs();
        friction_coeff.resize(n);
        ok(dpmd_compute_rigid_pair_observables(ctx_, pair_ids.xptr(), pair_ids.yptr(), nullptr, nullptr, nullptr,
                                               nullptr, nullptr, friction_coeff.ptr(), n));
    }

private:
    int size_pairs() {
        int64_t tot = 0;
        ok(dpmd_particle_neighbors(ctx_, nullptr, nullptr, 0, &tot));
        pair_ids.resize((int)tot);
        return (int)tot;
    }
};
}  // namespace rigid_bumpy

}  // namespace md
