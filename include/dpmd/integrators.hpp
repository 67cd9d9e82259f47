// md::integrators -- VelocityVerlet[Wall], DampedVelocityVerlet[Wall] and FIRE[Wall] over the
// C ABI.
//
// Reference interface: include/integrators/base_integrator.hpp:8-32,
// include/integrators/velocity_verlet.cuh:10-78, damped_velocity_verlet.cuh:10-87,
// include/integrators/fire.cuh:44-178.  Each
// step() is one fused device step (drift, device-side rebuild check, forces); the particle
// object is held by reference and must outlive the integrator, as in the reference.
#pragma once

#include "particles.hpp"

namespace md {
namespace integrators {

template <class ParticleT, int Mode>
class BaseVelocityVerlet {
public:
    BaseVelocityVerlet(ParticleT& p, double dt) : p_(p), dt_(p.per_system(dt)) {}
    BaseVelocityVerlet(ParticleT& p, const df::DeviceField1D<double>& dt) : p_(p), dt_(dt) {
        if (dt.size() != p.n_systems())
            throw std::invalid_argument("BaseVelocityVerlet: dt_init must have the same size as the number of systems");
    }
    BaseVelocityVerlet(const BaseVelocityVerlet&) = delete;
    BaseVelocityVerlet& operator=(const BaseVelocityVerlet&) = delete;

    void init() { ok(dpmd_vv_init(p_.ctx(), dt_.ptr(), Mode)); }
    void step() { ok(dpmd_vv_step(p_.ctx(), 1)); }
    // n device-driven steps (lean batches, optionally graph captured)
    void run(int n, bool use_graph = true) { ok(dpmd_vv_run(p_.ctx(), n, use_graph ? 1 : 0)); }

protected:
    void ok(int st) { df::detail::check(st, p_.ctx()); }
    ParticleT& p_;
    df::DeviceField1D<double> dt_;
};

template <class ParticleT>
class VelocityVerlet final : public BaseVelocityVerlet<ParticleT, DPMD_FORCE_PBC> {
    using BaseVelocityVerlet<ParticleT, DPMD_FORCE_PBC>::BaseVelocityVerlet;
};
template <class ParticleT>
class VelocityVerletWall final : public BaseVelocityVerlet<ParticleT, DPMD_FORCE_WALL> {
    using BaseVelocityVerlet<ParticleT, DPMD_FORCE_WALL>::BaseVelocityVerlet;
};
// addition: per-system boundary (ParticleT::boundary), for batches that mix box types
template <class ParticleT>
class VelocityVerletPerSystem final : public BaseVelocityVerlet<ParticleT, DPMD_FORCE_PER_SYSTEM> {
    using BaseVelocityVerlet<ParticleT, DPMD_FORCE_PER_SYSTEM>::BaseVelocityVerlet;
};

// DampedVelocityVerlet[Wall] (damped_velocity_verlet.cuh:10-87): velocity Verlet with
// compute_damping_forces(damping) after the forces of every step.
template <class ParticleT, int Mode>
class BaseDampedVelocityVerlet {
public:
    BaseDampedVelocityVerlet(ParticleT& p, double dt, double damping)
        : p_(p), dt_(p.per_system(dt)), damping_(p.per_system(damping)) {}
    BaseDampedVelocityVerlet(ParticleT& p, const df::DeviceField1D<double>& dt, const df::DeviceField1D<double>& damping)
        : p_(p), dt_(dt), damping_(damping) {
        if (dt.size() != p.n_systems())
            throw std::invalid_argument("BaseDampedVelocityVerlet: dt_init must have the same size as the number of systems");
        if (damping.size() != p.n_systems())
            throw std::invalid_argument(
                "BaseDampedVelocityVerlet: damping_init must have the same size as the number of systems");
    }
    BaseDampedVelocityVerlet(const BaseDampedVelocityVerlet&) = delete;
    BaseDampedVelocityVerlet& operator=(const BaseDampedVelocityVerlet&) = delete;

    void init() { ok(dpmd_damped_vv_init(p_.ctx(), dt_.ptr(), damping_.ptr(), Mode)); }
    void step() { ok(dpmd_vv_step(p_.ctx(), 1)); }
    void run(int n, bool use_graph = true) { ok(dpmd_vv_run(p_.ctx(), n, use_graph ? 1 : 0)); }

protected:
    void ok(int st) { df::detail::check(st, p_.ctx()); }
    ParticleT& p_;
    df::DeviceField1D<double> dt_, damping_;
};

template <class ParticleT>
class DampedVelocityVerlet final : public BaseDampedVelocityVerlet<ParticleT, DPMD_FORCE_PBC> {
    using BaseDampedVelocityVerlet<ParticleT, DPMD_FORCE_PBC>::BaseDampedVelocityVerlet;
};
template <class ParticleT>
class DampedVelocityVerletWall final : public BaseDampedVelocityVerlet<ParticleT, DPMD_FORCE_WALL> {
    using BaseDampedVelocityVerlet<ParticleT, DPMD_FORCE_WALL>::BaseDampedVelocityVerlet;
};

template <class ParticleT, int Mode>
class BaseFIRE {
public:
    BaseFIRE(ParticleT& p, double dt) : p_(p), dt_(p.per_system(dt)) {}
    BaseFIRE(ParticleT& p, const df::DeviceField1D<double>& dt) : p_(p), dt_(dt) {}
    BaseFIRE(const BaseFIRE&) = delete;
    BaseFIRE& operator=(const BaseFIRE&) = delete;

    void init() { ok(dpmd_fire_init(p_.ctx(), dt_.ptr(), Mode)); }
    void step() { ok(dpmd_fire_step(p_.ctx())); }
    // fire.cuh:98-109 (argument order as declared there)
    bool converged(double avg_pe_diff_target, double avg_pe_target) {
        int all = 0;
        ok(dpmd_fire_converged(p_.ctx(), avg_pe_diff_target, avg_pe_target, &all));
        return all != 0;
    }
    // addition: stop a system once every |F_i| < tol
    bool converged_max_force(double tol) {
        int all = 0;
        ok(dpmd_fire_converged_max_force(p_.ctx(), tol, &all));
        return all != 0;
    }
    df::DeviceField1D<double> get_dt() {
        df::DeviceField1D<double> out(p_.n_systems());
        ok(dpmd_fire_get_dt(p_.ctx(), out.ptr()));
        ok(dpmd_synchronize(p_.ctx()));
        return out;
    }

protected:
    void ok(int st) { df::detail::check(st, p_.ctx()); }
    ParticleT& p_;
    df::DeviceField1D<double> dt_;
};

template <class ParticleT>
class FIRE final : public BaseFIRE<ParticleT, DPMD_FORCE_PBC> {
    using BaseFIRE<ParticleT, DPMD_FORCE_PBC>::BaseFIRE;
};
template <class ParticleT>
class FIREWall final : public BaseFIRE<ParticleT, DPMD_FORCE_WALL> {
    using BaseFIRE<ParticleT, DPMD_FORCE_WALL>::BaseFIRE;
};
template <class ParticleT>
class FIREPerSystem final : public BaseFIRE<ParticleT, DPMD_FORCE_PER_SYSTEM> {
    using BaseFIRE<ParticleT, DPMD_FORCE_PER_SYSTEM>::BaseFIRE;
};

}  // namespace integrators
}  // namespace md
