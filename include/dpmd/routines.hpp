// md::routines -- minimize_fire family, the device-driven FIRE run, the bisection jammer and
// the rescale-NVT compression driver.
//
// Reference interface: include/routines/minimizers.cuh:10-121, include/routines/jammers.cuh:60-256,
// tests/nvt_rescale_disk_pbc_compress.cu:14-68.  The host loops keep the
// reference's call pattern (step(), then converged() with its arguments in the order the
// reference passes them: minimizers.cuh:25 hands (avg_pe_target, avg_pe_diff_target) to
// converged(avg_pe_diff_target, avg_pe_target); every caller passes equal values).
// minimize_fire_device runs the same minimisation with the stopping test evaluated per
// system on the GPU (no host round trip per step).
#pragma once

#include <algorithm>
#include <iostream>

#include "integrators.hpp"

namespace md {
namespace routines {

namespace detail {
template <class Fire>
int loop(Fire& fire, int max_steps, double avg_pe_target, double avg_pe_diff_target) {
    fire.init();
    bool converged = false;
    int step = 0;
    while (step < max_steps && !converged) {
        fire.step();
        converged = fire.converged(avg_pe_target, avg_pe_diff_target);
        step++;
    }
    if (!converged) std::cout << "Did not converge in " << max_steps << " steps" << std::endl;
    return step;
}
}  // namespace detail

template <class ParticleT>
int minimize_fire(ParticleT& P, df::DeviceField1D<double>& dt, int max_steps, double avg_pe_target,
                  double avg_pe_diff_target) {
    md::integrators::FIRE<ParticleT> fire(P, dt);
    return detail::loop(fire, max_steps, avg_pe_target, avg_pe_diff_target);
}

template <class ParticleT>
int minimize_fire_copy_dt(ParticleT& P, df::DeviceField1D<double>& dt, int max_steps, double avg_pe_target,
                          double avg_pe_diff_target) {
    md::integrators::FIRE<ParticleT> fire(P, dt);
    const int steps = detail::loop(fire, max_steps, avg_pe_target, avg_pe_diff_target);
    dt = fire.get_dt();
    P.scale_velocities(0.0);
    return steps;
}

template <class ParticleT>
int minimize_fire_wall(ParticleT& P, df::DeviceField1D<double>& dt, int max_steps, double avg_pe_target,
                       double avg_pe_diff_target) {
    md::integrators::FIREWall<ParticleT> fire(P, dt);
    return detail::loop(fire, max_steps, avg_pe_target, avg_pe_diff_target);
}

template <class ParticleT>
int minimize_fire_wall_copy_dt(ParticleT& P, df::DeviceField1D<double>& dt, int max_steps, double avg_pe_target,
                               double avg_pe_diff_target) {
    md::integrators::FIREWall<ParticleT> fire(P, dt);
    const int steps = detail::loop(fire, max_steps, avg_pe_target, avg_pe_diff_target);
    dt = fire.get_dt();
    P.scale_velocities(0.0);
    return steps;
}

// Device-driven FIRE (dpmd_fire_minimize): predicate DPMD_FIRE_PRED_REFERENCE with
// (a, b) = (avg_pe_diff_target, avg_pe_target), or DPMD_FIRE_PRED_MAX_FORCE with a = tol.
// mode: DPMD_FORCE_PBC / _WALL / _PER_SYSTEM.  Returns steps taken; *converged is set.
template <class ParticleT>
int minimize_fire_device(ParticleT& P, const df::DeviceField1D<double>& dt, int max_steps, int predicate, double a,
                         double b, int mode, bool use_graph = true, bool* converged = nullptr) {
    int steps = 0, conv = 0;
    df::detail::check(dpmd_fire_minimize(P.ctx(), dt.ptr(), max_steps, predicate, a, b, mode, use_graph ? 1 : 0, 0,
                                         &steps, &conv),
                      P.ctx());
    if (converged) *converged = conv != 0;
    return steps;
}

// jam_binary_search_pbc / _wall (jammers.cuh:60-256): the reference's bisection jammer, run
// by dpmd_jam_binary_search with every particle operation on the device.
template <class ParticleT>
int jam_binary_search_pbc(ParticleT& P, df::DeviceField1D<double> dt, int max_compression_steps,
                          int max_minimization_steps, double avg_pe_target, double avg_pe_diff_target,
                          double phi_increment, double phi_tolerance) {
    int rounds = 0;
    df::detail::check(dpmd_jam_binary_search(P.ctx(), dt.ptr(), max_compression_steps, max_minimization_steps,
                                             avg_pe_target, avg_pe_diff_target, phi_increment, phi_tolerance,
                                             DPMD_FORCE_PBC, 1, nullptr, nullptr, &rounds),
                      P.ctx());
    return rounds;
}
template <class ParticleT>
int jam_binary_search_wall(ParticleT& P, df::DeviceField1D<double> dt, int max_compression_steps,
                           int max_minimization_steps, double avg_pe_target, double avg_pe_diff_target,
                           double phi_increment, double phi_tolerance) {
    int rounds = 0;
    df::detail::check(dpmd_jam_binary_search(P.ctx(), dt.ptr(), max_compression_steps, max_minimization_steps,
                                             avg_pe_target, avg_pe_diff_target, phi_increment, phi_tolerance,
                                             DPMD_FORCE_WALL, 1, nullptr, nullptr, &rounds),
                      P.ctx());
    return rounds;
}

// Rescale-thermostat NVT with affine compression over the integrator last initialised
// (VelocityVerlet or DampedVelocityVerlet): the step loop of the reference's compression
// drivers (tests/nvt_rescale_disk_pbc_compress.cu:41-68,
// tests/damped_dynamics_compress_to_packing_fraction.cu:41-52).
template <class ParticleT>
void nvt_rescale_run(ParticleT& P, int n_steps, const df::DeviceField1D<double>* phi_increment, int compress_every,
                     const df::DeviceField1D<double>* temperature_target, int temperature_every, int step0 = 0,
                     bool remove_average_velocity = true, bool use_graph = true) {
    df::detail::check(dpmd_nvt_rescale_run(P.ctx(), n_steps, step0, phi_increment ? phi_increment->ptr() : nullptr,
                                           compress_every, temperature_target ? temperature_target->ptr() : nullptr,
                                           temperature_every, remove_average_velocity ? 1 : 0, use_graph ? 1 : 0),
                      P.ctx());
}

// The two phases of tests/nvt_rescale_disk_pbc_compress.cu:14-68: n_steps with phi_increment
// spread over compressions every min(n_steps, 100) steps and the thermostat every
// min(n_steps, 10) steps, then n_steps with the thermostat only.
template <class ParticleT>
void nvt_rescale_compress(ParticleT& P, const df::DeviceField1D<double>& dt, int n_steps, double phi_increment,
                          double temperature_target) {
    const int cf = std::min(n_steps, 100);
    const int tf = std::min(n_steps, 10);
    const double phi_step = phi_increment / (n_steps / cf);
    md::integrators::VelocityVerlet<ParticleT> vv(P, dt);
    vv.init();
    df::DeviceField1D<double> inc = P.per_system(phi_step), T = P.per_system(temperature_target);
    nvt_rescale_run(P, n_steps, &inc, cf, &T, tf, 0);
    nvt_rescale_run(P, n_steps, nullptr, 0, &T, tf, n_steps);
}

}  // namespace routines
}  // namespace md
