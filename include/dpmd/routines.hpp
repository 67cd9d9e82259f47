// md::routines -- minimize_fire family and the device-driven FIRE run.
//
// Reference interface: include/routines/minimizers.cuh:10-121.  The host loops keep the
// reference's call pattern (step(), then converged() with its arguments in the order the
// reference passes them: minimizers.cuh:25 hands (avg_pe_target, avg_pe_diff_target) to
// converged(avg_pe_diff_target, avg_pe_target); every caller passes equal values).
// minimize_fire_device runs the same minimisation with the stopping test evaluated per
// system on the GPU (no host round trip per step).
#pragma once

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

}  // namespace routines
}  // namespace md
