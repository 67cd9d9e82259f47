// Rigid bumpy polygons (placeholder until the vertex kernels land).
#include "internal.hpp"

namespace dpmd {
[[noreturn]] static void no_rigid() { throw StateError("rigid bumpy kernels are not built yet"); }
void rigid_compute_forces(dpmd_ctx*, int) { no_rigid(); }
void rigid_update_positions(dpmd_ctx*, const double*, double) { no_rigid(); }
void rigid_update_velocities(dpmd_ctx*, const double*, double) { no_rigid(); }
void rigid_scale_velocities(dpmd_ctx*, const double*) { no_rigid(); }
void rigid_scale_positions(dpmd_ctx*, const double*) { no_rigid(); }
void rigid_mix(dpmd_ctx*, const double*) { no_rigid(); }
void rigid_stress(dpmd_ctx*) { no_rigid(); }
void rigid_ke(dpmd_ctx*) { no_rigid(); }
void rigid_fire_drift(dpmd_ctx*, Gate) { no_rigid(); }
void rigid_fire_force(dpmd_ctx*, Gate, int, int, int, double, double) { no_rigid(); }
void rigid_vv_drift(dpmd_ctx*, Gate) { no_rigid(); }
void rigid_vv_force(dpmd_ctx*, Gate) { no_rigid(); }
}  // namespace dpmd
