"""dpmd_amd: MI355X-native (gfx950, HIP) batched 2D granular-MD hot path.

The physics runs in libdpmd_amd.so (C ABI: include/dpmd/dpmd.h); this package is the host
mirror of the reference's particle / integrator / routine API plus workload generators.
"""
from . import _lib
from .system import (FIRE, DeviceArray, Disk, RigidBumpy, VelocityVerlet, minimize_fire, run_fire)

__all__ = ["Disk", "RigidBumpy", "FIRE", "VelocityVerlet", "minimize_fire", "run_fire", "DeviceArray", "_lib"]
