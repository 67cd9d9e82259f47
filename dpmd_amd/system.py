"""Python host mirror of the reference particle / integrator / routine API over the C ABI.

Names follow the reference: ``Disk`` (md::disk::Disk, include/particles/disk.cuh:22),
``RigidBumpy`` (md::rigid_bumpy::RigidBumpy, include/particles/rigid_bumpy.cuh:24),
``VelocityVerlet`` / ``FIRE`` (include/integrators/{velocity_verlet,fire}.cuh) and
``minimize_fire`` (include/routines/minimizers.cuh:10-31).  Every call goes through
libdpmd_amd.so; nothing here computes physics on the CPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from ._lib import F, check, lib


class DeviceArray:
    """A small device buffer (the role of df::DeviceField1D for per-system parameters)."""

    def __init__(self, values, dtype=np.float64):
        a = np.ascontiguousarray(np.asarray(values, dtype=dtype))
        self.dtype = a.dtype
        self.n = a.size
        p = C.c_void_p()
        check(lib().dpmd_malloc(C.byref(p), max(1, a.nbytes)))
        self.ptr = p.value
        if a.nbytes:
            check(lib().dpmd_memcpy(self.ptr, a.ctypes.data, a.nbytes, 0))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if out.nbytes:
            check(lib().dpmd_memcpy(out.ctypes.data, self.ptr, out.nbytes, 1))
        return out

    def fill(self, v) -> None:
        a = np.full(self.n, v, self.dtype)
        check(lib().dpmd_memcpy(self.ptr, a.ctypes.data, a.nbytes, 0))

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().dpmd_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def _per_system(x, S: int, dtype=np.float64) -> DeviceArray:
    if isinstance(x, DeviceArray):
        return x
    a = np.asarray(x, dtype=dtype)
    if a.ndim == 0:
        a = np.full(S, a, dtype)
    if a.size != S:
        raise ValueError(f"expected {S} per-system values, got {a.size}")
    return DeviceArray(a, dtype)


class _Particles:
    kind = L.DISK

    def __init__(self, stream: Optional[int] = None):
        p = C.c_void_p()
        check(lib().dpmd_ctx_create(C.byref(p), self.kind, stream))
        self.ctx = p.value
        self._keep = []
        self.neighbor_method = L.NEIGHBOR_NAIVE

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                lib().dpmd_ctx_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    def _c(self, status: int) -> None:
        check(status, self.ctx)

    # ---- allocation / fields ------------------------------------------------------------
    def allocate_systems(self, S: int):
        self._c(lib().dpmd_allocate_systems(self.ctx, S))

    def allocate_particles(self, N: int):
        self._c(lib().dpmd_allocate_particles(self.ctx, N))

    def allocate_vertices(self, Nv: int):
        self._c(lib().dpmd_allocate_vertices(self.ctx, Nv))

    def set_neighbor_method(self, m: int):
        self._c(lib().dpmd_set_neighbor_method(self.ctx, m))
        self.neighbor_method = m

    def field_len(self, name: str) -> int:
        n = C.c_int64()
        self._c(lib().dpmd_field(self.ctx, F[name], None, C.byref(n)))
        return n.value

    def field_ptr(self, name: str) -> int:
        p = C.c_void_p()
        self._c(lib().dpmd_field(self.ctx, F[name], C.byref(p), None))
        return p.value

    def set(self, name: str, values) -> None:
        dt = np.int32 if name in L.INT_FIELDS else np.float64
        a = np.ascontiguousarray(np.asarray(values, dtype=dt))
        self._c(lib().dpmd_set_field(self.ctx, F[name], a.ctypes.data, a.size))

    def get(self, name: str, order: int = L.ORDER_CURRENT) -> np.ndarray:
        dt = np.int32 if name in L.INT_FIELDS else np.float64
        out = np.empty(self.field_len(name), dt)
        self._c(lib().dpmd_get_field(self.ctx, F[name], out.ctypes.data, out.size, order))
        return out

    def sizes(self):
        S, N, Nv, Cc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._c(lib().dpmd_sizes(self.ctx, C.byref(S), C.byref(N), C.byref(Nv), C.byref(Cc)))
        return S.value, N.value, Nv.value, Cc.value

    def n_systems(self): return self.sizes()[0]
    def n_particles(self): return self.sizes()[1]
    def n_vertices(self): return self.sizes()[2]
    def n_cells(self): return self.sizes()[3]

    def sync(self):
        self._c(lib().dpmd_sync(self.ctx))

    # ---- neighbors ------------------------------------------------------------------------
    def init_neighbors(self): self._c(lib().dpmd_init_neighbors(self.ctx))
    def update_neighbors(self): self._c(lib().dpmd_update_neighbors(self.ctx))
    def check_neighbors(self, force_update: bool = False):
        self._c(lib().dpmd_check_neighbors(self.ctx, int(force_update)))

    def set_neighbor_capacity(self, k: int): self._c(lib().dpmd_set_neighbor_capacity(self.ctx, k))

    def neighbor_info(self):
        cap, mx, tot = C.c_int(), C.c_int(), C.c_int64()
        self._c(lib().dpmd_neighbor_info(self.ctx, C.byref(cap), C.byref(mx), C.byref(tot)))
        return cap.value, mx.value, tot.value

    def n_neighbors(self) -> int: return self.neighbor_info()[2]

    def neighbors_csr(self):
        tot = self.n_neighbors()
        n = self.n_vertices() if self.kind == L.RIGID_BUMPY else self.n_particles()
        start = np.empty(n + 1, np.int32)
        ids = np.empty(max(1, tot), np.int32)
        self._c(lib().dpmd_get_neighbors_csr(self.ctx, start.ctypes.data, ids.ctypes.data, ids.size))
        return start, ids[:tot]

    # ---- primitives -------------------------------------------------------------------------
    def _ps(self, x, dtype=np.float64):
        d = _per_system(x, self.n_systems(), dtype)
        self._keep = [d]
        return d

    def compute_forces(self): self._c(lib().dpmd_compute_forces(self.ctx, L.FORCE_PBC))
    def compute_wall_forces(self): self._c(lib().dpmd_compute_forces(self.ctx, L.FORCE_WALL))
    def compute_forces_mode(self, mode: int): self._c(lib().dpmd_compute_forces(self.ctx, mode))

    def update_positions(self, scale, scale2: float = 1.0):
        self._c(lib().dpmd_update_positions(self.ctx, self._ps(scale).ptr, scale2))

    def update_velocities(self, scale, scale2: float = 1.0):
        self._c(lib().dpmd_update_velocities(self.ctx, self._ps(scale).ptr, scale2))

    # ---- contact observables (base_particle.hpp:649-702; disk.cu:434-700) -------------------
    def compute_contacts(self): self._c(lib().dpmd_compute_contacts(self.ctx))
    def compute_overlaps(self): self._c(lib().dpmd_compute_overlaps(self.ctx))
    def compute_n_contacts_total(self): self._c(lib().dpmd_compute_n_contacts_total(self.ctx))

    def pair_observables(self, ids: bool = True, dist: bool = True, forces: bool = True, hessian: bool = True):
        """compute_pair_dist / compute_pair_forces / compute_hessian in one pass.  Returns a dict
        of host arrays indexed by neighbor entry (CSR order of neighbors_csr()): pair_i, pair_j
        (canonical ids), dist, fx, fy and the Hessian columns h_ix_ix, h_ix_jx, h_ix_iy, h_ix_jy,
        h_iy_ix, h_iy_jx, h_iy_iy, h_iy_jy."""
        n = self.n_neighbors()
        bufs = {}
        def dev(name, dtype=np.float64):
            bufs[name] = DeviceArray(np.zeros(max(1, n), dtype), dtype)
            return bufs[name].ptr
        pi = dev("pair_i", np.int32) if ids else None
        pj = dev("pair_j", np.int32) if ids else None
        d = dev("dist") if dist else None
        fx = dev("fx") if forces else None
        fy = dev("fy") if forces else None
        h = None
        if hessian:
            names = ("h_ix_ix", "h_ix_jx", "h_ix_iy", "h_ix_jy", "h_iy_ix", "h_iy_jx", "h_iy_iy", "h_iy_jy")
            h = (C.c_void_p * 8)(*[dev(k) for k in names])
        self._c(lib().dpmd_compute_pair_observables(self.ctx, pi, pj, d, fx, fy, h, n))
        return {k: v.to_host()[:n] for k, v in bufs.items()}

    def compute_damping_forces(self, scale):
        self._c(lib().dpmd_compute_damping_forces(self.ctx, self._ps(scale).ptr))

    def scale_velocities(self, scale): self._c(lib().dpmd_scale_velocities(self.ctx, self._ps(scale).ptr))
    def scale_positions(self, scale): self._c(lib().dpmd_scale_positions(self.ctx, self._ps(scale).ptr))

    def mix_velocities_and_forces(self, alpha):
        self._c(lib().dpmd_mix_velocities_and_forces(self.ctx, self._ps(alpha).ptr))

    def zero_velocities(self): self._c(lib().dpmd_zero_velocities(self.ctx))

    def compute_ke(self): self._c(lib().dpmd_compute_ke(self.ctx))
    def compute_ke_total(self): self._c(lib().dpmd_compute_ke_total(self.ctx))
    def compute_pe_total(self): self._c(lib().dpmd_compute_pe_total(self.ctx))
    def compute_fpower_total(self): self._c(lib().dpmd_compute_fpower_total(self.ctx))
    def compute_temperature(self): self._c(lib().dpmd_compute_temperature(self.ctx))
    def compute_packing_fraction(self): self._c(lib().dpmd_compute_packing_fraction(self.ctx))
    def compute_stress_tensor(self): self._c(lib().dpmd_compute_stress_tensor(self.ctx))
    def compute_pressure(self): self._c(lib().dpmd_compute_pressure(self.ctx))
    def compute_max_force(self): self._c(lib().dpmd_compute_max_force(self.ctx))

    def set_average_velocity(self, vx=0.0, vy=0.0):
        ax, ay = self._ps(vx), _per_system(vy, self.n_systems())
        self._keep = [ax, ay]
        self._c(lib().dpmd_set_average_velocity(self.ctx, ax.ptr, ay.ptr))

    def set_temperature(self, T, remove_average_velocity: bool = True):
        self._c(lib().dpmd_set_temperature(self.ctx, self._ps(T).ptr, int(remove_average_velocity)))

    def increment_packing_fraction(self, inc):
        self._c(lib().dpmd_increment_packing_fraction(self.ctx, self._ps(inc).ptr))

    def save_state(self, flag=None, true_val: int = 1):
        f = self._ps(1 if flag is None else flag, np.int32)
        self._c(lib().dpmd_save_state(self.ctx, f.ptr, true_val))

    def restore_state(self, flag=None, true_val: int = 1):
        f = self._ps(1 if flag is None else flag, np.int32)
        self._c(lib().dpmd_restore_state(self.ctx, f.ptr, true_val))

    def rand_uniform_2d(self, fx: str, fy: str, seed: int, call: int, lox, hix, loy, hiy):
        self._c(lib().dpmd_rand_uniform_2d(self.ctx, F[fx], F[fy], seed, call, lox, hix, loy, hiy))

    def rand_normal_2d(self, fx: str, fy: str, seed: int, call: int, mx=0.0, sx=1.0, my=0.0, sy=1.0):
        self._c(lib().dpmd_rand_normal_2d(self.ctx, F[fx], F[fy], seed, call, mx, sx, my, sy))

    def synchronize(self): self._c(lib().dpmd_synchronize(self.ctx))


class Disk(_Particles):
    kind = L.DISK


class RigidBumpy(_Particles):
    kind = L.RIGID_BUMPY

    def particle_neighbors(self):
        """build_particle_neighbors (base_poly_particle.hpp:299-375): (start[N+1], ids)."""
        tot = C.c_int64()
        self._c(lib().dpmd_particle_neighbors(self.ctx, None, None, 0, C.byref(tot)))
        st = np.zeros(self.n_particles() + 1, np.int32)
        ids = np.zeros(max(1, tot.value), np.int32)
        self._c(lib().dpmd_particle_neighbors(self.ctx, st.ctypes.data, ids.ctypes.data, ids.size, C.byref(tot)))
        return st, ids[:tot.value]

    def hessian(self):
        """compute_hessian per particle-neighbor entry (dpmd_compute_rigid_hessian): dict with
        'h_i{a}_i{b}' and 'h_i{a}_j{b}' for a, b in x, y, t (theta), plus pair_i / pair_j."""
        st, ids = self.particle_neighbors()
        n = len(ids)
        cols = [DeviceArray(np.zeros(max(1, n))) for _ in range(18)]
        arr = (C.c_void_p * 18)(*[d.ptr for d in cols])
        self._c(lib().dpmd_compute_rigid_hessian(self.ctx, arr, n))
        out = {"pair_i": np.repeat(np.arange(self.n_particles(), dtype=np.int32), np.diff(st)), "pair_j": ids}
        q = "xyt"
        for m in range(9):
            a, b = q[m // 3], q[m % 3]
            out[f"h_i{a}_i{b}"] = cols[2 * m].to_host()[:n]
            out[f"h_i{a}_j{b}"] = cols[2 * m + 1].to_host()[:n]
        return out

    def pair_observables(self):
        """Rigid contacts / pair distances / friction / pair forces per particle-neighbor entry
        (dpmd_compute_rigid_pair_observables).  Returns a dict of host arrays."""
        tot = C.c_int64()
        self._c(lib().dpmd_particle_neighbors(self.ctx, None, None, 0, C.byref(tot)))
        n = tot.value
        names = (("pair_i", np.int32), ("pair_j", np.int32), ("dist", np.float64), ("fx", np.float64),
                 ("fy", np.float64), ("vertex_contacts_i", np.int32), ("vertex_contacts_j", np.int32),
                 ("friction", np.float64))
        bufs = {k: DeviceArray(np.zeros(max(1, n), t), t) for k, t in names}
        self._c(lib().dpmd_compute_rigid_pair_observables(self.ctx, *[bufs[k].ptr for k, _ in names], n))
        return {k: v.to_host()[:n] for k, v in bufs.items()}


class VelocityVerlet:
    """VelocityVerlet / VelocityVerletWall (include/integrators/velocity_verlet.cuh:53-78)."""

    def __init__(self, P: _Particles, dt, wall: bool = False, mode: Optional[int] = None):
        self.P = P
        self.mode = mode if mode is not None else (L.FORCE_WALL if wall else L.FORCE_PBC)
        self.dt = _per_system(dt, P.n_systems())

    def init(self): self.P._c(lib().dpmd_vv_init(self.P.ctx, self.dt.ptr, self.mode))
    def step(self, n: int = 1): self.P._c(lib().dpmd_vv_step(self.P.ctx, n))
    def run(self, n: int, use_graph: bool = True): self.P._c(lib().dpmd_vv_run(self.P.ctx, n, int(use_graph)))


class DampedVelocityVerlet:
    """DampedVelocityVerlet / DampedVelocityVerletWall (include/integrators/damped_velocity_verlet.cuh:62-87):
    velocity Verlet with compute_damping_forces(damping) after the forces of every step."""

    def __init__(self, P: _Particles, dt, damping, wall: bool = False, mode: Optional[int] = None):
        self.P = P
        self.mode = mode if mode is not None else (L.FORCE_WALL if wall else L.FORCE_PBC)
        self.dt = _per_system(dt, P.n_systems())
        self.damping = _per_system(damping, P.n_systems())

    def init(self):
        self.P._c(lib().dpmd_damped_vv_init(self.P.ctx, self.dt.ptr, self.damping.ptr, self.mode))

    def step(self, n: int = 1): self.P._c(lib().dpmd_vv_step(self.P.ctx, n))
    def run(self, n: int, use_graph: bool = True): self.P._c(lib().dpmd_vv_run(self.P.ctx, n, int(use_graph)))


class FIRE:
    """FIRE / FIREWall (include/integrators/fire.cuh:152-178)."""

    def __init__(self, P: _Particles, dt, wall: bool = False, mode: Optional[int] = None):
        self.P = P
        self.mode = mode if mode is not None else (L.FORCE_WALL if wall else L.FORCE_PBC)
        self.dt = _per_system(dt, P.n_systems())

    def init(self): self.P._c(lib().dpmd_fire_init(self.P.ctx, self.dt.ptr, self.mode))
    def step(self): self.P._c(lib().dpmd_fire_step(self.P.ctx))

    def converged(self, avg_pe_diff_target: float, avg_pe_target: float) -> bool:
        a = C.c_int()
        self.P._c(lib().dpmd_fire_converged(self.P.ctx, avg_pe_diff_target, avg_pe_target, C.byref(a)))
        return bool(a.value)

    def converged_max_force(self, tol: float) -> bool:
        a = C.c_int()
        self.P._c(lib().dpmd_fire_converged_max_force(self.P.ctx, tol, C.byref(a)))
        return bool(a.value)

    def get_dt(self) -> np.ndarray:
        out = DeviceArray(np.zeros(self.P.n_systems()))
        self.P._c(lib().dpmd_fire_get_dt(self.P.ctx, out.ptr))
        return out.to_host()

    def state(self):
        S = self.P.n_systems()
        dt, al = np.empty(S), np.empty(S)
        ng, nb = np.empty(S, np.int32), np.empty(S, np.int32)
        self.P._c(lib().dpmd_fire_state(self.P.ctx, dt.ctypes.data, al.ctypes.data, ng.ctypes.data, nb.ctypes.data))
        return dt, al, ng, nb


def minimize_fire(P: _Particles, dt, max_steps: int, avg_pe_target: float, avg_pe_diff_target: float,
                  wall: bool = False) -> int:
    """minimize_fire[_wall] (include/routines/minimizers.cuh:10-31, 68-88), host-driven.

    The reference passes (avg_pe_target, avg_pe_diff_target) to converged(diff, target):
    the two are swapped at that call (minimizers.cuh:25 vs fire.cuh:98); kept as written.
    Returns the number of steps taken.
    """
    fire = FIRE(P, dt, wall=wall)
    fire.init()
    step, conv = 0, False
    while step < max_steps and not conv:
        fire.step()
        conv = fire.converged(avg_pe_target, avg_pe_diff_target)
        step += 1
    return step


def run_fire(P: _Particles, dt, max_steps: int, predicate: int, a: float, b: float = 0.0,
             mode: int = L.FORCE_PBC, use_graph: bool = True, poll: int = 200):
    """Device-driven FIRE minimisation (dpmd_fire_minimize). Returns (steps, converged)."""
    d = _per_system(dt, P.n_systems())
    steps, conv = C.c_int(), C.c_int()
    P._c(lib().dpmd_fire_minimize(P.ctx, d.ptr, max_steps, predicate, a, b, mode, int(use_graph), poll,
                                  C.byref(steps), C.byref(conv)))
    return steps.value, bool(conv.value)


def jam_compress(P: _Particles, dt, dphi: float, p_target: float, max_compressions: int, fire_max_steps: int,
                 predicate: int = L.FIRE_PRED_MAX_FORCE, a: float = 1e-10, b: float = 0.0,
                 mode: int = L.FORCE_PER_SYSTEM, use_graph: bool = True, retries: int = 0):
    """Pressure-target jamming by compression (dpmd_jam_compress).

    retries: times a round that hit fire_max_steps is run again (FIRE restarted, no
    compression) before the system is dropped (DPMD_OPT_JAM_RETRIES).

    Returns (jam_state[S], n_compressions[S], work) where jam_state is 1 jammed, 0 never run
    or below p_target after max_compressions, -1 dropped (a minimisation did not stop), and
    work = sum of N_s x FIRE steps that did work.  P.get("FIRE_STEPS") holds the per-system
    step counts."""
    S = P.n_systems()
    d = _per_system(dt, S)
    st = np.zeros(S, np.int32)
    nc = np.zeros(S, np.int32)
    work = C.c_int64()
    P._c(lib().dpmd_set_option(P.ctx, L.OPT_JAM_RETRIES, int(retries)))
    P._c(lib().dpmd_jam_compress(P.ctx, d.ptr, dphi, p_target, max_compressions, fire_max_steps, predicate, a, b,
                                 mode, int(use_graph), st.ctypes.data_as(C.POINTER(C.c_int)),
                                 nc.ctypes.data_as(C.POINTER(C.c_int)), C.byref(work)))
    return st, nc, work.value


def jam_binary_search(P: _Particles, dt, max_compression_steps: int, max_minimization_steps: int,
                      avg_pe_target: float, avg_pe_diff_target: float, phi_increment: float, phi_tolerance: float,
                      wall: bool = False, mode: Optional[int] = None, use_graph: bool = True):
    """jam_binary_search_pbc / _wall (include/routines/jammers.cuh:60-256) on the device
    (dpmd_jam_binary_search).  Returns (action[S], phi[S], compression_rounds)."""
    S = P.n_systems()
    d = _per_system(dt, S)
    m = mode if mode is not None else (L.FORCE_WALL if wall else L.FORCE_PBC)
    act = np.zeros(S, np.int32)
    phi = np.zeros(S)
    rounds = C.c_int()
    P._c(lib().dpmd_jam_binary_search(P.ctx, d.ptr, max_compression_steps, max_minimization_steps, avg_pe_target,
                                      avg_pe_diff_target, phi_increment, phi_tolerance, m, int(use_graph),
                                      act.ctypes.data, phi.ctypes.data, C.byref(rounds)))
    return act, phi, rounds.value


def nvt_rescale_run(P: _Particles, n_steps: int, phi_increment=None, compress_every: int = 0,
                    temperature_target=None, temperature_every: int = 0, step0: int = 0,
                    remove_average_velocity: bool = True, use_graph: bool = True):
    """Rescale-thermostat NVT with affine compression (dpmd_nvt_rescale_run) over the
    integrator set up by VelocityVerlet.init() or DampedVelocityVerlet.init(): for global step
    k = step0 + i, one step, then increment_packing_fraction if k % compress_every == 0, then
    set_temperature if k % temperature_every == 0 (tests/nvt_rescale_disk_pbc_compress.cu:41-68)."""
    S = P.n_systems()
    inc = _per_system(phi_increment, S) if phi_increment is not None else None
    T = _per_system(temperature_target, S) if temperature_target is not None else None
    P._c(lib().dpmd_nvt_rescale_run(P.ctx, n_steps, step0, inc.ptr if inc else None, compress_every,
                                    T.ptr if T else None, temperature_every, int(remove_average_velocity),
                                    int(use_graph)))


def nvt_rescale_compress(P: _Particles, dt, n_steps: int, phi_increment: float, temperature_target: float,
                         use_graph: bool = True):
    """The reference's two-phase compression driver (tests/nvt_rescale_disk_pbc_compress.cu:14-68):
    n_steps of velocity Verlet with phi_increment spread over compressions every
    min(n_steps, 100) steps and set_temperature every min(n_steps, 10) steps, then n_steps more
    with the thermostat only."""
    cf = min(n_steps, 100)
    tf = min(n_steps, 10)
    phi_step = phi_increment / (n_steps // cf)
    vv = VelocityVerlet(P, dt)
    vv.init()
    nvt_rescale_run(P, n_steps, phi_step, cf, temperature_target, tf, 0, True, use_graph)
    nvt_rescale_run(P, n_steps, None, 0, temperature_target, tf, n_steps, True, use_graph)
