"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker / the timed CPU baseline.  The product (dpmd_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

vp, i32, f64, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64

_PROTOS = {
    "or_philox4x32_10": (None, [vp, vp, vp]),
    "or_rand_uniform_2d": (None, [i32, u64, u64, f64, f64, f64, f64, vp, vp]),
    "or_rand_uniform_1d": (None, [i32, u64, u64, f64, f64, vp]),
    "or_sys_sum": (f64, [vp, i32]),
    "or_set_threads": (None, [i32]),
    "or_disk_create": (vp, [i32, i32, i32, i32] + [vp] * 14),
    "or_disk_destroy": (None, [vp]),
    "or_disk_get": (None, [vp, i32, vp]),
    "or_disk_set": (None, [vp, i32, vp]),
    "or_disk_get_static_index": (None, [vp, vp]),
    "or_disk_get_cells": (None, [vp, vp, vp]),
    "or_disk_max_count": (i32, [vp]),
    "or_disk_n_neighbors": (i32, [vp]),
    "or_disk_get_neighbors": (None, [vp, vp, vp]),
    "or_disk_flags": (None, [vp, vp]),
    "or_disk_check_neighbors": (None, [vp, i32]),
    "or_disk_compute_forces": (None, [vp, i32]),
    "or_disk_update_positions": (None, [vp, vp, f64]),
    "or_disk_update_velocities": (None, [vp, vp, f64]),
    "or_disk_scale_velocities": (None, [vp, vp]),
    "or_disk_scale_positions": (None, [vp, vp]),
    "or_disk_mix": (None, [vp, vp]),
    "or_disk_pe_total": (None, [vp, vp]),
    "or_disk_ke_total": (None, [vp, vp]),
    "or_disk_fpower_total": (None, [vp, vp]),
    "or_disk_temperature": (None, [vp, vp]),
    "or_disk_pressure": (None, [vp, vp]),
    "or_disk_packing_fraction": (None, [vp, vp]),
    "or_disk_max_force2": (None, [vp, vp]),
    "or_disk_stress": (None, [vp, vp, vp, vp, vp]),
    "or_disk_set_temperature": (None, [vp, vp, i32]),
    "or_disk_remove_mean_velocity": (None, [vp]),
    "or_disk_increment_packing_fraction": (None, [vp, vp]),
    "or_disk_save_state": (None, [vp, vp, i32]),
    "or_disk_restore_state": (None, [vp, vp, i32]),
    "or_disk_fire_init": (None, [vp, vp, i32]),
    "or_disk_damped_vv_init": (None, [vp, vp, vp, i32]),
    "or_disk_jam_bisect": (i32, [vp, vp, i32, i32, f64, f64, f64, f64, i32, vp, vp]),
    "or_disk_contacts": (None, [vp, vp, vp, vp]),
    "or_disk_pair_obs": (None, [vp, vp, vp, vp, vp, vp, vp]),
    "or_rigid_jam_bisect": (i32, [vp, vp, i32, i32, f64, f64, f64, f64, i32, vp, vp]),
    "or_rigid_particle_neighbors": (i32, [vp, vp, vp]),
    "or_rigid_pair_obs": (None, [vp] * 10),
    "or_rigid_overlaps": (None, [vp, vp, vp]),
    "or_rigid_hessian": (None, [vp, vp]),
    "or_rigid_packing_fraction": (None, [vp, vp]),
    "or_rigid_increment_packing_fraction": (None, [vp, vp]),
    "or_rigid_save_state": (None, [vp, vp, i32]),
    "or_rigid_restore_state": (None, [vp, vp, i32]),
    "or_disk_nvt_rescale_run": (None, [vp, i32, i32, vp, i32, vp, i32, i32, i32]),
    "or_disk_damping": (None, [vp, vp]),
    "or_rigid_damped_vv_init": (None, [vp, vp, vp, i32]),
    "or_rigid_damping": (None, [vp, vp]),
    "or_disk_fire_step": (None, [vp, i32]),
    "or_disk_fire_converged": (i32, [vp, f64, f64, i32, f64]),
    "or_disk_fire_get": (None, [vp, vp, vp, vp, vp]),
    "or_fire_update_scalar": (None, [vp, vp, vp, vp, f64, f64, f64, vp, vp]),
    "or_disk_fire_minimize": (i32, [vp, vp, i32, f64, f64, i32, f64, i32, vp]),
    "or_disk_jam_compress": (C.c_longlong, [vp, vp, f64, f64, i32, i32, i32, f64, f64, i32, vp, vp, vp, i32]),
    "or_disk_vv_init": (None, [vp, vp, i32]),
    "or_disk_vv_run": (None, [vp, i32, i32]),
    "or_disk_brute_neighbors": (i32, [vp, i32, vp]),
    "or_disk_get_cell_ids": (None, [vp, vp]),
    "or_disk_get_cell_system_start": (None, [vp, vp]),
    "or_cell_layout": (None, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "or_box_inv": (None, [i32, vp, vp, vp, vp]),
}


def cell_layout(Lx, Ly, nx, ny):
    Lx, Ly = np.ascontiguousarray(Lx, np.float64), np.ascontiguousarray(Ly, np.float64)
    nx, ny = np.ascontiguousarray(nx, np.int32), np.ascontiguousarray(ny, np.int32)
    S = Lx.size
    o = [np.empty(S) for _ in range(4)]
    start = np.empty(S + 1, np.int32)
    lib().or_cell_layout(S, _p(Lx), _p(Ly), _p(nx), _p(ny), *[_p(a) for a in o], _p(start))
    return (*o, start)


def box_inv(Lx, Ly):
    Lx, Ly = np.ascontiguousarray(Lx, np.float64), np.ascontiguousarray(Ly, np.float64)
    ix, iy = np.empty(Lx.size), np.empty(Lx.size)
    lib().or_box_inv(Lx.size, _p(Lx), _p(Ly), _p(ix), _p(iy))
    return ix, iy

FIELDS = dict(x=0, y=1, vx=2, vy=3, fx=4, fy=5, pe=6, rad=7, mass=8, area=9, lastx=10, lasty=11, Lx=12, Ly=13)

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        for n, (r, a) in _PROTOS.items():
            fn = getattr(L, n)
            fn.restype = r
            fn.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def rand_uniform_2d(n, seed, call, lox=0.0, hix=1.0, loy=0.0, hiy=1.0):
    ox, oy = np.empty(n), np.empty(n)
    lib().or_rand_uniform_2d(n, seed, call, lox, hix, loy, hiy, _p(ox), _p(oy))
    return ox, oy


def sys_sum(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().or_sys_sum(_p(v), v.size)


class OracleDisk:
    """Host restatement of a batch of disk systems (see dpmd_oracle.cpp)."""

    def __init__(self, b, K: int = 32, method: int = 1):
        self.S, self.N = b.S, b.N
        arrs = [np.ascontiguousarray(b.off, np.int32), b.Lx, b.Ly, b.e, b.cutoff, b.thresh2,
                np.ascontiguousarray(b.cdx, np.int32), np.ascontiguousarray(b.cdy, np.int32),
                np.ascontiguousarray(b.boundary, np.int32), b.x, b.y, b.rad, b.mass, b.area]
        arrs = [np.ascontiguousarray(a, a.dtype if a.dtype == np.int32 else np.float64) for a in arrs]
        self._arrs = arrs
        ptrs = [_p(a) for a in arrs]
        self.h = lib().or_disk_create(self.S, self.N, K, method, *ptrs[:9], *ptrs[9:])

    def __del__(self):
        try:
            if self.h:
                lib().or_disk_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def get(self, name):
        n = self.S if name in ("Lx", "Ly") else self.N
        out = np.empty(n)
        lib().or_disk_get(self.h, FIELDS[name], _p(out))
        return out

    def set(self, name, v):
        v = np.ascontiguousarray(v, np.float64)
        lib().or_disk_set(self.h, FIELDS[name], _p(v))

    def static_index(self):
        out = np.empty(self.N, np.int32)
        lib().or_disk_get_static_index(self.h, _p(out))
        return out

    def cells(self):
        nx, ny = np.empty(self.S, np.int32), np.empty(self.S, np.int32)
        lib().or_disk_get_cells(self.h, _p(nx), _p(ny))
        return nx, ny

    def neighbors_csr(self):
        tot = lib().or_disk_n_neighbors(self.h)
        start = np.empty(self.N + 1, np.int32)
        ids = np.empty(max(1, tot), np.int32)
        lib().or_disk_get_neighbors(self.h, _p(start), _p(ids))
        return start, ids[:tot]

    def max_count(self):
        return lib().or_disk_max_count(self.h)

    def cell_ids(self):
        out = np.empty(self.N, np.int32)
        lib().or_disk_get_cell_ids(self.h, _p(out))
        return out

    def cell_system_start(self):
        out = np.empty(self.S + 1, np.int32)
        lib().or_disk_get_cell_system_start(self.h, _p(out))
        return out

    def brute(self, i):
        out = np.empty(self.N, np.int32)
        n = lib().or_disk_brute_neighbors(self.h, i, _p(out))
        return out[:n]

    def flags(self):
        out = np.empty(self.S, np.uint32)
        lib().or_disk_flags(self.h, _p(out))
        return out

    def _ps(self, v, dt=np.float64):
        a = np.asarray(v, dt)
        if a.ndim == 0:
            a = np.full(self.S, a, dt)
        return np.ascontiguousarray(a)

    def check_neighbors(self, force=False): lib().or_disk_check_neighbors(self.h, int(force))
    def compute_forces(self, mode=0): lib().or_disk_compute_forces(self.h, mode)

    def update_positions(self, sc, s2=1.0):
        a = self._ps(sc); lib().or_disk_update_positions(self.h, _p(a), s2)

    def update_velocities(self, sc, s2=1.0):
        a = self._ps(sc); lib().or_disk_update_velocities(self.h, _p(a), s2)

    def scale_velocities(self, sc):
        a = self._ps(sc); lib().or_disk_scale_velocities(self.h, _p(a))

    def scale_positions(self, sc):
        a = self._ps(sc); lib().or_disk_scale_positions(self.h, _p(a))

    def mix(self, alpha):
        a = self._ps(alpha); lib().or_disk_mix(self.h, _p(a))

    def _per_sys(self, fn):
        out = np.empty(self.S)
        fn(self.h, _p(out))
        return out

    def pe_total(self): return self._per_sys(lib().or_disk_pe_total)
    def ke_total(self): return self._per_sys(lib().or_disk_ke_total)
    def fpower_total(self): return self._per_sys(lib().or_disk_fpower_total)
    def temperature(self): return self._per_sys(lib().or_disk_temperature)
    def pressure(self): return self._per_sys(lib().or_disk_pressure)
    def packing_fraction(self): return self._per_sys(lib().or_disk_packing_fraction)
    def max_force2(self): return self._per_sys(lib().or_disk_max_force2)

    def stress(self):
        o = [np.empty(self.N) for _ in range(4)]
        lib().or_disk_stress(self.h, *[_p(a) for a in o])
        return o

    def set_temperature(self, T, remove_avg=True):
        a = self._ps(T); lib().or_disk_set_temperature(self.h, _p(a), int(remove_avg))

    def remove_mean_velocity(self): lib().or_disk_remove_mean_velocity(self.h)

    def increment_packing_fraction(self, inc):
        a = self._ps(inc); lib().or_disk_increment_packing_fraction(self.h, _p(a))

    def save_state(self, flag=1, tv=1):
        a = self._ps(flag, np.int32); lib().or_disk_save_state(self.h, _p(a), tv)

    def restore_state(self, flag=1, tv=1):
        a = self._ps(flag, np.int32); lib().or_disk_restore_state(self.h, _p(a), tv)

    def fire_init(self, dt, mode=0):
        a = self._ps(dt); lib().or_disk_fire_init(self.h, _p(a), mode)

    def fire_step(self, mode=0): lib().or_disk_fire_step(self.h, mode)

    def fire_converged(self, diff, tgt, pred=0, ftol=0.0):
        return bool(lib().or_disk_fire_converged(self.h, diff, tgt, pred, ftol))

    def fire_state(self):
        dt, al = np.empty(self.S), np.empty(self.S)
        ng, nb = np.empty(self.S, np.int32), np.empty(self.S, np.int32)
        lib().or_disk_fire_get(self.h, _p(dt), _p(al), _p(ng), _p(nb))
        return dt, al, ng, nb

    def fire_minimize(self, dt, max_steps, diff=0.0, tgt=0.0, pred=0, ftol=0.0, mode=0):
        a = self._ps(dt)
        conv = np.zeros(1, np.int32)
        steps = lib().or_disk_fire_minimize(self.h, _p(a), max_steps, diff, tgt, pred, ftol, mode, _p(conv))
        return steps, bool(conv[0])

    def jam_compress(self, dt, dphi, p_target, max_comp, fire_max, pred=1, a=1e-10, b=0.0, mode=0, retries=0):
        """Returns (jam_state[S], n_compressions[S], fire_steps[S], work)."""
        d = self._ps(dt)
        st, nc, ns = (np.zeros(self.S, np.int32) for _ in range(3))
        work = lib().or_disk_jam_compress(self.h, _p(d), dphi, p_target, max_comp, fire_max, pred, a, b, mode,
                                          _p(st), _p(nc), _p(ns), int(retries))
        return st, nc, ns, int(work)

    def vv_init(self, dt, mode=0):
        a = self._ps(dt); lib().or_disk_vv_init(self.h, _p(a), mode)

    def damped_vv_init(self, dt, damping, mode=0):
        a, d = self._ps(dt), self._ps(damping); lib().or_disk_damped_vv_init(self.h, _p(a), _p(d), mode)

    def contacts(self):
        """(contacts[N], overlap[N], contact_count[N]) in current order."""
        c = np.zeros(self.N, np.int32); ov = np.zeros(self.N); cc = np.zeros(self.N)
        lib().or_disk_contacts(self.h, _p(c), _p(ov), _p(cc))
        return c, ov, cc

    def pair_observables(self):
        """Per CSR entry of the current lists: pair ids (canonical), dist, fx, fy and the 8
        Hessian columns (same names as dpmd_amd.system's pair_observables)."""
        n = lib().or_disk_n_neighbors(self.h)
        pi = np.zeros(max(1, n), np.int32); pj = np.zeros(max(1, n), np.int32)
        d, fx, fy = np.zeros(max(1, n)), np.zeros(max(1, n)), np.zeros(max(1, n))
        hs = np.zeros(8 * max(1, n))
        lib().or_disk_pair_obs(self.h, _p(pi), _p(pj), _p(d), _p(fx), _p(fy), _p(hs))
        out = dict(pair_i=pi[:n], pair_j=pj[:n], dist=d[:n], fx=fx[:n], fy=fy[:n])
        names = ("h_ix_ix", "h_ix_jx", "h_ix_iy", "h_ix_jy", "h_iy_ix", "h_iy_jx", "h_iy_iy", "h_iy_jy")
        for m, k in enumerate(names):
            out[k] = hs[m * n:(m + 1) * n].copy()
        return out

    def jam_bisect(self, dt, max_comp, max_min, pe_target, pe_diff_target, dphi, phi_tol, mode=0):
        """jam_binary_search_pbc / _wall; returns (action[S], phi[S], rounds)."""
        d = self._ps(dt)
        act = np.zeros(self.S, np.int32)
        phi = np.zeros(self.S)
        r = lib().or_disk_jam_bisect(self.h, _p(d), max_comp, max_min, pe_target, pe_diff_target, dphi, phi_tol,
                                     mode, _p(act), _p(phi))
        return act, phi, int(r)

    def nvt_rescale_run(self, n, inc=None, cf=0, T=None, tf=0, step0=0, rm=1, mode=0):
        a = self._ps(inc if inc is not None else 0.0)
        t = self._ps(T if T is not None else 0.0)
        lib().or_disk_nvt_rescale_run(self.h, n, step0, _p(a), cf, _p(t), tf, rm, mode)

    def compute_damping_forces(self, scale):
        a = self._ps(scale); lib().or_disk_damping(self.h, _p(a))

    def vv_run(self, n, mode=0): lib().or_disk_vv_run(self.h, n, mode)


def fire_update_scalar(dt, alpha, ngood, nbad, power, dt_max, dt_min):
    d = np.array([dt]); a = np.array([alpha])
    g = np.array([ngood], np.int32); b = np.array([nbad], np.int32)
    dr = np.zeros(1); vs = np.zeros(1)
    lib().or_fire_update_scalar(_p(d), _p(a), _p(g), _p(b), power, dt_max, dt_min, _p(dr), _p(vs))
    return dict(dt=d[0], alpha=a[0], n_good=int(g[0]), n_bad=int(b[0]), dt_rev=dr[0], vscale=vs[0])


_RPROTOS = {
    "or_sincos": (None, [f64, vp, vp]),
    "or_rigid_create": (vp, [i32, i32, i32] + [vp] * 20),
    "or_rigid_destroy": (None, [vp]),
    "or_rigid_get": (None, [vp, i32, vp]),
    "or_rigid_set": (None, [vp, i32, vp]),
    "or_rigid_get_static_index": (None, [vp, vp]),
    "or_rigid_get_cells": (None, [vp, vp, vp]),
    "or_rigid_n_neighbors": (i32, [vp]),
    "or_rigid_max_count": (i32, [vp]),
    "or_rigid_get_neighbors": (None, [vp, vp, vp]),
    "or_rigid_check_neighbors": (None, [vp, i32]),
    "or_rigid_compute_forces": (None, [vp, i32]),
    "or_rigid_update_positions": (None, [vp, vp, f64]),
    "or_rigid_update_velocities": (None, [vp, vp, f64]),
    "or_rigid_scale_velocities": (None, [vp, vp]),
    "or_rigid_scale_positions": (None, [vp, vp]),
    "or_rigid_mix": (None, [vp, vp]),
    "or_rigid_pe_total": (None, [vp, vp]),
    "or_rigid_ke_total": (None, [vp, vp]),
    "or_rigid_fpower_total": (None, [vp, vp]),
    "or_rigid_pressure": (None, [vp, vp]),
    "or_rigid_temperature": (None, [vp, vp]),
    "or_rigid_n_dof": (i32, [vp, i32]),
    "or_rigid_max_force2": (None, [vp, vp]),
    "or_rigid_fire_init": (None, [vp, vp, i32]),
    "or_rigid_fire_step": (None, [vp, i32]),
    "or_rigid_fire_converged": (i32, [vp, f64, f64, i32, f64]),
    "or_rigid_fire_minimize": (i32, [vp, vp, i32, f64, f64, i32, f64, i32, vp]),
    "or_rigid_vv_init": (None, [vp, vp, i32]),
    "or_rigid_vv_run": (None, [vp, i32, i32]),
}
_PROTOS.update(_RPROTOS)

RFIELDS = dict(x=0, y=1, angle=2, vx=3, vy=4, w=5, fx=6, fy=7, torque=8, pe=9,
               vpx=20, vpy=21, vfx=22, vfy=23, vpe=24, vrad=25)


def sincos(x):
    s, c = np.zeros(1), np.zeros(1)
    lib().or_sincos(float(x), _p(s), _p(c))
    return s[0], c[0]


class OracleRigid:
    """Host restatement of a batch of rigid bumpy-polygon systems (see dpmd_oracle.cpp)."""

    def __init__(self, b):
        self.S, self.N, self.Nv = b.S, b.N, b.Nv
        ints = [np.ascontiguousarray(a, np.int32) for a in (b.off,)]
        arrs = [ints[0], b.Lx, b.Ly, b.e, b.cutoff, b.thresh2, np.ascontiguousarray(b.cdx, np.int32),
                np.ascontiguousarray(b.cdy, np.int32), np.ascontiguousarray(b.boundary, np.int32),
                np.ascontiguousarray(b.nvpp, np.int32), b.x, b.y, b.angle, b.mass, b.inertia, b.area, b.vx, b.vy,
                b.vrad, b.vmass]
        arrs = [np.ascontiguousarray(a, a.dtype if a.dtype == np.int32 else np.float64) for a in arrs]
        self._arrs = arrs
        self.h = lib().or_rigid_create(self.S, self.N, self.Nv, *[_p(a) for a in arrs])

    def __del__(self):
        try:
            if self.h:
                lib().or_rigid_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def get(self, name):
        n = self.Nv if RFIELDS[name] >= 20 else self.N
        out = np.empty(n)
        lib().or_rigid_get(self.h, RFIELDS[name], _p(out))
        return out

    def set(self, name, v):
        v = np.ascontiguousarray(v, np.float64)
        lib().or_rigid_set(self.h, RFIELDS[name], _p(v))

    def static_index(self):
        out = np.empty(self.Nv, np.int32)
        lib().or_rigid_get_static_index(self.h, _p(out))
        return out

    def cells(self):
        nx, ny = np.empty(self.S, np.int32), np.empty(self.S, np.int32)
        lib().or_rigid_get_cells(self.h, _p(nx), _p(ny))
        return nx, ny

    def neighbors_csr(self):
        tot = lib().or_rigid_n_neighbors(self.h)
        start = np.empty(self.Nv + 1, np.int32)
        ids = np.empty(max(1, tot), np.int32)
        lib().or_rigid_get_neighbors(self.h, _p(start), _p(ids))
        return start, ids[:tot]

    def max_count(self): return lib().or_rigid_max_count(self.h)

    def _ps(self, v, dt=np.float64):
        a = np.asarray(v, dt)
        if a.ndim == 0:
            a = np.full(self.S, a, dt)
        return np.ascontiguousarray(a)

    def _per_sys(self, fn):
        out = np.empty(self.S)
        fn(self.h, _p(out))
        return out

    def check_neighbors(self, force=False): lib().or_rigid_check_neighbors(self.h, int(force))
    def compute_forces(self, mode=0): lib().or_rigid_compute_forces(self.h, mode)

    def update_positions(self, sc, s2=1.0):
        a = self._ps(sc); lib().or_rigid_update_positions(self.h, _p(a), s2)

    def update_velocities(self, sc, s2=1.0):
        a = self._ps(sc); lib().or_rigid_update_velocities(self.h, _p(a), s2)

    def scale_velocities(self, sc):
        a = self._ps(sc); lib().or_rigid_scale_velocities(self.h, _p(a))

    def scale_positions(self, sc):
        a = self._ps(sc); lib().or_rigid_scale_positions(self.h, _p(a))

    def mix(self, alpha):
        a = self._ps(alpha); lib().or_rigid_mix(self.h, _p(a))

    def pe_total(self): return self._per_sys(lib().or_rigid_pe_total)
    def ke_total(self): return self._per_sys(lib().or_rigid_ke_total)
    def fpower_total(self): return self._per_sys(lib().or_rigid_fpower_total)
    def pressure(self): return self._per_sys(lib().or_rigid_pressure)
    def temperature(self): return self._per_sys(lib().or_rigid_temperature)
    def n_dof(self): return np.array([lib().or_rigid_n_dof(self.h, s) for s in range(self.S)])
    def max_force2(self): return self._per_sys(lib().or_rigid_max_force2)

    def fire_init(self, dt, mode=0):
        a = self._ps(dt); lib().or_rigid_fire_init(self.h, _p(a), mode)

    def fire_step(self, mode=0): lib().or_rigid_fire_step(self.h, mode)

    def fire_converged(self, diff, tgt, pred=0, ftol=0.0):
        return bool(lib().or_rigid_fire_converged(self.h, diff, tgt, pred, ftol))

    def fire_minimize(self, dt, max_steps, diff=0.0, tgt=0.0, pred=0, ftol=0.0, mode=0):
        a = self._ps(dt)
        conv = np.zeros(1, np.int32)
        steps = lib().or_rigid_fire_minimize(self.h, _p(a), max_steps, diff, tgt, pred, ftol, mode, _p(conv))
        return steps, bool(conv[0])

    def vv_init(self, dt, mode=0):
        a = self._ps(dt); lib().or_rigid_vv_init(self.h, _p(a), mode)

    def damped_vv_init(self, dt, damping, mode=0):
        a, d = self._ps(dt), self._ps(damping); lib().or_rigid_damped_vv_init(self.h, _p(a), _p(d), mode)

    def packing_fraction(self): return self._per_sys(lib().or_rigid_packing_fraction)

    def particle_neighbors(self):
        """build_particle_neighbors: (start[N+1], ids) of particle-level lists."""
        st = np.zeros(self.N + 1, np.int32)
        tot = lib().or_rigid_particle_neighbors(self.h, _p(st), None)
        ids = np.zeros(max(1, tot), np.int32)
        lib().or_rigid_particle_neighbors(self.h, _p(st), _p(ids))
        return st, ids[:tot]

    def pair_observables(self):
        """Per particle-neighbor entry: pair_i, pair_j, dist, fx, fy, vertex_contacts_i/_j,
        friction; plus contacts[N]."""
        st, ids = self.particle_neighbors()
        n = max(1, len(ids))
        out = dict(pair_i=np.zeros(n, np.int32), pair_j=np.zeros(n, np.int32), dist=np.zeros(n), fx=np.zeros(n),
                   fy=np.zeros(n), vertex_contacts_i=np.zeros(n, np.int32), vertex_contacts_j=np.zeros(n, np.int32),
                   friction=np.zeros(n))
        contacts = np.zeros(self.N, np.int32)
        lib().or_rigid_pair_obs(self.h, *[_p(out[k]) for k in ("pair_i", "pair_j", "dist", "fx", "fy",
                                                                 "vertex_contacts_i", "vertex_contacts_j",
                                                                 "friction")], _p(contacts))
        out = {k: v[:len(ids)] for k, v in out.items()}
        out["contacts"] = contacts
        return out

    def hessian(self):
        """compute_hessian per particle-neighbor entry: dict 'h_ia_ib' / 'h_ia_jb' for a, b in
        x, y, t (theta), plus pair_i / pair_j."""
        st, ids = self.particle_neighbors()
        n = len(ids)
        hs = np.zeros(18 * max(1, n))
        lib().or_rigid_hessian(self.h, _p(hs))
        out = {"pair_i": np.repeat(np.arange(self.N, dtype=np.int32), np.diff(st)), "pair_j": ids}
        q = "xyt"
        for m in range(9):
            a, b = q[m // 3], q[m % 3]
            out[f"h_i{a}_i{b}"] = hs[(2 * m) * n:(2 * m + 1) * n].copy()
            out[f"h_i{a}_j{b}"] = hs[(2 * m + 1) * n:(2 * m + 2) * n].copy()
        return out

    def overlaps(self):
        ov, cc = np.zeros(self.Nv), np.zeros(self.Nv)
        lib().or_rigid_overlaps(self.h, _p(ov), _p(cc))
        return ov, cc

    def increment_packing_fraction(self, inc):
        a = self._ps(inc); lib().or_rigid_increment_packing_fraction(self.h, _p(a))

    def save_state(self, flag=1, tv=1):
        f = self._ps(flag, np.int32); lib().or_rigid_save_state(self.h, _p(f), tv)

    def restore_state(self, flag=1, tv=1):
        f = self._ps(flag, np.int32); lib().or_rigid_restore_state(self.h, _p(f), tv)

    def jam_bisect(self, dt, max_comp, max_min, pe_target, pe_diff_target, dphi, phi_tol, mode=0):
        """jam_binary_search_pbc / _wall on rigid bodies; returns (action[S], phi[S], rounds)."""
        d = self._ps(dt)
        act = np.zeros(self.S, np.int32)
        phi = np.zeros(self.S)
        r = lib().or_rigid_jam_bisect(self.h, _p(d), max_comp, max_min, pe_target, pe_diff_target, dphi, phi_tol,
                                      mode, _p(act), _p(phi))
        return act, phi, int(r)

    def compute_damping_forces(self, scale):
        a = self._ps(scale); lib().or_rigid_damping(self.h, _p(a))

    def vv_run(self, n, mode=0): lib().or_rigid_vv_run(self.h, n, mode)
