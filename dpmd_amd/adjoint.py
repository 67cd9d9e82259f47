"""Adjoints of the disk contact model over the C ABI (dpmd_adjoint_*, dpmd_newton_minimize,
dpmd_vv_run_checkpointed, dpmd_vv_backprop).  All particle vectors are numpy arrays in
canonical order; per-system values are length-S arrays.

New functionality (the reference has no derivatives of its dynamics); the pair Hessian block
is the one of the reference's Hessian observable, src/particles/disk.cu:565-661.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from ._lib import lib
from .system import DeviceArray, Disk


def _dev(a: Optional[np.ndarray]):
    return None if a is None else DeviceArray(np.ascontiguousarray(a, np.float64))


def _ptr(d: Optional[DeviceArray]):
    return None if d is None else d.ptr


def set_canonical(P: Disk, name: str, values) -> None:
    """Write a per-particle field given in canonical order (fields are stored in current order)."""
    si = P.get("STATIC_INDEX")
    cur = np.empty(len(si))
    cur[si] = np.asarray(values, np.float64)
    P.set(name, cur)


def hvp(P: Disk, wx, wy, mode: int = L.FORCE_PBC):
    """(H w)_x, (H w)_y at the current configuration."""
    N = P.n_particles()
    a, b = _dev(wx), _dev(wy)
    ox, oy = DeviceArray(np.zeros(N)), DeviceArray(np.zeros(N))
    P._c(lib().dpmd_adjoint_hvp(P.ctx, mode, a.ptr, b.ptr, ox.ptr, oy.ptr))
    return ox.to_host(), oy.to_host()


def vjp(P: Disk, mode: int = L.FORCE_PBC, gU=None, gP=None, gF=None):
    """d/dx, d/dy, d/dr (canonical, length N) and d/dLx, d/dLy (length S) of
    sum_s gU_s U_s + gP_s P_s + sum_i gF_i . F_i.  gF is a pair (gfx, gfy) or None."""
    N, S = P.n_particles(), P.n_systems()
    dU, dP = _dev(gU), _dev(gP)
    gfx, gfy = (None, None) if gF is None else (_dev(gF[0]), _dev(gF[1]))
    out = {k: DeviceArray(np.zeros(N)) for k in ("dx", "dy", "dr")}
    out.update({k: DeviceArray(np.zeros(S)) for k in ("dLx", "dLy")})
    P._c(lib().dpmd_adjoint_vjp(P.ctx, mode, _ptr(dU), _ptr(dP), _ptr(gfx), _ptr(gfy), out["dx"].ptr,
                                out["dy"].ptr, out["dr"].ptr, out["dLx"].ptr, out["dLy"].ptr))
    return {k: v.to_host() for k, v in out.items()}


def solve(P: Disk, bx, by, mode: int = L.FORCE_PBC, tol: float = 1e-10, max_iter: int = 20000):
    """Per-system CG solve of H x = b.  Returns (xx, xy, iterations, max relative residual)."""
    N = P.n_particles()
    a, b = _dev(bx), _dev(by)
    ox, oy = DeviceArray(np.zeros(N)), DeviceArray(np.zeros(N))
    it, res = C.c_int(), C.c_double()
    P._c(lib().dpmd_adjoint_solve(P.ctx, mode, a.ptr, b.ptr, ox.ptr, oy.ptr, tol, max_iter, C.byref(it),
                                  C.byref(res)))
    return ox.to_host(), oy.to_host(), it.value, res.value


def pressure_grad(P: Disk, mode: int = L.FORCE_PBC, tol: float = 1e-10, max_iter: int = 20000):
    """Implicit derivative of each system's pressure through its minimum.
    Returns (dP/dr [N, canonical], dP/dLx [S], dP/dLy [S], CG iterations, relative residual)."""
    N, S = P.n_particles(), P.n_systems()
    dr = DeviceArray(np.zeros(N))
    dLx, dLy = DeviceArray(np.zeros(S)), DeviceArray(np.zeros(S))
    it, res = C.c_int(), C.c_double()
    P._c(lib().dpmd_adjoint_pressure(P.ctx, mode, tol, max_iter, dr.ptr, dLx.ptr, dLy.ptr, C.byref(it),
                                     C.byref(res)))
    return dr.to_host(), dLx.to_host(), dLy.to_host(), it.value, res.value


def newton_minimize(P: Disk, mode: int = L.FORCE_PBC, max_newton: int = 20, ftol: float = 1e-14,
                    cg_tol: float = 1e-10, cg_max_iter: int = 20000) -> float:
    """Newton polish of a configuration near a minimum; returns the final max |F_i|."""
    f = C.c_double()
    P._c(lib().dpmd_newton_minimize(P.ctx, mode, max_newton, ftol, cg_tol, cg_max_iter, C.byref(f)))
    return f.value


def vv_run_checkpointed(P: Disk, n_steps: int, every: int = 50, use_graph: bool = True) -> None:
    P._c(lib().dpmd_vv_run_checkpointed(P.ctx, n_steps, every, int(use_graph)))


def vv_backprop(P: Disk, gx=None, gy=None, gvx=None, gvy=None, use_graph: bool = True):
    """Cotangents of the final (x, y, vx, vy) -> cotangents of the initial ones (canonical)."""
    N = P.n_particles()
    ins = [_dev(a) for a in (gx, gy, gvx, gvy)]
    outs = [DeviceArray(np.zeros(N)) for _ in range(4)]
    P._c(lib().dpmd_vv_backprop(P.ctx, *[_ptr(a) for a in ins], *[o.ptr for o in outs], int(use_graph)))
    return tuple(o.to_host() for o in outs)
