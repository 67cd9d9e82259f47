"""torch.autograd.Function front ends of the disk adjoints.

Tensors are CPU float64 tensors; all device work runs in libdpmd_amd through the C ABI
(torch is only the autograd bookkeeping here and never touches the GPU).  The library is
loaded before torch is imported, so one HIP runtime serves the process.

- DiskObservables:         (x, y, r, Lx, Ly) -> (U_s, P_s, Fx, Fy) at fixed velocities;
- MinimizedPressure:       r -> P_s at the energy minimum reached from given positions
                           (FIRE, then Newton), differentiated implicitly (CG on H);
- NVEMeanSquareDisplacement: (vx0, vy0) -> MSD_s after n velocity-Verlet steps, reverse mode
                           with checkpoints.

Each backward runs on the state its forward left in the context; run forward/backward pairs
on one context in order.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L

L.lib()   # load libdpmd_amd (and its HIP runtime) before torch

import torch  # noqa: E402

from . import adjoint as A  # noqa: E402
from .system import Disk, VelocityVerlet, run_fire  # noqa: E402


def _np(t) -> np.ndarray:
    return t.detach().cpu().numpy().astype(np.float64)


def _sid(P: Disk) -> np.ndarray:
    off = P.get("SYSTEM_OFFSET")
    return np.repeat(np.arange(len(off) - 1), np.diff(off))


def _set_state(P: Disk, x, y, r=None, Lx=None, Ly=None) -> None:
    if Lx is not None:
        P.set("BOX_SIZE_X", Lx)
        P.set("BOX_SIZE_Y", Ly)
        P.sync()
    if r is not None:
        A.set_canonical(P, "RAD", r)
        A.set_canonical(P, "AREA", np.pi * r * r)
    A.set_canonical(P, "POS_X", x)
    A.set_canonical(P, "POS_Y", y)
    if Lx is not None:
        P.init_neighbors()
    else:
        P.check_neighbors(True)


class DiskObservables(torch.autograd.Function):
    """U_s, P_s and the forces at (x, y, r, Lx, Ly) with the context's velocities."""

    @staticmethod
    def forward(ctx, x, y, r, Lx, Ly, system: Disk, mode: int = L.FORCE_PBC):
        P = system
        _set_state(P, _np(x), _np(y), _np(r), _np(Lx), _np(Ly))
        P.compute_forces_mode(mode)
        P.compute_pe_total()
        P.compute_pressure()
        ctx.P, ctx.mode = P, mode
        fx, fy = P.get("FORCE_X", L.ORDER_CANONICAL), P.get("FORCE_Y", L.ORDER_CANONICAL)
        return (torch.from_numpy(P.get("PE_TOTAL")), torch.from_numpy(P.get("PRESSURE")), torch.from_numpy(fx),
                torch.from_numpy(fy))

    @staticmethod
    def backward(ctx, gU, gP, gfx, gfy):
        g = A.vjp(ctx.P, ctx.mode, gU=_np(gU), gP=_np(gP), gF=(_np(gfx), _np(gfy)))
        return (torch.from_numpy(g["dx"]), torch.from_numpy(g["dy"]), torch.from_numpy(g["dr"]),
                torch.from_numpy(g["dLx"]), torch.from_numpy(g["dLy"]), None, None)


class MinimizedPressure(torch.autograd.Function):
    """P_s(r): pressure of system s at the minimum reached from (x0, y0) by FIRE (max|F| <
    ftol) and Newton polishing; backward solves H lambda = dP/dx per system to cg_tol."""

    @staticmethod
    def forward(ctx, r, system: Disk, x0, y0, dt: float = 1e-2, fire_steps: int = 200000,
                ftol: float = 1e-12, mode: int = L.FORCE_PBC, cg_tol: float = 1e-10):
        P = system
        _set_state(P, np.asarray(x0, np.float64), np.asarray(y0, np.float64), _np(r))
        run_fire(P, dt, fire_steps, L.FIRE_PRED_MAX_FORCE, ftol, mode=mode)
        A.newton_minimize(P, mode)
        P.zero_velocities()
        P.compute_pressure()
        ctx.P, ctx.mode, ctx.cg_tol = P, mode, cg_tol
        return torch.from_numpy(P.get("PRESSURE"))

    @staticmethod
    def backward(ctx, gP):
        dr, _, _, _, _ = A.pressure_grad(ctx.P, ctx.mode, ctx.cg_tol)
        return (torch.from_numpy(_np(gP)[_sid(ctx.P)] * dr),) + (None,) * 8


class NVEMeanSquareDisplacement(torch.autograd.Function):
    """MSD_s = mean_i |x_i(T) - x_i(0)|^2 over system s after n_steps velocity-Verlet steps from
    (x0, y0) with initial velocities (vx0, vy0); checkpoints every `every` steps."""

    @staticmethod
    def forward(ctx, vx0, vy0, system: Disk, x0, y0, dt: float = 1e-2, n_steps: int = 1000, every: int = 50,
                mode: int = L.FORCE_PBC):
        P = system
        x0 = np.asarray(x0, np.float64); y0 = np.asarray(y0, np.float64)
        _set_state(P, x0, y0)
        A.set_canonical(P, "VEL_X", _np(vx0))
        A.set_canonical(P, "VEL_Y", _np(vy0))
        VelocityVerlet(P, dt, mode=mode).init()
        A.vv_run_checkpointed(P, n_steps, every)
        xT, yT = P.get("POS_X", L.ORDER_CANONICAL), P.get("POS_Y", L.ORDER_CANONICAL)
        sid = _sid(P)
        cnt = np.bincount(sid).astype(np.float64)
        d2 = (xT - x0) ** 2 + (yT - y0) ** 2
        msd = np.bincount(sid, weights=d2) / cnt
        ctx.P, ctx.sid, ctx.cnt = P, sid, cnt
        ctx.dx, ctx.dy = xT - x0, yT - y0
        return torch.from_numpy(msd)

    @staticmethod
    def backward(ctx, gm):
        g = _np(gm)[ctx.sid] * 2.0 / ctx.cnt[ctx.sid]
        _, _, gvx, gvy = A.vv_backprop(ctx.P, g * ctx.dx, g * ctx.dy, None, None)
        return (torch.from_numpy(gvx), torch.from_numpy(gvy)) + (None,) * 7
