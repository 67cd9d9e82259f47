"""Fused FIRE batches of one disk system (DPMD_OPT_FUSED_FIRE) against the three-launch
batches and the oracle, bit for bit.

A fused launch computes step k's forces, kick and power partials and then drifts step k + 1
with the scalars that F.v > 0 would give; the settle launch folds the partials, checks that
guess and redoes the drift (or, when the system stopped, writes step k's final state) when it
was wrong.  The
systems below start from random positions, so FIRE meets F.v <= 0 many times (counted on the
oracle, which must see at least one such step per case so the fixup path is exercised), and
rebuild halts land on steps of both parities.
"""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import FIRE, Disk, run_fire

pytestmark = pytest.mark.gpu

NAMES = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y", "pe": "PE"}


def load(b, method, fused, persistent=0):
    P = Disk()
    W.load_into(P, b, method)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_FUSED_FIRE, fused), P.ctx)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, persistent), P.ctx)
    P.init_neighbors()
    return P


def uphill_steps(b, method, mode, n, tol):
    """Steps (after the first) whose committed N_good is 0, i.e. F.v <= 0 at the step before."""
    ref = orc.OracleDisk(b, method=method)
    ref.fire_init(1e-2, mode)
    ups = 0
    for k in range(n):
        ref.fire_step(mode)
        if ref.fire_converged(0.0, 0.0, pred=1, ftol=tol):
            break
        if k >= 1 and ref.fire_state()[2][0] == 0:
            ups += 1
    return ups


CASES = [
    # (name, builder, neighbor method, force mode, max steps, max|F| tolerance)
    ("pbc-cell", lambda: W.single_disk_system(4096, 0.80, seed=5), L.NEIGHBOR_CELL, L.FORCE_PBC, 4000, 1e-9),
    ("wall-cell", lambda: W.single_disk_system(3000, 0.80, seed=6, boundary=1), L.NEIGHBOR_CELL, L.FORCE_WALL,
     10000, 1e-9),
    ("pbc-naive", lambda: W.single_disk_system(400, 0.86, seed=7), L.NEIGHBOR_NAIVE, L.FORCE_PER_SYSTEM, 20000, 1e-9),
    ("dense-cell", lambda: W.single_disk_system(4096, 0.85, seed=5), L.NEIGHBOR_CELL, L.FORCE_PBC, 20000, 1e-4),
    ("pbc-cap", lambda: W.single_disk_system(4096, 0.85, seed=8), L.NEIGHBOR_CELL, L.FORCE_PBC, 333, 1e-9),
]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fused_fire_equals_three_launch_fire_and_oracle(case, graph):
    name, build, method, mode, max_steps, tol = case
    b = build()
    print(f"case {name} graph={graph} N={b.N}", flush=True)
    out = []
    for fused in (1, 0):
        P = load(b, method, fused)
        steps, conv = run_fire(P, 1e-2, max_steps, L.FIRE_PRED_MAX_FORCE, tol, mode=mode, use_graph=graph)
        state = FIRE(P, 1e-2, mode=mode).state()
        out.append((P, steps, conv, state))
    ref = orc.OracleDisk(b, method=method)
    rsteps, rconv = ref.fire_minimize(1e-2, max_steps, pred=1, ftol=tol, mode=mode)
    rstate = ref.fire_state()
    for P, steps, conv, state in out:
        assert (steps, conv) == (rsteps, rconv), name
        for f, key in NAMES.items():
            g, r = P.get(key), ref.get(f)
            assert np.array_equal(g, r), f"{name}: {f} differs by {np.max(np.abs(g - r)):.3e}"
        for g, r in zip(state, rstate):
            assert np.array_equal(g, r), f"{name}: FIRE scalars differ"
    if name == "pbc-cap":
        assert not rconv and rsteps == max_steps
    else:
        assert rconv
    if not graph and name != "pbc-cap":   # the capped run is too short to meet F.v <= 0
        ups = uphill_steps(b, method, mode, rsteps, tol)
        print(f"{name}: {ups} steps after F.v <= 0 in {rsteps}", flush=True)
        assert ups >= 1


def test_fused_fire_from_a_relaxed_start_then_again():
    """Two runs in a row on one context: the second starts from the first one's minimum (few
    steps, converging inside a batch), and both equal the oracle."""
    b = W.single_disk_system(20000, 0.80, seed=9)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()    # default options: 20000 particles take the fused batches, not the persistent engine
    ref = orc.OracleDisk(b)
    for tol in (1e-6, 1e-9):
        steps, conv = run_fire(P, 1e-2, 6000, L.FIRE_PRED_MAX_FORCE, tol)
        rsteps, rconv = ref.fire_minimize(1e-2, 6000, pred=1, ftol=tol)
        print(f"tol {tol}: {steps} steps, converged {conv}", flush=True)
        assert (steps, conv) == (rsteps, rconv)
        for f, key in NAMES.items():
            assert np.array_equal(P.get(key), ref.get(f)), f"tol {tol}: {f} differs"
