"""GPU edge cases against the oracle, bit for bit: systems of one, two and three disks (no
neighbors, a single pair, a triangle) next to larger ones, and systems whose sizes sit on
either side of the 256-particle chunk boundary of the per-system sums."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu
NAMES = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y", "pe": "PE"}


def fire_then_nve(b, method, steps=600, nve=200):
    P = Disk()
    W.load_into(P, b, method)
    P.init_neighbors()
    ref = orc.OracleDisk(b, method=method)
    print(f"sizes={np.diff(b.off).tolist()} method={method}", flush=True)
    n, conv = run_fire(P, 1e-2, steps, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=L.FORCE_PER_SYSTEM)
    rn, rconv = ref.fire_minimize(1e-2, steps, pred=1, ftol=1e-12, mode=L.FORCE_PER_SYSTEM)
    assert (n, bool(conv)) == (rn, bool(rconv))
    for f, g in NAMES.items():
        assert np.array_equal(P.get(g), ref.get(f)), ("FIRE", f)
    rng = np.random.default_rng(7)
    v = rng.standard_normal((2, b.N)) * 0.01
    P.set("VEL_X", v[0]); P.set("VEL_Y", v[1]); ref.set("vx", v[0]); ref.set("vy", v[1])
    VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM).init()
    ref.vv_init(1e-2, L.FORCE_PER_SYSTEM)
    VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM).run(nve)
    ref.vv_run(nve, L.FORCE_PER_SYSTEM)
    for f, g in NAMES.items():
        assert np.array_equal(P.get(g), ref.get(f)), ("NVE", f)
    P.compute_pe_total(); P.compute_ke_total()
    return P


def test_tiny_systems_naive_lists():
    b = W.disk_batch([1, 2, 3, 5, 40], [0.30, 0.60, 0.70, 0.80, 0.84], [0, 1, 1, 0, 1], seed=19)
    P = fire_then_nve(b, L.NEIGHBOR_NAIVE)
    # a lone disk in a periodic box feels no force and has no energy
    assert P.get("PE_TOTAL")[0] == 0.0 and P.get("FORCE_X")[0] == 0.0


def test_chunk_boundary_sizes_cell_lists():
    b = W.disk_batch([255, 256, 257, 511, 512, 513], [0.84, 0.845, 0.83, 0.84, 0.85, 0.82], [0, 1, 0, 1, 0, 1],
                     seed=23)
    fire_then_nve(b, L.NEIGHBOR_CELL)
