"""Instrumentation (dpmd_profile kinds 2-4: HIP event pairs around the fused velocity-Verlet
launches of unhalted batches, around the fused FIRE steps, and around the rebuild that
finishes every halted batch): the events are recorded, and the runs keep their results bit
for bit."""
import ctypes as C

import numpy as np
import pytest

from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu


def run(kind):
    b = W.single_disk_system(20000, 0.84, seed=12)
    P = Disk()
    W.load_into(P, b)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_REBUILD_ENGINE, 2), P.ctx)
    P.init_neighbors()
    lib = L.lib()
    L.check(lib.dpmd_profile(P.ctx, kind), P.ctx)
    run_fire(P, 1e-2, 400, L.FIRE_PRED_MAX_FORCE, 0.0, use_graph=False)
    rng = np.random.default_rng(3)
    v = rng.standard_normal((2, b.N)) * 1e-2
    P.set("VEL_X", v[0]); P.set("VEL_Y", v[1])
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    vv.run(600, use_graph=False)
    ms, cnt = C.c_double(), C.c_int()
    L.check(lib.dpmd_profile_read(P.ctx, C.byref(ms), C.byref(cnt)), P.ctx)
    L.check(lib.dpmd_profile(P.ctx, 0), P.ctx)
    return {k: P.get(k) for k in ("POS_X", "POS_Y", "VEL_X", "VEL_Y", "FORCE_X", "FORCE_Y")}, ms.value, cnt.value


def test_profile_kinds_record_and_keep_results():
    ref, _, n0 = run(0)
    assert n0 == 0
    for kind in (2, 3, 4):
        got, ms, n = run(kind)
        print(f"kind {kind}: {n} covered, {ms:.3f} ms", flush=True)
        assert n > 0 and ms > 0.0, kind
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (kind, k)


def test_profile_kind_out_of_range_is_rejected():
    P = Disk()
    W.load_into(P, W.single_disk_system(256, 0.84, seed=1))
    assert L.lib().dpmd_profile(P.ctx, 5) == 1   # DPMD_ERR_ARGUMENT
