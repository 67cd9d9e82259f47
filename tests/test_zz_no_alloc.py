"""The allocation-free Hessian-vector product (dpmd_adjoint_hvp_noalloc) against
dpmd_adjoint_hvp, through the C ABI: two calls in a row on two contexts, the second on fewer
particles, bit for bit (the same kernels on the same inputs); and the size limit."""
import numpy as np
import pytest

from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import DeviceArray, Disk, run_fire

pytestmark = pytest.mark.gpu
MODE = L.FORCE_PER_SYSTEM


def packed(sizes, seed):
    b = W.disk_batch(sizes, [0.9] * len(sizes), [s % 2 for s in range(len(sizes))], seed=seed)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    run_fire(P, 1e-2, 300, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=MODE)
    return b, P


def both(P, n, rng):
    lib = L.lib()
    w = [DeviceArray(rng.standard_normal(n)) for _ in range(2)]
    out = [DeviceArray(np.full(n, np.nan)) for _ in range(4)]
    L.check(lib.dpmd_adjoint_hvp_noalloc(P.ctx, MODE, w[0].ptr, w[1].ptr, out[0].ptr, out[1].ptr), P.ctx)
    P.synchronize()
    L.check(lib.dpmd_adjoint_hvp(P.ctx, MODE, w[0].ptr, w[1].ptr, out[2].ptr, out[3].ptr), P.ctx)
    P.synchronize()
    return [o.to_host() for o in out]


def test_noalloc_hvp_equals_hvp_on_two_calls_second_smaller():
    rng = np.random.default_rng(1)
    for sizes, seed in (([3000, 1200], 4), ([500], 5)):
        b, P = packed(sizes, seed)
        print(f"hvp_noalloc N={b.N}", flush=True)
        nx, ny, rx, ry = both(P, b.N, rng)
        assert np.any(rx != 0.0) and np.all(np.isfinite(rx))
        assert np.array_equal(nx, rx) and np.array_equal(ny, ry)


def test_noalloc_hvp_rejects_more_than_its_maximum():
    b = W.single_disk_system((1 << 20) + 1, 0.5, seed=2)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    w = DeviceArray(np.zeros(b.N))
    lib = L.lib()
    assert lib.dpmd_adjoint_hvp_noalloc(P.ctx, MODE, w.ptr, w.ptr, w.ptr, w.ptr) == 1   # DPMD_ERR_ARGUMENT
    assert b"scratch maximum" in lib.dpmd_last_error(P.ctx)
