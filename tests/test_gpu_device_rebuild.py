"""Step batches that rebuild on the device (DPMD_OPT_DEVICE_REBUILD = 2: every step runs the
rebuild chain gated by the rebuild flags, so batches never halt) against the oracle, bit for
bit, for disks and rigid bodies, FIRE and velocity Verlet; and the automatic choice (0)
against the halting batches (1).  Each case runs with both rebuild engines: one workgroup
per system (DPMD_OPT_REBUILD_ENGINE = 1) and the device-wide chain with the tiled neighbor
build (2)."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, RigidBumpy, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu

DNAMES = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y"}
RNAMES = {"x": "POS_X", "y": "POS_Y", "angle": "ANGLE", "vx": "VEL_X", "w": "ANGULAR_VEL", "vpx": "VERTEX_POS_X"}


def disk_pair(dev, engine):
    b = W.disk_batch([64, 200, 97, 333, 150, 64], [0.80, 0.86, 0.83, 0.85, 0.84, 0.81], [0, 1, 0, 1, 0, 1], seed=11)
    P = Disk()
    W.load_into(P, b)
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, dev))
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_REBUILD_ENGINE, engine))
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, 0))
    P.init_neighbors()
    return b, P, orc.OracleDisk(b)


def same(P, ref, names, what):
    for f, g in names.items():
        a, r = P.get(g), ref.get(f)
        assert np.array_equal(a, r), f"{what}: {f} max |diff| {np.max(np.abs(a - r)):.3e}"


@pytest.mark.parametrize("engine", [1, 2])
@pytest.mark.parametrize("dev", [2, 0])
def test_disk_fire_and_vv_rebuild_on_device(dev, engine):
    b, P, ref = disk_pair(dev, engine)
    print(f"disk dev={dev} engine={engine} N={b.N}", flush=True)
    steps, _ = run_fire(P, 1e-2, 800, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=L.FORCE_PER_SYSTEM)
    rsteps, _ = ref.fire_minimize(1e-2, 800, pred=1, ftol=1e-12, mode=2)
    assert steps == rsteps
    same(P, ref, DNAMES, "FIRE")
    rng = np.random.default_rng(4)
    v = rng.standard_normal((2, b.N)) * 0.05          # hot: a rebuild every few steps
    P.set("VEL_X", v[0]); P.set("VEL_Y", v[1]); ref.set("vx", v[0]); ref.set("vy", v[1])
    VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM).init()
    ref.vv_init(1e-2, 2)
    VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM).run(600)
    ref.vv_run(600, 2)
    same(P, ref, DNAMES, "VV")
    assert np.array_equal(P.get("STATIC_INDEX"), ref.static_index())


@pytest.mark.parametrize("engine", [1, 2])
@pytest.mark.parametrize("dev", [2, 0])
def test_rigid_fire_and_vv_rebuild_on_device(dev, engine):
    rb = W.rigid_batch(S=4, n_per=32, seed=3, boundary=1, phi=0.70)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_DEVICE_REBUILD, dev))
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_REBUILD_ENGINE, engine))
    P.init_neighbors()
    ref = orc.OracleRigid(rb)
    print(f"rigid dev={dev} engine={engine} N={rb.N} Nv={rb.Nv}", flush=True)
    steps, _ = run_fire(P, 1e-2, 400, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL)
    rsteps, _ = ref.fire_minimize(1e-2, 400, pred=1, ftol=1e-9, mode=1)
    assert steps == rsteps
    same(P, ref, RNAMES, "rigid FIRE")
    rng = np.random.default_rng(5)
    for f, r, sc in (("VEL_X", "vx", 0.05), ("VEL_Y", "vy", 0.05), ("ANGULAR_VEL", "w", 0.1)):
        v = rng.standard_normal(rb.N) * sc
        P.set(f, v); ref.set(r, v)
    VelocityVerlet(P, 1e-2, wall=True).init()
    ref.vv_init(1e-2, 1)
    VelocityVerlet(P, 1e-2, wall=True).run(400)
    ref.vv_run(400, 1)
    same(P, ref, RNAMES, "rigid VV")


@pytest.mark.parametrize("engine", [1, 2])
def test_engines_lists_in_order_edge_systems(engine):
    """Systems of one, two and three disks beside larger ones, walls and PBC: both engines give
    the oracle's lists in order, static_index and cell starts."""
    b = W.disk_batch([1, 2, 3, 40, 257, 64], [0.3, 0.3, 0.3, 0.8, 0.85, 0.84], [0, 1, 0, 1, 0, 1], seed=2)
    P = Disk()
    W.load_into(P, b)
    P._c(L.lib().dpmd_set_option(P.ctx, L.OPT_REBUILD_ENGINE, engine))
    P.init_neighbors()
    ref = orc.OracleDisk(b)
    print(f"edge systems engine={engine}", flush=True)
    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)
    assert np.array_equal(P.get("STATIC_INDEX"), ref.static_index())
    for f, g in (("x", "POS_X"), ("rad", "RAD")):
        assert np.array_equal(P.get(g), ref.get(f))
