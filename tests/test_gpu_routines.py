"""GPU parity tests of damped velocity Verlet, the rescale-NVT compression routine and the
bisection jammer: HIP library (through the C ABI) vs the oracle, bit for bit (positions,
velocities, forces, boxes, jammer decisions and packing fractions)."""
import math

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import (DampedVelocityVerlet, Disk, RigidBumpy, VelocityVerlet, jam_binary_search,
                             nvt_rescale_run, run_fire)

pytestmark = pytest.mark.gpu

NAMES = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y",
         "pe": "PE"}


def make(b, method=L.NEIGHBOR_CELL):
    P = Disk()
    W.load_into(P, b, method)
    P.init_neighbors()
    return P, orc.OracleDisk(b, method=method)


def assert_same(P, ref, fields=("x", "y", "vx", "vy", "fx", "fy", "pe"), what=""):
    for f in fields:
        g, r = P.get(NAMES[f]), ref.get(f)
        if not np.array_equal(g, r):
            raise AssertionError(f"{what}: field {f} differs, max |diff| = {np.max(np.abs(g - r)):.3e}")


def batch_small():
    return W.disk_batch([64, 200, 97, 333, 150, 64], [0.80, 0.86, 0.83, 0.85, 0.84, 0.81],
                        [0, 1, 0, 1, 0, 1], seed=11)


def random_velocities(P, ref, N, seed=0, scale=1e-2):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((2, N)) * scale
    P.set("VEL_X", v[0]); P.set("VEL_Y", v[1])
    ref.set("vx", v[0]); ref.set("vy", v[1])


@pytest.mark.parametrize("mode", [L.FORCE_PBC, L.FORCE_WALL, L.FORCE_PER_SYSTEM])
def test_damped_vv_matches_oracle(mode):
    b = batch_small()
    P, ref = make(b)
    random_velocities(P, ref, b.N)
    damping = np.linspace(0.05, 0.5, b.S)
    dvv = DampedVelocityVerlet(P, 1e-2, damping, mode=mode)
    dvv.init()
    ref.damped_vv_init(1e-2, damping, mode)
    assert_same(P, ref, what="damped VV init")
    dvv.step(30)                     # per-step launches
    ref.vv_run(30, mode)
    assert_same(P, ref, what="damped VV steps")
    dvv.run(500, use_graph=True)     # fused batches, graphs, rebuild hand-offs
    ref.vv_run(500, mode)
    assert_same(P, ref, what="damped VV run")


def test_damped_vv_then_plain_vv():
    """dpmd_vv_init after a damped run switches the drag off."""
    b = W.single_disk_system(1024, 0.85, seed=2)
    P, ref = make(b)
    random_velocities(P, ref, b.N, seed=3)
    DampedVelocityVerlet(P, 1e-2, 0.3).init()
    ref.damped_vv_init(1e-2, 0.3)
    lib_run = lambda n: P._c(L.lib().dpmd_vv_run(P.ctx, n, 1))
    lib_run(64)
    ref.vv_run(64)
    VelocityVerlet(P, 1e-2).init()
    ref.vv_init(1e-2)
    lib_run(100)
    ref.vv_run(100)
    assert_same(P, ref, what="plain VV after damped VV")


def test_damping_primitive_matches_oracle():
    b = batch_small()
    P, ref = make(b)
    random_velocities(P, ref, b.N, seed=5)
    P.compute_forces_mode(L.FORCE_PER_SYSTEM)
    ref.compute_forces(2)
    sc = np.linspace(0.1, 1.1, b.S)
    P.compute_damping_forces(sc)
    ref.compute_damping_forces(sc)
    assert_same(P, ref, ("fx", "fy"), "compute_damping_forces")


def test_damped_rigid_vv_matches_oracle():
    b = W.rigid_batch(S=3, n_per=48, seed=9, boundary=0, phi=0.72)
    P = RigidBumpy()
    W.load_rigid_into(P, b)
    P.init_neighbors()
    ref = orc.OracleRigid(b)
    run_fire(P, 1e-2, 300, L.FIRE_PRED_MAX_FORCE, 1e-9)
    ref.fire_minimize(1e-2, 300, pred=1, ftol=1e-9)
    rng = np.random.default_rng(1)
    for f, r in (("VEL_X", "vx"), ("VEL_Y", "vy"), ("ANGULAR_VEL", "w")):
        v = rng.standard_normal(b.N) * 1e-2
        P.set(f, v); ref.set(r, v)
    damping = [0.1, 0.4, 1.0]
    dvv = DampedVelocityVerlet(P, 5e-3, damping)
    dvv.init()
    dvv.step(50)
    dvv.run(300)
    ref.damped_vv_init(5e-3, damping)
    ref.vv_run(350)
    for f, r in (("POS_X", "x"), ("POS_Y", "y"), ("ANGLE", "angle"), ("VEL_X", "vx"), ("ANGULAR_VEL", "w"),
                 ("TORQUE", "torque"), ("VERTEX_POS_X", "vpx")):
        g, o = P.get(f), ref.get(r)
        assert np.array_equal(g, o), f"rigid damped VV: {f} max |diff| {np.max(np.abs(g - o)):.3e}"


@pytest.mark.parametrize("damped", [False, True])
def test_nvt_rescale_run_matches_oracle(damped):
    b = W.disk_batch([300, 180, 256], [0.70, 0.74, 0.72], [0, 1, 0], seed=21)
    P, ref = make(b)
    random_velocities(P, ref, b.N, seed=7, scale=3e-2)
    mode = L.FORCE_PER_SYSTEM
    if damped:
        DampedVelocityVerlet(P, 1e-2, 0.05, mode=mode).init()
        ref.damped_vv_init(1e-2, 0.05, mode)
    else:
        VelocityVerlet(P, 1e-2, mode=mode).init()
        ref.vv_init(1e-2, mode)
    inc = np.array([1e-3, 5e-4, 2e-3])
    T = np.array([1e-4, 3e-4, 2e-4])
    nvt_rescale_run(P, 450, inc, 100, T, 10)
    ref.nvt_rescale_run(450, inc, 100, T, 10, mode=mode)
    nvt_rescale_run(P, 150, None, 0, T, 10, step0=450)
    ref.nvt_rescale_run(150, None, 0, T, 10, step0=450, mode=mode)
    assert_same(P, ref, what="NVT rescale run")
    assert np.array_equal(P.get("BOX_SIZE_X"), ref.get("Lx"))
    P.compute_packing_fraction()
    assert np.array_equal(P.get("PACKING_FRACTION"), ref.packing_fraction())


def test_jam_binary_search_matches_oracle():
    b = W.disk_batch([128, 160, 140, 128], [0.78, 0.80, 0.82, 0.79], [0, 0, 0, 0], seed=31)
    P, ref = make(b)
    args = (1e-2, 12, 4000, 1e-10, 1e-10, 5e-3, 1e-4)
    act, phi, rounds = jam_binary_search(P, *args[:7], mode=L.FORCE_PBC)
    ract, rphi, rrounds = ref.jam_bisect(*args, mode=0)
    assert rounds == rrounds
    assert np.array_equal(act, ract) and np.array_equal(phi, rphi)
    assert_same(P, ref, ("x", "y", "fx", "fy", "pe"), "jam_binary_search")
    assert np.array_equal(P.get("BOX_SIZE_X"), ref.get("Lx"))


def test_jam_binary_search_one_disk_walls():
    """Closed form: one disk of radius r in a walled square box jams at L = 2r, phi = pi / 4."""
    r, phi0 = 0.7, 0.5
    b = W.disk_batch([1, 1], [phi0, phi0], [1, 1], seed=1)
    b.rad[:] = r
    b.area[:] = math.pi * r * r
    b.Lx[:] = b.Ly[:] = math.sqrt(b.area[0] / phi0)
    P, ref = make(b, method=L.NEIGHBOR_NAIVE)
    tol = 1e-7
    act, phi, rounds = jam_binary_search(P, 1e-2, 300, 20000, 1e-16, 1e-16, 0.02, tol, wall=True)
    ract, rphi, rrounds = ref.jam_bisect(1e-2, 300, 20000, 1e-16, 1e-16, 0.02, tol, mode=1)
    assert np.all(act == 0) and rounds == rrounds
    assert np.array_equal(phi, rphi)
    assert np.all(np.abs(phi / (math.pi / 4) - 1) < 2 * tol)


RNAMES = {"x": "POS_X", "y": "POS_Y", "angle": "ANGLE", "vx": "VEL_X", "vy": "VEL_Y", "w": "ANGULAR_VEL",
          "fx": "FORCE_X", "fy": "FORCE_Y", "torque": "TORQUE", "pe": "PE", "vpx": "VERTEX_POS_X",
          "vpy": "VERTEX_POS_Y"}


def make_rigid(b):
    P = RigidBumpy()
    W.load_rigid_into(P, b)
    P.init_neighbors()
    return P, orc.OracleRigid(b)


def assert_same_rigid(P, ref, fields, what):
    for f in fields:
        g, r = P.get(RNAMES[f]), ref.get(f)
        if not np.array_equal(g, r):
            raise AssertionError(f"{what}: {f} differs, max |diff| = {np.max(np.abs(g - r)):.3e}")


def test_rigid_save_restore_matches_oracle():
    b = W.rigid_batch(S=2, n_per=24, seed=4, boundary=0, phi=0.6)
    P, ref = make_rigid(b)
    P.save_state([1, 0], 1)
    ref.save_state([1, 0], 1)
    rng = np.random.default_rng(2)
    for f, r, s in (("VEL_X", "vx", 0.05), ("VEL_Y", "vy", 0.05), ("ANGULAR_VEL", "w", 0.1)):
        v = rng.normal(size=b.N) * s
        P.set(f, v); ref.set(r, v)
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    vv.run(300)
    ref.vv_init(1e-2); ref.vv_run(300)
    P.restore_state([1, 0], 1)
    ref.restore_state([1, 0], 1)
    assert np.array_equal(P.get("STATIC_INDEX"), ref.static_index())
    assert_same_rigid(P, ref, ("x", "y", "angle", "vpx", "vpy"), "rigid restore_state")
    P.compute_forces()
    ref.compute_forces(0)
    assert_same_rigid(P, ref, ("fx", "fy", "torque", "pe"), "forces after restore")


def test_rigid_jam_binary_search_matches_oracle():
    b = W.rigid_batch(S=2, n_per=32, seed=6, boundary=0, phi=0.62)
    P, ref = make_rigid(b)
    args = (1e-2, 5, 3000, 1e-10, 1e-10, 1e-2, 1e-4)
    act, phi, rounds = jam_binary_search(P, *args, mode=L.FORCE_PBC)
    ract, rphi, rrounds = ref.jam_bisect(*args, mode=0)
    assert rounds == rrounds and np.array_equal(act, ract) and np.array_equal(phi, rphi)
    assert_same_rigid(P, ref, ("x", "y", "angle", "vpx", "vpy", "fx", "fy", "torque"), "rigid jam")


def test_rigid_jam_binary_search_one_body_walls():
    """Closed form: a regular bumpy 12-gon jams in a square wall box of side
    2 R cos(pi / 12) + 2 r_b, phi_J = area / L_J^2."""
    from test_oracle_routines import one_rigid_body_walls
    n_v, R = 12, 0.5
    b = one_rigid_body_walls(n_v, R)
    P, ref = make_rigid(b)
    act, phi, rounds = jam_binary_search(P, 1e-2, 300, 40000, 1e-16, 1e-16, 0.02, 1e-6, wall=True)
    ract, rphi, rrounds = ref.jam_bisect(1e-2, 300, 40000, 1e-16, 1e-16, 0.02, 1e-6, mode=1)
    assert rounds == rrounds and np.array_equal(phi, rphi)
    LJ = 2 * R * math.cos(math.pi / n_v) + 2 * R * math.sin(math.pi / n_v)
    assert act[0] == 0 and abs(phi[0] / (b.area[0] / LJ ** 2) - 1) < 1e-5
