"""GPU parity tests of the rigid bumpy-polygon path: HIP library vs the oracle, bit for bit
(vertex neighbor lists in order, vertex forces, centre-of-mass force / torque / energy,
FIRE and velocity-Verlet trajectories including the vertex rotation)."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import FIRE, DampedVelocityVerlet, RigidBumpy, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu

PNAMES = {"x": "POS_X", "y": "POS_Y", "angle": "ANGLE", "vx": "VEL_X", "vy": "VEL_Y", "w": "ANGULAR_VEL",
          "fx": "FORCE_X", "fy": "FORCE_Y", "torque": "TORQUE", "pe": "PE"}
VNAMES = {"vpx": "VERTEX_POS_X", "vpy": "VERTEX_POS_Y", "vfx": "VERTEX_FORCE_X", "vfy": "VERTEX_FORCE_Y",
          "vpe": "VERTEX_PE"}


def make(b):
    P = RigidBumpy()
    W.load_rigid_into(P, b)
    P.init_neighbors()
    return P, orc.OracleRigid(b)


def assert_same(P, ref, fields, what=""):
    for f in fields:
        g = P.get(PNAMES.get(f) or VNAMES[f])
        r = ref.get(f)
        if not np.array_equal(g, r):
            raise AssertionError(f"{what}: {f} differs, max |diff| = {np.max(np.abs(g - r)):.3e}")


def small(boundary=1, seed=3):
    return W.rigid_batch(S=3, n_per=48, seed=seed, boundary=boundary, phi=0.72)


@pytest.mark.parametrize("boundary", [0, 1])
def test_rigid_lists_and_forces(boundary):
    b = small(boundary)
    P, ref = make(b)
    assert np.array_equal(P.get("STATIC_INDEX"), ref.static_index())
    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)
    mode = L.FORCE_WALL if boundary else L.FORCE_PBC
    P.compute_forces_mode(mode)
    ref.compute_forces(mode)
    assert_same(P, ref, ("vfx", "vfy", "vpe", "fx", "fy", "torque", "pe"), "rigid forces")


def test_rigid_fire_api_and_fused():
    b = small(1)
    P, ref = make(b)
    fire = FIRE(P, 1e-2, mode=L.FORCE_WALL)
    fire.init()
    ref.fire_init(1e-2, 1)
    for _ in range(200):
        fire.step()
        c1 = fire.converged_max_force(1e-9)
        ref.fire_step(1)
        c2 = ref.fire_converged(0.0, 0.0, 1, 1e-9)
        assert c1 == c2
    assert_same(P, ref, ("x", "y", "angle", "vx", "vy", "w", "vpx", "vpy", "fx", "fy", "torque"), "rigid FIRE API")
    Q, ref2 = make(b)
    steps, conv = run_fire(Q, 1e-2, 3000, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_WALL, use_graph=True)
    rsteps, rconv = ref2.fire_minimize(1e-2, 3000, pred=1, ftol=1e-9, mode=1)
    assert (steps, conv) == (rsteps, rconv)
    assert_same(Q, ref2, ("x", "y", "angle", "vpx", "vpy", "fx", "fy", "torque", "pe"), "rigid fused FIRE")


def test_rigid_vv_and_energies():
    b = small(0, seed=9)
    P, ref = make(b)
    run_fire(P, 1e-2, 400, L.FIRE_PRED_MAX_FORCE, 1e-9)
    ref.fire_minimize(1e-2, 400, pred=1, ftol=1e-9)
    rng = np.random.default_rng(0)
    for f, s in (("VEL_X", 1e-2), ("VEL_Y", 1e-2), ("ANGULAR_VEL", 1e-2)):
        v = rng.standard_normal(b.N) * s
        P.set(f, v)
        ref.set({"VEL_X": "vx", "VEL_Y": "vy", "ANGULAR_VEL": "w"}[f], v)
    vv = VelocityVerlet(P, 5e-3)
    vv.init()
    vv.step(100)
    vv.run(400, use_graph=True)
    ref.vv_init(5e-3)
    ref.vv_run(500)
    assert_same(P, ref, ("x", "y", "angle", "vx", "vy", "w", "vpx", "vpy"), "rigid VV")
    P.compute_pe_total(); P.compute_ke_total(); P.compute_pressure(); P.compute_fpower_total()
    assert np.array_equal(P.get("PE_TOTAL"), ref.pe_total())
    assert np.array_equal(P.get("KE_TOTAL"), ref.ke_total())
    assert np.array_equal(P.get("FPOWER_TOTAL"), ref.fpower_total())
    np.testing.assert_allclose(P.get("PRESSURE"), ref.pressure(), rtol=1e-12)
    # set_n_dof_kernel (rigid_bumpy.cu:743-757) and T = 2 KE / n_dof
    P.compute_temperature()
    assert np.array_equal(P.get("N_DOF"), ref.n_dof())
    assert np.array_equal(P.get("TEMPERATURE"), ref.temperature())


def test_rigid_n_dof_counts_only_bodies_with_inertia():
    b = small(1, seed=5)
    b.inertia[::7] = 0.0
    P, ref = make(b)
    rng = np.random.default_rng(3)
    for f, r in (("VEL_X", "vx"), ("VEL_Y", "vy"), ("ANGULAR_VEL", "w")):
        v = rng.standard_normal(b.N) * 1e-2
        P.set(f, v); ref.set(r, v)
    P.compute_temperature()
    nd = [int(np.sum(2 + (b.inertia[b.off[s]:b.off[s + 1]] > 0))) for s in range(b.S)]
    assert list(P.get("N_DOF")) == nd == list(ref.n_dof())
    assert np.array_equal(P.get("TEMPERATURE"), ref.temperature())


@pytest.mark.parametrize("damped", [0, 1])
@pytest.mark.parametrize("graph", [True, False])
def test_rigid_fused_vv_equals_two_launch_vv(graph, damped):
    """Fused rigid velocity-Verlet batches (one launch closes a step and opens the next) against
    one drift and one force launch per step and against the oracle, hot enough that rebuilds
    halt batches on many steps; plain and damped."""
    b = small(1, seed=4)
    rng = np.random.default_rng(6)
    vel = {f: rng.standard_normal(b.N) * sc for f, sc in (("vx", 0.08), ("vy", 0.08), ("w", 0.2))}
    key = {"vx": "VEL_X", "vy": "VEL_Y", "w": "ANGULAR_VEL"}
    runs = []
    for fused in (1, 0):
        P = RigidBumpy()
        W.load_rigid_into(P, b)
        L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_FUSED_VV, fused), P.ctx)
        P.init_neighbors()
        for f, v in vel.items():
            P.set(key[f], v)
        vv = DampedVelocityVerlet(P, 1e-2, 0.3, wall=True) if damped else VelocityVerlet(P, 1e-2, wall=True)
        vv.init()
        vv.run(300, use_graph=graph)
        runs.append(P)
    ref = orc.OracleRigid(b)
    for f, v in vel.items():
        ref.set(f, v)
    if damped:
        ref.damped_vv_init(1e-2, 0.3, 1)
    else:
        ref.vv_init(1e-2, 1)
    ref.vv_run(300, 1)
    fields = ("x", "y", "angle", "vx", "vy", "w", "fx", "fy", "torque", "pe", "vpx", "vpy")
    for f in fields:
        g0 = runs[0].get(PNAMES.get(f) or VNAMES[f])
        g1 = runs[1].get(PNAMES.get(f) or VNAMES[f])
        assert np.array_equal(g0, g1), f
    assert_same(runs[0], ref, fields, "rigid fused VV vs oracle")
