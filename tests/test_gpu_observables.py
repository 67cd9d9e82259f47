"""GPU parity tests of the disk contact observables (HIP library vs the oracle, bit for bit),
and the assembled Hessian against the adjoint's Hessian-vector product."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.adjoint import hvp
from dpmd_amd.system import Disk, run_fire
from test_oracle_observables import assemble

pytestmark = pytest.mark.gpu


def relaxed(b, method=L.NEIGHBOR_CELL, mode=L.FORCE_PER_SYSTEM):
    P = Disk()
    W.load_into(P, b, method)
    P.init_neighbors()
    ref = orc.OracleDisk(b, method=method)
    run_fire(P, 1e-2, 2000, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=mode)
    ref.fire_minimize(1e-2, 2000, pred=1, ftol=1e-12, mode=mode)
    P.compute_forces_mode(mode)
    ref.compute_forces(mode)
    return P, ref


@pytest.mark.parametrize("method", [L.NEIGHBOR_CELL, L.NEIGHBOR_NAIVE])
def test_contact_observables_match_oracle(method):
    b = W.disk_batch([64, 200, 97, 150], [0.88, 0.86, 0.90, 0.84], [0, 1, 0, 1], seed=13)
    P, ref = relaxed(b, method)
    P.compute_contacts()
    c, ov, cc = ref.contacts()
    assert np.array_equal(P.get("CONTACTS"), c)
    assert np.array_equal(P.get("OVERLAP"), ov)
    assert np.array_equal(P.get("CONTACT_COUNT"), cc)
    P.compute_n_contacts_total()
    tot = [int(c[b.off[s]:b.off[s + 1]].sum()) for s in range(b.S)]   # canonical and current agree per system
    assert np.array_equal(P.get("N_CONTACTS_TOTAL"), tot)
    g = P.pair_observables()
    r = ref.pair_observables()
    for k in r:
        assert np.array_equal(g[k], r[k]), k


def test_hessian_times_vector_matches_adjoint_hvp():
    b = W.disk_batch([128, 96], [0.90, 0.88], [0, 0], seed=14)
    P, _ = relaxed(b, mode=L.FORCE_PBC)
    H = assemble(P.pair_observables(), b.N)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((b.N, 2))
    yx, yy = hvp(P, w[:, 0], w[:, 1], mode=L.FORCE_PBC)
    y = H @ w.reshape(-1)
    np.testing.assert_allclose(np.stack([yx, yy], 1).reshape(-1), y, rtol=0, atol=1e-12 * np.abs(y).max())


@pytest.mark.parametrize("boundary", [0, 1])
def test_rigid_contact_observables_match_oracle(boundary):
    from dpmd_amd.system import RigidBumpy
    rb = W.rigid_batch(S=3, n_per=32, seed=8, boundary=boundary, phi=0.76)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P.init_neighbors()
    ref = orc.OracleRigid(rb)
    mode = L.FORCE_WALL if boundary else L.FORCE_PBC
    run_fire(P, 1e-2, 300, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=mode)
    ref.fire_minimize(1e-2, 300, pred=1, ftol=1e-9, mode=mode)
    gs, gi = P.particle_neighbors()
    rs, ri = ref.particle_neighbors()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)
    g = P.pair_observables()
    r = ref.pair_observables()
    for k in g:
        assert np.array_equal(g[k], r[k]), k
    P.compute_contacts()
    assert np.array_equal(P.get("CONTACTS"), r["contacts"])
    P.compute_n_contacts_total()
    assert np.array_equal(P.get("N_CONTACTS_TOTAL"),
                          [int(r["contacts"][rb.off[s]:rb.off[s + 1]].sum()) for s in range(rb.S)])
    P.compute_overlaps()
    ov, cc = ref.overlaps()
    assert np.array_equal(P.get("OVERLAP"), ov) and np.array_equal(P.get("CONTACT_COUNT"), cc)


def test_rigid_hessian_matches_oracle():
    from dpmd_amd.system import RigidBumpy
    rb = W.rigid_batch(S=2, n_per=24, seed=4, boundary=0, phi=0.82)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P.init_neighbors()
    ref = orc.OracleRigid(rb)
    run_fire(P, 1e-2, 40, L.FIRE_PRED_MAX_FORCE, 1e-9, mode=L.FORCE_PBC)
    ref.fire_minimize(1e-2, 40, pred=1, ftol=1e-9, mode=0)
    g, r = P.hessian(), ref.hessian()
    assert set(g) == set(r)
    for k in r:
        assert np.array_equal(g[k], r[k]), k
