"""GPU tests of the autotuner and of the launch options it turns.

Bars: the workgroup size and hipGraph batch length change no bit of a trajectory (checked
against the oracle); a tuned configuration keeps the neighbor lists complete (canonical
sorted sets equal the oracle's at the same cutoff and grid) and the forces within 1e-12
relative; trials leave positions and velocities bit-identical; velocity Verlet after tuning
agrees with the untuned run to 1e-9 absolute over 100 steps (only the force summation order
differs); FIRE polished from a common near-minimum reaches the same energy within 1e-10
relative (jammed systems; unjammed ones stop at energies below 1e-16 on both sides); a cache
hit applies the stored configuration without trials."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import autotune as T
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu

NAMES = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y", "pe": "PE"}


def batch_small():
    return W.disk_batch([64, 200, 97, 333, 150, 64], [0.80, 0.86, 0.83, 0.85, 0.84, 0.81],
                        [0, 1, 0, 1, 0, 1], seed=11)


def make(b):
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    return P


def set_canonical_velocities(P, v):
    si = P.get("STATIC_INDEX")
    cvx, cvy = np.empty(P.n_particles()), np.empty(P.n_particles())
    cvx[si], cvy[si] = v[0], v[1]
    P.set("VEL_X", cvx)
    P.set("VEL_Y", cvy)


def canon(P, f):
    return P.get(NAMES[f], L.ORDER_CANONICAL)


def state(P):
    return {f: canon(P, f) for f in ("x", "y", "vx", "vy")}


@pytest.mark.parametrize("block,max_batch,fused", [(64, 128, 1), (128, 16, 1), (512, 1, 1), (512, 32, 0)])
def test_launch_options_leave_vv_bit_identical(block, max_batch, fused):
    b = batch_small()
    v = np.random.default_rng(12).standard_normal((2, b.N)) * 0.3
    P = make(b)
    for key, val in ((L.OPT_BLOCK, block), (L.OPT_MAX_BATCH, max_batch), (L.OPT_FUSED_VV, fused)):
        L.check(L.lib().dpmd_set_option(P.ctx, key, val), P.ctx)
        got = L.C.c_int()
        L.check(L.lib().dpmd_get_option(P.ctx, key, L.C.byref(got)), P.ctx)
        assert got.value == val
    set_canonical_velocities(P, v)
    vv = VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM)
    vv.init()
    vv.run(400, use_graph=True)
    ref = orc.OracleDisk(b)
    si = ref.static_index()
    cvx, cvy = np.empty(b.N), np.empty(b.N)
    cvx[si], cvy[si] = v[0], v[1]
    ref.set("vx", cvx); ref.set("vy", cvy)
    ref.vv_init(1e-2, 2)
    ref.vv_run(400, 2)
    for f in ("x", "y", "vx", "vy", "fx", "fy", "pe"):
        assert np.array_equal(P.get(NAMES[f]), ref.get(f)), f


@pytest.mark.parametrize("block", [64, 512])
def test_fire_block_size_bit_identical(block):
    b = batch_small()
    P = make(b)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_BLOCK, block), P.ctx)
    steps, _ = run_fire(P, 1e-2, 600, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=L.FORCE_PER_SYSTEM)
    ref = orc.OracleDisk(b)
    rsteps, _ = ref.fire_minimize(1e-2, 600, pred=1, ftol=1e-12, mode=2)
    assert steps == rsteps
    for f in ("x", "y", "vx", "vy", "fx", "fy", "pe"):
        assert np.array_equal(P.get(NAMES[f]), ref.get(f)), f


def test_option_validation():
    P = make(batch_small())
    for key, bad in ((L.OPT_BLOCK, 96), (L.OPT_BLOCK, 1024), (L.OPT_MAX_BATCH, 0), (L.OPT_MAX_BATCH, 48),
                     (L.OPT_MAX_BATCH, 256)):
        assert L.lib().dpmd_set_option(P.ctx, key, bad) == 1


def test_apply_rejects_grids_that_break_the_stencil():
    b = batch_small()
    P = make(b)
    before = (P.get("CELL_DIM_X").copy(), P.get("NEIGHBOR_CUTOFF").copy())
    # the N=64, phi=0.80 box is 9.6 wide: cells of 2 x (2 x 0.7 + 0.3) leave only 2 per axis
    with pytest.raises(L.DpmdError) as e:
        T.apply(P, dict(cell_factor=2.0))
    assert e.value.status == 1 and "system 0" in str(e.value)
    with pytest.raises(L.DpmdError):
        T.apply(P, dict(skin=4.0))        # cutoff 4 -> cells of 5.4, one per axis
    assert np.array_equal(P.get("CELL_DIM_X"), before[0])
    assert np.array_equal(P.get("NEIGHBOR_CUTOFF"), before[1])
    P.compute_forces_mode(L.FORCE_PER_SYSTEM)


def test_applied_config_keeps_lists_complete_and_forces():
    b = batch_small()
    P = make(b)
    P.compute_forces_mode(L.FORCE_PER_SYSTEM)
    f0 = (canon(P, "fx"), canon(P, "fy"))
    cfg = dict(skin=0.6, cell_factor=1.25, block=128, max_batch=32, fused_vv=1)
    T.apply(P, cfg)
    got = T.current(P)
    assert got["skin"] == pytest.approx(0.6) and got["cell_factor"] >= 1.25
    assert (got["block"], got["max_batch"], got["fused_vv"]) == (128, 32, 1)
    # the oracle at the same cutoff, threshold and grid
    b2 = W.disk_batch([64, 200, 97, 333, 150, 64], [0.80, 0.86, 0.83, 0.85, 0.84, 0.81], [0, 1, 0, 1, 0, 1], seed=11)
    for s in range(b2.S):
        lo, hi = b2.off[s], b2.off[s + 1]
        b2.cutoff[s] = 0.6 * 2 * b2.rad[lo:hi].min()
        b2.thresh2[s] = (0.5 * b2.cutoff[s]) ** 2
        b2.cdx[s], b2.cdy[s] = T.cell_grid(b2.Lx[s], b2.Ly[s], b2.rad[lo:hi].max(), b2.cutoff[s], 1.25)
    assert np.array_equal(P.get("CELL_DIM_X"), b2.cdx) and np.array_equal(P.get("CELL_DIM_Y"), b2.cdy)
    assert np.allclose(P.get("THRESH2"), b2.thresh2, rtol=0, atol=0)
    ref = orc.OracleDisk(b2)

    def sets(start, ids, si_cur_of_canon):
        canon_of_cur = np.empty_like(si_cur_of_canon)
        canon_of_cur[si_cur_of_canon] = np.arange(si_cur_of_canon.size)
        out = {}
        for i in range(start.size - 1):
            out[canon_of_cur[i]] = sorted(canon_of_cur[ids[start[i]:start[i + 1]]].tolist())
        return out

    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert sets(gs, gi, P.get("STATIC_INDEX")) == sets(rs, ri, ref.static_index())
    P.compute_forces_mode(L.FORCE_PER_SYSTEM)
    scale = max(np.abs(f0[0]).max(), np.abs(f0[1]).max())
    assert np.max(np.abs(canon(P, "fx") - f0[0])) <= 1e-12 * scale
    assert np.max(np.abs(canon(P, "fy") - f0[1])) <= 1e-12 * scale


def test_autotune_fire_restores_state_and_caches(tmp_path):
    b = W.uneven_batch(S=16, nmin=64, nmax=1024, seed=4)
    cache = str(tmp_path / "tune.json")
    P = make(b)
    s0 = state(P)
    res = T.autotune(P, T.FIRE, 1e-2, mode=L.FORCE_PER_SYSTEM, steps=64, cache=cache)
    assert len(res.trials) >= 10 and res.rate > 0
    c = res.config
    assert c["skin"] in (0.1, 0.15, 0.2, 0.3, 0.4, 0.6) and c["cell_factor"] in (1.0, 1.25, 1.5)   # start: 0.3
    assert c["block"] in (64, 128, 256, 512) and c["max_batch"] in (16, 32, 64, 128)
    assert res.rate >= res.trials[0][1]   # the first trial times the starting configuration
    s1 = state(P)
    for k in s0:
        assert np.array_equal(s0[k], s1[k]), k
    now = T.current(P)
    assert now["block"] == c["block"] and now["max_batch"] == c["max_batch"]
    assert now["skin"] == pytest.approx(c["skin"], rel=1e-15)
    hit = T.cache_load(cache, res.signature)
    assert hit is not None and hit[0] == c
    # the tuned context is in the state of a fresh one with the configuration applied: same
    # particle order and neighbor lists, so both run the same arithmetic from here
    R = make(b)
    T.apply(R, c)
    assert np.array_equal(P.get("STATIC_INDEX"), R.get("STATIC_INDEX"))
    for a1, a2 in zip(P.neighbors_csr(), R.neighbors_csr()):
        assert np.array_equal(a1, a2)
    # a fresh context with the same workload: cache hit, no trials, same configuration applied
    Q = make(b)
    res2 = T.autotune(Q, T.FIRE, 1e-2, mode=L.FORCE_PER_SYSTEM, steps=64, cache=cache)
    assert res2.trials == [] and res2.config == c and res2.signature == res.signature
    assert T.current(Q)["block"] == c["block"]


def test_tuned_fire_reaches_the_same_minimum():
    """Polish one near-minimum with the default and a tuned configuration."""
    b = W.uneven_batch(S=8, nmin=64, nmax=512, seed=6)
    P = make(b)
    run_fire(P, 1e-2, 100000, L.FIRE_PRED_MAX_FORCE, 1e-6, mode=L.FORCE_PER_SYSTEM)
    x0, y0 = canon(P, "x"), canon(P, "y")
    results = []
    for cfg in (None, dict(skin=0.15, cell_factor=1.25, block=64, max_batch=16)):
        Q = make(b)
        si = Q.get("STATIC_INDEX")
        cx, cy = np.empty(b.N), np.empty(b.N)
        cx[si], cy[si] = x0, y0
        Q.set("POS_X", cx); Q.set("POS_Y", cy)
        Q.check_neighbors(True)
        if cfg:
            T.apply(Q, cfg)
        steps, conv = run_fire(Q, 1e-2, 200000, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=L.FORCE_PER_SYSTEM)
        assert conv
        Q.compute_pe_total()
        results.append(Q.get("PE_TOTAL"))
    e0, e1 = results
    # jammed systems: same minimum energy to 1e-10 relative; unjammed ones stop at max|F| <
    # 1e-12 with an energy of order N (1e-12)^2 / k, zero to this bar on both sides
    jammed = e0 > 1e-12
    assert jammed.any() and (~jammed).any()
    assert np.all(np.abs(e1 - e0)[jammed] <= 1e-10 * e0[jammed]), (e0, e1)
    assert np.all(e0[~jammed] < 1e-16) and np.all(e1[~jammed] < 1e-16), (e0, e1)


def test_autotune_nve_and_trajectory_equivalence():
    b = batch_small()
    v = np.random.default_rng(5).standard_normal((2, b.N)) * 0.05
    runs = []
    for tune in (False, True):
        P = make(b)
        set_canonical_velocities(P, v)
        if tune:
            s0 = state(P)
            res = T.autotune(P, T.NVE, 1e-2, mode=L.FORCE_PER_SYSTEM, steps=64)
            assert res.config["fused_vv"] in (0, 1) and len(res.trials) >= 10
            s1 = state(P)
            for k in s0:
                assert np.array_equal(s0[k], s1[k]), k
        vv = VelocityVerlet(P, 1e-2, mode=L.FORCE_PER_SYSTEM)
        vv.init()
        vv.run(100, use_graph=True)
        runs.append(state(P))
    for k in runs[0]:
        assert np.max(np.abs(runs[0][k] - runs[1][k])) <= 1e-9, k
