"""GPU tests of the persistent per-system FIRE kernel (one workgroup per system for the whole
run) against the graph-batched per-step launches and the oracle.

Bar: bit-identical.  Same step count and convergence flag, same per-system FIRE step counts,
FIRE scalars, particle order, positions, velocities, forces and energies, for every
workgroup size, for runs where every system stops (the last stop sets the run length, and
systems that stopped earlier take one more step) and runs cut by the step cap."""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, jam_compress, run_fire

pytestmark = pytest.mark.gpu

FIELDS = ("POS_X", "POS_Y", "VEL_X", "VEL_Y", "FORCE_X", "FORCE_Y", "PE", "STATIC_INDEX", "FIRE_STEPS",
          "LAST_POS_X", "LAST_POS_Y", "NEIGHBOR_COUNT")


def batch_mixed():
    return W.disk_batch([64, 200, 97, 333, 150, 64, 1500, 700], [0.80, 0.86, 0.83, 0.85, 0.84, 0.81, 0.845, 0.82],
                        [0, 1, 0, 1, 0, 1, 0, 1], seed=11)


def run(b, persist, block=512, max_steps=20000, ftol=1e-12, phase=2000, lds=1):
    """persist: 0 per-step launches, 1 persistent phases with hand-off, 2 persistent only;
    lds: 1 lets launches whose systems have at most 1536 particles hold them in LDS."""
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, int(persist)), P.ctx)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_BLOCK, block), P.ctx)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_PHASE, phase), P.ctx)
    L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_LDS, lds), P.ctx)
    print(f"persist={persist} block={block} lds={lds} phase={phase} S={b.S} N={b.N} max_steps={max_steps}",
          flush=True)
    steps, conv = run_fire(P, 1e-2, max_steps, L.FIRE_PRED_MAX_FORCE, ftol, mode=L.FORCE_PER_SYSTEM)
    S = P.n_systems()
    st = [np.empty(S), np.empty(S), np.empty(S, np.int32), np.empty(S, np.int32)]
    L.check(L.lib().dpmd_fire_state(P.ctx, *[a.ctypes.data for a in st]), P.ctx)
    return P, steps, conv, st


@pytest.fixture(scope="module")
def lean_converged():
    # every system stops (the oracle takes 65822 steps at this tolerance)
    return run(batch_mixed(), 0, max_steps=100000, ftol=1e-8)


@pytest.mark.parametrize("block,lds", [(256, 0), (512, 0), (1024, 0), (512, 1)])
def test_persistent_equals_graph_batches_to_convergence(lean_converged, block, lds):
    b = batch_mixed()
    Q, qsteps, qconv, qst = lean_converged
    P, steps, conv, st = run(b, 2, block, max_steps=100000, ftol=1e-8, phase=7000, lds=lds)
    assert (steps, conv) == (qsteps, qconv) and conv
    for a, c in zip(st, qst):
        assert np.array_equal(a, c)
    for f in FIELDS:
        assert np.array_equal(P.get(f), Q.get(f)), f
    # systems stop at different steps: the run length is set by the last one
    ns = P.get("FIRE_STEPS")
    assert ns.max() == steps and ns.min() < steps


@pytest.mark.parametrize("lds", [0, 1])
def test_persistent_equals_graph_batches_at_step_cap(lds):
    b = batch_mixed()
    Q, qsteps, qconv, qst = run(b, 0, max_steps=700, ftol=1e-12)
    P, steps, conv, st = run(b, 2, max_steps=700, ftol=1e-12, phase=300, lds=lds)
    assert (steps, conv) == (qsteps, qconv) == (700, 0)
    for a, c in zip(st, qst):
        assert np.array_equal(a, c)
    for f in FIELDS:
        assert np.array_equal(P.get(f), Q.get(f)), f


@pytest.mark.parametrize("lds", [0, 1])
def test_persistent_matches_oracle(lds):
    b = batch_mixed()
    P, steps, conv, _ = run(b, 2, max_steps=3000, ftol=1e-11, phase=1000, lds=lds)
    ref = orc.OracleDisk(b)
    rsteps, rconv = ref.fire_minimize(1e-2, 3000, pred=1, ftol=1e-11, mode=2)
    assert P.get("FIRE_STEPS").min() < steps   # some systems stopped within a phase
    assert steps == rsteps and bool(conv) == bool(rconv)
    for f, n in (("x", "POS_X"), ("y", "POS_Y"), ("vx", "VEL_X"), ("vy", "VEL_Y"), ("fx", "FORCE_X"),
                 ("fy", "FORCE_Y"), ("pe", "PE")):
        assert np.array_equal(P.get(n), ref.get(f)), f


def test_persistent_jam_equals_graph_batches():
    b = W.uneven_batch(S=12, nmin=48, nmax=400, seed=2, phi_lo=0.80, phi_hi=0.84)
    out = []
    for persist in (False, True):
        P = Disk()
        W.load_into(P, b)
        P.init_neighbors()
        L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, 2 if persist else 0), P.ctx)
        L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_PHASE, 700), P.ctx)
        print(f"jam persist={persist}", flush=True)
        st, nc, work = jam_compress(P, 1e-2, 4e-3, 1e-7, 30, 20000, L.FIRE_PRED_FORCE_OR_PE, 1e-9, 1e-16,
                                    mode=L.FORCE_PER_SYSTEM)
        out.append((st, nc, work, {f: P.get(f) for f in FIELDS}))
    (s0, n0, w0, f0), (s1, n1, w1, f1) = out
    assert np.array_equal(s0, s1) and np.array_equal(n0, n1) and w0 == w1
    for f in FIELDS:
        assert np.array_equal(f0[f], f1[f]), f


def test_persistent_option_validation():
    P = Disk()
    W.load_into(P, batch_mixed())
    for bad in (128, 768, 2048):
        assert L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_BLOCK, bad) == 1
    v = C.c_int()
    L.check(L.lib().dpmd_get_option(P.ctx, L.OPT_PERSISTENT_FIRE, C.byref(v)), P.ctx)
    assert v.value == 1
    for bad in (-1, 3):
        assert L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, bad) == 1
    assert L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_PHASE, 0) == 1
    L.check(L.lib().dpmd_get_option(P.ctx, L.OPT_PERSIST_LDS, C.byref(v)), P.ctx)
    assert v.value == 1
    assert L.lib().dpmd_set_option(P.ctx, L.OPT_PERSIST_LDS, 2) == 1


def test_lds_engine_single_system_from_random_start():
    """One 1024-disk system from uniform random positions (the plain-method workload): many
    rebuilds inside the LDS-resident run, each storing the state, rebuilding and reloading.
    Equal to the global-memory engine and to the oracle."""
    b = W.single_disk_system(1024, 0.85, seed=1)
    Q, qsteps, qconv, qst = run(b, 2, max_steps=4000, ftol=1e-12, phase=4000, lds=0)
    P, steps, conv, st = run(b, 2, max_steps=4000, ftol=1e-12, phase=4000, lds=1)
    assert (steps, conv) == (qsteps, qconv)
    for x, y in zip(st, qst):
        assert np.array_equal(x, y)
    for f in FIELDS:
        assert np.array_equal(P.get(f), Q.get(f)), f
    ref = orc.OracleDisk(b)
    rsteps, rconv = ref.fire_minimize(1e-2, 4000, pred=1, ftol=1e-12, mode=L.FORCE_PER_SYSTEM)
    assert steps == rsteps and bool(conv) == bool(rconv)
    for f, n in (("x", "POS_X"), ("vx", "VEL_X"), ("fy", "FORCE_Y"), ("pe", "PE")):
        assert np.array_equal(P.get(n), ref.get(f)), f


def test_hand_off_to_per_step_launches_mid_run():
    """200 systems of 512 disks and one of 2048: persistent phases while the small systems
    run, then (once they have stopped) the rest of the run as per-step launches.  Equal to
    the per-step launches alone and to persistent phases alone."""
    b = W.disk_batch([512] * 200 + [2048], [0.80] * 200 + [0.845], [0] * 100 + [1] * 100 + [0], seed=5)
    ref = run(b, 0, max_steps=30000, ftol=1e-8)
    for persist in (1, 2):
        P, steps, conv, st = run(b, persist, max_steps=30000, ftol=1e-8, phase=500)
        assert (steps, conv) == (ref[1], ref[2])
        for x, y in zip(st, ref[3]):
            assert np.array_equal(x, y)
        for f in FIELDS:
            assert np.array_equal(P.get(f), ref[0].get(f)), (persist, f)
    ns = ref[0].get("FIRE_STEPS")
    assert ns[:200].max() < ns[200]   # the large system outlasts the small ones


def test_jamming_hand_off_shrinks_to_one_fused_system():
    """jam_compress (frozen systems masked) on 60 systems of 256 disks and one of 2048: the
    rounds are handed to per-step launches over the running systems, whose set shrinks as
    they freeze, and a system left running alone takes the fused FIRE step on its slice of
    the batch (counted with HIP events).  Equal, bit for bit, to the same jamming with
    per-step launches over the whole batch."""
    b = W.disk_batch([256] * 60 + [2048], [0.80] * 60 + [0.845], [0] * 30 + [1] * 30 + [0], seed=5)
    out = []
    for persist in (0, 1):
        P = Disk()
        W.load_into(P, b)
        P.init_neighbors()
        L.check(L.lib().dpmd_set_option(P.ctx, L.OPT_PERSISTENT_FIRE, persist), P.ctx)
        print(f"jam hand-off persist={persist} S={b.S} N={b.N}", flush=True)
        # the hand-off run is launched directly, with HIP events around its fused FIRE steps
        L.check(L.lib().dpmd_profile(P.ctx, 3 if persist else 0), P.ctx)
        st, nc, work = jam_compress(P, 1e-2, 4e-3, 1e-7, 3, 30000, L.FIRE_PRED_MAX_FORCE, 1e-9,
                                    mode=L.FORCE_PER_SYSTEM, use_graph=not persist)
        ms, cnt = C.c_double(), C.c_int()
        L.check(L.lib().dpmd_profile_read(P.ctx, C.byref(ms), C.byref(cnt)), P.ctx)
        out.append((st, nc, work, {f: P.get(f) for f in FIELDS}, cnt.value))
    (s0, n0, w0, f0, _), (s1, n1, w1, f1, fused_steps) = out
    print(f"fused FIRE steps of the last system: {fused_steps}", flush=True)
    assert fused_steps > 0, "the last running system should take the fused FIRE step"
    assert np.array_equal(s0, s1) and np.array_equal(n0, n1) and w0 == w1
    for f in FIELDS:
        assert np.array_equal(f0[f], f1[f]), f
