"""GPU parity tests of the batched jamming driver and of frozen-system skipping: the HIP
library (through the C ABI) against the oracle, bit for bit, and the batched run against
one context per system (the plain path) at identical final states."""
import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, jam_compress, run_fire

pytestmark = pytest.mark.gpu

DPHI, PT, FTOL, FMAX, CMAX = 4e-3, 1e-7, 1e-9, 20000, 30
FIELDS = (("x", "POS_X"), ("y", "POS_Y"), ("fx", "FORCE_X"), ("fy", "FORCE_Y"), ("pe", "PE"))


def small_batch():
    return W.uneven_batch(S=4, nmin=48, nmax=64, seed=2, phi_lo=0.78, phi_hi=0.81)


def load(b):
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    return P


def final_state(P):
    P.compute_pe_total()
    P.compute_pressure()
    pos = {f: P.get(n, L.ORDER_CANONICAL) for f, n in FIELDS[:2]}
    return pos | {"L": P.get("BOX_SIZE_X"), "PE_TOTAL": P.get("PE_TOTAL"), "P": P.get("PRESSURE")}


@pytest.fixture(scope="module")
def oracle_run():
    orc.lib().or_set_threads(1)
    b = small_batch()
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, CMAX, FMAX, pred=1, a=FTOL, mode=2)
    return b, o, st, nc, ns, work


@pytest.mark.parametrize("use_graph", [True, False])
def test_jam_matches_oracle(oracle_run, use_graph):
    b, o, st, nc, ns, work = oracle_run
    P = load(b)
    print(f"jam S={b.S} N={b.N} graph={use_graph}", flush=True)
    gst, gnc, gwork = jam_compress(P, 1e-2, DPHI, PT, CMAX, FMAX, L.FIRE_PRED_MAX_FORCE, FTOL,
                                   mode=L.FORCE_PER_SYSTEM, use_graph=use_graph)
    assert np.array_equal(gst, st) and np.array_equal(gnc, nc)
    assert np.array_equal(P.get("FIRE_STEPS"), ns) and gwork == work
    for f, n in FIELDS:
        assert np.array_equal(P.get(n), o.get(f)), f
    assert np.array_equal(P.get("BOX_SIZE_X"), o.get("Lx"))
    P.compute_pressure()
    assert np.array_equal(P.get("PRESSURE"), o.pressure())


def test_jam_batched_equals_plain_per_system():
    """While a system is compressing, its rebuild history is the same in the batch and alone, so
    its trajectory is bit-identical.  After it jams, the batch still force-rebuilds it at the
    other systems' compressions (increment_packing_fraction rebuilds every system), which
    re-sorts its particles: positions stay bit-identical (compared in canonical order) while
    forces are summed in another neighbor order, so energies and pressure agree to rounding
    (bar: 1e-12 relative for jammed systems, 1e-14 absolute for the others, whose
    pressure is a near-cancelling sum; the workload's bar is 1e-10 relative)."""
    b = small_batch()
    P = load(b)
    jst, jnc, _ = jam_compress(P, 1e-2, DPHI, PT, CMAX, FMAX, L.FIRE_PRED_MAX_FORCE, FTOL, mode=L.FORCE_PER_SYSTEM)
    batched = final_state(P)
    off = b.off
    for s in range(b.S):
        Q = load(W.system_of(b, s))
        qst, qnc, _ = jam_compress(Q, 1e-2, DPHI, PT, CMAX, FMAX, L.FIRE_PRED_MAX_FORCE, FTOL,
                                   mode=L.FORCE_PER_SYSTEM)
        assert qst[0] == jst[s] and qnc[0] == jnc[s]
        one = final_state(Q)
        for f in ("x", "y"):
            assert np.array_equal(one[f], batched[f][off[s]:off[s + 1]]), (s, f)
        assert one["L"][0] == batched["L"][s]
        # a system that is not jammed has P ~ 0 from cancelling virial terms: an absolute floor
        atol = 0.0 if jst[s] == 1 else 1e-14
        for k in ("PE_TOTAL", "P"):
            np.testing.assert_allclose(one[k][0], batched[k][s], rtol=1e-12, atol=atol, err_msg=f"system {s} {k}")


def test_frozen_systems_skip_work_exactly():
    """An uneven batch whose systems stop at very different steps: the fused run (which skips
    frozen systems) ends bit-identical to the oracle, and each system is charged
    min(stop + 1, total) steps."""
    b = W.disk_batch([64, 300, 97, 200], [0.70, 0.845, 0.72, 0.85], [0, 1, 0, 0], seed=5)
    P = load(b)
    steps, conv = run_fire(P, 1e-2, 30000, L.FIRE_PRED_MAX_FORCE, 1e-11, mode=L.FORCE_PER_SYSTEM)
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, -1.0, 0, 30000, pred=1, a=1e-11, mode=2)
    assert conv and steps == int(ns.max())
    assert np.array_equal(P.get("FIRE_STEPS"), ns)
    assert len(set(ns.tolist())) > 1, "systems should stop at different steps"
    for f, n in FIELDS:
        assert np.array_equal(P.get(n), o.get(f)), f


def test_force_or_pe_predicate_matches_oracle():
    """Jamming with stops at max|F| < 1e-9 or <pe> < 1e-16: library and oracle bit for bit."""
    orc.lib().or_set_threads(1)
    b = small_batch()
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, CMAX, FMAX, pred=2, a=FTOL, b=1e-16, mode=2)
    P = load(b)
    gst, gnc, gwork = jam_compress(P, 1e-2, DPHI, PT, CMAX, FMAX, L.FIRE_PRED_FORCE_OR_PE, FTOL, 1e-16,
                                   mode=L.FORCE_PER_SYSTEM)
    assert np.array_equal(gst, st) and np.array_equal(gnc, nc)
    assert np.array_equal(P.get("FIRE_STEPS"), ns) and gwork == work
    for f, n in FIELDS:
        assert np.array_equal(P.get(n), o.get(f)), f


@pytest.mark.parametrize("use_graph", [True, False])
def test_jam_with_round_retries_matches_oracle(use_graph):
    """Rounds that hit a short step cap run again (DPMD_OPT_JAM_RETRIES = 2): decisions,
    compressions, step counts and the final state bit for bit with the oracle's driver."""
    orc.lib().or_set_threads(1)
    cap = 1500
    b = small_batch()
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, CMAX, cap, pred=1, a=FTOL, mode=2, retries=2)
    P = load(b)
    print(f"jam retries S={b.S} graph={use_graph}", flush=True)
    gst, gnc, gwork = jam_compress(P, 1e-2, DPHI, PT, CMAX, cap, L.FIRE_PRED_MAX_FORCE, FTOL,
                                   mode=L.FORCE_PER_SYSTEM, use_graph=use_graph, retries=2)
    assert np.array_equal(gst, st) and np.array_equal(gnc, nc)
    assert np.array_equal(P.get("FIRE_STEPS"), ns) and gwork == work
    for f, n in FIELDS:
        assert np.array_equal(P.get(n), o.get(f)), f


def test_jam_retries_batched_equals_plain_per_system():
    """With retries, a system's rounds (and the rebuild at each round boundary) do not depend
    on the other systems: positions bit-identical alone and in the batch."""
    cap = 1500
    b = small_batch()
    P = load(b)
    jst, jnc, _ = jam_compress(P, 1e-2, DPHI, PT, CMAX, cap, L.FIRE_PRED_MAX_FORCE, FTOL, mode=L.FORCE_PER_SYSTEM,
                               retries=2)
    batched = final_state(P)
    for s in range(b.S):
        Q = load(W.system_of(b, s))
        qst, qnc, _ = jam_compress(Q, 1e-2, DPHI, PT, CMAX, cap, L.FIRE_PRED_MAX_FORCE, FTOL,
                                   mode=L.FORCE_PER_SYSTEM, retries=2)
        assert qst[0] == jst[s] and qnc[0] == jnc[s]
        one = final_state(Q)
        for f in ("x", "y"):
            assert np.array_equal(one[f], batched[f][b.off[s]:b.off[s + 1]]), (s, f)
