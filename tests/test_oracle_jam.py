"""Pins the oracle's pressure-target jamming driver (CPU): packing-fraction bookkeeping
(phi_final = phi_0 + n_compressions * dphi), the jammed / unjammed decision against the
final pressure, the per-system work count, and the whole driver against an independent
loop over the oracle's pinned primitives (fire_minimize, pressure, increment_packing_fraction)."""
import os

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import workloads as W

# A coarser compression step and force tolerance than the 512-system workload (1e-3, 1e-10)
# keep the CPU cost to seconds; the logic under test is the same.
DPHI, PT, FTOL, FMAX, CMAX = 4e-3, 1e-7, 1e-9, 20000, 30


def small_batch(S=4, nmax=64, seed=2, phi_lo=0.78, phi_hi=0.81):
    return W.uneven_batch(S=S, nmin=48, nmax=nmax, seed=seed, phi_lo=phi_lo, phi_hi=phi_hi)


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    orc.lib().or_set_threads(1)          # systems of ~60 disks: threads only add overhead
    yield
    orc.lib().or_set_threads(os.cpu_count() or 1)


@pytest.fixture(scope="module")
def jammed():
    b = small_batch()
    o = orc.OracleDisk(b)
    phi0 = o.packing_fraction()
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, CMAX, FMAX, pred=1, a=FTOL, mode=2)
    return b, o, phi0, st, nc, ns, work


def test_phi_bookkeeping_and_decision(jammed):
    b, o, phi0, st, nc, ns, work = jammed
    phi = o.packing_fraction()
    np.testing.assert_allclose(phi, phi0 + nc * DPHI, rtol=0, atol=1e-12)
    P = o.pressure()
    assert np.all(P[st == 1] >= PT)
    assert np.all(P[(st == 0) & (nc == CMAX)] < PT)
    # this batch holds a system that jams after compressions and one jammed from the start;
    # a system is dropped only when its initial minimisation hits the step cap
    assert np.any((st == 1) & (nc > 0)) and np.any((st == 1) & (nc == 0))
    assert np.all(nc[st == -1] == 0)
    sizes = np.diff(b.off)
    assert work == int(np.sum(sizes.astype(np.int64) * ns))


def test_driver_equals_loop_over_primitives(jammed):
    """Re-run the rounds with the oracle's primitives from Python; the driver must agree bit for bit."""
    b, o, phi0, st, nc, ns, work = jammed
    q = orc.OracleDisk(small_batch())
    S = q.S
    state = np.zeros(S, int)
    ncomp = np.zeros(S, int)
    while True:
        active = state == 0
        if not active.any():
            break
        dts = np.where(active, 1e-2, 0.0)
        q.fire_minimize(dts, FMAX, pred=1, ftol=FTOL, mode=2)
        fdt = q.fire_state()[0]
        P = q.pressure()
        inc = np.zeros(S)
        for s in np.flatnonzero(active):
            if fdt[s] != 0.0 and ncomp[s] == 0:   # the initial minimisation did not stop
                state[s] = -1
            elif P[s] >= PT:
                state[s] = 1
            elif ncomp[s] >= CMAX:
                state[s] = 2
            else:
                inc[s] = DPHI
                ncomp[s] += 1
        if not inc.any():
            break
        q.increment_packing_fraction(inc)
    assert np.array_equal(np.where(state == 2, 0, state), st)
    assert np.array_equal(ncomp, nc)
    for f in ("x", "y", "fx", "fy"):
        assert np.array_equal(q.get(f), o.get(f)), f
    assert np.array_equal(q.pressure(), o.pressure())


def test_counted_steps_stop_one_after_stop():
    """A system costs min(stop_step + 1, total) steps: one uneven pair of systems where one stops early."""
    b = small_batch(S=2, nmax=80, seed=3)
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, 0, FMAX, pred=1, a=FTOL, mode=2)
    q = orc.OracleDisk(b)
    q.fire_init(1e-2, 2)
    stop = [None, None]
    k = 0
    while not all(v is not None for v in stop) and k < FMAX:
        q.fire_step(2)
        q.fire_converged(0.0, 0.0, 1, FTOL)
        k += 1
        dt = q.fire_state()[0]
        for s in range(2):
            if stop[s] is None and dt[s] == 0.0:
                stop[s] = k
    assert all(v is not None for v in stop)
    total = max(stop)
    assert list(ns) == [min(v + 1, total) for v in stop]


def test_force_or_pe_predicate_stops_at_the_earlier_test():
    """FORCE_OR_PE (pred 2) on one system stops at min(step of max|F| < a alone, step of
    <pe> < b alone); the pe-only run is the reference predicate with a diff target of 0,
    which never fires.  Two systems: one below jamming (pe test first), one above (force
    test first)."""
    for phi, expect_pe_first in ((0.78, True), (0.88, False)):
        b = W.disk_batch([60], [phi], [0], seed=9)
        runs = {}
        for key, kw in (("f", dict(pred=1, ftol=1e-11)), ("pe", dict(pred=0, diff=0.0, tgt=1e-16)),
                        ("both", dict(pred=2, ftol=1e-11, tgt=1e-16))):
            o = orc.OracleDisk(b)
            steps, conv = o.fire_minimize(1e-2, 60000, mode=0, **kw)
            runs[key] = (steps, conv, o.get("x"))
        assert runs["both"][1]
        first = min(runs["f"][0] if runs["f"][1] else 10**9, runs["pe"][0] if runs["pe"][1] else 10**9)
        assert runs["both"][0] == first
        src = "pe" if expect_pe_first else "f"
        assert runs[src][0] == first
        assert np.array_equal(runs["both"][2], runs[src][2])


def test_retried_rounds_equal_loop_over_primitives():
    """Round retries (a round that hit the step cap runs again, FIRE restarted, no
    compression, every system rebuilt at the round boundary; after `retries` retries the
    system is dropped if no compression was applied yet, else decided on the state reached):
    the driver against the same rule written out over the primitives, with a short cap so
    that rounds do hit it, also after compressions, bit for bit."""
    cap, retries = 1500, 2
    b = small_batch()
    o = orc.OracleDisk(b)
    st, nc, ns, work = o.jam_compress(1e-2, DPHI, PT, CMAX, cap, pred=1, a=FTOL, mode=2, retries=retries)
    q = orc.OracleDisk(small_batch())
    S = q.S
    state = np.zeros(S, int)
    ncomp = np.zeros(S, int)
    nret = np.zeros(S, int)
    n_retried = n_capped_decided = 0
    while True:
        active = state == 0
        if not active.any():
            break
        dts = np.where(active, 1e-2, 0.0)
        q.fire_minimize(dts, cap, pred=1, ftol=FTOL, mode=2)
        fdt = q.fire_state()[0]
        P = q.pressure()
        inc = np.zeros(S)
        retry = False
        for s in np.flatnonzero(active):
            if fdt[s] != 0.0 and nret[s] < retries:
                nret[s] += 1; retry = True; n_retried += 1
            elif fdt[s] != 0.0 and ncomp[s] == 0:
                state[s] = -1
            else:
                n_capped_decided += int(fdt[s] != 0.0)
                if P[s] >= PT:
                    state[s] = 1
                elif ncomp[s] >= CMAX:
                    state[s] = 2
                else:
                    inc[s] = DPHI
                    ncomp[s] += 1
        if not inc.any() and not retry:
            break
        q.increment_packing_fraction(inc)
    assert n_retried > 0, "the cap should make some round run again"
    assert n_capped_decided > 0, "some capped round after a compression should be decided on its state"
    print(f"retried rounds {n_retried}, capped rounds decided after a compression {n_capped_decided}")
    assert np.array_equal(np.where(state == 2, 0, state), st)
    assert np.array_equal(ncomp, nc)
    for f in ("x", "y", "fx", "fy"):
        assert np.array_equal(q.get(f), o.get(f)), f
    # a retry gives a capped round more steps: no system is dropped with fewer retries used
    o0 = orc.OracleDisk(small_batch())
    st0, *_ = o0.jam_compress(1e-2, DPHI, PT, CMAX, cap, pred=1, a=FTOL, mode=2, retries=0)
    assert np.sum(st == -1) <= np.sum(st0 == -1)
