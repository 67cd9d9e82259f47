"""Pins the oracle's disk contact observables (CPU): contact counts and overlaps against a
brute-force O(N^2) minimum-image scan; pair distances against numpy; pair forces summing to
the particle forces bit for bit; the assembled Hessian against central finite differences of
the forces and against its own symmetry."""
import os

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import workloads as W


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    orc.lib().or_set_threads(1)
    yield
    orc.lib().or_set_threads(os.cpu_count() or 1)


@pytest.fixture(scope="module")
def packed():
    b = W.disk_batch([40, 56], [0.90, 0.88], [0, 0], seed=12)
    o = orc.OracleDisk(b)
    o.fire_minimize(1e-2, 3000, pred=1, ftol=1e-12)
    o.compute_forces(0)
    return b, o


def canonical(o, f):
    return o.get(f)[o.static_index()]


def min_image(d, L):
    return d - L * np.round(d / L)


def test_contacts_and_overlaps_brute_force(packed):
    b, o = packed
    c, ov, cc = o.contacts()
    si = o.static_index()
    x, y, r = canonical(o, "x"), canonical(o, "y"), canonical(o, "rad")
    for s in range(b.S):
        lo, hi = b.off[s], b.off[s + 1]
        L = b.Lx[s]
        dx = min_image(x[lo:hi, None] - x[None, lo:hi], L)
        dy = min_image(y[lo:hi, None] - y[None, lo:hi], L)
        d = np.hypot(dx, dy)
        rs = r[lo:hi, None] + r[None, lo:hi]
        np.fill_diagonal(d, np.inf)
        touch = d < rs
        assert np.array_equal(c[si[lo:hi]], touch.sum(1))
        np.testing.assert_allclose(ov[si[lo:hi]], np.where(touch, rs - d, 0.0).sum(1), rtol=1e-12, atol=1e-15)
    assert np.array_equal(cc, c.astype(float))
    assert c.sum() > 2 * b.N          # a jammed packing: well above two contacts per disk


def test_pair_distances_and_forces(packed):
    b, o = packed
    po = o.pair_observables()
    x, y = canonical(o, "x"), canonical(o, "y")
    i, j = po["pair_i"], po["pair_j"]
    sid = np.repeat(np.arange(b.S), np.diff(b.off))
    L = b.Lx[sid[i]]
    d = np.hypot(min_image(x[i] - x[j], L), min_image(y[i] - y[j], L))
    np.testing.assert_allclose(po["dist"], d, rtol=1e-14)
    # the pair forces of particle i, added in list order, are the force kernel's sum
    start, _ = o.neighbors_csr()
    fx, fy = o.get("fx"), o.get("fy")
    for q in range(b.N):
        sx, sy = 0.0, 0.0
        for k in range(start[q], start[q + 1]):
            sx += po["fx"][k]; sy += po["fy"][k]
        assert sx == fx[q] and sy == fy[q]
    # Newton's third law between the two directed entries of each contact
    key = {(a, c): k for k, (a, c) in enumerate(zip(i, j))}
    for k, (a, c) in enumerate(zip(i, j)):
        kk = key[(c, a)]
        assert po["fx"][k] == -po["fx"][kk] and po["fy"][k] == -po["fy"][kk]


def assemble(po, N):
    H = np.zeros((2 * N, 2 * N))
    cols = {"h_ix_ix": (0, 0, 0), "h_ix_jx": (0, 0, 1), "h_ix_iy": (0, 1, 0), "h_ix_jy": (0, 1, 1),
            "h_iy_ix": (1, 0, 0), "h_iy_jx": (1, 0, 1), "h_iy_iy": (1, 1, 0), "h_iy_jy": (1, 1, 1)}
    for name, (a, bb, other) in cols.items():
        rows = 2 * po["pair_i"] + a
        cj = 2 * (po["pair_j"] if other else po["pair_i"]) + bb
        np.add.at(H, (rows, cj), po[name])
    return H


def test_hessian_matches_finite_differences(packed):
    b, o = packed
    N = b.N
    H = assemble(o.pair_observables(), N)
    np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-12 * np.abs(H).max())
    si = o.static_index()
    x0, y0 = o.get("x").copy(), o.get("y").copy()
    h = 1e-6
    J = np.zeros((2 * N, 2 * N))
    for c in range(N):
        for a, (name, base) in enumerate((("x", x0), ("y", y0))):
            cols = []
            for sgn in (1.0, -1.0):
                v = base.copy()
                v[si[c]] += sgn * h
                o.set(name, v)
                o.compute_forces(0)
                cols.append(np.stack([o.get("fx")[si], o.get("fy")[si]], 1).reshape(-1))
            o.set(name, base)
            J[:, 2 * c + a] = -(cols[0] - cols[1]) / (2 * h)
    o.compute_forces(0)
    np.testing.assert_allclose(H, J, rtol=0, atol=1e-6 * np.abs(H).max())
    # positive semi-definite up to the translation modes at a FIRE minimum (small negative
    # eigenvalues from the -k delta / d tangential terms are allowed)
    ev = np.linalg.eigvalsh(H)
    assert ev.min() > -0.05 * ev.max()
