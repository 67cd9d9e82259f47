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


# ---- rigid bumpy polygons -------------------------------------------------------------
@pytest.fixture(scope="module")
def rigid_packed():
    rb = W.rigid_batch(S=2, n_per=24, seed=8, boundary=0, phi=0.78)
    o = orc.OracleRigid(rb)
    o.fire_minimize(1e-2, 300, pred=1, ftol=1e-9)
    o.check_neighbors(True)      # lists exact at the current positions (Verlet lists are supersets)
    o.compute_forces(0)
    return rb, o


def rigid_brute(rb, o):
    """Vertex positions (canonical), particle ids, pair sets from an O(Nv^2) scan."""
    si = o.static_index()
    vx, vy = o.get("vpx")[si], o.get("vpy")[si]
    vr = rb.vrad
    pid = np.repeat(np.arange(rb.N), rb.nvpp)
    sid_p = np.repeat(np.arange(rb.S), np.diff(rb.off))
    vsid = sid_p[pid]
    near, touch = set(), {}
    for s in range(rb.S):
        idx = np.nonzero(vsid == s)[0]
        L = rb.Lx[s]
        dx = min_image(vx[idx, None] - vx[None, idx], L)
        dy = min_image(vy[idx, None] - vy[None, idx], L)
        d2 = dx * dx + dy * dy
        rs = vr[idx, None] + vr[None, idx]
        same = pid[idx, None] == pid[None, idx]
        nb = (d2 < (rs + rb.cutoff[s]) ** 2) & ~same
        ov = (d2 < rs * rs) & ~same
        for a, c in zip(*np.nonzero(nb)):
            near.add((pid[idx[a]], pid[idx[c]]))
        for a, c in zip(*np.nonzero(ov)):
            touch.setdefault((pid[idx[a]], pid[idx[c]]), set()).add(idx[a])
    return near, touch


def test_rigid_particle_neighbors_contacts_brute_force(rigid_packed):
    rb, o = rigid_packed
    st, ids = o.particle_neighbors()
    pairs = {(i, int(j)) for i in range(rb.N) for j in ids[st[i]:st[i + 1]]}
    near, touch = rigid_brute(rb, o)
    assert pairs == near
    for i in range(rb.N):
        assert np.all(np.diff(ids[st[i]:st[i + 1]]) > 0)          # ascending, no repeats
    po = o.pair_observables()
    for k, (i, j) in enumerate(zip(po["pair_i"], po["pair_j"])):
        assert po["vertex_contacts_i"][k] == len(touch.get((i, j), ()))
        assert po["vertex_contacts_j"][k] == len(touch.get((j, i), ()))
    want = np.zeros(rb.N, int)
    for (i, j) in touch:
        want[i] += 1
    assert np.array_equal(po["contacts"], want)
    assert want.sum() > rb.N                                       # a packing with contacts


def test_rigid_pair_forces_sum_to_particle_forces(rigid_packed):
    rb, o = rigid_packed
    po = o.pair_observables()
    fx = np.zeros(rb.N); fy = np.zeros(rb.N)
    np.add.at(fx, po["pair_i"], po["fx"]); np.add.at(fy, po["pair_i"], po["fy"])
    scale = np.abs(o.get("fx")).max()
    np.testing.assert_allclose(fx, o.get("fx"), rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(fy, o.get("fy"), rtol=0, atol=1e-12 * scale)
    key = {(a, c): k for k, (a, c) in enumerate(zip(po["pair_i"], po["pair_j"]))}
    for k, (a, c) in enumerate(zip(po["pair_i"], po["pair_j"])):
        kk = key[(c, a)]
        assert abs(po["fx"][k] + po["fx"][kk]) <= 1e-12 * scale and abs(po["fy"][k] + po["fy"][kk]) <= 1e-12 * scale
    # centre distances
    sid_p = np.repeat(np.arange(rb.S), np.diff(rb.off))
    x, y = o.get("x"), o.get("y")
    L = rb.Lx[sid_p[po["pair_i"]]]
    d = np.hypot(min_image(x[po["pair_i"]] - x[po["pair_j"]], L), min_image(y[po["pair_i"]] - y[po["pair_j"]], L))
    np.testing.assert_allclose(po["dist"], d, rtol=1e-14)


def test_rigid_overlaps_brute_force(rigid_packed):
    rb, o = rigid_packed
    ov, cc = o.overlaps()
    si = o.static_index()
    vx, vy = o.get("vpx"), o.get("vpy")
    pid = np.repeat(np.arange(rb.N), rb.nvpp)
    pid_cur = np.empty_like(pid); pid_cur[si] = pid
    vr = np.empty(rb.Nv); vr[si] = rb.vrad
    sid_p = np.repeat(np.arange(rb.S), np.diff(rb.off))
    for s in range(rb.S):
        idx = np.nonzero(sid_p[pid_cur] == s)[0]
        L = rb.Lx[s]
        d = np.hypot(min_image(vx[idx, None] - vx[None, idx], L), min_image(vy[idx, None] - vy[None, idx], L))
        rs = vr[idx, None] + vr[None, idx]
        t = (d < rs) & (pid_cur[idx, None] != pid_cur[None, idx])
        np.testing.assert_allclose(ov[idx], np.where(t, rs - d, 0.0).sum(1), rtol=1e-12, atol=1e-15)
        assert np.array_equal(cc[idx], t.sum(1).astype(float))


def test_rigid_friction_closed_form():
    """One vertex contact whose normal makes angle theta with the x axis: the friction ratio is
    |tan| of the angle between the contact force and the centre line."""
    r, delta, theta = 0.3, 0.05, 0.4
    a, b = (2 * r - delta) * np.cos(theta), (2 * r - delta) * np.sin(theta)
    x0, y0 = 5.0, 5.0
    v1x, v1y = x0 + 0.5 + a, y0 + b
    x1, y1 = v1x + 0.5, v1y
    rb = W.RigidBatch(
        off=np.array([0, 2], np.int32), nvpp=np.array([2, 2], np.int32),
        x=np.array([x0, x1]), y=np.array([y0, y1]), angle=np.zeros(2), mass=np.ones(2),
        inertia=np.array([0.25, 0.25]), area=np.ones(2),
        vx=np.array([x0 + 0.5, x0 - 0.5, v1x, x1 + 0.5]), vy=np.array([y0, y0, v1y, y1]),
        vrad=np.full(4, r), vmass=np.full(4, 0.5), Lx=np.array([20.0]), Ly=np.array([20.0]), e=np.array([1.0]),
        cutoff=np.array([0.2]), thresh2=np.array([0.01]), boundary=np.array([0], np.int32),
        cdx=np.array([0], np.int32), cdy=np.array([0], np.int32))
    o = orc.OracleRigid(rb)
    po = o.pair_observables()
    assert list(po["pair_i"]) == [0, 1] and list(po["vertex_contacts_i"]) == [1, 1]
    want = abs(np.tan(theta - np.arctan2(y1 - y0, x1 - x0)))
    np.testing.assert_allclose(po["friction"], [want, want], rtol=1e-12)
    np.testing.assert_allclose(np.hypot(po["fx"][0], po["fy"][0]), delta, rtol=1e-12)


def assemble_rigid(po, N):
    H = np.zeros((3 * N, 3 * N))
    q = "xyt"
    for a in range(3):
        for bb in range(3):
            rows = 3 * po["pair_i"] + a
            np.add.at(H, (rows, 3 * po["pair_i"] + bb), po[f"h_i{q[a]}_i{q[bb]}"])
            np.add.at(H, (rows, 3 * po["pair_j"] + bb), po[f"h_i{q[a]}_j{q[bb]}"])
    return H


def test_rigid_hessian_matches_finite_differences():
    """H over (x, y, theta) per body against central differences of the centre forces and
    torques (rigid moves of one body's vertices); periodic box, no walls."""
    rb = W.rigid_batch(S=1, n_per=16, seed=4, boundary=0, phi=0.82)
    o = orc.OracleRigid(rb)
    o.fire_minimize(1e-2, 40, pred=1, ftol=1e-9, mode=0)    # a few steps: contacts, not a minimum
    o.check_neighbors(True)
    o.compute_forces(0)
    N = rb.N
    H = assemble_rigid(o.hessian(), N)
    si = o.static_index()
    x0, y0, vx0, vy0 = o.get("x").copy(), o.get("y").copy(), o.get("vpx").copy(), o.get("vpy").copy()
    off = np.concatenate([[0], np.cumsum(rb.nvpp)])
    h = 1e-6
    J = np.zeros((3 * N, 3 * N))
    for c in range(N):
        vs = si[off[c]:off[c + 1]]
        for a in range(3):
            cols = []
            for sgn in (1.0, -1.0):
                x, y, vx, vy = x0.copy(), y0.copy(), vx0.copy(), vy0.copy()
                if a == 0:
                    x[c] += sgn * h; vx[vs] += sgn * h
                elif a == 1:
                    y[c] += sgn * h; vy[vs] += sgn * h
                else:
                    cs, sn = np.cos(sgn * h), np.sin(sgn * h)
                    dx, dy = vx0[vs] - x0[c], vy0[vs] - y0[c]
                    vx[vs] = x0[c] + cs * dx - sn * dy
                    vy[vs] = y0[c] + sn * dx + cs * dy
                o.set("x", x); o.set("y", y); o.set("vpx", vx); o.set("vpy", vy)
                o.compute_forces(0)
                cols.append(np.stack([o.get("fx"), o.get("fy"), o.get("torque")], 1).reshape(-1))
            J[:, 3 * c + a] = -(cols[0] - cols[1]) / (2 * h)
    o.set("x", x0); o.set("y", y0); o.set("vpx", vx0); o.set("vpy", vy0)
    o.compute_forces(0)
    assert np.abs(H).max() > 0.1
    np.testing.assert_allclose(H, J, rtol=0, atol=1e-6 * np.abs(H).max())
