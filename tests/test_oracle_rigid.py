"""Pins the rigid bumpy-polygon restatement of the oracle (CPU): the R0 recipe closed forms,
own sincos against libm, rigid rotation, r x F torque, vertex neighbor exclusion vs brute
force, Newton's third law and momentum/rotation conservation."""
import math

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import workloads as W


def test_own_sincos_matches_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-1e-3, 1e-3, 3000), rng.uniform(-4, 4, 3000), rng.uniform(-60, 60, 3000),
                         [0.0, math.pi / 4, math.pi / 2, math.pi, -math.pi / 2, 1e-300]])
    for x in xs:
        s, c = orc.sincos(x)
        assert abs(s - math.sin(x)) <= 2 * math.ulp(max(abs(math.sin(x)), 1e-300)) or abs(s - math.sin(x)) < 1e-16
        assert abs(c - math.cos(x)) <= 2 * math.ulp(max(abs(math.cos(x)), 1e-300)) or abs(c - math.cos(x)) < 1e-16
        assert s * s + c * c == pytest.approx(1.0, abs=4e-16)


def test_recipe_closed_forms():
    b = W.rigid_batch(S=2, n_per=16, seed=2)
    assert b.nvpp.min() >= 12 and b.nvpp.max() <= 26
    off = np.concatenate([[0], np.cumsum(b.nvpp)])
    for i in range(b.N):
        vx, vy = b.vx[off[i]:off[i + 1]], b.vy[off[i]:off[i + 1]]
        n = b.nvpp[i]
        R = math.hypot(vx[0] - b.x[i], vy[0] - b.y[i])
        assert R in (pytest.approx(0.5), pytest.approx(0.7))
        spacing = math.hypot(vx[1] - vx[0], vy[1] - vy[0])
        assert b.vrad[off[i]] == pytest.approx(0.5 * spacing, rel=1e-12)           # bump = half spacing
        m_v = b.vmass[off[i]:off[i + 1]]
        assert m_v.sum() == pytest.approx(b.mass[i], rel=1e-14)
        I = np.sum(m_v * ((vx - b.x[i]) ** 2 + (vy - b.y[i]) ** 2))
        assert b.inertia[i] == pytest.approx(I, rel=1e-12)                          # I = m R^2
        # union area: polygon + (pi + 2pi/n) r_b^2 / 2 per vertex; independent Monte-Carlo check
    i = 0
    n, R, rb = b.nvpp[i], 0.5 if b.inertia[i] < 0.3 else 0.7, b.vrad[0]
    rng = np.random.default_rng(1)
    pts = rng.uniform(-R - rb, R + rb, size=(400000, 2))
    vx, vy = b.vx[:n] - b.x[0], b.vy[:n] - b.y[0]
    inside = np.zeros(len(pts), bool)
    for k in range(n):
        inside |= (pts[:, 0] - vx[k]) ** 2 + (pts[:, 1] - vy[k]) ** 2 < rb * rb
    ang = np.arctan2(vy, vx)
    # polygon membership: inside all edges (convex, counter-clockwise)
    inpoly = np.ones(len(pts), bool)
    for k in range(n):
        x1, y1, x2, y2 = vx[k], vy[k], vx[(k + 1) % n], vy[(k + 1) % n]
        inpoly &= (x2 - x1) * (pts[:, 1] - y1) - (y2 - y1) * (pts[:, 0] - x1) >= 0
    mc = (inside | inpoly).mean() * (2 * (R + rb)) ** 2
    assert b.area[0] == pytest.approx(mc, rel=0.01)


def one_vertex_pair(d, r=0.3, L=20.0, boundary=0):
    """Two particles with two vertices each; vertex 0 of particle 0 touches vertex 0 of particle 1."""
    b = W.RigidBatch(
        off=np.array([0, 2], np.int32), nvpp=np.array([2, 2], np.int32),
        x=np.array([5.0, 5.0 + 1.0 + d]), y=np.array([5.0, 5.0]), angle=np.zeros(2), mass=np.ones(2),
        inertia=np.array([0.5, 0.5]), area=np.ones(2),
        vx=np.array([5.5, 4.5, 5.5 + d, 6.5 + d]), vy=np.array([5.0, 5.0, 5.0, 5.0]),
        vrad=np.full(4, r), vmass=np.full(4, 0.5), Lx=np.array([L]), Ly=np.array([L]), e=np.array([2.0]),
        cutoff=np.array([0.2]), thresh2=np.array([0.01]), boundary=np.array([boundary], np.int32),
        cdx=np.array([0], np.int32), cdy=np.array([0], np.int32))
    return b


def test_vertex_contact_force_and_torque_closed_form():
    d, r, k = 0.5, 0.3, 2.0
    b = one_vertex_pair(d, r)
    # rotate the contact off the line of centres: move vertex 0 of particle 0 up by h
    h = 0.1
    b.vy[0] += h
    o = orc.OracleRigid(b)
    o.compute_forces(0)
    dx, dy = (5.5) - (5.5 + d), h
    dist = math.hypot(dx, dy)
    delta = 2 * r - dist
    fx0, fy0 = k * delta * dx / dist, k * delta * dy / dist
    fx, fy, tq = o.get("fx"), o.get("fy"), o.get("torque")
    np.testing.assert_allclose([fx[0], fy[0]], [fx0, fy0], rtol=1e-13)
    np.testing.assert_allclose([fx[1], fy[1]], [-fx0, -fy0], rtol=1e-13)
    # torque on particle 0 = (r_v - R) x F_v with r_v - R = (0.5, h)
    np.testing.assert_allclose(tq[0], 0.5 * fy0 - h * fx0, rtol=1e-13)
    np.testing.assert_allclose(o.pe_total(), [0.5 * k * delta ** 2], rtol=1e-13)


def test_rigid_rotation_is_rigid_and_exact_angle():
    b = W.rigid_batch(S=1, n_per=8, seed=4, phi=0.3)
    o = orc.OracleRigid(b)
    w = np.linspace(-2.0, 2.0, b.N)
    o.set("w", w)
    before = o.get("vpx"), o.get("vpy")
    si = o.static_index()
    o.update_positions(0.05)
    vx, vy = o.get("vpx")[si], o.get("vpy")[si]
    bx, by = before[0][si], before[1][si]
    off = np.concatenate([[0], np.cumsum(b.nvpp)])
    for i in range(b.N):
        sl = slice(off[i], off[i + 1])
        a = w[i] * 0.05
        ex = b.x[i] + math.cos(a) * (bx[sl] - b.x[i]) - math.sin(a) * (by[sl] - b.y[i])
        ey = b.y[i] + math.sin(a) * (bx[sl] - b.x[i]) + math.cos(a) * (by[sl] - b.y[i])
        np.testing.assert_allclose(vx[sl], ex, atol=2e-15)
        np.testing.assert_allclose(vy[sl], ey, atol=2e-15)
    np.testing.assert_allclose(o.get("angle"), b.angle + w * 0.05, rtol=1e-15)


def test_vertex_lists_exclude_own_particle_and_match_brute_force():
    b = W.rigid_batch(S=2, n_per=40, seed=6, boundary=0)
    o = orc.OracleRigid(b)
    st, ids = o.neighbors_csr()
    si = o.static_index()
    inv = np.empty_like(si); inv[si] = np.arange(len(si))
    pid = np.repeat(np.arange(b.N), b.nvpp)            # canonical vertex -> particle
    vx, vy, vr = o.get("vpx"), o.get("vpy"), o.get("vrad")
    voff = np.concatenate([[0], np.cumsum(b.nvpp)])[b.off]
    for cur in range(0, b.Nv, 11):
        c = inv[cur]
        s = 0 if c < voff[1] else 1
        L = b.Lx[s]
        got = sorted(inv[ids[st[cur]:st[cur + 1]]])
        assert all(pid[g] != pid[c] for g in got)
        brute = []
        for cc in range(voff[s], voff[s + 1]):
            if pid[cc] == pid[c]:
                continue
            j = si[cc]
            dx = vx[cur] - vx[j]; dy = vy[cur] - vy[j]
            dx -= L * np.rint(dx * (1.0 / L)); dy -= L * np.rint(dy * (1.0 / L))
            cut = vr[cur] + vr[j] + b.cutoff[s]
            if dx * dx + dy * dy < cut * cut:
                brute.append(cc)
        assert got == sorted(brute)


def test_newton_and_momentum_conservation_pbc():
    b = W.rigid_batch(S=1, n_per=64, seed=7, boundary=0, phi=0.75)
    o = orc.OracleRigid(b)
    o.compute_forces(0)
    fx, fy = o.get("fx"), o.get("fy")
    assert abs(fx.sum()) < 1e-12 * np.abs(fx).sum() and abs(fy.sum()) < 1e-12 * np.abs(fy).sum()
    rng = np.random.default_rng(2)
    o.set("vx", rng.standard_normal(b.N) * 0.01); o.set("vy", rng.standard_normal(b.N) * 0.01)
    p0 = np.array([o.get("vx").sum(), o.get("vy").sum()])
    o.vv_init(1e-3)
    o.vv_run(200)
    p1 = np.array([o.get("vx").sum(), o.get("vy").sum()])
    np.testing.assert_allclose(p1, p0, atol=1e-12)


def test_rigid_fire_reaches_minimum_in_walls():
    b = W.rigid_batch(S=2, n_per=32, seed=8)
    o = orc.OracleRigid(b)
    o.compute_forces(1)
    pe0 = o.pe_total()
    steps, conv = o.fire_minimize(1e-2, 20000, pred=1, ftol=1e-9, mode=1)
    assert conv and np.all(o.pe_total() < pe0)
    assert np.all(np.sqrt(o.max_force2()) < 1e-9)
