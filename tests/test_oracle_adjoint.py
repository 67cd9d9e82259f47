"""Pins the FP64 host reference of the adjoints (oracle/adjoint_ref.py) on the CPU.

- its energy, forces and pressure equal the C++ oracle's (the restatement of the reference);
- its Hessian, and every first derivative of energy, pressure and forces with respect to
  positions, radii and box, match central finite differences of the C++ oracle;
- the implicit derivative of the pressure at a minimum matches finite differences through
  re-minimisation (closed form for a three-disk ring, numerically for small packings);
- reverse mode through velocity Verlet matches finite differences of the trajectory.
"""
import numpy as np
import pytest

import adjoint_ref as R
import oracle as orc
from dpmd_amd import workloads as W


def one_system(n, phi, seed, wall):
    b = W.disk_batch([n], [phi], [int(wall)], seed=seed)
    return b


def oracle_eval(b, vx=None, vy=None):
    """Canonical forces, U and P of a one-system batch from the C++ oracle."""
    o = orc.OracleDisk(b)
    mode = int(b.boundary[0])
    o.compute_forces(mode)
    si = o.static_index()
    if vx is not None:
        cur_vx, cur_vy = np.empty(b.N), np.empty(b.N)
        cur_vx[si], cur_vy[si] = vx, vy
        o.set("vx", cur_vx); o.set("vy", cur_vy)
    return o.get("fx")[si], o.get("fy")[si], o.pe_total()[0], o.pressure()[0]


def with_(b, **kw):
    c = W.systems_of(b, 0, 1)
    for k, v in kw.items():
        setattr(c, k, np.asarray(v, np.float64).copy())
    return c


def args(b, vx, vy):
    return b.x, b.y, b.rad, b.mass, vx, vy, b.Lx[0], b.Ly[0], b.e[0], bool(b.boundary[0])


@pytest.fixture(params=[False, True], ids=["pbc", "wall"])
def sysb(request):
    b = one_system(24, 0.92, 7, request.param)
    rng = np.random.default_rng(3)
    return b, rng.standard_normal(b.N) * 0.1, rng.standard_normal(b.N) * 0.1


def test_forward_equals_oracle(sysb):
    b, vx, vy = sysb
    fx, fy, U, P = oracle_eval(b, vx, vy)
    x, y, r, m, _, _, Lx, Ly, k, wall = args(b, vx, vy)
    rfx, rfy = R.forces(x, y, r, Lx, Ly, k, wall)
    np.testing.assert_allclose(rfx, fx, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(rfy, fy, rtol=1e-12, atol=1e-14)
    assert R.energy(x, y, r, Lx, Ly, k, wall) == pytest.approx(U, rel=1e-12)
    assert R.pressure(x, y, r, m, vx, vy, Lx, Ly, k, wall) == pytest.approx(P, rel=1e-12)


def test_hessian_and_gradients_match_fd_of_oracle(sysb):
    b, vx, vy = sysb
    g = R.grads(*args(b, vx, vy))
    x, y, r, m, _, _, Lx, Ly, k, wall = args(b, vx, vy)
    H = R.hessian(x, y, r, Lx, Ly, k, wall)
    h = 1e-6
    rng = np.random.default_rng(5)
    # Hessian-vector product: -(F(x + h w) - F(x - h w)) / 2h
    w = rng.standard_normal(2 * b.N)
    fp = oracle_eval(with_(b, x=x + h * w[0::2], y=y + h * w[1::2]))
    fm = oracle_eval(with_(b, x=x - h * w[0::2], y=y - h * w[1::2]))
    fd = np.empty(2 * b.N)
    fd[0::2] = -(fp[0] - fm[0]) / (2 * h)
    fd[1::2] = -(fp[1] - fm[1]) / (2 * h)
    np.testing.assert_allclose(H @ w, fd, rtol=1e-6, atol=1e-7)
    # radii: dU/dr, dP/dr, dF/dr (a few components)
    for i in (0, 5, 17):
        e = np.zeros(b.N); e[i] = h
        p = oracle_eval(with_(b, rad=r + e), vx, vy)
        q = oracle_eval(with_(b, rad=r - e), vx, vy)
        assert (p[2] - q[2]) / (2 * h) == pytest.approx(g["dUdr"][i], rel=1e-6, abs=1e-9)
        assert (p[3] - q[3]) / (2 * h) == pytest.approx(g["dPdr"][i], rel=1e-6, abs=1e-11)
        np.testing.assert_allclose((p[0] - q[0]) / (2 * h), g["dFdr"][0::2, i], rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose((p[1] - q[1]) / (2 * h), g["dFdr"][1::2, i], rtol=1e-6, atol=1e-8)
    # positions: dP/dx
    for i in (2, 11):
        e = np.zeros(b.N); e[i] = h
        p = oracle_eval(with_(b, x=x + e), vx, vy)
        q = oracle_eval(with_(b, x=x - e), vx, vy)
        assert (p[3] - q[3]) / (2 * h) == pytest.approx(g["dPdx"][2 * i], rel=1e-6, abs=1e-11)
    # box at fixed reduced coordinates
    for ax in (0, 1):
        sc = np.array([1.0, 1.0]); sc[ax] += h / (Lx if ax == 0 else Ly)
        def scaled(f):
            s = np.array([1.0, 1.0]) + f * (sc - 1.0)
            return oracle_eval(with_(b, x=x * s[0], y=y * s[1], Lx=[Lx * s[0]], Ly=[Ly * s[1]]), vx, vy)
        p, q = scaled(1.0), scaled(-1.0)
        assert (p[2] - q[2]) / (2 * h) == pytest.approx(g["dUdL"][ax], rel=1e-6, abs=1e-9)
        assert (p[3] - q[3]) / (2 * h) == pytest.approx(g["dPdL"][ax], rel=1e-6, abs=1e-11)
        np.testing.assert_allclose((p[0] - q[0]) / (2 * h), g["dFdL"][0::2, ax], rtol=1e-6, atol=1e-8)
        np.testing.assert_allclose((p[1] - q[1]) / (2 * h), g["dFdL"][1::2, ax], rtol=1e-6, atol=1e-8)


def test_ring_implicit_pressure_closed_form():
    """Three disks on a line in a periodic box (Lx = 3, Ly = 10; no contacts in y) form a ring
    of three springs with rest lengths l = r_i + r_j.  At equilibrium the tension is uniform,
    T = k (2 sum r - Lx) / 3, and the virial W = sum k delta d = T Lx, so
    P = k (2 sum r - Lx) / (6 Ly):  dP/dr_m = k / (3 Ly) for every disk (relaxation included;
    the unrelaxed partial k (d_left + d_right) / (2 Lx Ly) differs for the middle disk),
    dP/dLx = -k / (6 Ly) and dP/dLy = -P / Ly.  Every disk has two contacts, so the contact
    mask keeps them all (a rule that drops particles with fewer than three contacts, the 2-D
    rigid backbone, would lose the whole ring)."""
    Lx, Ly, k = 3.0, 10.0, 1.0
    r = np.array([0.55, 0.6, 0.5])
    x = np.array([0.5, 1.55, 2.55]); y = np.full(3, 5.0); m = np.ones(3); z = np.zeros(3)
    fx, fy = R.forces(x, y, r, Lx, Ly, k, False)
    assert np.max(np.abs(fx)) < 1e-14 and np.max(np.abs(fy)) == 0.0
    P = R.pressure(x, y, r, m, z, z, Lx, Ly, k, False)
    assert P == pytest.approx(k * (2 * r.sum() - Lx) / (6 * Ly), rel=1e-13)
    assert R.touching(x, y, r, Lx, Ly, False).all()
    dr, dL, lam = R.implicit_pressure_grad(x, y, r, m, z, z, Lx, Ly, k, False)
    np.testing.assert_allclose(dr, [k / (3 * Ly)] * 3, rtol=1e-12)
    np.testing.assert_allclose(dL, [-k / (6 * Ly), -P / Ly], rtol=1e-12)
    g = R.grads(x, y, r, m, z, z, Lx, Ly, k, False)
    assert abs(g["dPdr"][1] - dr[1]) > 1e-4          # relaxation matters for the middle disk
    h = 1e-6
    Pm = lambda rr: R.pressure(*R.newton_minimum(x, y, rr, Lx, Ly, k, False), rr, m, z, z, Lx, Ly, k, False)
    e = np.array([0.0, h, 0.0])
    assert (Pm(r + e) - Pm(r - e)) / (2 * h) == pytest.approx(dr[1], rel=1e-8)


@pytest.mark.parametrize("wall", [False, True], ids=["pbc", "wall"])
def test_implicit_pressure_matches_fd_through_reminimisation(wall):
    b = one_system(32, 0.90, 11, wall)
    o = orc.OracleDisk(b)
    o.fire_minimize(1e-2, 200000, pred=1, ftol=1e-12, mode=int(wall))
    si = o.static_index()
    x0, y0 = o.get("x")[si], o.get("y")[si]
    r, m, Lx, Ly, k = b.rad, b.mass, b.Lx[0], b.Ly[0], b.e[0]
    x, y = R.newton_minimum(x0, y0, r, Lx, Ly, k, wall)
    z = np.zeros(b.N)
    dr, dL, _ = R.implicit_pressure_grad(x, y, r, m, z, z, Lx, Ly, k, wall)
    h = 1e-6

    def P(rr=r, LL=(Lx, Ly)):
        sx, sy = LL[0] / Lx, LL[1] / Ly
        xm, ym = R.newton_minimum(x * sx, y * sy, rr, LL[0], LL[1], k, wall)
        return R.pressure(xm, ym, rr, m, z, z, LL[0], LL[1], k, wall)

    P0 = P()
    assert P0 > 1e-6
    checked = 0
    for i in np.argsort(-np.abs(dr))[:4]:
        e = np.zeros(b.N); e[i] = h
        fd = (P(r + e) - P(r - e)) / (2 * h)
        assert fd == pytest.approx(dr[i], rel=1e-6), i
        checked += 1
    for ax in (0, 1):
        LL = np.array([Lx, Ly]); d = np.zeros(2); d[ax] = h
        fd = (P(LL=LL + d) - P(LL=LL - d)) / (2 * h)
        assert fd == pytest.approx(dL[ax], rel=1e-6)
    assert checked == 4


def test_implicit_pressure_through_struts_matches_fd():
    """A packing of the differentiation workload (system 11 of the 128 x 1024 batch at
    phi = 0.86, FIRE to max|F| < 1e-12 on the oracle) holds struts: particles between two
    contacts that carry almost no force.  In the linear response such a contact is a spring of
    stiffness k, so dP/dr_i goes through them.  The implicit derivative over the particles
    with contacts matches central differences through re-minimisation (h = 1e-6, 1e-6
    relative) where the contact network (contacts of overlap above 1e-10) is unchanged at +-h;
    dropping every particle with
    fewer than three contacts (the rigid backbone) misses the struts and is 2 % off here."""
    b = W.system_of(W.disk_batch([1024] * 128, [0.86] * 128, [0] * 128, seed=1), 11)
    o = orc.OracleDisk(b)
    assert o.fire_minimize(1e-2, 300000, pred=1, ftol=1e-12)[1]
    si = o.static_index()
    x1, y1 = o.get("x")[si], o.get("y")[si]
    r, m, Lx, Ly, k = b.rad, b.mass, b.Lx[0], b.Ly[0], b.e[0]
    z = np.zeros(b.N)
    i, h = 960, 1e-6

    def minimum(rr):   # FIRE to max|F| < 1e-13 from the packing
        q = W.system_of(b, 0)
        q.rad, q.x, q.y = rr.copy(), x1.copy(), y1.copy()
        oq = orc.OracleDisk(q)
        oq.fire_minimize(1e-2, 300000, pred=1, ftol=1e-13)
        s2 = oq.static_index()
        xq, yq = oq.get("x")[s2], oq.get("y")[s2]
        I, J, _, _, _, de = R.contacts(xq, yq, rr, Lx, Ly, False)
        firm = de > 1e-10   # contacts that do not open within the minimiser's tolerance
        return xq, yq, R.pressure(xq, yq, rr, m, z, z, Lx, Ly, k, False), set(zip(I[firm].tolist(), J[firm].tolist()))

    x, y, _, c0 = minimum(r)
    dr, _, _ = R.implicit_pressure_grad(x, y, r, m, z, z, Lx, Ly, k, False)
    e = np.zeros(b.N); e[i] = h
    (_, _, pp, cp), (_, _, pm, cm) = minimum(r + e), minimum(r - e)
    assert cp == cm == c0
    fd = (pp - pm) / (2 * h)
    assert fd == pytest.approx(dr[i], rel=1e-6)
    I0, J0, *_ = R.contacts(x, y, r, Lx, Ly, False)
    cnt = np.bincount(I0, minlength=b.N) + np.bincount(J0, minlength=b.N)
    live3 = np.ones(b.N, bool)
    while True:   # the rigid backbone (3-core), for contrast
        c3 = np.bincount(I0[live3[J0]], minlength=b.N) + np.bincount(J0[live3[I0]], minlength=b.N)
        nxt = live3 & (c3 >= 3)
        if np.array_equal(nxt, live3):
            break
        live3 = nxt
    assert (cnt > 0).sum() > live3.sum()
    dr3, _, _ = R.implicit_pressure_grad(x, y, r, m, z, z, Lx, Ly, k, False, live=live3)
    assert abs(dr3[i] - fd) > 1e-2 * abs(fd)


def test_vv_backprop_matches_fd():
    b = one_system(16, 0.88, 13, False)
    rng = np.random.default_rng(8)
    vx0, vy0 = rng.standard_normal(b.N) * 0.05, rng.standard_normal(b.N) * 0.05
    x, y, r, m, _, _, Lx, Ly, k, wall = args(b, vx0, vy0)
    dt, n = 1e-2, 150
    # numpy trajectory equals the oracle's velocity Verlet
    xs, (xT, yT, vxT, vyT) = R.vv_trajectory(x, y, vx0, vy0, r, m, Lx, Ly, k, wall, dt, n)
    o = orc.OracleDisk(b)
    si = o.static_index()
    cvx, cvy = np.empty(b.N), np.empty(b.N)
    cvx[si], cvy[si] = vx0, vy0
    o.set("vx", cvx); o.set("vy", cvy)
    o.vv_init(dt)
    o.vv_run(n)
    si = o.static_index()
    np.testing.assert_allclose(o.get("x")[si], xT, rtol=0, atol=1e-12)
    np.testing.assert_allclose(o.get("vx")[si], vxT, rtol=0, atol=1e-12)

    def msd(vx, vy):
        _, (a, c, _, _) = R.vv_trajectory(x, y, vx, vy, r, m, Lx, Ly, k, wall, dt, n)
        return np.mean((a - x) ** 2 + (c - y) ** 2)

    gx, gy = 2 * (xT - x) / b.N, 2 * (yT - y) / b.N
    _, _, gvx, gvy = R.vv_backprop(xs, gx, gy, np.zeros(b.N), np.zeros(b.N), r, m, Lx, Ly, k, wall, dt)
    h = 1e-6
    for i in (0, 7, 15):
        e = np.zeros(b.N); e[i] = h
        assert (msd(vx0 + e, vy0) - msd(vx0 - e, vy0)) / (2 * h) == pytest.approx(gvx[i], rel=1e-6, abs=1e-12)
        assert (msd(vx0, vy0 + e) - msd(vx0, vy0 - e)) / (2 * h) == pytest.approx(gvy[i], rel=1e-6, abs=1e-12)


@pytest.mark.parametrize("wall", [False, True])
def test_sparse_forms_equal_dense(wall):
    """The k-d-tree forms used at full size: the same contact set and order, forces, Hessian
    products, trajectory and reverse mode as the dense restatement (pinned above)."""
    b = one_system(90, 0.88, 4, wall)
    x, y, r, m = b.x, b.y, b.rad, b.mass
    L_ = b.Lx[0]
    D = R.contacts(x, y, r, L_, L_, wall)
    Sp = R.contacts_sparse(x, y, r, L_, L_, wall)
    for a, c in zip(D, Sp):
        assert np.array_equal(a, c)
    fd, fs = R.forces(x, y, r, L_, L_, 1.0, wall), R.forces_sparse(x, y, r, L_, L_, 1.0, wall)
    assert np.array_equal(fd[0], fs[0]) and np.array_equal(fd[1], fs[1])
    rng = np.random.default_rng(3)
    w = rng.standard_normal(2 * b.N)
    H = R.hessian(x, y, r, L_, L_, 1.0, wall)
    np.testing.assert_allclose(R.hvp_sparse(x, y, r, L_, L_, 1.0, wall, w), H @ w, rtol=0,
                               atol=1e-13 * np.max(np.abs(H @ w)))
    vx, vy = rng.standard_normal(b.N) * 0.05, rng.standard_normal(b.N) * 0.05
    xs_d, fin_d = R.vv_trajectory(x, y, vx, vy, r, m, L_, L_, 1.0, wall, 1e-2, 30)
    xs_s, fin_s = R.vv_trajectory_sparse(x, y, vx, vy, r, m, L_, L_, 1.0, wall, 1e-2, 30)
    for a, c in zip(fin_d, fin_s):
        assert np.array_equal(a, c)
    gx, gy = rng.standard_normal(b.N), rng.standard_normal(b.N)
    z = np.zeros(b.N)
    bd = R.vv_backprop(xs_d, gx, gy, z, z, r, m, L_, L_, 1.0, wall, 1e-2)
    bs = R.vv_backprop_sparse(xs_s, gx, gy, z, z, r, m, L_, L_, 1.0, wall, 1e-2)
    for a, c in zip(bd, bs):
        np.testing.assert_allclose(c, a, rtol=0, atol=1e-12 * np.max(np.abs(a)))
