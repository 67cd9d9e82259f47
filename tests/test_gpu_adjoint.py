"""GPU tests of the disk adjoints (HIP kernels through the C ABI) against the FP64 host
reference (oracle/adjoint_ref.py, pinned by finite differences of the oracle in
tests/test_oracle_adjoint.py) and against central finite differences (h = 1e-6) on the GPU.

Bars (written per check): Hessian-vector and vector-Jacobian products 1e-11 relative to the
largest entry (summation order differs from the dense host matrices); implicit pressure
gradients 1e-7 relative (CG to 1e-12 vs a dense least-squares solve); finite differences
1e-6 relative.
"""
import numpy as np
import pytest

import adjoint_ref as R
from dpmd_amd import _lib as L
from dpmd_amd import adjoint as A
from dpmd_amd import workloads as W
from dpmd_amd.system import Disk, VelocityVerlet, run_fire

pytestmark = pytest.mark.gpu
MODE = L.FORCE_PER_SYSTEM


def make(b):
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    return P


def canon(P, name):
    return P.get(name, L.ORDER_CANONICAL)


def per_system(b, P, vx, vy):
    x, y = canon(P, "POS_X"), canon(P, "POS_Y")
    for s in range(b.S):
        sl = slice(int(b.off[s]), int(b.off[s + 1]))
        yield s, sl, (x[sl], y[sl], b.rad[sl], b.mass[sl], vx[sl], vy[sl], b.Lx[s], b.Ly[s], b.e[s],
                      bool(b.boundary[s]))


def close(a, b, rel):
    scale = max(np.max(np.abs(b)), 1e-300)
    np.testing.assert_allclose(a, b, rtol=0, atol=rel * scale)


def test_hvp_and_vjp_match_host_reference():
    b = W.disk_batch([40, 60, 50], [0.92, 0.90, 0.93], [0, 1, 0], seed=21)
    P = make(b)
    print(f"adjoint hvp/vjp S={b.S} N={b.N}", flush=True)
    rng = np.random.default_rng(4)
    vx, vy = rng.standard_normal(b.N) * 0.1, rng.standard_normal(b.N) * 0.1
    A.set_canonical(P, "VEL_X", vx); A.set_canonical(P, "VEL_Y", vy)
    wx, wy = rng.standard_normal(b.N), rng.standard_normal(b.N)
    yx, yy = A.hvp(P, wx, wy, MODE)
    gU, gP = rng.standard_normal(b.S), rng.standard_normal(b.S)
    gfx, gfy = rng.standard_normal(b.N), rng.standard_normal(b.N)
    out = A.vjp(P, MODE, gU=gU, gP=gP, gF=(gfx, gfy))
    for s, sl, args in per_system(b, P, vx, vy):
        x, y, r, m, svx, svy, Lx, Ly, k, wall = args
        H = R.hessian(x, y, r, Lx, Ly, k, wall)
        w = np.empty(2 * len(x)); w[0::2], w[1::2] = wx[sl], wy[sl]
        ref = H @ w
        close(yx[sl], ref[0::2], 1e-11); close(yy[sl], ref[1::2], 1e-11)
        g = R.grads(*args)
        gf = np.empty(2 * len(x)); gf[0::2], gf[1::2] = gfx[sl], gfy[sl]
        dx = gU[s] * g["dUdx"] + gP[s] * g["dPdx"] - H @ gf
        dr = gU[s] * g["dUdr"] + gP[s] * g["dPdr"] + gf @ g["dFdr"]
        dL = gU[s] * g["dUdL"] + gP[s] * g["dPdL"] + gf @ g["dFdL"]
        close(out["dx"][sl], dx[0::2], 1e-11); close(out["dy"][sl], dx[1::2], 1e-11)
        close(out["dr"][sl], dr, 1e-11)
        close(np.array([out["dLx"][s], out["dLy"][s]]), dL, 1e-11)


@pytest.fixture(scope="module")
def minimum():
    b = W.disk_batch([64, 96, 80, 72], [0.88, 0.90, 0.87, 0.90], [0, 1, 0, 1], seed=5)
    P = make(b)
    print(f"adjoint minimum S={b.S} N={b.N}", flush=True)
    steps, conv = run_fire(P, 1e-2, 200000, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=MODE)
    assert conv
    f = A.newton_minimize(P, MODE)
    assert f < 1e-13
    P.zero_velocities()
    return b, P


def test_cg_solution_and_pressure_gradient_match_host(minimum):
    b, P = minimum
    z = np.zeros(b.N)
    dr, dLx, dLy, it, res = A.pressure_grad(P, MODE, tol=1e-12)
    assert res <= 1e-10, res
    for s, sl, args in per_system(b, P, z, z):
        hdr, hdL, _ = R.implicit_pressure_grad(*args)
        close(dr[sl], hdr, 1e-7)
        close(np.array([dLx[s], dLy[s]]), hdL, 1e-7)


def test_pressure_gradient_full_size_systems_match_host():
    """Systems of the differentiation workload's size (1024 disks at phi = 0.86, one periodic,
    one walled): the GPU implicit gradient (MINRES over the particles with contacts) against
    the dense host reference (least squares over the same particles)."""
    b = W.disk_batch([1024, 1024], [0.86, 0.86], [0, 1], seed=1)
    P = make(b)
    print(f"adjoint full-size S={b.S} N={b.N}", flush=True)
    steps, conv = run_fire(P, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-12, mode=MODE)
    assert conv
    P.zero_velocities()
    dr, dLx, dLy, it, res = A.pressure_grad(P, MODE, tol=1e-11, max_iter=50000)
    assert res <= 1e-10, res
    z = np.zeros(b.N)
    for s, sl, args in per_system(b, P, z, z):
        hdr, hdL, _ = R.implicit_pressure_grad(*args)
        close(dr[sl], hdr, 1e-7)
        close(np.array([dLx[s], dLy[s]]), hdL, 1e-7)


def test_pressure_gradient_matches_gpu_finite_differences(minimum):
    """Perturb one radius by +-h, re-minimise on the GPU (Newton from the minimum), difference
    the pressures."""
    b, P = minimum
    dr, dLx, dLy, it, res = A.pressure_grad(P, MODE, tol=1e-12)
    x0, y0, r0 = canon(P, "POS_X"), canon(P, "POS_Y"), b.rad.copy()
    h = 1e-6
    P.compute_pressure()
    base = P.get("PRESSURE")

    def pressure_at(r):
        A.set_canonical(P, "RAD", r)
        A.set_canonical(P, "POS_X", x0); A.set_canonical(P, "POS_Y", y0)
        P.check_neighbors(True)
        assert A.newton_minimize(P, MODE) < 1e-13
        P.compute_pressure()
        return P.get("PRESSURE")

    for s in (0, 1):
        sl = slice(int(b.off[s]), int(b.off[s + 1]))
        i = int(b.off[s]) + int(np.argmax(np.abs(dr[sl])))
        e = np.zeros(b.N); e[i] = h
        fd = (pressure_at(r0 + e)[s] - pressure_at(r0 - e)[s]) / (2 * h)
        assert fd == pytest.approx(dr[i], rel=1e-6), (s, i)
    np.testing.assert_allclose(pressure_at(r0), base, rtol=1e-12)


def test_vv_backprop_matches_host_and_fd():
    b = W.disk_batch([32, 48], [0.86, 0.86], [0, 1], seed=3)
    P = make(b)
    print(f"adjoint vv S={b.S} N={b.N}", flush=True)
    run_fire(P, 1e-2, 100000, L.FIRE_PRED_MAX_FORCE, 1e-10, mode=MODE)   # start from a packing
    rng = np.random.default_rng(9)
    vx0, vy0 = rng.standard_normal(b.N) * 0.05, rng.standard_normal(b.N) * 0.05
    x0, y0 = canon(P, "POS_X"), canon(P, "POS_Y")
    dt, n, every = 1e-2, 200, 50
    sid = np.repeat(np.arange(b.S), np.diff(b.off))
    cnt = np.diff(b.off).astype(float)

    def run(vx, vy, ck=False):
        A.set_canonical(P, "POS_X", x0); A.set_canonical(P, "POS_Y", y0)
        A.set_canonical(P, "VEL_X", vx); A.set_canonical(P, "VEL_Y", vy)
        P.check_neighbors(True)
        VelocityVerlet(P, dt, mode=MODE).init()
        if ck:
            A.vv_run_checkpointed(P, n, every)
        else:
            VelocityVerlet(P, dt, mode=MODE).run(n)
        return canon(P, "POS_X"), canon(P, "POS_Y")

    xP, yP = run(vx0, vy0)
    xT, yT = run(vx0, vy0, ck=True)
    # checkpointed segments (step batches replayed across segments) = one plain run, bit for bit
    assert np.array_equal(xT, xP) and np.array_equal(yT, yP)
    gx, gy = 2 * (xT - x0) / cnt[sid], 2 * (yT - y0) / cnt[sid]
    gx0, gy0, gvx, gvy = A.vv_backprop(P, gx, gy, None, None)
    assert np.array_equal(canon(P, "POS_X"), xT)          # the context is left at the final state
    # graph-captured replays (cached across segments) = direct launches, bit for bit
    direct = A.vv_backprop(P, gx, gy, None, None, use_graph=False)
    for u, v in zip((gx0, gy0, gvx, gvy), direct):
        assert np.array_equal(u, v)
    for s in range(b.S):
        sl = slice(int(b.off[s]), int(b.off[s + 1]))
        args = (x0[sl], y0[sl], vx0[sl], vy0[sl], b.rad[sl], b.mass[sl], b.Lx[s], b.Ly[s], b.e[s],
                bool(b.boundary[s]), dt, n)
        xs, (hx, hy, _, _) = R.vv_trajectory(*args)
        np.testing.assert_allclose(xT[sl], hx, rtol=0, atol=1e-11)
        ref = R.vv_backprop(xs, gx[sl], gy[sl], np.zeros(len(hx)), np.zeros(len(hx)), b.rad[sl], b.mass[sl],
                            b.Lx[s], b.Ly[s], b.e[s], bool(b.boundary[s]), dt)
        close(gx0[sl], ref[0], 1e-8); close(gvx[sl], ref[2], 1e-8); close(gvy[sl], ref[3], 1e-8)

    def msd(vx, vy):
        ax, ay = run(vx, vy)
        return np.bincount(sid, weights=(ax - x0) ** 2 + (ay - y0) ** 2) / cnt

    h = 1e-6
    for i in (3, int(b.off[1]) + 7):
        e = np.zeros(b.N); e[i] = h
        fd = (msd(vx0 + e, vy0) - msd(vx0 - e, vy0))[sid[i]] / (2 * h)
        assert fd == pytest.approx(gvx[i], rel=1e-6), i


def test_torch_autograd_functions():
    torch = pytest.importorskip("torch")
    from dpmd_amd.autograd import DiskObservables, MinimizedPressure, NVEMeanSquareDisplacement

    b = W.disk_batch([48, 40], [0.89, 0.89], [0, 0], seed=17)
    P = make(b)
    x0, y0 = canon(P, "POS_X"), canon(P, "POS_Y")
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)
    # observables: backward equals the C ABI vector-Jacobian product
    x, y, r, Lx, Ly = t(x0), t(y0), t(b.rad), t(b.Lx), t(b.Ly)
    U, Pr, fx, fy = DiskObservables.apply(x, y, r, Lx, Ly, P, L.FORCE_PBC)
    wf = torch.linspace(-1.0, 1.0, b.N, dtype=torch.float64)
    (U.sum() + 3.0 * Pr.sum() + (wf * fx).sum() - (wf * fy).sum()).backward()
    g = A.vjp(P, L.FORCE_PBC, gU=np.ones(b.S), gP=np.full(b.S, 3.0), gF=(wf.numpy(), -wf.numpy()))
    assert np.array_equal(r.grad.numpy(), g["dr"]) and np.array_equal(x.grad.numpy(), g["dx"])
    assert np.array_equal(Lx.grad.numpy(), g["dLx"])
    # minimised pressure: one component against finite differences through the Function
    r = t(b.rad)
    p = MinimizedPressure.apply(r, P, x0, y0, 1e-2, 200000, 1e-12, L.FORCE_PBC, 1e-12)
    p.sum().backward()
    grad = r.grad.numpy().copy()
    i = int(np.argmax(np.abs(grad)))
    h = 1e-6
    rp, rm = b.rad.copy(), b.rad.copy()
    rp[i] += h; rm[i] -= h
    with torch.no_grad():
        fp = MinimizedPressure.apply(torch.tensor(rp), P, x0, y0, 1e-2, 200000, 1e-12, L.FORCE_PBC, 1e-12)
        fm = MinimizedPressure.apply(torch.tensor(rm), P, x0, y0, 1e-2, 200000, 1e-12, L.FORCE_PBC, 1e-12)
    assert float((fp.sum() - fm.sum()) / (2 * h)) == pytest.approx(grad[i], rel=1e-6)
    # NVE mean-square displacement
    rng = np.random.default_rng(2)
    v = rng.standard_normal((2, b.N)) * 0.05
    vx, vy = t(v[0]), t(v[1])
    msd = NVEMeanSquareDisplacement.apply(vx, vy, P, x0, y0, 1e-2, 200, 50, L.FORCE_PBC)
    msd.sum().backward()
    j = 5
    e = np.zeros(b.N); e[j] = h
    with torch.no_grad():
        mp = NVEMeanSquareDisplacement.apply(torch.tensor(v[0] + e), torch.tensor(v[1]), P, x0, y0, 1e-2, 200, 50,
                                             L.FORCE_PBC)
        mm = NVEMeanSquareDisplacement.apply(torch.tensor(v[0] - e), torch.tensor(v[1]), P, x0, y0, 1e-2, 200, 50,
                                             L.FORCE_PBC)
    assert float((mp.sum() - mm.sum()) / (2 * h)) == pytest.approx(float(vx.grad[j]), rel=1e-6)
