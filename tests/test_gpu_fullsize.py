"""GPU tests at the benchmark workloads' full sizes, through the C ABI.

  * 2^20 bidisperse disks at phi = 0.84 (PBC, skin 0.3): 20 FIRE steps from the random
    start (a rebuild on nearly every step, device-wide chain with the tiled neighbor build),
    then 20 velocity-Verlet steps (fused launches), bit for bit against the oracle; lists in
    order; sampled particles' lists against brute force over the whole system.
  * 256 wall boxes x 256 rigid bumpy polygons (about 1.24 M vertices): 10 FIRE + 10 NVE steps,
    bit for bit against the oracle.
  * The 512-system uneven batch (N_s in [64, 4096], phi_s in [0.80, 0.86], half PBC, half
    walls): a capped pressure-target jamming run on the whole batch, compared per system with
    the oracle on 32 systems spread over the batch.
  * The reference's energy-conservation assertion at its own scale (tests/test_disk.cu:356-624:
    1000 systems x 20 disks, phi = 0.5, dt in {1e-1 ... 1e-3} for {1e3 ... 1e5} steps, log-log
    slope of the total-energy fluctuation against dt = 2 +- 0.1), PBC and walls, naive and cell
    lists.  The reference's cell variants keep its 4 x 4 grid (cells of side 1.40) with a skin
    of 2r = 1.0, so cells are narrower than 2r + skin = 2.0 and the 3 x 3 stencil can miss a
    pair that closes in between rebuilds; the cell variants here use a skin of 0.4, the largest
    the 4 x 4 grid covers (the naive variants keep 1.0).
"""
import math

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import DampedVelocityVerlet, Disk, RigidBumpy, VelocityVerlet, jam_compress, run_fire

pytestmark = pytest.mark.gpu

DN = {"x": "POS_X", "y": "POS_Y", "vx": "VEL_X", "vy": "VEL_Y", "fx": "FORCE_X", "fy": "FORCE_Y", "pe": "PE"}


def same(P, ref, names, what):
    for f, g in names.items():
        a, r = P.get(g), ref.get(f)
        assert np.array_equal(a, r), f"{what}: {f} differs, max |diff| {np.max(np.abs(a - r)):.3e}"


def test_single_large_system_fire_and_nve_bit_exact():
    orc.lib().or_set_threads(16)
    b = W.single_disk_system(1 << 20, 0.84, seed=1, skin=0.3)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    ref = orc.OracleDisk(b)
    print(f"2^20 disks: lists", flush=True)
    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)
    steps, conv = run_fire(P, 1e-2, 20, L.FIRE_PRED_MAX_FORCE, 0.0)
    rsteps, _ = ref.fire_minimize(1e-2, 20, pred=1, ftol=0.0)
    print(f"2^20 disks: FIRE {steps} steps", flush=True)
    assert steps == rsteps == 20 and not conv
    same(P, ref, DN, "FIRE")
    assert np.array_equal(P.get("STATIC_INDEX"), ref.static_index())
    rng = np.random.default_rng(11)
    v = rng.standard_normal((2, b.N)) * 1e-2
    P.set("VEL_X", v[0]); P.set("VEL_Y", v[1]); ref.set("vx", v[0]); ref.set("vy", v[1])
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    vv.run(20)
    ref.vv_init(1e-2)
    ref.vv_run(20)
    print("2^20 disks: NVE 20 steps", flush=True)
    same(P, ref, DN, "NVE")
    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)
    # sampled lists against an O(N) scan over the whole system (minimum image)
    x, y, r = P.get("POS_X"), P.get("POS_Y"), P.get("RAD")
    L_ = b.Lx[0]
    for i in np.linspace(0, b.N - 1, 64).astype(int):
        dx = x[i] - x; dy = y[i] - y
        dx -= L_ * np.rint(dx * (1.0 / L_)); dy -= L_ * np.rint(dy * (1.0 / L_))
        cut = r[i] + r + b.cutoff[0]
        brute = np.nonzero(dx * dx + dy * dy < cut * cut)[0]
        assert sorted(gi[gs[i]:gs[i + 1]]) == [j for j in brute if j != i]


def test_rigid_256x256_fire_and_nve_bit_exact():
    rb = W.rigid_batch(S=256, n_per=256, seed=1)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P.init_neighbors()
    ref = orc.OracleRigid(rb)
    print(f"rigid 256x256: Nv = {rb.Nv}", flush=True)
    steps, conv = run_fire(P, 1e-2, 10, L.FIRE_PRED_MAX_FORCE, 0.0, mode=L.FORCE_WALL)
    rsteps, _ = ref.fire_minimize(1e-2, 10, pred=1, ftol=0.0, mode=1)
    assert steps == rsteps == 10
    names = {"x": "POS_X", "y": "POS_Y", "angle": "ANGLE", "vx": "VEL_X", "vy": "VEL_Y", "w": "ANGULAR_VEL",
             "fx": "FORCE_X", "fy": "FORCE_Y", "torque": "TORQUE", "vpx": "VERTEX_POS_X", "vpy": "VERTEX_POS_Y"}
    same(P, ref, names, "rigid FIRE")
    rng = np.random.default_rng(2)
    for f, rr, sc in (("VEL_X", "vx", 1e-2), ("VEL_Y", "vy", 1e-2), ("ANGULAR_VEL", "w", 1e-2)):
        v = rng.standard_normal(rb.N) * sc
        P.set(f, v); ref.set(rr, v)
    VelocityVerlet(P, 1e-2, wall=True).init()
    ref.vv_init(1e-2, 1)
    VelocityVerlet(P, 1e-2, wall=True).run(10)
    ref.vv_run(10, 1)
    print("rigid 256x256: NVE 10 steps", flush=True)
    same(P, ref, names, "rigid NVE")
    gs, gi = P.neighbors_csr()
    rs, ri = ref.neighbors_csr()
    assert np.array_equal(gs, rs) and np.array_equal(gi, ri)


def test_uneven_512_batch_jam_round_per_system_vs_oracle():
    """The whole 512-system batch through the pressure-target driver with a short FIRE cap
    (1500 steps per round, at most 2 compressions), against the oracle run of 32 of its
    systems (every 16th).  Positions (canonical), jam decisions, compression and step counts
    are compared exactly; energies and pressures to 1e-10 relative (a system that stopped
    earlier is re-sorted in the batch at the other systems' compressions, which changes its
    summation order only)."""
    orc.lib().or_set_threads(16)
    b = W.uneven_batch(S=512)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    print(f"uneven batch: S={b.S} N={b.N}", flush=True)
    args = (1e-2, 1e-3, 1e-7, 2, 1500)
    gst, gnc, gwork = jam_compress(P, *args, L.FIRE_PRED_MAX_FORCE, 1e-10, mode=L.FORCE_PER_SYSTEM)
    # the batch zeroes a stopped system's velocities when other systems run on (its next FIRE
    # step has dt = 0); the sample alone may end first: compare the configurational pressure
    P.zero_velocities()
    P.compute_pe_total()
    P.compute_pressure()
    gsteps = P.get("FIRE_STEPS")
    gx, gy = P.get("POS_X", L.ORDER_CANONICAL), P.get("POS_Y", L.ORDER_CANONICAL)
    gpe, gp, gL = P.get("PE_TOTAL"), P.get("PRESSURE"), P.get("BOX_SIZE_X")
    pick = list(range(0, 512, 16))
    sub = W.disk_batch(np.diff(b.off)[pick], [0.8] * len(pick), b.boundary[pick], seed=1)
    # the sample's own inputs: the systems' positions, boxes and radii
    lo = [int(b.off[s]) for s in pick]; hi = [int(b.off[s + 1]) for s in pick]
    sub.x = np.concatenate([b.x[a:c] for a, c in zip(lo, hi)])
    sub.y = np.concatenate([b.y[a:c] for a, c in zip(lo, hi)])
    sub.rad = np.concatenate([b.rad[a:c] for a, c in zip(lo, hi)])
    sub.area = np.concatenate([b.area[a:c] for a, c in zip(lo, hi)])
    for f in ("Lx", "Ly", "e", "cutoff", "thresh2", "cdx", "cdy"):
        setattr(sub, f, np.ascontiguousarray(getattr(b, f)[pick]))
    o = orc.OracleDisk(sub)
    ost, onc, ons, _ = o.jam_compress(*args, pred=1, a=1e-10, mode=2)
    o.set("vx", np.zeros(sub.N)); o.set("vy", np.zeros(sub.N))
    print(f"oracle sample: {len(pick)} systems, jammed {int(np.sum(ost == 1))}", flush=True)
    assert np.array_equal(gst[pick], ost) and np.array_equal(gnc[pick], onc)
    assert np.array_equal(gsteps[pick], ons)
    sst = o.static_index()                          # canonical -> current
    ox = np.empty(sub.N); oy = np.empty(sub.N)
    ox[:] = o.get("x")[sst]; oy[:] = o.get("y")[sst]
    k = 0
    for s, a, c in zip(pick, lo, hi):
        n = c - a
        assert np.array_equal(gx[a:c], ox[k:k + n]) and np.array_equal(gy[a:c], oy[k:k + n]), s
        k += n
    assert np.array_equal(gL[pick], o.get("Lx"))
    np.testing.assert_allclose(gpe[pick], o.pe_total(), rtol=1e-10, atol=1e-300)
    np.testing.assert_allclose(gp[pick], o.pressure(), rtol=1e-10, atol=1e-16)
    assert gwork > 0


def _slope(dts, y, abs_floor=1e-18, rel_floor=1e-14):
    """fit_loglog_slope_filtered (tests/test_disk.cu:36-61)."""
    ymax = max(y)
    pts = [(math.log(d), math.log(v)) for d, v in zip(dts, y) if v > abs_floor and v > rel_floor * ymax]
    xs, ys = np.array([p[0] for p in pts]), np.array([p[1] for p in pts])
    n = len(xs)
    return (n * np.sum(xs * ys) - xs.sum() * ys.sum()) / (n * np.sum(xs * xs) - xs.sum() ** 2)


@pytest.fixture(scope="module")
def slope_start():
    """tests/test_disk.cu:62-324: 1000 systems of 20 disks (r = 0.5, m = 1, k = 1) at phi = 0.5,
    4 x 4 cells, skin 2r; uniform positions, 10 x 500 damped wall steps (damping 1, dt 1e-2)
    with forces and velocities cleared after each, then velocities N(-1, 1) x 1e-3."""
    S, n, r = 1000, 20, 0.5
    L_ = math.sqrt(n * math.pi * r * r / 0.5)
    b = W.disk_batch([n] * S, [0.5] * S, [1] * S, seed=0, r_small=r, r_large=r, skin=2 * r)
    b.Lx[:] = L_; b.Ly[:] = L_
    b.area[:] = math.pi * r * r
    ux, uy = W.uniform_2d(b.N, 0, 0)
    b.x, b.y = L_ * ux, L_ * uy
    b.cdx[:] = 4; b.cdy[:] = 4
    P = Disk()
    W.load_into(P, b, L.NEIGHBOR_NAIVE)
    P.init_neighbors()
    for _ in range(10):
        d = DampedVelocityVerlet(P, 1e-2, 1.0, wall=True)
        d.init()
        d.run(500)
        P.set("FORCE_X", np.zeros(b.N)); P.set("FORCE_Y", np.zeros(b.N))
        P.set("VEL_X", np.zeros(b.N)); P.set("VEL_Y", np.zeros(b.N))
    P.rand_normal_2d("VEL_X", "VEL_Y", 0, 1, -1.0, 1.0, -1.0, 1.0)
    x, y = P.get("POS_X"), P.get("POS_Y")
    vx, vy = P.get("VEL_X") * 1e-3, P.get("VEL_Y") * 1e-3
    return b, x, y, vx, vy


@pytest.mark.parametrize("wall", [False, True])
@pytest.mark.parametrize("method", [L.NEIGHBOR_NAIVE, L.NEIGHBOR_CELL])
def test_reference_energy_conservation_slope(slope_start, method, wall):
    b, x, y, vx, vy = slope_start
    q = W.DiskBatch(**{**b.__dict__, "x": x.copy(), "y": y.copy()})
    if method == L.NEIGHBOR_CELL:
        q.cutoff = np.full(b.S, 0.4); q.thresh2 = np.full(b.S, 0.2 ** 2)
    P = Disk()
    W.load_into(P, q, method)
    P.init_neighbors()
    dts = [1e-1, 5e-2, 1e-2, 5e-3, 1e-3]
    nsteps = [1000, 5000, 10000, 50000, 100000]
    rel = []
    for dt, n in zip(dts, nsteps):
        # identical disks: positions and velocities are set together in the current order, so
        # a permutation of the particles by earlier rebuilds changes nothing physical
        P.set("POS_X", x); P.set("POS_Y", y)
        P.set("VEL_X", vx); P.set("VEL_Y", vy)
        P.check_neighbors(True)
        vv = VelocityVerlet(P, dt, wall=wall)
        vv.init()
        every = n // 1000                              # 1000 samples of the total energy per dt
        te = []
        for _ in range(1000):
            vv.run(every)
            P.compute_pe_total(); P.compute_ke_total()
            te.append(float(np.sum(P.get("PE_TOTAL")) + np.sum(P.get("KE_TOTAL"))))
        te = np.array(te)
        rel.append(float(te.std() / abs(te.mean())))
    p = _slope(dts, rel)
    print(f"energy slope method={method} wall={wall}: p = {p:.3f}, fluctuations {rel}", flush=True)
    assert abs(p - 2.0) < 0.1


def test_adjoint_pressure_gradient_128x1024_vs_host():
    """Differentiation workload (a): 128 periodic systems x 1024 disks at phi = 0.86, FIRE to
    max|F| < 1e-12, d(pressure)/d(radii) by implicit differentiation (MINRES over the
    particles with contacts to a 1e-10 residual), against the dense host reference (least
    squares over the same particles) for three systems of the batch, one of them system 11,
    whose packing holds struts (particles between two contacts of almost no force) that the
    response goes through; bar 1e-7 relative to the largest entry."""
    import adjoint_ref as R
    from dpmd_amd import adjoint as A
    b = W.disk_batch([1024] * 128, [0.86] * 128, [0] * 128, seed=1)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    steps, conv = run_fire(P, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-12)
    print(f"adjoint 128x1024: FIRE {steps} steps conv={conv}", flush=True)
    assert conv
    P.zero_velocities()
    dr, dLx, dLy, it, res = A.pressure_grad(P, L.FORCE_PBC, tol=1e-11, max_iter=50000)
    print(f"adjoint 128x1024: MINRES {it} iterations, residual {res:.2e}", flush=True)
    assert res <= 1e-10
    x, y = P.get("POS_X", L.ORDER_CANONICAL), P.get("POS_Y", L.ORDER_CANONICAL)
    z = np.zeros(1024)
    for s in (0, 11, 63, 127):
        sl = slice(int(b.off[s]), int(b.off[s + 1]))
        hdr, hdL, _ = R.implicit_pressure_grad(x[sl], y[sl], b.rad[sl], b.mass[sl], z, z, b.Lx[s], b.Ly[s],
                                               b.e[s], False)
        sc = np.max(np.abs(hdr))
        np.testing.assert_allclose(dr[sl], hdr, rtol=0, atol=1e-7 * sc, err_msg=f"system {s}")
        np.testing.assert_allclose([dLx[s], dLy[s]], hdL, rtol=0, atol=1e-7 * np.max(np.abs(hdL)))


def test_adjoint_nve_1024_disks_1000_steps_vs_host():
    """Differentiation workload (b) on one of its systems: 1024 disks at phi = 0.86 from a FIRE
    packing, velocities at T ~ 1e-4, 1000 velocity-Verlet steps (dt = 1e-2) with checkpoints
    every 50; the checkpointed run equals the plain run bit for bit, its final state the host
    trajectory (1e-10), and the reverse-mode gradient of the mean-square displacement with
    respect to the initial positions and velocities the host reverse mode (1e-8 relative to
    the largest entry)."""
    import adjoint_ref as R
    from dpmd_amd import adjoint as A
    b = W.disk_batch([1024], [0.86], [0], seed=2)
    P = Disk()
    W.load_into(P, b)
    P.init_neighbors()
    run_fire(P, 1e-2, 300000, L.FIRE_PRED_MAX_FORCE, 1e-10)
    rng = np.random.default_rng(7)
    vx0, vy0 = rng.standard_normal(b.N) * 1e-2, rng.standard_normal(b.N) * 1e-2
    x0, y0 = P.get("POS_X", L.ORDER_CANONICAL), P.get("POS_Y", L.ORDER_CANONICAL)
    dt, n, every = 1e-2, 1000, 50

    def run(ck):
        A.set_canonical(P, "POS_X", x0); A.set_canonical(P, "POS_Y", y0)
        A.set_canonical(P, "VEL_X", vx0); A.set_canonical(P, "VEL_Y", vy0)
        P.check_neighbors(True)
        VelocityVerlet(P, dt).init()
        if ck:
            A.vv_run_checkpointed(P, n, every)
        else:
            VelocityVerlet(P, dt).run(n)
        return P.get("POS_X", L.ORDER_CANONICAL), P.get("POS_Y", L.ORDER_CANONICAL)

    xP, yP = run(False)
    xT, yT = run(True)
    assert np.array_equal(xT, xP) and np.array_equal(yT, yP)
    gx, gy = 2 * (xT - x0) / b.N, 2 * (yT - y0) / b.N
    gx0, gy0, gvx, gvy = A.vv_backprop(P, gx, gy, None, None)
    print("adjoint NVE 1024 x 1000: GPU backprop done", flush=True)
    xs, (hx, hy, _, _) = R.vv_trajectory_sparse(x0, y0, vx0, vy0, b.rad, b.mass, b.Lx[0], b.Ly[0], b.e[0], False,
                                                dt, n)
    np.testing.assert_allclose(xT, hx, rtol=0, atol=1e-10)
    np.testing.assert_allclose(yT, hy, rtol=0, atol=1e-10)
    z = np.zeros(b.N)
    ref = R.vv_backprop_sparse(xs, gx, gy, z, z, b.rad, b.mass, b.Lx[0], b.Ly[0], b.e[0], False, dt)
    for got, want, name in zip((gx0, gy0, gvx, gvy), ref, ("x0", "y0", "vx0", "vy0")):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-8 * np.max(np.abs(want)), err_msg=name)
