"""Pins the oracle's damped velocity Verlet, rescale-NVT compression loop and bisection jammer
(CPU) against closed forms:

- damped VV of a free disk / rigid body: the half-step velocity obeys h_{n+1} = (1 - 2a) h_n
  with a = gamma dt / (2 m) (drag at the half-step velocity, added after the forces, before
  the closing half kick; init without drag), so v_n = v_0 (1 - 2a)^(n-1) (1 - a) and
  x_n = x_0 + dt v_0 (1 - (1 - 2a)^n) / (2a);
- a damped two-disk contact loses energy every step (a damped oscillator);
- the NVT loop: phi grows by phi_increment at every compression step k % cf == 0, the
  temperature equals the target after a thermostat step and the momentum is zero;
- the bisection jammer: one disk in a walled square box jams when the box side reaches 2r,
  phi_J = pi r^2 / (2r)^2 = pi / 4; two equal disks jam at opposite corners,
  L = r (2 + sqrt 2), phi_J = 2 pi / (2 + sqrt 2)^2.
"""
import math
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


def free_disk(L=50.0, r=0.5):
    b = W.disk_batch([1], [math.pi * r * r / (L * L)], [0], seed=3)
    b.rad[:] = r
    b.area[:] = math.pi * r * r
    b.x[:] = 0.3 * L
    b.y[:] = 0.6 * L
    return b


@pytest.mark.parametrize("gamma,dt,m", [(0.3, 1e-2, 1.0), (2.0, 5e-2, 3.0)])
def test_damped_vv_free_disk_closed_form(gamma, dt, m):
    b = free_disk()
    b.mass[:] = m
    o = orc.OracleDisk(b, method=0)
    v0x, v0y = 0.7, -0.4
    o.set("vx", [v0x]); o.set("vy", [v0y])
    x0, y0 = o.get("x")[0], o.get("y")[0]
    o.damped_vv_init(dt, gamma)
    n = 37
    o.vv_run(n)
    a = gamma * dt / (2.0 * m)
    q = 1.0 - 2.0 * a
    vx, vy = o.get("vx")[0], o.get("vy")[0]
    np.testing.assert_allclose([vx, vy], [v0x * q ** (n - 1) * (1 - a), v0y * q ** (n - 1) * (1 - a)], rtol=1e-12)
    xs = dt * (1.0 - q ** n) / (2.0 * a)
    np.testing.assert_allclose([o.get("x")[0], o.get("y")[0]], [x0 + v0x * xs, y0 + v0y * xs], rtol=1e-12)
    # the stored force is the drag at the last half-step velocity
    h = v0x * q ** (n - 1)
    np.testing.assert_allclose(o.get("fx")[0], -gamma * h, rtol=1e-12)


def test_damped_vv_zero_damping_is_vv():
    b = W.disk_batch([64], [0.8], [0], seed=5)
    p, q = orc.OracleDisk(b), orc.OracleDisk(b)
    v = np.linspace(-0.01, 0.01, 64)
    for o in (p, q):
        o.set("vx", v); o.set("vy", v[::-1].copy())
    p.vv_init(1e-2)
    q.damped_vv_init(1e-2, 0.0)
    p.vv_run(200); q.vv_run(200)
    for f in ("x", "y", "vx", "vy", "fx", "fy"):
        assert np.array_equal(p.get(f), q.get(f)), f


def test_damped_rigid_free_body_closed_form():
    rb = W.rigid_batch(S=1, n_per=1, phi=0.001, boundary=0, seed=2)
    o = orc.OracleRigid(rb)
    gamma, dt = 0.5, 1e-2
    m, I = rb.mass[0], rb.inertia[0]
    o.set("vx", [0.2]); o.set("vy", [0.1]); o.set("w", [0.9])
    o.damped_vv_init(dt, gamma)
    n = 25
    o.vv_run(n)
    for name, v0, inert in (("vx", 0.2, m), ("vy", 0.1, m), ("w", 0.9, I)):
        a = gamma * dt / (2.0 * inert)
        q = 1.0 - 2.0 * a
        np.testing.assert_allclose(o.get(name)[0], v0 * q ** (n - 1) * (1 - a), rtol=1e-12, err_msg=name)


def test_damped_pair_loses_energy():
    b = W.disk_batch([2], [0.01], [0], seed=1)
    L = b.Lx[0]
    b.x[:] = [0.5 * L - 0.5, 0.5 * L + 0.5]     # radii 0.5 and 0.7: overlap 0.2
    b.y[:] = 0.5 * L
    o = orc.OracleDisk(b, method=0)
    o.damped_vv_init(1e-2, 0.2)
    e_prev = o.pe_total()[0] + o.ke_total()[0]
    for _ in range(300):
        o.vv_run(1)
        e = o.pe_total()[0] + o.ke_total()[0]
        assert e <= e_prev * (1 + 1e-14)
        e_prev = e
    assert e_prev < 0.5 * 0.5 * 0.2 ** 2


def test_nvt_rescale_loop_phi_temperature_momentum():
    b = W.disk_batch([80, 120], [0.6, 0.65], [0, 0], seed=4)
    o = orc.OracleDisk(b)
    rng = np.random.default_rng(0)
    o.set("vx", rng.normal(size=b.N) * 1e-2); o.set("vy", rng.normal(size=b.N) * 1e-2)
    phi0 = o.packing_fraction()
    o.vv_init(1e-2)
    n, cf, tf, inc, T = 250, 100, 10, np.array([1e-3, 2e-3]), np.array([1e-3, 2e-4])
    o.nvt_rescale_run(n, inc, cf, T, tf)
    ncomp = len(range(0, n, cf))                       # steps 0, 100, 200
    np.testing.assert_allclose(o.packing_fraction(), phi0 + ncomp * inc, rtol=0, atol=1e-12)
    # last thermostat step was k = 240; 9 unthermostatted steps follow, so compare one more
    o.nvt_rescale_run(1, None, 0, T, 1, step0=n)        # k = 250: thermostat
    np.testing.assert_allclose(o.temperature(), T, rtol=1e-12)
    for s in range(2):
        sl = slice(b.off[s], b.off[s + 1])
        assert abs(np.sum(o.get("vx")[sl])) < 1e-14 and abs(np.sum(o.get("vy")[sl])) < 1e-14


def test_nvt_rescale_loop_event_order():
    """Compression and thermostat at the same step commute: running the step loop with the
    events split over two calls (step0 offsets) gives the same bits as one call."""
    b = W.disk_batch([100], [0.7], [0], seed=6)
    p, q = orc.OracleDisk(b), orc.OracleDisk(b)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(2, b.N)) * 1e-2
    for o in (p, q):
        o.set("vx", v[0]); o.set("vy", v[1]); o.vv_init(1e-2)
    p.nvt_rescale_run(60, 5e-4, 20, 1e-4, 7)
    q.nvt_rescale_run(23, 5e-4, 20, 1e-4, 7)
    q.nvt_rescale_run(37, 5e-4, 20, 1e-4, 7, step0=23)
    for f in ("x", "y", "vx", "vy"):
        assert np.array_equal(p.get(f), q.get(f)), f


def single_disk_walls(r=0.7, phi0=0.5, seed=1):
    b = W.disk_batch([1], [phi0], [1], seed=seed)
    b.rad[:] = r
    b.area[:] = math.pi * r * r
    L = math.sqrt(b.area[0] / phi0)
    b.Lx[:] = L; b.Ly[:] = L
    return b


@pytest.mark.parametrize("r", [0.5, 0.7])
def test_jam_bisect_one_disk_in_walls(r):
    b = single_disk_walls(r=r)
    o = orc.OracleDisk(b, method=0)
    tol = 1e-7
    act, phi, rounds = o.jam_bisect(1e-2, 300, 20000, 1e-16, 1e-16, 0.02, tol, mode=1)
    assert act[0] == 0 and rounds < 300
    assert abs(phi[0] / (math.pi / 4) - 1) < 2 * tol
    # the final box is the last one tested, an end of the final bracket around phi
    assert abs(o.packing_fraction()[0] / phi[0] - 1) < tol


def test_jam_bisect_two_disks_in_walls():
    b = W.disk_batch([2], [0.3], [1], seed=2)
    b.rad[:] = 0.5
    b.area[:] = math.pi * 0.25
    L = math.sqrt(2 * b.area[0] / 0.3)
    b.Lx[:] = L; b.Ly[:] = L
    b.x[:] = [0.3 * L, 0.6 * L]; b.y[:] = [0.35 * L, 0.7 * L]
    o = orc.OracleDisk(b, method=0)
    act, phi, rounds = o.jam_bisect(1e-2, 300, 200000, 1e-16, 1e-16, 0.02, 1e-6, mode=1)
    phi_j = 2 * math.pi / (2 + math.sqrt(2)) ** 2
    assert act[0] == 0
    assert abs(phi[0] / phi_j - 1) < 1e-5


def test_jam_bisect_drops_unstopped_systems():
    """A first minimisation that hits the step cap leaves dt != 0: the system is dropped
    (action 0) and never moves again (jammers.cu:7-17)."""
    b = W.disk_batch([64, 64], [0.9, 0.5], [0, 0], seed=7)
    o = orc.OracleDisk(b)
    act, phi, rounds = o.jam_bisect(1e-2, 3, 5, 1e-16, 1e-16, 0.01, 1e-6, mode=0)
    # 5 FIRE steps cannot relax the phi = 0.9 system: dropped with its phi unchanged
    assert act[0] == 0
    np.testing.assert_allclose(o.packing_fraction()[0], phi[0], rtol=1e-12)


def one_rigid_body_walls(n_v=12, R=0.5, phi0=0.35, theta=0.3, pos=(0.45, 0.55)):
    """One regular bumpy n_v-gon (R0 recipe) in a walled square box."""
    rb = R * math.sin(math.pi / n_v)
    area = 0.5 * n_v * R * R * math.sin(2 * math.pi / n_v) + 0.5 * n_v * rb * rb * (math.pi + 2 * math.pi / n_v)
    L = math.sqrt(area / phi0)
    x, y = pos[0] * L, pos[1] * L
    ang = theta + 2 * math.pi * np.arange(n_v) / n_v
    return W.RigidBatch(off=np.array([0, 1], np.int32), nvpp=np.array([n_v], np.int32), x=np.array([x]),
                        y=np.array([y]), angle=np.array([theta]), mass=np.array([1.0]), inertia=np.array([R * R]),
                        area=np.array([area]), vx=x + R * np.cos(ang), vy=y + R * np.sin(ang),
                        vrad=np.full(n_v, rb), vmass=np.full(n_v, 1.0 / n_v), Lx=np.array([L]), Ly=np.array([L]),
                        e=np.array([1.0]), cutoff=np.array([0.05]), thresh2=np.array([0.025 ** 2]),
                        boundary=np.array([1], np.int32), cdx=np.array([3], np.int32), cdy=np.array([3], np.int32))


def test_rigid_save_restore_round_trip():
    b = W.rigid_batch(S=2, n_per=24, seed=4, boundary=0, phi=0.6)
    o = orc.OracleRigid(b)
    o.save_state([1, 0], 1)                 # system 0 only
    snap = {f: o.get(f).copy() for f in ("x", "y", "angle", "vpx", "vpy")}
    si0 = o.static_index().copy()
    rng = np.random.default_rng(2)
    o.set("vx", rng.normal(size=b.N) * 0.05); o.set("vy", rng.normal(size=b.N) * 0.05)
    o.set("w", rng.normal(size=b.N) * 0.1)
    o.vv_init(1e-2); o.vv_run(300)          # rebuilds reorder the vertices
    moved = {f: o.get(f).copy() for f in ("x", "y", "angle")}
    o.restore_state([1, 0], 1)
    n0 = b.off[1]
    for f in ("x", "y", "angle"):
        assert np.array_equal(o.get(f)[:n0], snap[f][:n0]), f             # restored
        assert np.array_equal(o.get(f)[n0:], moved[f][n0:]), f            # untouched
    # canonical vertex positions of system 0 are back (vertex order may differ after the rebuild)
    si = o.static_index()
    nv0 = int(np.sum(b.nvpp[:n0]))
    for f in ("vpx", "vpy"):
        assert np.array_equal(o.get(f)[si[:nv0]], snap[f][si0[:nv0]]), f


def test_rigid_increment_packing_fraction_keeps_shapes():
    b = W.rigid_batch(S=2, n_per=32, seed=5, boundary=1, phi=0.5)
    o = orc.OracleRigid(b)
    phi0 = o.packing_fraction()
    np.testing.assert_allclose(phi0, 0.5, rtol=1e-12)
    o.increment_packing_fraction([0.01, 0.02])
    np.testing.assert_allclose(o.packing_fraction(), phi0 + [0.01, 0.02], rtol=0, atol=1e-12)
    # rigid shapes: every vertex keeps its distance R from its own centre
    si = o.static_index()
    vx, vy = o.get("vpx")[si], o.get("vpy")[si]
    pid = np.repeat(np.arange(b.N), b.nvpp)
    d = np.hypot(vx - o.get("x")[pid], vy - o.get("y")[pid])
    R = np.sqrt(b.inertia / b.mass)[pid]
    np.testing.assert_allclose(d, R, rtol=1e-12)


def test_rigid_jam_bisect_one_body_walls():
    """A regular bumpy 12-gon (a multiple of 4 vertices) fits a square wall box down to the side
    2 R cos(pi / n) + 2 r_b (edge-aligned in both axes), so phi_J = area / L_J^2."""
    n_v, R = 12, 0.5
    b = one_rigid_body_walls(n_v, R)
    o = orc.OracleRigid(b)
    tol = 1e-6
    act, phi, rounds = o.jam_bisect(1e-2, 300, 40000, 1e-16, 1e-16, 0.02, tol, mode=1)
    LJ = 2 * R * math.cos(math.pi / n_v) + 2 * R * math.sin(math.pi / n_v)
    phi_j = b.area[0] / LJ ** 2
    assert act[0] == 0
    assert abs(phi[0] / phi_j - 1) < 1e-5
