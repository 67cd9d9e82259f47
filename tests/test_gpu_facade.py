"""The C++ facade (include/dpmd/*.hpp) drives the library exactly like the Python binding:
the driver tests/cpp/facade_test.cpp, written in the reference's API, must reproduce bit for
bit the same call sequence issued through dpmd_amd.system and checked against the oracle."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
from dpmd_amd import _lib as L
from dpmd_amd import workloads as W
from dpmd_amd.system import (FIRE, DampedVelocityVerlet, Disk, VelocityVerlet, jam_binary_search, minimize_fire,
                             nvt_rescale_run)

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def python_sequence():
    S, n_per, phi = 2, 128, 0.80
    N = S * n_per
    rad = np.array(([0.5] * (n_per // 2) + [0.7] * (n_per // 2)) * S)
    area = math.pi * rad * rad
    a_sys = 0.0
    for v in area[:n_per]:          # sequential, as the C++ driver sums
        a_sys += float(v)
    L_box = math.sqrt(a_sys / phi)
    P = Disk()
    P.set_neighbor_method(L.NEIGHBOR_CELL)
    P.allocate_systems(S)
    P.allocate_particles(N)
    P.set("SYSTEM_OFFSET", [0, n_per, N])
    P.set("NEIGHBOR_CUTOFF", [0.3] * S)
    P.set("THRESH2", [0.15 * 0.15] * S)
    P.set("CELL_DIM_X", [0] * S); P.set("CELL_DIM_Y", [0] * S)
    P.set("BOX_SIZE_X", [L_box] * S); P.set("BOX_SIZE_Y", [L_box] * S)
    P.set("E_INTERACTION", [1.0] * S)
    P.set("RAD", rad); P.set("MASS", np.ones(N)); P.set("AREA", area)
    P.rand_uniform_2d("POS_X", "POS_Y", 11, 0, 0.0, L_box, 0.0, L_box)
    P.sync()
    P.init_neighbors()
    P.compute_forces(); P.compute_pe_total()
    pe0 = P.get("PE_TOTAL")
    steps = minimize_fire(P, 1e-2, 300, 1e-16, 1e-16)
    P.rand_normal_2d("VEL_X", "VEL_Y", 3, 1, 0.0, 1e-2, 0.0, 1e-2)
    P.set_average_velocity(0.0, 0.0)
    P.set_temperature(1e-4)
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    vv.step(50)
    vv.run(50)
    dvv = DampedVelocityVerlet(P, 1e-2, 0.2)
    dvv.init()
    dvv.step(20)
    dvv.run(20)
    nvt_rescale_run(P, 60, 2e-3, 25, 2e-4, 10)
    _, _, rounds = jam_binary_search(P, 1e-2, 3, 500, 1e-10, 1e-10, 5e-3, 1e-4)
    P.compute_pe_total(); P.compute_ke_total(); P.compute_pressure()
    c = L.ORDER_CANONICAL
    state = np.concatenate([P.get("POS_X", c), P.get("POS_Y", c), P.get("VEL_X", c), P.get("VEL_Y", c),
                            pe0, P.get("PE_TOTAL"), P.get("KE_TOTAL"), P.get("PRESSURE")])
    return state, steps, rounds, P


def test_cpp_facade_matches_python_binding(tmp_path):
    exe = ROOT / "tests" / "cpp" / "facade_test.bin"
    assert exe.exists(), "build() compiles tests/cpp/facade_test.cpp"
    out = tmp_path / "state.bin"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    cpp = np.fromfile(out, dtype=np.float64)
    py, steps, rounds, _ = python_sequence()
    assert f"fire_steps={steps} jam_rounds={rounds}" in r.stdout
    assert cpp.shape == py.shape
    assert np.array_equal(cpp, py)


def test_cpp_output_manager_matches_python_output_manager(tmp_path):
    """io::OutputManager (C++) and dpmd_amd.output.OutputManager write the same arrays, bit for
    bit, for the same run, in both containers; the frames are the fields read directly after
    those steps."""
    from dpmd_amd import h5, output
    with_h5 = h5.available()
    exe = ROOT / "tests" / "cpp" / "facade_test.bin"
    args = [str(exe), str(tmp_path / "state.bin"), str(tmp_path / "cpp")] + ([str(tmp_path / "cpp_h5")] if with_h5 else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    _, _, _, P = python_sequence()
    traj = ["pos", "vel", "pe_total", "ke_total", "temperature", "contacts", "n_contacts_total"]
    managers = []
    for fmt, d in (("npy", "py"), ("h5", "py_h5"))[: 2 if with_h5 else 1]:
        om = output.OutputManager(P, tmp_path / d, max_workers=2, format=fmt)
        om.set_trajectory_fields(traj)
        om.set_trajectory_interval(7)
        om.set_restart_interval(25)
        om.set_extra_init_fields(["packing_fraction"])
        om.set_extra_final_fields(["pe", "pressure", "stress_tensor_total_x"])
        om.set_queue_limit(3)
        om.initialize()
        managers.append(om)
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    direct = {}
    for k in range(1, 61):
        vv.step(1)
        if k % 7 == 0:
            direct[k] = P.get("POS_X", L.ORDER_CANONICAL).copy()
        for om in managers:
            om.step(k)
    for om in managers:
        om.finalize()
    runs = [output.load(tmp_path / d) for d in (("cpp", "py", "cpp_h5", "py_h5") if with_h5 else ("cpp", "py"))]
    a = runs[0]
    assert np.array_equal(a["timestep"], np.arange(7, 61, 7))
    for b in runs[1:]:
        assert np.array_equal(a["timestep"], b["timestep"])
        for g in ("static", "init", "restart", "final", "trajectory"):
            names = set(a[g]) - {"step"}
            assert names == set(b[g]) - {"step"}, g
            for n in names:
                x, y = a[g][n], b[g][n]
                if isinstance(x, np.ndarray):
                    assert x.dtype == y.dtype and np.array_equal(x, y), (g, n)
                else:
                    assert x == y, (g, n)
    assert set(a["trajectory"]) == set(traj)
    assert a["trajectory"]["pos"].shape == (8, P.n_particles(), 2)
    for t, k in enumerate(a["timestep"]):
        assert np.array_equal(a["trajectory"]["pos"][t, :, 0], direct[int(k)])
    assert int(a["restart"]["step"][0]) == 60 and int(a["final"]["step"][0]) == 60
    assert np.array_equal(a["final"]["pos"][:, 0], P.get("POS_X", L.ORDER_CANONICAL))
    assert np.array_equal(a["static"]["rad"], P.get("RAD", L.ORDER_CANONICAL))
    assert "cell_dim" in a["static"] and a["static"]["box_size"].shape == (2, 2)


def test_python_output_manager_rigid(tmp_path):
    """Rigid bodies: the default groups are the reference's (vertex fields, angles), per-particle
    fields keep the load order, per-vertex fields come back in canonical vertex order."""
    from dpmd_amd import output
    from dpmd_amd.system import RigidBumpy
    rb = W.rigid_batch(S=2, n_per=16, seed=5, boundary=0, phi=0.70)
    P = RigidBumpy()
    W.load_rigid_into(P, rb)
    P.init_neighbors()
    om = output.OutputManager(P, tmp_path / "rb", max_workers=2)
    om.set_trajectory_fields(["pos", "angle", "vertex_pos", "pe_total"])
    om.set_trajectory_interval(10)
    om.initialize()
    vv = VelocityVerlet(P, 1e-2)
    vv.init()
    for k in range(1, 31):
        vv.step(1)
        om.step(k)
    om.finalize()
    d = output.load(tmp_path / "rb")
    assert np.array_equal(d["timestep"], [10, 20, 30])
    assert d["trajectory"]["vertex_pos"].shape == (3, rb.Nv, 2)
    assert np.array_equal(d["trajectory"]["angle"][-1], P.get("ANGLE"))
    assert np.array_equal(d["final"]["vertex_pos"][:, 1], P.get("VERTEX_POS_Y", L.ORDER_CANONICAL))
    assert {"vertex_pos", "pos", "vel", "force", "torque", "angular_vel", "angle"} <= set(d["final"])
    assert {"vertex_mass", "vertex_rad", "n_vertices_per_particle", "particle_offset", "moment_inertia"} <= \
        set(d["static"])
