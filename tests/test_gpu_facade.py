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
    return state, steps, rounds


def test_cpp_facade_matches_python_binding(tmp_path):
    exe = ROOT / "tests" / "cpp" / "facade_test.bin"
    assert exe.exists(), "build() compiles tests/cpp/facade_test.cpp"
    out = tmp_path / "state.bin"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    cpp = np.fromfile(out, dtype=np.float64)
    py, steps, rounds = python_sequence()
    assert f"fire_steps={steps} jam_rounds={rounds}" in r.stdout
    assert cpp.shape == py.shape
    assert np.array_equal(cpp, py)
