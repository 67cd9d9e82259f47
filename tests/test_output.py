"""Host-side tests of trajectory output: the C++ writer's .npy format and log policies
(include/dpmd/output_manager.hpp, through a host-only driver), and dpmd_amd.output's
OutputManager and reader over a stand-in particle object (no GPU)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dpmd_amd import _lib as L
from dpmd_amd import output

ROOT = Path(__file__).resolve().parent.parent


def test_cpp_npy_writer_and_log_policies(tmp_path):
    exe = tmp_path / "oft"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/output_format_test.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, check=True)
    a = np.load(tmp_path / "f8_5x2.npy", allow_pickle=False)
    assert a.dtype == np.float64 and a.shape == (5, 2)
    assert np.array_equal(a.ravel(), 0.5 * np.arange(10) - 1.0)
    b = np.load(tmp_path / "i4_7.npy", allow_pickle=False)
    assert b.dtype == np.int32 and np.array_equal(b, np.arange(7) ** 2 - 3)
    assert np.load(tmp_path / "empty.npy", allow_pickle=False).shape == (0,)
    lines = dict(ln.split(":", 1) for ln in r.stdout.strip().splitlines())
    # the Python policy gives the same steps: Linear logs multiples of the interval; Logarithmic
    # logs the first step at or past its next step and then moves that past the step
    lin, lg = output.LogPolicy(output.LINEAR, 4), output.LogPolicy(output.LOGARITHMIC, 3)
    assert lines["linear"].split() == [str(s) for s in range(21) if lin.due(s)] == ["0", "4", "8", "12", "16", "20"]
    assert lines["logarithmic"].split() == [str(s) for s in range(0, 21, 2) if lg.due(s)] == \
        ["4", "6", "10", "12", "16", "18"]
    assert lines["disabled"].split() == []


class FakeParticles:
    """Stands in for a Disk: field f at step t, canonical entry k, is f + k/4 + t/1000 (ints:
    f + k + t), the same closed form as the stand-in C ABI of tests/cpp/output_format_test.cpp."""
    kind = L.DISK

    def __init__(self, N=12, S=2, cell=True):
        self.N, self.S, self.t = N, S, 0
        self.neighbor_method = L.NEIGHBOR_CELL if cell else L.NEIGHBOR_NAIVE
        self.calls = []

    def value(self, name):
        f = L.F[name]
        n = self.S + 1 if name in ("SYSTEM_OFFSET", "CELL_SYSTEM_START") else self.S if 32 <= f < 80 else self.N
        if name in L.INT_FIELDS:
            return (f + np.arange(n) + self.t).astype(np.int32)
        return f + 0.25 * np.arange(n) + 0.001 * self.t

    def get(self, name, order=L.ORDER_CURRENT):
        assert order == L.ORDER_CANONICAL
        return self.value(name)

    def compute_pe_total(self): self.calls.append(("pe_total", self.t))
    def compute_packing_fraction(self): self.calls.append(("packing_fraction", self.t))


def run(tmp, workers, steps=30, append=False, fp=None):
    P = fp or FakeParticles()
    om = output.OutputManager(P, tmp, max_workers=workers, append_mode=append)
    om.set_trajectory_fields(["pos", "pe_total", "system_offset"])
    om.set_trajectory_interval(5)
    om.set_restart_log_policy(output.LOGARITHMIC, 8)
    om.set_extra_final_fields(["packing_fraction"])
    om.set_queue_limit(2)
    om.initialize()
    for k in range(1, steps + 1):
        P.t = k
        om.step(k)
    om.finalize()
    return P


@pytest.mark.parametrize("workers", [1, 3])
def test_python_output_manager_round_trip(tmp_path, workers):
    P = run(tmp_path / "run", workers)
    d = output.load(tmp_path / "run")
    assert np.array_equal(d["timestep"], [5, 10, 15, 20, 25, 30])
    pos = d["trajectory"]["pos"]
    assert pos.shape == (6, P.N, 2) and pos.dtype == np.float64
    for t, k in enumerate(d["timestep"]):
        P.t = int(k)
        assert np.array_equal(pos[t, :, 0], P.value("POS_X")) and np.array_equal(pos[t, :, 1], P.value("POS_Y"))
    assert d["trajectory"]["pe_total"].shape == (6, P.S)
    assert np.array_equal(d["trajectory"]["system_offset"][2], [64 + 15, 65 + 15, 66 + 15])
    # preprocess hooks run before each capture, in the caller's thread, at the step captured
    assert [t for n, t in P.calls if n == "pe_total"] == [5, 10, 15, 20, 25, 30]
    assert ("packing_fraction", 30) in P.calls
    assert set(d["static"]) == {"system_id", "system_size", "system_offset", "box_size", "cell_dim",
                                "cell_system_start", "neighbor_cutoff", "thresh2", "e_interaction", "mass", "rad",
                                "area"}
    assert d["static"]["box_size"].shape == (P.S, 2) and d["static"]["system_offset"].dtype == np.int32
    assert set(d["init"]) == {"pos", "vel", "force"}
    assert set(d["final"]) == {"pos", "vel", "force", "packing_fraction", "step"}
    assert int(d["final"]["step"][0]) == 30 and int(d["restart"]["step"][0]) == 30
    P.t = 30
    assert np.array_equal(d["restart"]["vel"][:, 1], P.value("VEL_Y"))


def test_worker_counts_write_identical_files(tmp_path):
    run(tmp_path / "a", 1)
    run(tmp_path / "b", 4)
    for f in sorted((tmp_path / "a").rglob("*")):
        if f.is_file():
            assert f.read_bytes() == (tmp_path / "b" / f.relative_to(tmp_path / "a")).read_bytes(), f


def test_append_mode_extends_trajectory(tmp_path):
    run(tmp_path / "r", 2, steps=10)
    fp = FakeParticles()
    om = output.OutputManager(fp, tmp_path / "r", max_workers=2, append_mode=True)
    om.set_trajectory_fields(["pos", "pe_total"])
    om.set_trajectory_interval(5)
    om.initialize()
    for k in range(11, 21):
        fp.t = k
        om.step(k)
    om.finalize()
    d = output.load(tmp_path / "r")
    assert np.array_equal(d["timestep"], [5, 10, 15, 20])
    assert d["trajectory"]["pos"].shape == (4, fp.N, 2)


def test_unknown_field_is_an_error(tmp_path):
    om = output.OutputManager(FakeParticles(), tmp_path / "x")
    om.set_trajectory_fields(["no_such_field"])
    om.set_trajectory_interval(1)
    om.initialize()
    with pytest.raises(ValueError, match="no_such_field"):
        om.step(1)


@pytest.mark.parametrize("workers", [1, 4])
def test_cpp_output_manager_matches_python_over_stand_in_abi(tmp_path, workers):
    """The C++ writer (threads, bounded queue, ordered writes) over a stand-in C ABI writes
    the same arrays as the Python writer over the same closed-form fields."""
    exe = tmp_path / "oft"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/output_format_test.cpp"),
                    "-o", str(exe), "-lpthread"], check=True)
    r = subprocess.run([str(exe), str(tmp_path / "cpp"), "run", str(workers)], capture_output=True, text=True,
                       check=True, timeout=60)
    calls = r.stdout.strip().splitlines()
    assert calls[-1] == "error: OutputManager: unknown field 'no_such_field'"
    P = run(tmp_path / "py", workers)
    assert calls[:-1] == [f"{n}:{t}" for n, t in P.calls]
    a, b = output.load(tmp_path / "cpp"), output.load(tmp_path / "py")
    assert np.array_equal(a["timestep"], [5, 10, 15, 20, 25, 30])
    for g in ("static", "init", "restart", "final"):
        assert set(a[g]) == set(b[g]), g
        for n in a[g]:
            assert a[g][n].dtype == b[g][n].dtype and np.array_equal(a[g][n], b[g][n]), (g, n)
    assert set(a["trajectory"]) == set(b["trajectory"])
    for n in a["trajectory"]:
        assert a["trajectory"][n].dtype == b["trajectory"][n].dtype
        assert np.array_equal(a["trajectory"][n], b["trajectory"][n]), n
