"""Host-side tests of trajectory output: the C++ writer's .npy format and log policies
(include/dpmd/output_manager.hpp, through a host-only driver), dpmd_amd.output's
OutputManager and reader over a stand-in particle object, and both writers' HDF5 container
(libdpmd_h5) checked with HDF5's own h5dump where it is installed (no GPU)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dpmd_amd import _lib as L
from dpmd_amd import h5, output

ROOT = Path(__file__).resolve().parent.parent
FORMATS = ["npy", pytest.param("h5", marks=pytest.mark.skipif(not h5.available(), reason="libdpmd_h5 not built"))]


def same(a, b):
    if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
        return np.asarray(a).dtype == np.asarray(b).dtype and np.array_equal(a, b)
    return a == b


def step_of(x) -> int:
    return int(np.asarray(x).reshape(-1)[0])


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

    def n_systems(self): return self.S
    def n_particles(self): return self.N
    def n_vertices(self): return 0

    def compute_pe_total(self): self.calls.append(("pe_total", self.t))
    def compute_packing_fraction(self): self.calls.append(("packing_fraction", self.t))


def run(tmp, workers, steps=30, append=False, fp=None, fmt="npy"):
    P = fp or FakeParticles()
    om = output.OutputManager(P, tmp, max_workers=workers, append_mode=append, format=fmt)
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


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("workers", [1, 3])
def test_python_output_manager_round_trip(tmp_path, workers, fmt):
    P = run(tmp_path / "run", workers, fmt=fmt)
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
    info = {"class_name": "Disk", "neighbor_method": "Cell", "n_particles": P.N, "n_systems": P.S, "n_vertices": 0}
    assert set(d["static"]) == {"system_id", "system_size", "system_offset", "box_size", "cell_dim",
                                "cell_system_start", "neighbor_cutoff", "thresh2", "e_interaction", "mass", "rad",
                                "area"} | set(info)
    assert all(d["static"][k] == v for k, v in info.items())
    assert d["static"]["box_size"].shape == (P.S, 2) and d["static"]["system_offset"].dtype == np.int32
    assert set(d["init"]) == {"pos", "vel", "force"}
    assert set(d["final"]) - {"step"} == {"pos", "vel", "force", "packing_fraction"}
    assert step_of(d["restart"]["step"]) == 30
    P.t = 30
    assert np.array_equal(d["restart"]["vel"][:, 1], P.value("VEL_Y"))
    assert np.array_equal(d["final"]["pos"][:, 0], P.value("POS_X"))


def test_worker_counts_write_identical_files(tmp_path):
    run(tmp_path / "a", 1)
    run(tmp_path / "b", 4)
    for f in sorted((tmp_path / "a").rglob("*")):
        if f.is_file():
            assert f.read_bytes() == (tmp_path / "b" / f.relative_to(tmp_path / "a")).read_bytes(), f


@pytest.mark.parametrize("fmt", FORMATS)
def test_append_mode_extends_trajectory(tmp_path, fmt):
    run(tmp_path / "r", 2, steps=10, fmt=fmt)
    fp = FakeParticles()
    om = output.OutputManager(fp, tmp_path / "r", max_workers=2, append_mode=True, format=fmt)
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
    fp.t = 15
    assert np.array_equal(d["trajectory"]["pos"][2, :, 1], fp.value("POS_Y"))
    # append mode keeps the first run's static and init groups
    assert step_of(d["restart"]["step"]) == 20


@pytest.mark.parametrize("fmt", FORMATS)
def test_ragged_frames(tmp_path, fmt):
    """A frame whose length changes goes to ragged/<field>/step_<k>, as the reference does."""
    fp = FakeParticles()
    om = output.OutputManager(fp, tmp_path / "g", format=fmt)
    om.set_trajectory_fields(["pos"])
    om.set_trajectory_interval(1)
    om.enable_meta = False
    om.initialize()
    for k in range(1, 5):
        fp.t = k
        fp.N = 12 if k != 3 else 9
        om.step(k)
    om.finalize()
    d = output.load(tmp_path / "g")
    assert np.array_equal(d["timestep"], [1, 2, 3, 4])
    assert d["trajectory"]["pos"].shape[1:] == (12, 2)
    assert d["ragged"]["pos"]["step_3"].shape == (9, 2)


@pytest.mark.parametrize("fmt", FORMATS)
def test_unknown_field_is_an_error(tmp_path, fmt):
    om = output.OutputManager(FakeParticles(), tmp_path / "x", format=fmt)
    om.set_trajectory_fields(["no_such_field"])
    om.set_trajectory_interval(1)
    om.initialize()
    with pytest.raises(ValueError, match="no_such_field"):
        om.step(1)
    om.backend.close()


def build_stand_in(tmp_path, fmt):
    exe = tmp_path / f"oft_{fmt}"
    cmd = ["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/output_format_test.cpp"),
           "-o", str(exe), "-lpthread"]
    if fmt == "h5":
        lib = ROOT / "dpmd_amd"
        cmd += ["-DDPMD_WITH_H5", f"-L{lib}", "-ldpmd_h5", "-ldpmd_amd", f"-Wl,-rpath,{lib}"]
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("workers", [1, 4])
def test_cpp_output_manager_matches_python_over_stand_in_abi(tmp_path, workers, fmt):
    """The C++ writer (threads, bounded queue, ordered writes) over a stand-in C ABI writes
    the same arrays as the Python writer over the same closed-form fields."""
    exe = build_stand_in(tmp_path, fmt)
    args = [str(exe), str(tmp_path / "cpp"), "run", str(workers)] + (["h5"] if fmt == "h5" else [])
    r = subprocess.run(args, capture_output=True, text=True, check=True, timeout=60)
    calls = r.stdout.strip().splitlines()
    assert calls[-1] == "error: OutputManager: unknown field 'no_such_field'"
    P = run(tmp_path / "py", workers, fmt=fmt)
    assert calls[:-1] == [f"{n}:{t}" for n, t in P.calls]
    a, b = output.load(tmp_path / "cpp"), output.load(tmp_path / "py")
    assert np.array_equal(a["timestep"], [5, 10, 15, 20, 25, 30])
    for g in ("static", "init", "restart", "final", "trajectory"):
        assert set(a[g]) == set(b[g]), g
        for n in a[g]:
            assert same(a[g][n], b[g][n]), (g, n)


H5DUMP = Path("/opt/conda/bin/h5dump")


@pytest.mark.skipif(not h5.available(), reason="libdpmd_h5 not built")
@pytest.mark.skipif(not H5DUMP.exists(), reason="h5dump not installed")
def test_h5_layout_read_by_h5dump(tmp_path):
    """HDF5's own tool sees the reference's layout: scalar int counts, UTF-8 variable-length
    strings, (N,) and (N, 2) native datasets, (T, 1) int timesteps and (T, N, 2) frames with an
    unlimited first dimension."""
    run(tmp_path / "r", 2, fmt="h5")
    meta = subprocess.run([str(H5DUMP), "-H", str(tmp_path / "r" / "meta.h5")], capture_output=True, text=True,
                          check=True).stdout
    traj = subprocess.run([str(H5DUMP), "-H", str(tmp_path / "r" / "trajectory.h5")], capture_output=True,
                          text=True, check=True).stdout

    def block(text, name):
        i = text.index(f'DATASET "{name}"')
        return text[i:text.index("}", text.index("DATASPACE", i)) + 1]

    assert "H5T_STD_I32LE" in block(meta, "n_particles") and "SCALAR" in block(meta, "n_particles")
    cn = block(meta, "class_name")
    assert "H5T_VARIABLE" in cn and "H5T_CSET_UTF8" in cn
    assert "H5T_IEEE_F64LE" in block(meta, "box_size") and "( 2, 2 )" in block(meta, "box_size")
    assert "H5T_STD_I32LE" in block(meta, "system_offset") and "( 3 )" in block(meta, "system_offset")
    assert "( 6, 1 ) / ( H5S_UNLIMITED, 1 )" in block(traj, "timestep")
    assert "( 6, 12, 2 ) / ( H5S_UNLIMITED, 12, 2 )" in block(traj, "pos")
    assert "GROUP \"restart\"" in meta and "GROUP \"final\"" in meta and "restart_tmp" not in meta
    vals = subprocess.run([str(H5DUMP), "-d", "/static/n_systems", "-d", "/restart/step",
                           str(tmp_path / "r" / "meta.h5")], capture_output=True, text=True, check=True).stdout
    assert "(0): 2" in vals and "(0): 30" in vals


@pytest.mark.skipif(not h5.available(), reason="libdpmd_h5 not built")
def test_h5_binding_errors(tmp_path):
    with h5.H5File(tmp_path / "x.h5", h5.CREATE) as f:
        f.write("/a/b", np.arange(3.0))
        with pytest.raises(ValueError, match="neither"):
            f.write("/a/c", np.zeros((2, 3)))
        with pytest.raises(ValueError, match="missing dataset"):
            f.read("/nope")
        assert f.write_frame("/fr", np.zeros(4), 0) and not f.write_frame("/fr", np.zeros(5), 1)
        assert f.read("/fr").shape == (1, 4)
    with pytest.raises(ValueError, match="no such file"):
        h5.H5File(tmp_path / "missing.h5")
