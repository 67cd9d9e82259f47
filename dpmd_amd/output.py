"""Trajectory and snapshot output, and its readers.

`OutputManager` mirrors the reference's io::OutputManager<P> (include/utils/output_manager.hpp:
124-1085) for the Python binding; include/dpmd/output_manager.hpp does the same for C++
callers.  Both write one of two containers:

* "h5" (dpmd_amd.h5, needs libdpmd_h5): the reference's meta.h5 / trajectory.h5, which
  BaseParticle::load_from_hdf5 and dpmd_amd.h5.load_from_hdf5 read back;
* "npy" (no dependencies): the same groups as a directory of NumPy files,

      <path>/meta/static/run.json                            class name, neighbor method, sizes
      <path>/meta/{static,init,restart,final}/<field>.npy   one array per field (+ step.npy)
      <path>/trajectory/<field>.bin, <field>.json            frames appended; dtype and frame shape
      <path>/trajectory/timestep.bin                         int64 step of every frame
      <path>/trajectory/ragged/<field>/step_<k>.npy          frames whose shape changed

`load` reads either.  Per-particle fields are written in canonical (load) order, the
reindexing the reference applies when the particles are cell-sorted (output_manager.hpp:
532-574).  Per-system totals are computed before they are read, as the reference's registry
preprocess hooks do.
"""
from __future__ import annotations

import json
import os
import queue
import shutil
import threading
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L

DISABLED, LINEAR, LOGARITHMIC = 0, 1, 2   # io::LogMode (output_manager.hpp:24)

# name -> (x field, y field or None, preprocess method or None)
FIELDS = {
    "pos": ("POS_X", "POS_Y", None), "vel": ("VEL_X", "VEL_Y", None), "force": ("FORCE_X", "FORCE_Y", None),
    "pe": ("PE", None, None), "ke": ("KE", None, "compute_ke"), "rad": ("RAD", None, None),
    "mass": ("MASS", None, None), "area": ("AREA", None, None), "angle": ("ANGLE", None, None),
    "angular_vel": ("ANGULAR_VEL", None, None), "torque": ("TORQUE", None, None),
    "moment_inertia": ("MOMENT_INERTIA", None, None),
    "stress_tensor_x": ("STRESS_XX", "STRESS_XY", "compute_stress_tensor"),
    "stress_tensor_y": ("STRESS_YX", "STRESS_YY", "compute_stress_tensor"),
    "overlaps": ("OVERLAP", None, "compute_overlaps"), "contact_count": ("CONTACT_COUNT", None, "compute_overlaps"),
    "box_size": ("BOX_SIZE_X", "BOX_SIZE_Y", None), "e_interaction": ("E_INTERACTION", None, None),
    "neighbor_cutoff": ("NEIGHBOR_CUTOFF", None, None), "thresh2": ("THRESH2", None, None),
    "packing_fraction": ("PACKING_FRACTION", None, "compute_packing_fraction"),
    "pe_total": ("PE_TOTAL", None, "compute_pe_total"), "ke_total": ("KE_TOTAL", None, "compute_ke_total"),
    "fpower_total": ("FPOWER_TOTAL", None, "compute_fpower_total"),
    "pressure": ("PRESSURE", None, "compute_pressure"), "temperature": ("TEMPERATURE", None, "compute_temperature"),
    "stress_tensor_total_x": ("STRESS_TOTAL_XX", "STRESS_TOTAL_XY", "compute_stress_tensor"),
    "stress_tensor_total_y": ("STRESS_TOTAL_YX", "STRESS_TOTAL_YY", "compute_stress_tensor"),
    "vertex_pos": ("VERTEX_POS_X", "VERTEX_POS_Y", None), "vertex_force": ("VERTEX_FORCE_X", "VERTEX_FORCE_Y", None),
    "vertex_pe": ("VERTEX_PE", None, None), "vertex_rad": ("VERTEX_RAD", None, None),
    "vertex_mass": ("VERTEX_MASS", None, None),
    "system_offset": ("SYSTEM_OFFSET", None, None), "system_size": ("SYSTEM_SIZE", None, None),
    "cell_dim": ("CELL_DIM_X", "CELL_DIM_Y", None), "cell_system_start": ("CELL_SYSTEM_START", None, None),
    "n_contacts_total": ("N_CONTACTS_TOTAL", None, "compute_n_contacts_total"),
    "system_id": ("SYSTEM_ID", None, None), "static_index": ("STATIC_INDEX", None, None),
    "contacts": ("CONTACTS", None, "compute_contacts"), "particle_offset": ("PARTICLE_OFFSET", None, None),
    "n_vertices_per_particle": ("N_VERTICES_PER_PARTICLE", None, None),
    "vertex_particle_id": ("VERTEX_PARTICLE_ID", None, None), "vertex_system_id": ("VERTEX_SYSTEM_ID", None, None),
    "vertex_system_offset": ("VERTEX_SYSTEM_OFFSET", None, None),
}


def static_names(P) -> List[str]:
    """Default static group (base_particle.hpp:746-755, base_point_particle.hpp:246-249,
    base_poly_particle.hpp:414-417, rigid_bumpy.cu:1887-1889)."""
    n = ["system_id", "system_size", "system_offset", "box_size"]
    if P.neighbor_method == L.NEIGHBOR_CELL:
        n += ["cell_dim", "cell_system_start", "neighbor_cutoff", "thresh2"]
    if P.kind == L.RIGID_BUMPY:
        return n + ["e_interaction", "vertex_mass", "vertex_rad", "n_vertices_per_particle", "particle_offset",
                    "vertex_particle_id", "vertex_system_id", "vertex_system_offset", "mass", "moment_inertia", "area"]
    return n + ["e_interaction", "mass", "rad", "area"]


def state_names(P) -> List[str]:
    """Default state group (base_point_particle.hpp:254-257, base_poly_particle.hpp:422-425,
    rigid_bumpy.cu:1891-1893)."""
    if P.kind == L.RIGID_BUMPY:
        return ["vertex_pos", "pos", "vel", "force", "torque", "angular_vel", "angle"]
    return ["pos", "vel", "force"]


def _union(*lists: Sequence[str]) -> List[str]:
    out: List[str] = []
    for lst in lists:
        for x in lst:
            if x not in out:
                out.append(x)
    return out


def capture(P, name: str) -> np.ndarray:
    """One field on the host in canonical order; (n,) or (n, 2)."""
    if name not in FIELDS:
        raise ValueError(f"OutputManager: unknown field '{name}'")
    fx, fy, pre = FIELDS[name]
    if pre:
        getattr(P, pre)()
    a = P.get(fx, L.ORDER_CANONICAL)
    return a if fy is None else np.stack([a, P.get(fy, L.ORDER_CANONICAL)], axis=1)


class LogPolicy:
    """output_manager.hpp:26-31, :621-637."""

    def __init__(self, mode: int = DISABLED, interval: int = 0):
        self.mode = mode if interval > 0 else DISABLED
        self.interval = max(0, interval)
        self.next_log_step = self.interval if mode == LOGARITHMIC else 0

    def due(self, step: int) -> bool:
        if self.mode == LINEAR:
            return self.interval > 0 and step % self.interval == 0
        if self.mode == LOGARITHMIC:
            if self.interval <= 0 or step < self.next_log_step:
                return False
            while True:
                self.next_log_step += self.interval
                if self.next_log_step > step:
                    return True
        return False


def _write_npy(path: Path, a: np.ndarray) -> None:
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as f:
        np.lib.format.write_array(f, np.ascontiguousarray(a), allow_pickle=False)
    os.replace(tmp, path)


def run_info(P) -> dict:
    """What the reference records beside the static fields (output_manager.hpp:248-256)."""
    S, N, Nv = P.n_systems(), P.n_particles(), (P.n_vertices() if P.kind == L.RIGID_BUMPY else 0)
    return {"class_name": "RigidBumpy" if P.kind == L.RIGID_BUMPY else "Disk",
            "neighbor_method": "Cell" if P.neighbor_method == L.NEIGHBOR_CELL else "Naive",
            "n_particles": N, "n_systems": S, "n_vertices": Nv}


class NpyBackend:
    """Directory of NumPy files.  Calls come from one writer at a time."""

    def open(self, path: Path, append: bool, meta: bool, traj: bool) -> bool:
        self.path = path
        existed = (path / "meta" / "static" / "run.json").exists()
        if meta:
            for g in ("static", "init", "final"):
                (path / "meta" / g).mkdir(parents=True, exist_ok=True)
        self.shapes: Dict[str, tuple] = {}
        if traj:
            t = path / "trajectory"
            if not append and t.exists():
                shutil.rmtree(t)
            t.mkdir(parents=True, exist_ok=True)
            for h in t.glob("*.json"):
                meta_ = json.loads(h.read_text())
                self.shapes[h.stem] = (meta_["dtype"], tuple(meta_["shape"]))
        return existed

    def write_info(self, info: dict):
        with open(self.path / "meta" / "static" / "run.json", "w") as f:
            json.dump(info, f)

    def write_group(self, g: str, fields):
        for n, a in fields:
            _write_npy(self.path / "meta" / g / f"{n}.npy", a)

    def write_restart(self, step: int, fields):
        tmp, dst = self.path / "meta" / "restart_tmp", self.path / "meta" / "restart"
        if tmp.exists():
            shutil.rmtree(tmp)
        tmp.mkdir(parents=True)
        for n, a in fields:
            _write_npy(tmp / f"{n}.npy", a)
        _write_npy(tmp / "step.npy", np.array([step], np.int64))
        if dst.exists():
            shutil.rmtree(dst)
        tmp.rename(dst)

    def write_final_step(self, step: int):
        _write_npy(self.path / "meta" / "final" / "step.npy", np.array([step], np.int64))

    def write_frame(self, step: int, fields):
        d = self.path / "trajectory"
        for n, a in fields:
            a = np.ascontiguousarray(a)
            if n not in self.shapes:
                self.shapes[n] = (a.dtype.str, a.shape)
                with open(d / f"{n}.json", "w") as f:
                    f.write(json.dumps({"dtype": a.dtype.str, "shape": list(a.shape)}) + "\n")
            if self.shapes[n] != (a.dtype.str, a.shape):
                (d / "ragged" / n).mkdir(parents=True, exist_ok=True)
                _write_npy(d / "ragged" / n / f"step_{step}.npy", a)
                continue
            with open(d / f"{n}.bin", "ab") as f:
                f.write(a.tobytes())
        with open(d / "timestep.bin", "ab") as f:
            f.write(np.int64(step).tobytes())

    def close(self):
        pass


class H5Backend:
    """The reference's meta.h5 / trajectory.h5 through libdpmd_h5 (output_manager.hpp:169-311,
    :704-932, :1000-1037)."""

    def open(self, path: Path, append: bool, meta: bool, traj: bool) -> bool:
        from . import h5
        self.h5 = h5
        path.mkdir(parents=True, exist_ok=True)
        existed = (path / "meta.h5").exists()
        self.meta = h5.H5File(path / "meta.h5", h5.READ_WRITE) if meta else None
        self.traj, self.rows = None, 0
        if traj:
            if not append and (path / "trajectory.h5").exists():
                (path / "trajectory.h5").unlink()
            self.traj = h5.H5File(path / "trajectory.h5", h5.READ_WRITE)
            if self.traj.exists("/timestep"):
                self.rows = self.traj.read("/timestep").shape[0]
        return existed

    def write_info(self, info: dict):
        for k in ("class_name", "neighbor_method", "n_particles", "n_systems", "n_vertices"):
            self.meta.write(f"/static/{k}", info[k])
        self.meta.flush()

    def write_group(self, g: str, fields):
        for n, a in fields:
            self.meta.write(f"/{g}/{n}", a)
        self.meta.flush()

    def write_restart(self, step: int, fields):
        self.meta.delete("/restart_tmp")
        self.meta.write("/restart_tmp/step", int(step))
        for n, a in fields:
            self.meta.write(f"/restart_tmp/{n}", a)
        self.meta.move("/restart_tmp", "/restart")
        self.meta.flush()

    def write_final_step(self, step: int):
        pass   # the reference's /final holds fields only

    def write_frame(self, step: int, fields):
        self.traj.write_frame("/timestep", np.array([step], np.int32), self.rows)
        for n, a in fields:
            if a.shape[0] == 0 or not self.traj.write_frame(f"/{n}", a, self.rows):
                self.traj.write(f"/ragged/{n}/step_{step}", a)
        self.rows += 1

    def close(self):
        for f in (self.meta, self.traj):
            if f is not None:
                f.close()
        self.meta = self.traj = None


class OutputManager:
    """io::OutputManager<P> (output_manager.hpp:124-445) over the Python binding; format
    "npy" or "h5"."""

    def __init__(self, particles, path, max_workers: int = 1, append_mode: bool = False, format: str = "npy"):
        if format not in ("npy", "h5"):
            raise ValueError(f"OutputManager: unknown format '{format}'")
        self.backend = NpyBackend() if format == "npy" else H5Backend()
        self.P, self.path = particles, Path(path)
        self.max_workers, self.append_mode = max(1, max_workers), append_mode
        self.traj_names: List[str] = []
        self.extra_init: List[str] = []
        self.extra_final: List[str] = []
        self.extra_static: List[str] = []
        self.extra_restart: List[str] = []
        self.console_names: List[str] = []
        self.traj, self.restart, self.console = LogPolicy(), LogPolicy(), LogPolicy()
        self.enable_meta, self.enable_traj, self.enable_console = True, True, False
        self.queue_limit: Optional[int] = None
        self.current_step = 0
        self._q: Optional[queue.Queue] = None
        self._workers: List[threading.Thread] = []
        self._turn = threading.Condition()   # tasks are written in the order they were captured
        self._seq = self._done = 0
        self._error: Optional[BaseException] = None

    # ---- policy (output_manager.hpp:134-157) ----
    def set_trajectory_fields(self, n): self.traj_names = list(n)
    def set_extra_init_fields(self, n): self.extra_init = list(n)
    def set_extra_final_fields(self, n): self.extra_final = list(n)
    def set_extra_static_fields(self, n): self.extra_static = list(n)
    def set_extra_restart_fields(self, n): self.extra_restart = list(n)
    def set_console_fields(self, n): self.console_names = list(n)
    def set_trajectory_log_policy(self, mode, interval): self.traj = LogPolicy(mode, interval)
    def set_restart_log_policy(self, mode, interval): self.restart = LogPolicy(mode, interval)
    def set_console_log_policy(self, mode, interval): self.console = LogPolicy(mode, interval)
    def set_trajectory_interval(self, v): self.set_trajectory_log_policy(LINEAR, v)
    def set_restart_interval(self, v): self.set_restart_log_policy(LINEAR, v)
    def set_console_interval(self, v): self.set_console_log_policy(LINEAR, v)
    def set_queue_limit(self, tasks): self.queue_limit = max(1, int(tasks))

    # ---- run ----
    def initialize(self):
        existed = self.backend.open(self.path, self.append_mode, self.enable_meta, self.enable_traj)
        if self.enable_meta:
            self.backend.write_info(run_info(self.P))
            if not (self.append_mode and existed):
                self.backend.write_group("static", self._capture(_union(static_names(self.P), self.extra_static)))
                self.backend.write_group("init", self._capture(_union(state_names(self.P), self.extra_init)))
        if self.max_workers > 1:
            self._q = queue.Queue(maxsize=self.queue_limit or self.max_workers)
            self._workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(self.max_workers)]
            for t in self._workers:
                t.start()

    def step(self, step_num: int):
        self.current_step = step_num
        if self.enable_console and self.console_names and self.console.due(step_num):
            vals = "  ".join(f"{n}={np.ravel(capture(self.P, n))[:4]}" for n in self.console_names)
            print(f"step {step_num}  {vals}", flush=True)
        if self.enable_traj and self.traj_names and self.traj.due(step_num):
            self._enqueue(("trajectory", step_num, self._capture(self.traj_names)))
        if self.enable_meta and self.restart.due(step_num):
            self._enqueue(("restart", step_num, self._capture(self._restart_names())))

    def finalize(self):
        if self.enable_meta:
            self._enqueue(("restart", self.current_step, self._capture(self._restart_names())))
            self._enqueue(("final", self.current_step, self._capture(_union(state_names(self.P), self.extra_final))))
        for _ in self._workers:
            self._q.put(None)
        for t in self._workers:
            t.join()
        self._workers = []
        self.backend.close()
        if self._error is not None:
            raise self._error

    # ---- internals ----
    def _capture(self, names):
        return [(n, capture(self.P, n)) for n in names]

    def _restart_names(self):
        return _union(static_names(self.P), state_names(self.P), self.extra_static, self.extra_restart)

    def _enqueue(self, task):
        task = (self._seq,) + task
        self._seq += 1
        if not self._workers:
            self._process(task)
        else:
            self._q.put(task)   # blocks while the queue is full

    def _worker(self):
        while True:
            task = self._q.get()
            if task is None:
                return
            try:
                self._process(task)
            except BaseException as e:   # reported by finalize()
                self._error = e

    def _process(self, task):
        seq, kind, step, fields = task
        with self._turn:
            self._turn.wait_for(lambda: self._done == seq)
        try:
            self._write(kind, step, fields)
        finally:
            with self._turn:
                self._done += 1
                self._turn.notify_all()

    def _write(self, kind, step, fields):
        if kind == "trajectory":
            self.backend.write_frame(step, fields)
        elif kind == "restart":
            self.backend.write_restart(step, fields)
        else:
            self.backend.write_group("final", fields)
            self.backend.write_final_step(step)


def load(path) -> dict:
    """Read a run written by either OutputManager, in either container: {"static": {...},
    "init": {...}, "restart": {...}, "final": {...}, "trajectory": {field: (T, ...) array},
    "timestep": (T,), "ragged": {field: {"step_<k>": array}}}.  Scalars of the static group
    (class_name, neighbor_method, n_particles, n_systems, n_vertices) are Python values."""
    path = Path(path)
    if (path / "meta.h5").exists() or (path / "trajectory.h5").exists():
        from . import h5
        return h5.read_run(path)
    out: dict = {}
    for g in ("static", "init", "restart", "final"):
        d = path / "meta" / g
        out[g] = {f.stem: np.load(f, allow_pickle=False) for f in sorted(d.glob("*.npy"))} if d.is_dir() else {}
    info = path / "meta" / "static" / "run.json"
    if info.exists():
        out["static"].update(json.loads(info.read_text()))
    t = path / "trajectory"
    out["trajectory"], out["ragged"], out["timestep"] = {}, {}, np.zeros(0, np.int64)
    if (t / "timestep.bin").exists():
        out["timestep"] = np.fromfile(t / "timestep.bin", dtype="<i8")
    for h in sorted(t.glob("*.json")):
        meta = json.loads(h.read_text())
        a = np.fromfile(t / f"{h.stem}.bin", dtype=np.dtype(meta["dtype"]))
        out["trajectory"][h.stem] = a.reshape([-1] + list(meta["shape"]))
    if (t / "ragged").is_dir():
        for d in sorted((t / "ragged").iterdir()):
            out["ragged"][d.name] = {f.stem: np.load(f, allow_pickle=False) for f in sorted(d.glob("*.npy"))}
    return out
