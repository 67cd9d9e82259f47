"""ctypes binding of libdpmd_h5 (include/dpmd/dpmd_h5.h): HDF5 files in the reference's
layout, the meta.h5 particle loader, and readers for meta.h5 / trajectory.h5.

libdpmd_h5.so is built only where an HDF5 C library is installed; `available()` says whether
it was.  Nothing here needs h5py.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib as L

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libdpmd_h5.so"
F64, I32 = 0, 1
CREATE, READ_WRITE, READ_ONLY = 0, 1, 2
_lib: Optional[C.CDLL] = None

_PROTOS = {
    "dpmd_h5_last_error": (C.c_char_p, []),
    "dpmd_h5_open": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int64)]),
    "dpmd_h5_close": (C.c_int, [C.c_int64]),
    "dpmd_h5_flush": (C.c_int, [C.c_int64]),
    "dpmd_h5_exists": (C.c_int, [C.c_int64, C.c_char_p, C.POINTER(C.c_int)]),
    "dpmd_h5_delete": (C.c_int, [C.c_int64, C.c_char_p]),
    "dpmd_h5_move": (C.c_int, [C.c_int64, C.c_char_p, C.c_char_p]),
    "dpmd_h5_write_array": (C.c_int, [C.c_int64, C.c_char_p, C.c_int, C.c_int64, C.c_int, C.c_void_p]),
    "dpmd_h5_write_int": (C.c_int, [C.c_int64, C.c_char_p, C.c_int]),
    "dpmd_h5_write_string": (C.c_int, [C.c_int64, C.c_char_p, C.c_char_p]),
    "dpmd_h5_shape": (C.c_int, [C.c_int64, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int)]),
    "dpmd_h5_read_array": (C.c_int, [C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_int64]),
    "dpmd_h5_read_int": (C.c_int, [C.c_int64, C.c_char_p, C.POINTER(C.c_int)]),
    "dpmd_h5_read_string": (C.c_int, [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64]),
    "dpmd_h5_list": (C.c_int, [C.c_int64, C.c_char_p, C.c_char_p, C.c_int64]),
    "dpmd_h5_write_frame": (C.c_int, [C.c_int64, C.c_char_p, C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_int64]),
    "dpmd_h5_trim_frames": (C.c_int, [C.c_int64, C.c_int64]),
    "dpmd_h5_load": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
}


def available() -> bool:
    """libdpmd_h5 is built and loads (its HDF5 runtime is installed on this machine)."""
    if not LIB_PATH.exists():
        return False
    try:
        lib()
    except OSError:
        return False
    return True


def lib() -> C.CDLL:
    """Loads libdpmd_amd first, so libdpmd_h5 shares its C++ runtime."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (no HDF5 installation was found at build time)")
        L.lib()
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(h, name)
            fn.restype, fn.argtypes = res, args
        _lib = h
    return _lib


class H5Error(RuntimeError):
    pass


def _check(st: int) -> None:
    if st != 0:
        msg = lib().dpmd_h5_last_error().decode()
        raise (ValueError if st == 1 else H5Error)(msg)


def _b(s: str) -> bytes:
    return s.encode()


class H5File:
    """One open HDF5 file (dpmd_h5_open / _close); a context manager."""

    def __init__(self, path, mode: int = READ_ONLY):
        h = C.c_int64()
        _check(lib().dpmd_h5_open(_b(str(path)), mode, C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            _check(lib().dpmd_h5_close(self.h))
            self.h = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def flush(self): _check(lib().dpmd_h5_flush(self.h))
    def delete(self, path: str): _check(lib().dpmd_h5_delete(self.h, _b(path)))
    def move(self, a: str, b: str): _check(lib().dpmd_h5_move(self.h, _b(a), _b(b)))

    def exists(self, path: str) -> bool:
        e = C.c_int()
        _check(lib().dpmd_h5_exists(self.h, _b(path), C.byref(e)))
        return bool(e.value)

    def list(self, group: str = "/"):
        """(name, kind) of each member, kind 'group' or 'dataset'."""
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            st = lib().dpmd_h5_list(self.h, _b(group), buf, cap)
            if st == 4:   # buffer too small
                cap *= 4
                continue
            _check(st)
            out = []
            for ln in buf.value.decode().splitlines():
                kind, name = ln.split(" ", 1)
                out.append((name, kind))
            return out

    def write(self, path: str, a) -> None:
        """(n,) or (n, 2) float64 / int32 array, an int scalar or a str (write_vector[_2d] /
        write_scalar)."""
        if isinstance(a, str):
            _check(lib().dpmd_h5_write_string(self.h, _b(path), _b(a)))
            return
        if isinstance(a, (int, np.integer)):
            _check(lib().dpmd_h5_write_int(self.h, _b(path), int(a)))
            return
        a = np.ascontiguousarray(a)
        t = I32 if a.dtype.kind in "iu" else F64
        a = np.ascontiguousarray(a, np.int32 if t == I32 else np.float64)
        if a.ndim == 1:
            cols = 0
        elif a.ndim == 2 and a.shape[1] == 2:
            cols = 2
        else:
            raise ValueError(f"write: shape {a.shape} is neither (n,) nor (n, 2)")
        _check(lib().dpmd_h5_write_array(self.h, _b(path), t, a.shape[0], cols, a.ctypes.data))

    def write_frame(self, path: str, a: np.ndarray, row: int) -> bool:
        """Frame `row` of a (T, n[, 2]) dataset; False if its shape differs (a ragged frame)."""
        a = np.ascontiguousarray(a)
        t = I32 if a.dtype.kind in "iu" else F64
        a = np.ascontiguousarray(a, np.int32 if t == I32 else np.float64)
        cols = 0 if a.ndim == 1 else 2
        st = lib().dpmd_h5_write_frame(self.h, _b(path), t, a.shape[0], cols, a.ctypes.data, row)
        if st == 1 and "frame shape differs" in lib().dpmd_h5_last_error().decode():
            return False
        _check(st)
        return True

    def trim_frames(self, rows: int): _check(lib().dpmd_h5_trim_frames(self.h, rows))

    def read(self, path: str):
        """A dataset as an ndarray (float64 or int32), an int (int scalar) or a str."""
        nd, t = C.c_int(), C.c_int()
        dims = (C.c_int64 * 3)()
        _check(lib().dpmd_h5_shape(self.h, _b(path), C.byref(nd), dims, C.byref(t)))
        if t.value < 0:
            cap = 1 << 12
            while True:
                buf = C.create_string_buffer(cap)
                st = lib().dpmd_h5_read_string(self.h, _b(path), buf, cap)
                if st == 0:
                    return buf.value.decode()
                if cap > (1 << 24):
                    _check(st)
                cap *= 16
        shape = tuple(int(dims[k]) for k in range(nd.value))
        if nd.value == 0 and t.value == I32:
            v = C.c_int()
            _check(lib().dpmd_h5_read_int(self.h, _b(path), C.byref(v)))
            return v.value
        out = np.empty(shape, np.int32 if t.value == I32 else np.float64)
        _check(lib().dpmd_h5_read_array(self.h, _b(path), t.value, out.ctypes.data, out.size))
        return out if nd.value else out.item()

    def read_group(self, group: str) -> dict:
        out = {}
        for name, kind in self.list(group):
            if kind == "dataset":
                out[name] = self.read(f"{group.rstrip('/')}/{name}")
        return out


def load_from_hdf5(P, run_dir, location: str = "init") -> None:
    """BaseParticle::load_from_hdf5(path, location) (base_particle.hpp:120-160) into a Disk or
    RigidBumpy of dpmd_amd.system."""
    st = lib().dpmd_h5_load(C.c_void_p(P.ctx), _b(str(run_dir)), _b(location))
    _check(st)
    m = C.c_int()
    P._c(L.lib().dpmd_ctx_info(P.ctx, None, C.byref(m)))
    P.neighbor_method = m.value


def read_run(run_dir) -> dict:
    """meta.h5 groups and trajectory.h5 datasets, in the shape dpmd_amd.output.load returns."""
    run_dir = Path(run_dir)
    out: dict = {g: {} for g in ("static", "init", "restart", "final")}
    out["trajectory"], out["ragged"], out["timestep"] = {}, {}, np.zeros(0, np.int64)
    if (run_dir / "meta.h5").exists():
        with H5File(run_dir / "meta.h5") as f:
            for g in out:
                if g in ("static", "init", "restart", "final") and f.exists("/" + g):
                    out[g] = f.read_group("/" + g)
    if (run_dir / "trajectory.h5").exists():
        with H5File(run_dir / "trajectory.h5") as f:
            for name, kind in f.list("/"):
                if name == "timestep":
                    out["timestep"] = f.read("/timestep").reshape(-1).astype(np.int64)
                elif kind == "dataset":
                    out["trajectory"][name] = f.read("/" + name)
                elif name == "ragged":
                    for field, _ in f.list("/ragged"):
                        out["ragged"][field] = f.read_group(f"/ragged/{field}")
    return out
