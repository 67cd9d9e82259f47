"""Build libdpmd_amd.so (HIP, gfx950) in-tree with hipcc.

Every .hip file under csrc/ is compiled separately (in parallel) and linked into one
shared library next to this file.  Floating-point contraction is off so every a*b+c is
rounded as written (the deterministic-trajectory design relies on it).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libdpmd_amd.so"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DPMD_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-z,defs", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


# ---- optional HDF5 I/O library (libdpmd_h5.so) -----------------------------------------
H5LIB = HERE / "libdpmd_h5.so"
H5DEPS = HERE / "h5lib"     # private links to the HDF5 runtime, so nothing else is taken from its prefix


def find_hdf5():
    """An HDF5 C installation (header and shared library): $DPMD_HDF5_ROOT, /usr, /opt/conda."""
    roots = [os.environ.get("DPMD_HDF5_ROOT", ""), "/usr", "/usr/lib/x86_64-linux-gnu/hdf5/serial", "/opt/conda"]
    for r in filter(None, roots):
        inc, lib = Path(r) / "include", Path(r) / "lib"
        for so in sorted(lib.glob("libhdf5.so.*")) + sorted((lib / "x86_64-linux-gnu").glob("libhdf5*.so.*")):
            if (inc / "hdf5.h").exists() and so.name.count(".") == 2 and "_" not in so.name:
                return inc, so
    return None


def build_h5(verbose: bool = True):
    """libdpmd_h5.so, linked to libdpmd_amd.so and the HDF5 runtime; skipped (and any stale
    copy removed) when no HDF5 installation is found."""
    found = find_hdf5()
    if found is None:
        H5LIB.unlink(missing_ok=True)
        if verbose:
            print("HDF5 not found: libdpmd_h5.so not built", file=sys.stderr)
        return None
    inc, so = found
    H5DEPS.mkdir(exist_ok=True)
    for dep in (so, *sorted(so.parent.glob("libz.so.1"))):
        link = H5DEPS / dep.name
        if link.is_symlink() or link.exists():
            link.unlink()
        link.symlink_to(dep)
    tmp = H5LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", f"-I{HERE.parent / 'include'}", f"-I{inc}",
           str(HERE / "h5" / "dpmd_h5.cpp"), "-o", str(tmp), f"-L{HERE}", "-ldpmd_amd", str(H5DEPS / so.name),
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/h5lib", "-Wl,-z,defs"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"libdpmd_h5 build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5LIB)
    if verbose:
        print(f"built {H5LIB} against {so}", file=sys.stderr)
    return H5LIB


if __name__ == "__main__":
    build()
    build_h5()
