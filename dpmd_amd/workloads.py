"""Synthetic workloads of the batched 2D granular-MD hot path.

Inputs are generated with the reference's stateless Philox4x32-10 recipe
(include/utils/device_fields.cuh:150-298: counter = (index, call, 0, 0), key = seed), here in
vectorised numpy so that host-built inputs are bit-identical to the library's device
generator (dpmd_rand_uniform_2d).

Workloads (sizes as in the benchmark definitions):
  * single system, N = 1024 bidisperse disks (radii 0.5 / 0.7, 50:50), phi = 0.85, PBC;
  * single large system, N = 2**20, phi = 0.84, PBC, skin 0.3;
  * 512 uneven systems, N_s in [64, 4096], phi_s in [0.80, 0.86], half PBC / half walls;
  * rigid bumpy polygons, 256 systems x 256 particles, walls, phi = 0.70.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10; all inputs uint32-valued (any int dtype), returns 4 uint64 arrays."""
    c0, c1, c2, c3 = (np.asarray(v, np.uint64) & _MASK for v in (c0, c1, c2, c3))
    k0 = np.asarray(k0, np.uint64) & _MASK
    k1 = np.asarray(k1, np.uint64) & _MASK
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def _u01d(hi, lo):
    bits = (hi << np.uint64(21)) | (lo >> np.uint64(11))
    return (bits.astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def uniform_2d(n: int, seed: int, call: int):
    """(ux, uy) in [0, 1), the draws of k_stateless_uniform_2d with lo=0, hi=1."""
    i = np.arange(n, dtype=np.uint64)
    c0, c1, c2, c3 = philox4x32_10(i, np.uint64(call), 0, 0, np.uint64(seed & 0xFFFFFFFF), np.uint64(seed >> 32))
    ux, uy = _u01d(c0, c1), _u01d(c2, c3)
    ux[ux == 1.0] = 0.0
    uy[uy == 1.0] = 0.0
    return ux, uy


def uniform_1d(n: int, seed: int, call: int):
    i = np.arange(n, dtype=np.uint64)
    c0, c1, _, _ = philox4x32_10(i, np.uint64(call), 0, 0, np.uint64(seed & 0xFFFFFFFF), np.uint64(seed >> 32))
    u = _u01d(c0, c1)
    u[u == 1.0] = 0.0
    return u


@dataclass
class DiskBatch:
    """Host description of S disk systems concatenated in canonical order."""
    off: np.ndarray
    rad: np.ndarray
    mass: np.ndarray
    area: np.ndarray
    Lx: np.ndarray
    Ly: np.ndarray
    e: np.ndarray
    cutoff: np.ndarray
    thresh2: np.ndarray
    boundary: np.ndarray
    x: np.ndarray
    y: np.ndarray
    cdx: np.ndarray = field(default=None)
    cdy: np.ndarray = field(default=None)

    @property
    def S(self): return len(self.Lx)

    @property
    def N(self): return int(self.off[-1])


def disk_batch(sizes, phis, boundaries, seed: int = 1, call: int = 0, r_small: float = 0.5, r_large: float = 0.7,
               k: float = 1.0, skin: float = 0.3, mass: float = 1.0) -> DiskBatch:
    """Bidisperse disks, 50:50 by canonical index within each system (small first).

    Box side L_s = sqrt(sum(pi r^2) / phi_s) (square boxes); positions U[0, L_s)^2 from
    Philox (seed, call); neighbor cutoff = skin; rebuild threshold = (skin / 2)^2.
    """
    sizes = np.asarray(sizes, np.int64)
    S = len(sizes)
    off = np.zeros(S + 1, np.int32)
    off[1:] = np.cumsum(sizes)
    N = int(off[-1])
    rad = np.empty(N)
    for s in range(S):
        n = int(sizes[s])
        rad[off[s]:off[s] + n // 2] = r_small
        rad[off[s] + n // 2:off[s + 1]] = r_large
    area = math.pi * rad * rad
    L = np.empty(S)
    for s in range(S):
        L[s] = math.sqrt(float(np.sum(area[off[s]:off[s + 1]])) / float(phis[s]))
    ux, uy = uniform_2d(N, seed, call)
    sid = np.repeat(np.arange(S), sizes)
    x = L[sid] * ux
    y = L[sid] * uy
    return DiskBatch(off=off, rad=rad, mass=np.full(N, mass), area=area, Lx=L.copy(), Ly=L.copy(),
                     e=np.full(S, k), cutoff=np.full(S, skin), thresh2=np.full(S, (0.5 * skin) ** 2),
                     boundary=np.asarray(boundaries, np.int32), x=x, y=y,
                     cdx=np.zeros(S, np.int32), cdy=np.zeros(S, np.int32))


def single_disk_system(N: int = 1024, phi: float = 0.85, seed: int = 1, skin: float = 0.3, boundary: int = 0):
    return disk_batch([N], [phi], [boundary], seed=seed, skin=skin)


def uneven_batch(S: int = 512, nmin: int = 64, nmax: int = 4096, phi_lo: float = 0.80, phi_hi: float = 0.86,
                 seed: int = 1, skin: float = 0.3) -> DiskBatch:
    """N_s = nmin + floor(u * (nmax - nmin + 1)) (call 2), phi_s = lo + (hi - lo) u (call 3);
    even systems periodic, odd systems walled."""
    un = uniform_1d(S, seed, 2)
    up = uniform_1d(S, seed, 3)
    sizes = nmin + np.floor(un * (nmax - nmin + 1)).astype(np.int64)
    phis = phi_lo + (phi_hi - phi_lo) * up
    boundaries = (np.arange(S) % 2).astype(np.int32)
    return disk_batch(sizes, phis, boundaries, seed=seed, skin=skin)


def systems_of(b: DiskBatch, s0: int, s1: int) -> DiskBatch:
    """Systems s0..s1-1 of a batch as a batch of their own (same particles, boxes and grids)."""
    lo, hi = int(b.off[s0]), int(b.off[s1])
    sl = lambda v: np.ascontiguousarray(v[s0:s1])
    return DiskBatch(off=np.ascontiguousarray(b.off[s0:s1 + 1] - lo).astype(np.int32), rad=b.rad[lo:hi].copy(),
                     mass=b.mass[lo:hi].copy(), area=b.area[lo:hi].copy(), Lx=sl(b.Lx), Ly=sl(b.Ly), e=sl(b.e),
                     cutoff=sl(b.cutoff), thresh2=sl(b.thresh2), boundary=sl(b.boundary), x=b.x[lo:hi].copy(),
                     y=b.y[lo:hi].copy(), cdx=None if b.cdx is None else sl(b.cdx),
                     cdy=None if b.cdy is None else sl(b.cdy))


def system_of(b: DiskBatch, s: int) -> DiskBatch:
    """System s of a batch as a batch of one."""
    return systems_of(b, s, s + 1)


def load_into(P, b: DiskBatch, method: int = 1) -> None:
    """Allocate and fill a Disk context (dpmd_amd.system.Disk) from a DiskBatch, then sync."""
    P.set_neighbor_method(method)
    P.allocate_systems(b.S)
    P.allocate_particles(b.N)
    P.set("SYSTEM_OFFSET", b.off)
    for name, v in (("BOX_SIZE_X", b.Lx), ("BOX_SIZE_Y", b.Ly), ("E_INTERACTION", b.e),
                    ("NEIGHBOR_CUTOFF", b.cutoff), ("THRESH2", b.thresh2)):
        P.set(name, v)
    P.set("BOUNDARY", b.boundary)
    P.set("CELL_DIM_X", b.cdx)
    P.set("CELL_DIM_Y", b.cdy)
    for name, v in (("POS_X", b.x), ("POS_Y", b.y), ("RAD", b.rad), ("MASS", b.mass), ("AREA", b.area)):
        P.set(name, v)
    P.sync()


@dataclass
class RigidBatch:
    """Host description of S rigid bumpy-polygon systems (canonical particle / vertex order)."""
    off: np.ndarray          # S+1 particle offsets
    nvpp: np.ndarray         # vertices per particle
    x: np.ndarray
    y: np.ndarray
    angle: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    area: np.ndarray
    vx: np.ndarray           # vertex positions
    vy: np.ndarray
    vrad: np.ndarray
    vmass: np.ndarray
    Lx: np.ndarray
    Ly: np.ndarray
    e: np.ndarray
    cutoff: np.ndarray
    thresh2: np.ndarray
    boundary: np.ndarray
    cdx: np.ndarray
    cdy: np.ndarray

    @property
    def S(self): return len(self.Lx)

    @property
    def N(self): return int(self.off[-1])

    @property
    def Nv(self): return int(np.sum(self.nvpp))


def rigid_batch(S: int = 256, n_per: int = 256, phi: float = 0.70, nv_min: int = 12, nv_max: int = 26,
                radii=(0.5, 0.7), seed: int = 1, k: float = 1.0, skin: float = 0.2, boundary: int = 1,
                mass: float = 1.0) -> RigidBatch:
    """Rigid bumpy polygons.  Per particle: n_v = nv_min + floor(u * (nv_max - nv_min + 1))
    (Philox call 4), circle radius R = radii[0] for the first half of each system and radii[1]
    for the second, orientation theta = 2 pi u (call 5).  Vertices at R (cos, sin)(theta + 2 pi
    k / n_v) around the centre; bump radius = half the vertex spacing = R sin(pi / n_v);
    vertex mass m / n_v, so I = sum m_v |r_v|^2 = m R^2; area of the union of polygon and
    bumps = n_v R^2 sin(2 pi / n_v) / 2 + n_v r_b^2 (pi + 2 pi / n_v) / 2.  Square box of side
    sqrt(sum(area) / phi); centres U[0, L)^2 (call 0)."""
    N = S * n_per
    un = uniform_1d(N, seed, 4)
    nvpp = (nv_min + np.floor(un * (nv_max - nv_min + 1))).astype(np.int32)
    R = np.empty(N)
    for s in range(S):
        R[s * n_per:s * n_per + n_per // 2] = radii[0]
        R[s * n_per + n_per // 2:(s + 1) * n_per] = radii[1]
    theta = 2.0 * math.pi * uniform_1d(N, seed, 5)
    rb = R * np.sin(math.pi / nvpp)
    area = 0.5 * nvpp * R * R * np.sin(2 * math.pi / nvpp) + 0.5 * nvpp * rb * rb * (math.pi + 2 * math.pi / nvpp)
    L = np.array([math.sqrt(float(np.sum(area[s * n_per:(s + 1) * n_per])) / phi) for s in range(S)])
    ux, uy = uniform_2d(N, seed, 0)
    sid = np.repeat(np.arange(S), n_per)
    x, y = L[sid] * ux, L[sid] * uy
    pid = np.repeat(np.arange(N), nvpp)
    kk = np.arange(int(nvpp.sum())) - np.repeat(np.concatenate([[0], np.cumsum(nvpp)[:-1]]), nvpp)
    ang = theta[pid] + 2.0 * math.pi * kk / nvpp[pid]
    vx = x[pid] + R[pid] * np.cos(ang)
    vy = y[pid] + R[pid] * np.sin(ang)
    off = np.arange(0, N + 1, n_per, dtype=np.int32)
    return RigidBatch(off=off, nvpp=nvpp, x=x, y=y, angle=theta, mass=np.full(N, mass), inertia=mass * R * R,
                      area=area, vx=vx, vy=vy, vrad=rb[pid], vmass=(mass / nvpp)[pid], Lx=L.copy(), Ly=L.copy(),
                      e=np.full(S, k), cutoff=np.full(S, skin), thresh2=np.full(S, (0.5 * skin) ** 2),
                      boundary=np.full(S, boundary, np.int32), cdx=np.zeros(S, np.int32), cdy=np.zeros(S, np.int32))


def load_rigid_into(P, b: RigidBatch, method: int = 1) -> None:
    """Allocate and fill a RigidBumpy context from a RigidBatch, then sync."""
    P.set_neighbor_method(method)
    P.allocate_systems(b.S)
    P.allocate_particles(b.N)
    P.allocate_vertices(b.Nv)
    P.set("SYSTEM_OFFSET", b.off)
    P.set("N_VERTICES_PER_PARTICLE", b.nvpp)
    for name, v in (("BOX_SIZE_X", b.Lx), ("BOX_SIZE_Y", b.Ly), ("E_INTERACTION", b.e),
                    ("NEIGHBOR_CUTOFF", b.cutoff), ("THRESH2", b.thresh2)):
        P.set(name, v)
    P.set("BOUNDARY", b.boundary)
    P.set("CELL_DIM_X", b.cdx)
    P.set("CELL_DIM_Y", b.cdy)
    for name, v in (("POS_X", b.x), ("POS_Y", b.y), ("ANGLE", b.angle), ("MASS", b.mass),
                    ("MOMENT_INERTIA", b.inertia), ("AREA", b.area), ("VERTEX_POS_X", b.vx),
                    ("VERTEX_POS_Y", b.vy), ("VERTEX_RAD", b.vrad), ("VERTEX_MASS", b.vmass)):
        P.set(name, v)
    P.sync()
