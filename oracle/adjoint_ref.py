"""FP64 host reference for the disk adjoints -- TEST INFRASTRUCTURE ONLY.

Dense numpy restatement: brute-force contact enumeration per system, the energy / force /
pressure of the harmonic disk model (src/particles/disk.cu:103-246 and the stress of
disk.cu:702-769 as restated in oracle/dpmd_oracle.cpp), the dense Hessian (pair block of the
reference's Hessian observable, disk.cu:565-661: c = k, t = -k delta), analytic gradients,
the implicit derivative through a minimum by a least-squares (minimum-norm) solve, and
reverse mode through velocity Verlet.  Every derivative here is checked against central
finite differences of the C++ oracle in tests/test_oracle_adjoint.py.

All arrays are in canonical order.  A system is given as (x, y, r, m, Lx, Ly, k, wall).
"""
from __future__ import annotations

import numpy as np


def contacts(x, y, r, Lx, Ly, wall):
    """All contacts (i < j) with r_ij < r_i + r_j: arrays i, j, qx, qy (q = x_i - x_j,
    minimum image unless walled), d, delta."""
    n = len(x)
    I, J = np.triu_indices(n, 1)
    qx = x[I] - x[J]
    qy = y[I] - y[J]
    if not wall:
        qx = qx - Lx * np.rint(qx * (1.0 / Lx))
        qy = qy - Ly * np.rint(qy * (1.0 / Ly))
    d2 = qx * qx + qy * qy
    rs = r[I] + r[J]
    on = d2 < rs * rs
    I, J, qx, qy, d2, rs = I[on], J[on], qx[on], qy[on], d2[on], rs[on]
    d = np.sqrt(d2)
    return I, J, qx, qy, d, rs - d


def wall_gaps(x, y, r, Lx, Ly):
    """Wall overlaps per particle and axis: (ox_lo, ox_hi, oy_lo, oy_hi), zero if none."""
    return (np.maximum(r - x, 0.0), np.maximum(r - (Lx - x), 0.0), np.maximum(r - y, 0.0),
            np.maximum(r - (Ly - y), 0.0))


def energy(x, y, r, Lx, Ly, k, wall):
    I, J, qx, qy, d, de = contacts(x, y, r, Lx, Ly, wall)
    U = 0.5 * k * np.sum(de * de)
    if wall:
        U += 0.5 * k * sum(np.sum(o * o) for o in wall_gaps(x, y, r, Lx, Ly))
    return U


def forces(x, y, r, Lx, Ly, k, wall):
    n = len(x)
    I, J, qx, qy, d, de = contacts(x, y, r, Lx, Ly, wall)
    fx, fy = np.zeros(n), np.zeros(n)
    np.add.at(fx, I, k * de * qx / d)
    np.add.at(fy, I, k * de * qy / d)
    np.add.at(fx, J, -k * de * qx / d)
    np.add.at(fy, J, -k * de * qy / d)
    if wall:
        olx, ohx, oly, ohy = wall_gaps(x, y, r, Lx, Ly)
        fx += k * olx - k * ohx
        fy += k * oly - k * ohy
    return fx, fy


def pressure(x, y, r, m, vx, vy, Lx, Ly, k, wall):
    """The reference's pressure: the virial always uses minimum-image contacts, also in
    walled boxes, and has no wall term (disk.cu:702-769)."""
    I, J, qx, qy, d, de = contacts(x, y, r, Lx, Ly, False)
    W = np.sum(k * de * d)
    return (np.sum(m * (vx * vx + vy * vy)) + W) / (2.0 * Lx * Ly)


def touching(x, y, r, Lx, Ly, wall):
    """Contact mask: the particles with at least one contact (pair or wall); the coordinates
    of the others are exact null directions of the Hessian.  Returns the boolean mask."""
    n = len(x)
    I, J, *_ = contacts(x, y, r, Lx, Ly, wall)
    wc = np.zeros(n, int)
    if wall:
        wc = sum((o > 0).astype(int) for o in wall_gaps(x, y, r, Lx, Ly))
    live = np.ones(n, bool)
    while True:
        cnt = wc + np.bincount(I[live[J]], minlength=n) + np.bincount(J[live[I]], minlength=n)
        new = live & (cnt >= 1)
        if np.array_equal(new, live):
            return live
        live = new


def hessian(x, y, r, Lx, Ly, k, wall, live=None):
    """Dense 2n x 2n Hessian of U, ordering (x0, y0, x1, y1, ...).  With a mask, restricted
    to the masked particles and their mutual contacts."""
    n = len(x)
    H = np.zeros((2 * n, 2 * n))
    I, J, qx, qy, d, de = contacts(x, y, r, Lx, Ly, wall)
    for a, b, ux, uy, dd, dl in zip(I, J, qx / d, qy / d, d, de):
        if live is not None and not (live[a] and live[b]):
            continue
        nn = np.array([[ux * ux, ux * uy], [uy * ux, uy * uy]])
        K = k * nn - (k * dl / dd) * (np.eye(2) - nn)
        for p, q, sgn in ((a, a, 1), (b, b, 1), (a, b, -1), (b, a, -1)):
            H[2 * p:2 * p + 2, 2 * q:2 * q + 2] += sgn * K
    if wall:
        olx, ohx, oly, ohy = wall_gaps(x, y, r, Lx, Ly)
        for i in range(n):
            if live is not None and not live[i]:
                continue
            H[2 * i, 2 * i] += k * ((olx[i] > 0) + (ohx[i] > 0))
            H[2 * i + 1, 2 * i + 1] += k * ((oly[i] > 0) + (ohy[i] > 0))
    return H


def grads(x, y, r, m, vx, vy, Lx, Ly, k, wall, live=None):
    """Analytic first derivatives: dU/dx (2n, interleaved), dU/dr, dU/dL (2,), dP/dx, dP/dr,
    dP/dL (box at fixed reduced coordinates, velocities fixed), dF/dr (2n x n), dF/dL (2n x 2).
    With a mask, dF/dr and dF/dL keep only the masked particles' mutual contacts and walls."""
    n = len(x)
    A = Lx * Ly
    I, J, qx, qy, d, de = contacts(x, y, r, Lx, Ly, wall)
    ux, uy = qx / d, qy / d
    dUdx = np.zeros(2 * n)
    fx, fy = forces(x, y, r, Lx, Ly, k, wall)
    dUdx[0::2], dUdx[1::2] = -fx, -fy
    dUdr = np.zeros(n)
    np.add.at(dUdr, I, k * de)
    np.add.at(dUdr, J, k * de)
    # box: dd/dLx = qx^2 / (d Lx)
    cx, cy = qx * qx / (d * Lx), qy * qy / (d * Ly)
    dUdL = np.array([np.sum(-k * de * cx), np.sum(-k * de * cy)])
    # pressure virial over minimum-image contacts (see pressure())
    Ip, Jp, pqx, pqy, pd, pde = contacts(x, y, r, Lx, Ly, False)
    pux, puy = pqx / pd, pqy / pd
    dWdx = np.zeros(2 * n)
    for arr, sgn in ((Ip, 1.0), (Jp, -1.0)):
        np.add.at(dWdx, 2 * arr, sgn * k * (pde - pd) * pux)
        np.add.at(dWdx, 2 * arr + 1, sgn * k * (pde - pd) * puy)
    dWdr = np.zeros(n)
    np.add.at(dWdr, Ip, k * pd)
    np.add.at(dWdr, Jp, k * pd)
    pcx, pcy = pqx * pqx / (pd * Lx), pqy * pqy / (pd * Ly)
    dWdL = np.array([np.sum(k * (pde - pd) * pcx), np.sum(k * (pde - pd) * pcy)])
    dFdr = np.zeros((2 * n, n))
    dFdL = np.zeros((2 * n, 2))
    for a, b, ex, ey, qqx, qqy, dd, dl in zip(I, J, ux, uy, qx, qy, d, de):
        if live is not None and not (live[a] and live[b]):
            continue
        nn = np.array([[ex * ex, ex * ey], [ey * ex, ey * ey]])
        K = k * nn - (k * dl / dd) * (np.eye(2) - nn)
        for p, sgn in ((a, 1.0), (b, -1.0)):
            dFdr[2 * p, a] += sgn * k * ex; dFdr[2 * p + 1, a] += sgn * k * ey
            dFdr[2 * p, b] += sgn * k * ex; dFdr[2 * p + 1, b] += sgn * k * ey
            dFdL[2 * p:2 * p + 2, 0] += sgn * (-K @ np.array([qqx / Lx, 0.0]))
            dFdL[2 * p:2 * p + 2, 1] += sgn * (-K @ np.array([0.0, qqy / Ly]))
    if wall:
        olx, ohx, oly, ohy = wall_gaps(x, y, r, Lx, Ly)
        dUdr += k * (olx + ohx + oly + ohy)
        dUdL += np.array([np.sum(-k * olx * x / Lx) + np.sum(-k * ohx * (Lx - x) / Lx),
                          np.sum(-k * oly * y / Ly) + np.sum(-k * ohy * (Ly - y) / Ly)])
        for i in range(n):
            if live is not None and not live[i]:
                continue
            dFdr[2 * i, i] += k * (float(olx[i] > 0) - float(ohx[i] > 0))
            dFdr[2 * i + 1, i] += k * (float(oly[i] > 0) - float(ohy[i] > 0))
            dFdL[2 * i, 0] += -k * x[i] / Lx * (olx[i] > 0) + k * (Lx - x[i]) / Lx * (ohx[i] > 0)
            dFdL[2 * i + 1, 1] += -k * y[i] / Ly * (oly[i] > 0) + k * (Ly - y[i]) / Ly * (ohy[i] > 0)
    P = pressure(x, y, r, m, vx, vy, Lx, Ly, k, wall)
    dPdx, dPdr = dWdx / (2 * A), dWdr / (2 * A)
    dPdL = dWdL / (2 * A) - P / np.array([Lx, Ly])
    return dict(dUdx=dUdx, dUdr=dUdr, dUdL=dUdL, dPdx=dPdx, dPdr=dPdr, dPdL=dPdL, dFdr=dFdr, dFdL=dFdL, P=P)


def implicit_pressure_grad(x, y, r, m, vx, vy, Lx, Ly, k, wall, live=None):
    """dP/dr and dP/dL through the minimum: dP/dq + lambda^T dF/dq, H lambda = dP/dx, over the
    particles with contacts (or the given mask; minimum-norm least squares, which removes
    translations and the coordinates outside the mask)."""
    if live is None:
        live = touching(x, y, r, Lx, Ly, wall)
    g = grads(x, y, r, m, vx, vy, Lx, Ly, k, wall, live)
    H = hessian(x, y, r, Lx, Ly, k, wall, live)
    lam = np.linalg.lstsq(H, g["dPdx"], rcond=1e-13)[0]
    return g["dPdr"] + lam @ g["dFdr"], g["dPdL"] + lam @ g["dFdL"], lam


def newton_minimum(x, y, r, Lx, Ly, k, wall, iters=20, ftol=1e-15):
    """Newton's method from near a minimum (dense, minimum-norm steps)."""
    x, y = x.copy(), y.copy()
    for _ in range(iters):
        fx, fy = forces(x, y, r, Lx, Ly, k, wall)
        if max(np.max(np.abs(fx)), np.max(np.abs(fy))) <= ftol:
            break
        H = hessian(x, y, r, Lx, Ly, k, wall)
        f = np.empty(2 * len(x)); f[0::2], f[1::2] = fx, fy
        dx = np.linalg.lstsq(H, f, rcond=1e-13)[0]
        x += dx[0::2]; y += dx[1::2]
    return x, y


def vv_trajectory(x, y, vx, vy, r, m, Lx, Ly, k, wall, dt, n):
    """Velocity Verlet (velocity_verlet.cuh:34-41) in numpy; returns positions x_0..x_n and the
    final (x, y, vx, vy)."""
    xs = [np.stack([x, y]).copy()]
    fx, fy = forces(x, y, r, Lx, Ly, k, wall)
    x, y, vx, vy = x.copy(), y.copy(), vx.copy(), vy.copy()
    h = dt * 0.5
    for _ in range(n):
        vx = vx + fx * (h / m); vy = vy + fy * (h / m)
        x = x + vx * dt; y = y + vy * dt
        fx, fy = forces(x, y, r, Lx, Ly, k, wall)
        vx = vx + fx * (h / m); vy = vy + fy * (h / m)
        xs.append(np.stack([x, y]).copy())
    return xs, (x, y, vx, vy)


def vv_backprop(xs, gx, gy, gvx, gvy, r, m, Lx, Ly, k, wall, dt):
    """Reverse mode through vv_trajectory with dense Hessians, step by step (each step's
    three updates written out separately, no merging)."""
    n = len(xs) - 1
    xb = np.empty(2 * len(r)); xb[0::2], xb[1::2] = gx, gy
    vb = np.empty(2 * len(r)); vb[0::2], vb[1::2] = gvx, gvy
    mm = np.repeat(m, 2)
    h = dt * 0.5
    for s in range(n - 1, -1, -1):
        Hn = hessian(xs[s + 1][0], xs[s + 1][1], r, Lx, Ly, k, wall)
        Ho = hessian(xs[s][0], xs[s][1], r, Lx, Ly, k, wall)
        # v' = v_h + h F(x')/m
        vhb = vb.copy()
        xpb = xb - h * (Hn @ (vb / mm))
        # x' = x + dt v_h
        xb = xpb.copy()
        vhb = vhb + dt * xpb
        # v_h = v + h F(x)/m
        vb = vhb.copy()
        xb = xb - h * (Ho @ (vhb / mm))
    return xb[0::2], xb[1::2], vb[0::2], vb[1::2]


# ---- sparse forms (the same functions over k-d-tree contact lists, for large systems) ----
def contacts_sparse(x, y, r, Lx, Ly, wall):
    """contacts() with candidate pairs from a k-d tree (periodic unless walled): the same
    contacts, in the same (i < j, lexicographic) order."""
    from scipy.spatial import cKDTree
    if wall:
        tree = cKDTree(np.stack([x, y], 1))
    else:
        u = x - Lx * np.floor(x / Lx); u[u >= Lx] = 0.0
        v = y - Ly * np.floor(y / Ly); v[v >= Ly] = 0.0
        tree = cKDTree(np.stack([u, v], 1), boxsize=[Lx, Ly])
    P = tree.query_pairs(2.0 * float(np.max(r)) * (1 + 1e-12), output_type="ndarray")
    if len(P) == 0:
        e = np.zeros(0, int)
        return e, e, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0)
    I, J = np.minimum(P[:, 0], P[:, 1]), np.maximum(P[:, 0], P[:, 1])
    o = np.lexsort((J, I))
    I, J = I[o], J[o]
    qx = x[I] - x[J]
    qy = y[I] - y[J]
    if not wall:
        qx = qx - Lx * np.rint(qx * (1.0 / Lx))
        qy = qy - Ly * np.rint(qy * (1.0 / Ly))
    d2 = qx * qx + qy * qy
    rs = r[I] + r[J]
    on = d2 < rs * rs
    I, J, qx, qy, d2, rs = I[on], J[on], qx[on], qy[on], d2[on], rs[on]
    d = np.sqrt(d2)
    return I, J, qx, qy, d, rs - d


def forces_sparse(x, y, r, Lx, Ly, k, wall):
    n = len(x)
    I, J, qx, qy, d, de = contacts_sparse(x, y, r, Lx, Ly, wall)
    fx, fy = np.zeros(n), np.zeros(n)
    np.add.at(fx, I, k * de * qx / d)
    np.add.at(fy, I, k * de * qy / d)
    np.add.at(fx, J, -k * de * qx / d)
    np.add.at(fy, J, -k * de * qy / d)
    if wall:
        olx, ohx, oly, ohy = wall_gaps(x, y, r, Lx, Ly)
        fx += k * olx - k * ohx
        fy += k * oly - k * ohy
    return fx, fy


def hvp_sparse(x, y, r, Lx, Ly, k, wall, w):
    """H w for an interleaved vector w (x0, y0, x1, y1, ...), block by block as hessian()."""
    n = len(x)
    I, J, qx, qy, d, de = contacts_sparse(x, y, r, Lx, Ly, wall)
    ux, uy = qx / d, qy / d
    c, t = k, -(k * de / d)
    wx, wy = w[0::2], w[1::2]
    dwx, dwy = wx[I] - wx[J], wy[I] - wy[J]
    # K = c n n^T + t (I - n n^T)
    un = ux * dwx + uy * dwy
    kx = c * ux * un + t * (dwx - ux * un)
    ky = c * uy * un + t * (dwy - uy * un)
    out = np.zeros(2 * n)
    ox, oy = np.zeros(n), np.zeros(n)
    np.add.at(ox, I, kx); np.add.at(oy, I, ky)
    np.add.at(ox, J, -kx); np.add.at(oy, J, -ky)
    if wall:
        olx, ohx, oly, ohy = wall_gaps(x, y, r, Lx, Ly)
        ox += k * ((olx > 0).astype(float) + (ohx > 0)) * wx
        oy += k * ((oly > 0).astype(float) + (ohy > 0)) * wy
    out[0::2], out[1::2] = ox, oy
    return out


def vv_trajectory_sparse(x, y, vx, vy, r, m, Lx, Ly, k, wall, dt, n):
    """vv_trajectory() with forces_sparse."""
    xs = [np.stack([x, y]).copy()]
    fx, fy = forces_sparse(x, y, r, Lx, Ly, k, wall)
    x, y, vx, vy = x.copy(), y.copy(), vx.copy(), vy.copy()
    h = dt * 0.5
    for _ in range(n):
        vx = vx + fx * (h / m); vy = vy + fy * (h / m)
        x = x + vx * dt; y = y + vy * dt
        fx, fy = forces_sparse(x, y, r, Lx, Ly, k, wall)
        vx = vx + fx * (h / m); vy = vy + fy * (h / m)
        xs.append(np.stack([x, y]).copy())
    return xs, (x, y, vx, vy)


def vv_backprop_sparse(xs, gx, gy, gvx, gvy, r, m, Lx, Ly, k, wall, dt):
    """vv_backprop() with hvp_sparse in place of the dense Hessians."""
    n = len(xs) - 1
    xb = np.empty(2 * len(r)); xb[0::2], xb[1::2] = gx, gy
    vb = np.empty(2 * len(r)); vb[0::2], vb[1::2] = gvx, gvy
    mm = np.repeat(m, 2)
    h = dt * 0.5
    for s in range(n - 1, -1, -1):
        vhb = vb.copy()
        xpb = xb - h * hvp_sparse(xs[s + 1][0], xs[s + 1][1], r, Lx, Ly, k, wall, vb / mm)
        xb = xpb.copy()
        vhb = vhb + dt * xpb
        vb = vhb.copy()
        xb = xb - h * hvp_sparse(xs[s][0], xs[s][1], r, Lx, Ly, k, wall, vhb / mm)
    return xb[0::2], xb[1::2], vb[0::2], vb[1::2]
