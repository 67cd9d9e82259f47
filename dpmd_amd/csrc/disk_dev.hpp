// Disk device code shared by the step kernels (disk.hip) and the persistent per-system
// FIRE kernel (fire_persist.hip): argument block, contact forces, velocity mixing and the
// FIRE drift of one particle.
//
// Reference: src/particles/disk.cu (kernels :20-335), src/integrators/fire.cu.  Every
// expression keeps the reference's operand order; the library is built without
// contraction, so each launch shape produces the same bits.
#pragma once
#include "internal.hpp"
#include "fire.hpp"

namespace dpmd {

struct DiskArgs {
    double* x; double* y; double* vx; double* vy; double* fx; double* fy; double* pe;
    const double* rad; const double* mass;
    const int* sid; const int* off;
    const double* Lx; const double* Ly; const double* iLx; const double* iLy; const double* e;
    const double* thresh2; const int* boundary;
    double* lx; double* ly; unsigned* flag; int* any;
    const int* nn; const int* nbr;
    int n; int cell;
};

inline DiskArgs disk_args(dpmd_ctx* c) {
    DiskArgs a;
    a.x = c->x.p; a.y = c->y.p; a.vx = c->vx.p; a.vy = c->vy.p; a.fx = c->fx.p; a.fy = c->fy.p; a.pe = c->pe.p;
    a.rad = c->rad.p; a.mass = c->mass.p; a.sid = c->sid.p; a.off = c->off.p;
    a.Lx = c->Lx.p; a.Ly = c->Ly.p; a.iLx = c->iLx.p; a.iLy = c->iLy.p; a.e = c->e.p;
    a.thresh2 = c->thresh2.p; a.boundary = c->boundary.p;
    a.lx = c->lx.p; a.ly = c->ly.p; a.flag = c->flag.p; a.any = c->run.p + RUN_ANY;
    a.nn = c->nn.p; a.nbr = c->nbr.p; a.n = c->N; a.cell = c->method == DPMD_NEIGHBOR_CELL;
    return a;
}

// Position sources of the contact sum: the context's SoA arrays, or an interleaved (x, y)
// buffer (the ping-pong buffers of fused velocity-Verlet batches), where a neighbor's
// coordinates are one 16-byte gather instead of two 8-byte ones.
struct SoAPos {
    const double* x; const double* y;
    __device__ __forceinline__ double2 at(int j) const { return make_double2(x[j], y[j]); }
    __device__ __forceinline__ double rad(const DiskArgs& a, int j) const { return a.rad[j]; }
};
struct PackedPos {
    const double2* p;
    __device__ __forceinline__ double2 at(int j) const { return p[j]; }
    __device__ __forceinline__ double rad(const DiskArgs& a, int j) const { return a.rad[j]; }
};
// One system's positions and radii held in LDS (index j - beg).
struct LdsPos {
    const double* x; const double* y; const double* r; int beg;
    __device__ __forceinline__ double2 at(int j) const { return make_double2(x[j - beg], y[j - beg]); }
    __device__ __forceinline__ double rad(const DiskArgs&, int j) const { return r[j - beg]; }
};

// compute_pair_forces_kernel (disk.cu:103-160) and compute_wall_forces_kernel
// (disk.cu:163-246).  In walled boxes the pair displacement is xi - xj (repulsive,
// consistent with the 0.5 k delta^2 energy); the reference writes xj - xi there (:218).
template <bool WALL, class Pos>
__device__ __forceinline__ void disk_force_pos(const DiskArgs& a, const Pos& pos, int i, int s, double& fxo,
                                               double& fyo, double& peo) {
    const double ei = a.e[s];
    const double2 pi = pos.at(i);
    const double xi = pi.x, yi = pi.y, ri = pos.rad(a, i);
    double fxi = 0.0, fyi = 0.0, pei = 0.0;
    double Lx = a.Lx[s], Ly = a.Ly[s];
    if (WALL) {
        if (xi < ri) { const double d = ri - xi; fxi += ei * d; pei += (0.5 * ei * d * d); }
        if (xi > Lx - ri) { const double d = ri - (Lx - xi); fxi -= ei * d; pei += (0.5 * ei * d * d); }
        if (yi < ri) { const double d = ri - yi; fyi += ei * d; pei += (0.5 * ei * d * d); }
        if (yi > Ly - ri) { const double d = ri - (Ly - yi); fyi -= ei * d; pei += (0.5 * ei * d * d); }
    }
    const double iLx = a.iLx[s], iLy = a.iLy[s];
    const int cnt = a.nn[i];
    const int* __restrict__ col = a.nbr + i;
    // The first 8 neighbor ids are requested at once (one round trip instead of one per
    // group); then neighbors go in groups of G: every coordinate of a group is requested
    // before any is used.  Contributions are added in list order, so the sums are those of
    // the one-at-a-time loop.
    constexpr int G = 4;
    auto group = [&](int k0, const int (&jj)[G]) {
        double xj[G], yj[G], rj[G];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const double2 q = pos.at(jj[u]);
            xj[u] = q.x; yj[u] = q.y; rj[u] = pos.rad(a, jj[u]);
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (k0 + u >= cnt) break;
            double dx = xi - xj[u], dy = yi - yj[u];
            if (!WALL) { dx = min_image(dx, Lx, iLx); dy = min_image(dy, Ly, iLy); }
            const double r2 = dx * dx + dy * dy;
            const double radsum = ri + rj[u];
            if (r2 >= radsum * radsum) continue;
            const double r = sqrt(r2);
            const double inv = 1.0 / r;
            const double nx = dx * inv, ny = dy * inv;
            const double delta = radsum - r;
            const double fmag = ei * delta;
            fxi += fmag * nx;
            fyi += fmag * ny;
            pei += (0.5 * ei * delta * delta) * 0.5;
        }
    };
    int pre[2 * G];
#pragma unroll
    for (int u = 0; u < 2 * G; ++u) pre[u] = (u < cnt) ? col[(size_t)u * a.n] : i;
    if (cnt > 0) { const int jj[G] = {pre[0], pre[1], pre[2], pre[3]}; group(0, jj); }
    if (cnt > G) { const int jj[G] = {pre[4], pre[5], pre[6], pre[7]}; group(G, jj); }
    for (int k0 = 2 * G; k0 < cnt; k0 += G) {
        int jj[G];
#pragma unroll
        for (int u = 0; u < G; ++u) jj[u] = (k0 + u < cnt) ? col[(size_t)(k0 + u) * a.n] : i;
        group(k0, jj);
    }
    fxo = fxi; fyo = fyi; peo = pei;
}

template <bool WALL>
__device__ __forceinline__ void disk_force_at(const DiskArgs& a, int i, int s, double& fxo, double& fyo, double& peo) {
    disk_force_pos<WALL>(a, SoAPos{a.x, a.y}, i, s, fxo, fyo, peo);
}

template <int MODE, class Pos>
__device__ __forceinline__ void disk_force_mode_pos(const DiskArgs& a, const Pos& pos, int i, int s, double& fx,
                                                    double& fy, double& pe) {
    if (MODE == DPMD_FORCE_PBC) disk_force_pos<false>(a, pos, i, s, fx, fy, pe);
    else if (MODE == DPMD_FORCE_WALL) disk_force_pos<true>(a, pos, i, s, fx, fy, pe);
    else {
        if (a.boundary[s]) disk_force_pos<true>(a, pos, i, s, fx, fy, pe);
        else disk_force_pos<false>(a, pos, i, s, fx, fy, pe);
    }
}

template <int MODE>
__device__ __forceinline__ void disk_force_mode(const DiskArgs& a, int i, int s, double& fx, double& fy, double& pe) {
    if (MODE == DPMD_FORCE_PBC) disk_force_at<false>(a, i, s, fx, fy, pe);
    else if (MODE == DPMD_FORCE_WALL) disk_force_at<true>(a, i, s, fx, fy, pe);
    else {
        if (a.boundary[s]) disk_force_at<true>(a, i, s, fx, fy, pe);
        else disk_force_at<false>(a, i, s, fx, fy, pe);
    }
}

// mix_velocities_and_forces_kernel (disk.cu:307-335): per-particle |v_i|/|F_i| mixing.
__device__ __forceinline__ void disk_mix_one(double& vx, double& vy, double fx, double fy, double a) {
    const double fn = sqrt(fx * fx + fy * fy);
    const double vn = sqrt(vx * vx + vy * vy);
    double ratio = 0.0;
    if (fn > 1e-16) ratio = vn / fn * a;
    else { vx = 0.0; vy = 0.0; ratio = 0.0; }
    vx = vx * (1 - a) + fx * ratio;
    vy = vy * (1 - a) + fy * ratio;
}

// One particle's FIRE drift on values: update_positions(dt_reverse), the skin test,
// scale_velocities(vscale), update_velocities(dt, 0.5), mixing with alpha,
// scale_velocities(vscale), update_positions(dt, 0.5), the skin test.  Returns the test.
__device__ __forceinline__ bool fire_drift_vals(double& x, double& y, double& vx, double& vy, double fx, double fy,
                                                double m, bool cell, double lx, double ly, double th2, double dt,
                                                double alpha, double dtrev, double vscale) {
    // update_positions(dt_reverse, 1.0)
    const double s1 = dtrev * 1.0;
    x += vx * s1; y += vy * s1;
    bool flag = false;
    if (cell) {
        const double dx = x - lx, dy = y - ly;
        flag = (dx * dx + dy * dy) > th2;
    }
    // scale_velocities(velocity_scale)
    vx *= vscale; vy *= vscale;
    // update_velocities(dt, 0.5)
    const double smi = dt * 0.5 / m;
    vx = vx + fx * smi; vy = vy + fy * smi;
    // mix_velocities_and_forces(alpha)
    disk_mix_one(vx, vy, fx, fy, alpha);
    // scale_velocities(velocity_scale)
    vx *= vscale; vy *= vscale;
    // update_positions(dt, 0.5)
    const double s2 = dt * 0.5;
    x += vx * s2; y += vy * s2;
    if (cell) {
        const double dx = x - lx, dy = y - ly;
        flag = flag || ((dx * dx + dy * dy) > th2);
    }
    return flag;
}

// One particle's FIRE drift with the step's scalars (next-step dt, alpha, reverse drift,
// velocity scale): writes x, y, vx, vy and returns the skin test.
__device__ __forceinline__ bool fire_drift_core(const DiskArgs& a, int i, int s, double dt, double alpha, double dtrev,
                                                double vscale) {
    double x = a.x[i], y = a.y[i], vx = a.vx[i], vy = a.vy[i];
    const double lx = a.cell ? a.lx[i] : 0.0, ly = a.cell ? a.ly[i] : 0.0, th2 = a.cell ? a.thresh2[s] : 0.0;
    const bool flag = fire_drift_vals(x, y, vx, vy, a.fx[i], a.fy[i], a.mass[i], a.cell, lx, ly, th2, dt, alpha, dtrev,
                                      vscale);
    a.x[i] = x; a.y[i] = y; a.vx[i] = vx; a.vy[i] = vy;
    return flag;
}

// k_fire_drift's body: the drift with the next-step scalars, which the system's first
// particle then commits as the live ones.
__device__ __forceinline__ bool fire_drift_one(const DiskArgs& a, const FireArgs& f, int i, int s) {
    const double dt = f.s_dt[s], alpha = f.s_alpha[s], dtrev = f.s_dtrev[s], vscale = f.s_vscale[s];
    const bool flag = fire_drift_core(a, i, s, dt, alpha, dtrev, vscale);
    if (i == a.off[s]) {
        f.dt[s] = dt; f.alpha[s] = alpha; f.ngood[s] = f.s_ngood[s]; f.nbad[s] = f.s_nbad[s];
    }
    return flag;
}

}  // namespace dpmd
