// Rigid bumpy polygons: vertex contact forces, the vertex -> centre-of-mass reduction
// (force, torque, energy), rigid-body drift with incremental vertex rotation, kicks with the
// moment of inertia, rotational FIRE mixing, and the fused FIRE / velocity-Verlet steps.
//
// Reference: src/particles/rigid_bumpy.cu (kernels :20-370, :507-629, :968-1127).  Particle
// arrays keep canonical order; vertex arrays are reordered by the cell rebuild and reached
// from a particle through static_index[particle_offset[i] + k], as in the reference.  The
// rotation uses own_sincos (common.hpp) so host and device agree to the bit.
#include "internal.hpp"
#include "fire.hpp"

namespace dpmd {
namespace {

struct RigidArgs {
    // particles
    double* x; double* y; double* th; double* vx; double* vy; double* w;
    double* fx; double* fy; double* tq; double* pe;
    const double* mass; const double* inertia;
    const int* sid; const int* off; const int* poff; double* lx; double* ly;
    int n;
    // vertices
    double* vpx; double* vpy; double* vfx; double* vfy; double* vpe; const double* vrad;
    const int* vsid; const int* vstatic; const int* nn; const int* nbr; int nv;
    // systems
    const double* Lx; const double* Ly; const double* iLx; const double* iLy; const double* e;
    const double* thresh2; const int* boundary; unsigned* flag; int* any;
    int cell;
    const int* frozen;   // systems whose FIRE state is final (skipped); null outside fused runs
};

RigidArgs rigid_args(dpmd_ctx* c) {
    RigidArgs a;
    a.x = c->x.p; a.y = c->y.p; a.th = c->angle.p; a.vx = c->vx.p; a.vy = c->vy.p; a.w = c->avel.p;
    a.fx = c->fx.p; a.fy = c->fy.p; a.tq = c->torque.p; a.pe = c->pe.p;
    a.mass = c->mass.p; a.inertia = c->inertia.p; a.sid = c->sid.p; a.off = c->off.p; a.poff = c->poff.p;
    a.lx = c->lx.p; a.ly = c->ly.p; a.n = c->N;
    a.vpx = c->vpx.p; a.vpy = c->vpy.p; a.vfx = c->vfx.p; a.vfy = c->vfy.p; a.vpe = c->vpe.p; a.vrad = c->vrad.p;
    a.vsid = c->vsid.p; a.vstatic = c->static_index.p; a.nn = c->nn.p; a.nbr = c->nbr.p; a.nv = c->Nv;
    a.Lx = c->Lx.p; a.Ly = c->Ly.p; a.iLx = c->iLx.p; a.iLy = c->iLy.p; a.e = c->e.p;
    a.thresh2 = c->thresh2.p; a.boundary = c->boundary.p; a.flag = c->flag.p; a.any = c->run.p + RUN_ANY;
    a.cell = c->method == DPMD_NEIGHBOR_CELL;
    a.frozen = c->mask_frozen ? c->f_frozen.p : nullptr;
    return a;
}

// compute_pair_forces_kernel (rigid_bumpy.cu:164-221) / compute_wall_forces_kernel (:224-307);
// in walled boxes the pair displacement is xi - xj (repulsive), the reference writes xj - xi.
// PE: also the vertex energy (the steps inside a fused velocity-Verlet batch skip it: their
// energies are never read, and the batch's last step and a halted step compute them).
template <bool WALL, bool PE = true>
__device__ __forceinline__ void vertex_force_at(const RigidArgs& a, int i, int s) {
    const double ei = a.e[s];
    const double xi = a.vpx[i], yi = a.vpy[i], ri = a.vrad[i];
    double fxi = 0.0, fyi = 0.0, pei = 0.0;
    const double Lx = a.Lx[s], Ly = a.Ly[s];
    if (WALL) {
        if (xi < ri) { const double d = ri - xi; fxi += ei * d; pei += (0.5 * ei * d * d); }
        if (xi > Lx - ri) { const double d = ri - (Lx - xi); fxi -= ei * d; pei += (0.5 * ei * d * d); }
        if (yi < ri) { const double d = ri - yi; fyi += ei * d; pei += (0.5 * ei * d * d); }
        if (yi > Ly - ri) { const double d = ri - (Ly - yi); fyi -= ei * d; pei += (0.5 * ei * d * d); }
    }
    const double iLx = a.iLx[s], iLy = a.iLy[s];
    const int cnt = a.nn[i];
    const int* __restrict__ col = a.nbr + i;
    // neighbors in groups of four, every load of a group issued before any is used; the
    // contributions are added in list order
    constexpr int G = 4;
    for (int k0 = 0; k0 < cnt; k0 += G) {
        int jj[G];
        double xj[G], yj[G], rj[G];
#pragma unroll
        for (int u = 0; u < G; ++u) jj[u] = (k0 + u < cnt) ? col[(size_t)(k0 + u) * a.nv] : i;
#pragma unroll
        for (int u = 0; u < G; ++u) { xj[u] = a.vpx[jj[u]]; yj[u] = a.vpy[jj[u]]; rj[u] = a.vrad[jj[u]]; }
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
    }
    a.vfx[i] = fxi; a.vfy[i] = fyi;
    if (PE) a.vpe[i] = pei;
}

template <int MODE, bool PE = true>
__global__ void __launch_bounds__(DPMD_BLOCK) k_vertex_force(RigidArgs a, Gate g) {
    if (!gate_force(g)) return;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * DPMD_BLOCK + threadIdx.x;
    if (i >= a.nv) return;
    const int s = a.vsid[i];
    if (a.frozen && a.frozen[s]) return;
    if (MODE == DPMD_FORCE_PBC) vertex_force_at<false, PE>(a, i, s);
    else if (MODE == DPMD_FORCE_WALL) vertex_force_at<true, PE>(a, i, s);
    else if (a.boundary[s]) vertex_force_at<true, PE>(a, i, s);
    else vertex_force_at<false, PE>(a, i, s);
}

// compute_particle_forces_kernel (rigid_bumpy.cu:331-370): canonical vertex order.
__device__ __forceinline__ void particle_force_one(const RigidArgs& a, int i, double& fxo, double& fyo, double& to,
                                                   double& peo) {
    const double px = a.x[i], py = a.y[i];
    double fxi = 0.0, fyi = 0.0, t = 0.0, p = 0.0;
    const int jb = a.poff[i], je = a.poff[i + 1];
    // vertices in groups of eight: the loads of a group are issued before the sums, which
    // run in canonical vertex order (a launch has one thread per particle, so few waves;
    // deeper groups mean fewer dependent round trips per particle)
    constexpr int G = 8;
    for (int j0 = jb; j0 < je; j0 += G) {
        int v[G];
        double fvx[G], fvy[G], pv[G], vx[G], vy[G];
#pragma unroll
        for (int u = 0; u < G; ++u) v[u] = a.vstatic[j0 + u < je ? j0 + u : j0];
#pragma unroll
        for (int u = 0; u < G; ++u) {
            fvx[u] = a.vfx[v[u]]; fvy[u] = a.vfy[v[u]]; pv[u] = a.vpe[v[u]];
            vx[u] = a.vpx[v[u]]; vy[u] = a.vpy[v[u]];
        }
#pragma unroll
        for (int u = 0; u < G; ++u) {
            if (j0 + u >= je) break;
            fxi += fvx[u];
            fyi += fvy[u];
            p += pv[u];
            const double dx = vx[u] - px, dy = vy[u] - py;
            t += fvy[u] * dx - fvx[u] * dy;
        }
    }
    fxo = fxi; fyo = fyi; to = t; peo = p;
}

__global__ void k_particle_force(RigidArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double fx, fy, t, p;
    particle_force_one(a, i, fx, fy, t, p);
    a.fx[i] = fx; a.fy[i] = fy; a.tq[i] = t; a.pe[i] = p;
}

// update_velocities_kernel (rigid_bumpy.cu:130-161) for one particle
__device__ __forceinline__ void kick_one(const RigidArgs& a, int i, double sc, double fx, double fy, double t,
                                         double& vx, double& vy, double& w) {
    const double I = a.inertia[i];
    const double smi = sc / a.mass[i];
    const double sii = I != 0.0 ? sc / I : 0.0;
    vx = vx + fx * smi;
    vy = vy + fy * smi;
    w = w + t * sii;
}

// update_positions_kernel (rigid_bumpy.cu:20-127) for one particle: COM and angle drift,
// vertices rotated by the angle increment about the old COM; returns the skin test of the
// first vertex.
__device__ __forceinline__ bool drift_one(const RigidArgs& a, int i, int s, double sc, double& x, double& y,
                                          double& th, double vx, double vy, double w) {
    const double x0 = x, y0 = y, a0 = th;
    const double nx = x0 + vx * sc, ny = y0 + vy * sc, na = a0 + w * sc;
    x = nx; y = ny; th = na;
    double sn, cs;
    own_sincos(na - a0, &sn, &cs);
    const int b = a.poff[i], e = a.poff[i + 1];
    for (int j = b; j < e; ++j) {
        const int v = a.vstatic[j];
        const double dx = a.vpx[v] - x0, dy = a.vpy[v] - y0;
        const double rx = cs * dx - sn * dy, ry = sn * dx + cs * dy;
        a.vpx[v] = nx + rx;
        a.vpy[v] = ny + ry;
    }
    if (!a.cell) return false;
    const int fv = a.vstatic[b];
    const double dx = a.vpx[fv] - a.lx[i], dy = a.vpy[fv] - a.ly[i];
    return (dx * dx + dy * dy) > a.thresh2[s];
}

__global__ void k_update_positions(RigidArgs a, const double* __restrict__ scale, double scale2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid) {
        double x = a.x[i], y = a.y[i], th = a.th[i];
        flag = drift_one(a, i, s, scale[s] * scale2, x, y, th, a.vx[i], a.vy[i], a.w[i]);
        a.x[i] = x; a.y[i] = y; a.th[i] = th;
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

__global__ void k_update_velocities(RigidArgs a, const double* __restrict__ scale, double scale2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
    kick_one(a, i, scale[a.sid[i]] * scale2, a.fx[i], a.fy[i], a.tq[i], vx, vy, w);
    a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
}

// compute_damping_forces_kernel (rigid_bumpy.cu:310-327): F += -v s, torque += -omega s
__global__ void k_damping(RigidArgs a, const double* __restrict__ scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double sc = scale[a.sid[i]];
    a.fx[i] += -a.vx[i] * sc;
    a.fy[i] += -a.vy[i] * sc;
    a.tq[i] += -a.w[i] * sc;
}

__global__ void k_scale_velocities(RigidArgs a, const double* __restrict__ scale) {   // :540-553
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double sc = scale[a.sid[i]];
    a.vx[i] *= sc; a.vy[i] *= sc; a.w[i] *= sc;
}

__global__ void k_scale_positions(RigidArgs a, const double* __restrict__ scale) {    // :555-583
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double px = a.x[i], py = a.y[i];
    const double sc = scale[a.sid[i]];
    const double nx = px * sc, ny = py * sc;
    a.x[i] = nx; a.y[i] = ny;
    for (int j = a.poff[i]; j < a.poff[i + 1]; ++j) {
        const int v = a.vstatic[j];
        a.vpx[v] += (nx - px);
        a.vpy[v] += (ny - py);
    }
}

// mix_velocities_and_forces_kernel (rigid_bumpy.cu:586-629): no-op when alpha == 0.
__device__ __forceinline__ void mix_one(double& vx, double& vy, double& w, double fx, double fy, double t, double a) {
    if (a == 0.0) return;
    double vxi = vx, vyi = vy, wi = w;
    const double fn = sqrt(fx * fx + fy * fy), vn = sqrt(vxi * vxi + vyi * vyi);
    double ratio = 0.0;
    if (fn > 1e-16) ratio = vn / fn * a;
    else { vxi = 0.0; vyi = 0.0; }
    const double tn = fabs(t), wn = fabs(wi);
    double tratio = 0.0;
    if (tn > 1e-16) tratio = wn / tn * a;
    else { wi = 0.0; tratio = 0.0; }
    vx = vxi * (1 - a) + fx * ratio;
    vy = vyi * (1 - a) + fy * ratio;
    w = wi * (1 - a) + t * tratio;
}

__global__ void k_mix(RigidArgs a, const double* __restrict__ alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
    mix_one(vx, vy, w, a.fx[i], a.fy[i], a.tq[i], alpha[a.sid[i]]);
    a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
}

__global__ void k_ke(RigidArgs a, double* ke) {   // compute_ke_kernel (rigid_bumpy.cu:507-524)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
    ke[i] = 0.5 * ((vx * vx + vy * vy) * a.mass[i] + w * w * a.inertia[i]);
}

// compute_stress_tensor_virial_kernel + _kinetic_kernel (rigid_bumpy.cu:968-1068): vertex
// force times COM separation / 2, summed per particle in canonical vertex order (the
// reference adds them with atomics), plus m v v, over the box area.
__global__ void k_stress(RigidArgs a, const int* vpid, double* sxx, double* sxy, double* syx, double* syy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double ei = a.e[s], Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xi = a.x[i], yi = a.y[i];
    double axx = 0.0, axy = 0.0, ayx = 0.0, ayy = 0.0;
    for (int jc = a.poff[i]; jc < a.poff[i + 1]; ++jc) {
        const int v = a.vstatic[jc];
        const double vx_ = a.vpx[v], vy_ = a.vpy[v], rv = a.vrad[v];
        double vxx = 0.0, vxy = 0.0, vyx = 0.0, vyy = 0.0;
        const int cnt = a.nn[v];
        for (int k = 0; k < cnt; ++k) {
            const int u = a.nbr[(size_t)k * a.nv + v];
            const double dx = min_image(vx_ - a.vpx[u], Lx, iLx);
            const double dy = min_image(vy_ - a.vpy[u], Ly, iLy);
            const int pu = vpid[u];
            const double dxp = min_image(xi - a.x[pu], Lx, iLx);
            const double dyp = min_image(yi - a.y[pu], Ly, iLy);
            const double r2 = dx * dx + dy * dy;
            const double radsum = rv + a.vrad[u];
            if (r2 >= radsum * radsum) continue;
            const double r = sqrt(r2);
            const double inv = 1.0 / r;
            const double delta = radsum - r;
            const double fmag = ei * delta;
            const double fxv = fmag * (dx * inv), fyv = fmag * (dy * inv);
            vxx += dxp * fxv / 2.0; vxy += dxp * fyv / 2.0;
            vyx += dyp * fxv / 2.0; vyy += dyp * fyv / 2.0;
        }
        axx += vxx; axy += vxy; ayx += vyx; ayy += vyy;
    }
    const double area_box = Lx * Ly, m = a.mass[i], vx = a.vx[i], vy = a.vy[i];
    sxx[i] = (axx + m * vx * vx) / area_box;
    sxy[i] = (axy + m * vx * vy) / area_box;
    syx[i] = (ayx + m * vy * vx) / area_box;
    syy[i] = (ayy + m * vy * vy) / area_box;
}

// ---- fused FIRE (fire.cuh:59-88) ---------------------------------------------------------
__global__ void __launch_bounds__(DPMD_BLOCK) k_fire_drift(RigidArgs a, FireArgs f, Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid && !fire_frozen(f, s)) {
        const double dt = f.s_dt[s], alpha = f.s_alpha[s], dtrev = f.s_dtrev[s], vscale = f.s_vscale[s];
        double x = a.x[i], y = a.y[i], th = a.th[i], vx = a.vx[i], vy = a.vy[i], w = a.w[i];
        const double fx = a.fx[i], fy = a.fy[i], t = a.tq[i];
        flag = drift_one(a, i, s, dtrev * 1.0, x, y, th, vx, vy, w);            // update_positions(dt_rev, 1)
        vx *= vscale; vy *= vscale; w *= vscale;                                  // scale_velocities
        kick_one(a, i, dt * 0.5, fx, fy, t, vx, vy, w);                           // update_velocities(dt, 0.5)
        mix_one(vx, vy, w, fx, fy, t, alpha);                                     // mix
        vx *= vscale; vy *= vscale; w *= vscale;                                  // scale_velocities
        flag = drift_one(a, i, s, dt * 0.5, x, y, th, vx, vy, w) || flag;       // update_positions(dt, 0.5)
        a.x[i] = x; a.y[i] = y; a.th[i] = th; a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
        if (i == a.off[s]) {
            f.dt[s] = dt; f.alpha[s] = alpha; f.ngood[s] = f.s_ngood[s]; f.nbad[s] = f.s_nbad[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_particle(RigidArgs a, FireArgs f, RedArgs r, int kick, Gate g) {
    __shared__ double lds[12];
    if (!gate_open(g)) return;
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    const int s = r.chunk_sys[c];
    const int beg = r.chunk_beg[c];
    if (fire_frozen(f, s)) return;   // whole workgroup: chunks never straddle systems
    const int end = min(beg + DPMD_BLOCK, a.off[s + 1]);
    const int i = beg + threadIdx.x;
    double vp = 0.0, vpe = 0.0, vf2 = 0.0;
    if (i < end) {
        double fx, fy, t, p;
        particle_force_one(a, i, fx, fy, t, p);
        a.fx[i] = fx; a.fy[i] = fy; a.tq[i] = t; a.pe[i] = p;
        double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
        if (kick) {
            kick_one(a, i, f.dt[s] * 0.5, fx, fy, t, vx, vy, w);
            a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
        }
        vp = fx * vx + fy * vy + t * w;
        vpe = p;
        vf2 = fx * fx + fy * fy;
    }
    // block_sum256 / block_max256 of the three values with one barrier
    vp = wave_sum(vp); vpe = wave_sum(vpe); vf2 = wave_max(vf2);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { lds[w] = vp; lds[4 + w] = vpe; lds[8 + w] = vf2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        r.part[c] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
        r.part[r.nch + c] = (lds[4] + lds[5]) + (lds[6] + lds[7]);
        r.part[2 * r.nch + c] = fmax(fmax(lds[8], lds[9]), fmax(lds[10], lds[11]));
    }
}

// ---- fused velocity Verlet (velocity_verlet.cuh:34-41) -----------------------------------
__global__ void __launch_bounds__(DPMD_BLOCK)
k_vv_drift(RigidArgs a, const double* __restrict__ dt, const double* __restrict__ dth, Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid) {
        double x = a.x[i], y = a.y[i], th = a.th[i], vx = a.vx[i], vy = a.vy[i], w = a.w[i];
        kick_one(a, i, dth[s] * 1.0, a.fx[i], a.fy[i], a.tq[i], vx, vy, w);
        flag = drift_one(a, i, s, dt[s] * 1.0, x, y, th, vx, vy, w);
        a.x[i] = x; a.y[i] = y; a.th[i] = th; a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

__global__ void __launch_bounds__(DPMD_BLOCK)
k_vv_particle(RigidArgs a, const double* __restrict__ dth, const double* __restrict__ damp, Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double fx, fy, t, p;
    particle_force_one(a, i, fx, fy, t, p);
    double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
    if (damp) {   // compute_damping_forces_kernel (rigid_bumpy.cu:310-327) at the half-step velocity
        const double sc = damp[a.sid[i]];
        fx += -vx * sc;
        fy += -vy * sc;
        t += -w * sc;
    }
    a.fx[i] = fx; a.fy[i] = fy; a.tq[i] = t; a.pe[i] = p;
    kick_one(a, i, dth[a.sid[i]] * 1.0, fx, fy, t, vx, vy, w);
    a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
}

// ---- velocity Verlet with one thread per vertex ------------------------------------------
// A workgroup takes G consecutive particles (G = DPMD_BLOCK / most vertices per particle, so
// their vertices fit one thread each).  Per-particle work (kicks, the centre and angle drift,
// the rotation's sin and cos, the final sums) runs on G threads; per-vertex work (rotating a
// vertex, loading its force and forming its torque term) on one thread per vertex, with the
// particle's values passed through LDS.  Sums still run over the vertices in canonical order
// on one thread, so the results are those of the one-thread-per-particle kernels.
constexpr int kRigidGroupMax = DPMD_BLOCK;

__global__ void __launch_bounds__(DPMD_BLOCK)
k_vv_drift_v(RigidArgs a, const double* __restrict__ dt, const double* __restrict__ dth, int G, Gate g) {
    if (!gate_open(g)) return;
    __shared__ double l_x0[kRigidGroupMax], l_y0[kRigidGroupMax], l_nx[kRigidGroupMax], l_ny[kRigidGroupMax];
    __shared__ double l_cs[kRigidGroupMax], l_sn[kRigidGroupMax];
    __shared__ int l_off[kRigidGroupMax + 1];
    const int p0 = blockIdx.x * G;
    const int np = min(G, a.n - p0);
    const int t = threadIdx.x;
    if (t < np) {
        const int i = p0 + t;
        const int s = a.sid[i];
        double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
        kick_one(a, i, dth[s] * 1.0, a.fx[i], a.fy[i], a.tq[i], vx, vy, w);
        const double sc = dt[s] * 1.0;
        const double x0 = a.x[i], y0 = a.y[i], a0 = a.th[i];
        const double nx = x0 + vx * sc, ny = y0 + vy * sc, na = a0 + w * sc;
        double sn, cs;
        own_sincos(na - a0, &sn, &cs);
        a.x[i] = nx; a.y[i] = ny; a.th[i] = na; a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
        l_x0[t] = x0; l_y0[t] = y0; l_nx[t] = nx; l_ny[t] = ny; l_cs[t] = cs; l_sn[t] = sn;
    }
    if (t <= np) l_off[t] = a.poff[p0 + t];
    __syncthreads();
    const int c = l_off[0] + t;
    const bool valid = c < l_off[np];
    int q = 0, s = -1;
    bool flag = false;
    if (valid) {
        while (l_off[q + 1] <= c) ++q;
        const int i = p0 + q;
        s = a.sid[i];
        const int v = a.vstatic[c];
        const double dx = a.vpx[v] - l_x0[q], dy = a.vpy[v] - l_y0[q];
        const double cs = l_cs[q], sn = l_sn[q];
        const double rx = cs * dx - sn * dy, ry = sn * dx + cs * dy;
        const double vxn = l_nx[q] + rx, vyn = l_ny[q] + ry;
        a.vpx[v] = vxn;
        a.vpy[v] = vyn;
        if (a.cell && c == l_off[q]) {   // the skin test uses the particle's first vertex
            const double ddx = vxn - a.lx[i], ddy = vyn - a.ly[i];
            flag = (ddx * ddx + ddy * ddy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// Centre force, torque and energy of G particles (particle_force_one's sums), then, with
// kick, the closing half kick (k_vv_particle).
__global__ void __launch_bounds__(DPMD_BLOCK)
k_particle_v(RigidArgs a, const double* __restrict__ dth, const double* __restrict__ damp, int G, int kick, Gate g) {
    if (!gate_open(g)) return;
    __shared__ double l_fx[DPMD_BLOCK], l_fy[DPMD_BLOCK], l_pe[DPMD_BLOCK], l_tq[DPMD_BLOCK];
    __shared__ int l_off[kRigidGroupMax + 1];
    const int p0 = blockIdx.x * G;
    const int np = min(G, a.n - p0);
    const int t = threadIdx.x;
    if (t <= np) l_off[t] = a.poff[p0 + t];
    __syncthreads();
    const int c = l_off[0] + t;
    if (c < l_off[np]) {
        int q = 0;
        while (l_off[q + 1] <= c) ++q;
        const int i = p0 + q;
        const int v = a.vstatic[c];
        const double fvx = a.vfx[v], fvy = a.vfy[v];
        const double dx = a.vpx[v] - a.x[i], dy = a.vpy[v] - a.y[i];
        l_fx[t] = fvx; l_fy[t] = fvy; l_pe[t] = a.vpe[v];
        l_tq[t] = fvy * dx - fvx * dy;
    }
    __syncthreads();
    if (t >= np) return;
    const int i = p0 + t;
    double fx = 0.0, fy = 0.0, tq = 0.0, pe = 0.0;
    for (int k = l_off[t] - l_off[0]; k < l_off[t + 1] - l_off[0]; ++k) {
        fx += l_fx[k];
        fy += l_fy[k];
        pe += l_pe[k];
        tq += l_tq[k];
    }
    double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
    if (damp) {   // compute_damping_forces_kernel (rigid_bumpy.cu:310-327) at the half-step velocity
        const double sc = damp[a.sid[i]];
        fx += -vx * sc;
        fy += -vy * sc;
        tq += -w * sc;
    }
    a.fx[i] = fx; a.fy[i] = fy; a.tq[i] = tq; a.pe[i] = pe;
    if (kick) {
        kick_one(a, i, dth[a.sid[i]] * 1.0, fx, fy, tq, vx, vy, w);
        a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
    }
}

// Velocity Verlet across a step boundary for G particles: k_particle_v's sums and closing half
// kick, then k_vv_drift_v's opening half kick, drift, vertex rotation and skin test, each in
// the same expressions and order, so a fused run equals the two-launch run bit for bit.  The
// particle energies of these steps are not formed (nothing reads them: a batch's last step
// and a halted step run k_particle_v, which does).  Each
// vertex position is loaded once, for its torque term and its rotation; a workgroup reads and
// writes only its own particles' vertices, so they are updated in place.
__global__ void __launch_bounds__(DPMD_BLOCK)
k_vv_fused_v(RigidArgs a, const double* __restrict__ dt, const double* __restrict__ dth,
             const double* __restrict__ damp, int G, Gate g) {
    if (!gate_open(g)) return;
    __shared__ double l_fx[DPMD_BLOCK], l_fy[DPMD_BLOCK], l_tq[DPMD_BLOCK];
    __shared__ double l_x0[kRigidGroupMax], l_y0[kRigidGroupMax], l_nx[kRigidGroupMax], l_ny[kRigidGroupMax];
    __shared__ double l_cs[kRigidGroupMax], l_sn[kRigidGroupMax];
    __shared__ int l_off[kRigidGroupMax + 1];
    const int p0 = blockIdx.x * G;
    const int np = min(G, a.n - p0);
    const int t = threadIdx.x;
    if (t <= np) l_off[t] = a.poff[p0 + t];
    __syncthreads();
    const int c = l_off[0] + t;
    const bool valid = c < l_off[np];
    int q = 0, v = 0;
    double vx0 = 0.0, vy0 = 0.0;
    if (valid) {
        while (l_off[q + 1] <= c) ++q;
        const int i = p0 + q;
        v = a.vstatic[c];
        const double fvx = a.vfx[v], fvy = a.vfy[v];
        vx0 = a.vpx[v]; vy0 = a.vpy[v];
        const double dx = vx0 - a.x[i], dy = vy0 - a.y[i];
        l_fx[t] = fvx; l_fy[t] = fvy;
        l_tq[t] = fvy * dx - fvx * dy;
    }
    __syncthreads();
    if (t < np) {
        const int i = p0 + t;
        const int s = a.sid[i];
        double fx = 0.0, fy = 0.0, tq = 0.0;
        for (int k = l_off[t] - l_off[0]; k < l_off[t + 1] - l_off[0]; ++k) {
            fx += l_fx[k];
            fy += l_fy[k];
            tq += l_tq[k];
        }
        double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
        if (damp) {
            const double sc = damp[s];
            fx += -vx * sc;
            fy += -vy * sc;
            tq += -w * sc;
        }
        a.fx[i] = fx; a.fy[i] = fy; a.tq[i] = tq;   // no energy: see rigid_vv_fused
        kick_one(a, i, dth[s] * 1.0, fx, fy, tq, vx, vy, w);     // end of this step
        kick_one(a, i, dth[s] * 1.0, fx, fy, tq, vx, vy, w);     // next step: update_velocities(dt_half)
        const double sc = dt[s] * 1.0;                           // update_positions(dt)
        const double x0 = a.x[i], y0 = a.y[i], a0 = a.th[i];
        const double nx = x0 + vx * sc, ny = y0 + vy * sc, na = a0 + w * sc;
        double sn, cs;
        own_sincos(na - a0, &sn, &cs);
        a.x[i] = nx; a.y[i] = ny; a.th[i] = na; a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
        l_x0[t] = x0; l_y0[t] = y0; l_nx[t] = nx; l_ny[t] = ny; l_cs[t] = cs; l_sn[t] = sn;
    }
    __syncthreads();
    int s = -1;
    bool flag = false;
    if (valid) {
        const int i = p0 + q;
        s = a.sid[i];
        const double dx = vx0 - l_x0[q], dy = vy0 - l_y0[q];
        const double cs = l_cs[q], sn = l_sn[q];
        const double rx = cs * dx - sn * dy, ry = sn * dx + cs * dy;
        const double vxn = l_nx[q] + rx, vyn = l_ny[q] + ry;
        a.vpx[v] = vxn;
        a.vpy[v] = vyn;
        if (a.cell && c == l_off[q]) {   // the skin test uses the particle's first vertex
            const double ddx = vxn - a.lx[i], ddy = vyn - a.ly[i];
            flag = (ddx * ddx + ddy * ddy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// ---- FIRE with one thread per vertex -------------------------------------------------------
// k_fire_drift's two rigid-body drifts (update_positions(dt_reverse) and, after the kick and
// mixing, update_positions(dt / 2)) for G particles: the centre, angle, kick and mixing on G
// threads, then each vertex thread applies both rotations in turn, exactly as the two
// drift_one calls do through memory.  Frozen systems are left untouched.
__global__ void __launch_bounds__(DPMD_BLOCK) k_fire_drift_v(RigidArgs a, FireArgs f, int G, Gate g) {
    if (!gate_open(g)) return;
    __shared__ double l_x0[kRigidGroupMax], l_y0[kRigidGroupMax], l_x1[kRigidGroupMax], l_y1[kRigidGroupMax];
    __shared__ double l_c1[kRigidGroupMax], l_s1[kRigidGroupMax], l_x2[kRigidGroupMax], l_y2[kRigidGroupMax];
    __shared__ double l_c2[kRigidGroupMax], l_s2[kRigidGroupMax];
    __shared__ int l_off[kRigidGroupMax + 1], l_live[kRigidGroupMax];
    const int p0 = blockIdx.x * G;
    const int np = min(G, a.n - p0);
    const int t = threadIdx.x;
    if (t < np) {
        const int i = p0 + t;
        const int s = a.sid[i];
        const bool live = !fire_frozen(f, s);
        if (live) {
            const double dt = f.s_dt[s], alpha = f.s_alpha[s], dtrev = f.s_dtrev[s], vscale = f.s_vscale[s];
            double vx = a.vx[i], vy = a.vy[i], w = a.w[i];
            const double fx = a.fx[i], fy = a.fy[i], tq = a.tq[i];
            const double x0 = a.x[i], y0 = a.y[i], a0 = a.th[i];
            const double sc1 = dtrev * 1.0;                                        // update_positions(dt_rev, 1)
            const double x1 = x0 + vx * sc1, y1 = y0 + vy * sc1, a1 = a0 + w * sc1;
            double s1, c1;
            own_sincos(a1 - a0, &s1, &c1);
            vx *= vscale; vy *= vscale; w *= vscale;                                 // scale_velocities
            kick_one(a, i, dt * 0.5, fx, fy, tq, vx, vy, w);                         // update_velocities(dt, 0.5)
            mix_one(vx, vy, w, fx, fy, tq, alpha);                                   // mix
            vx *= vscale; vy *= vscale; w *= vscale;                                 // scale_velocities
            const double sc2 = dt * 0.5;                                             // update_positions(dt, 0.5)
            const double x2 = x1 + vx * sc2, y2 = y1 + vy * sc2, a2 = a1 + w * sc2;
            double s2, c2;
            own_sincos(a2 - a1, &s2, &c2);
            a.x[i] = x2; a.y[i] = y2; a.th[i] = a2; a.vx[i] = vx; a.vy[i] = vy; a.w[i] = w;
            if (i == a.off[s]) {
                f.dt[s] = dt; f.alpha[s] = alpha; f.ngood[s] = f.s_ngood[s]; f.nbad[s] = f.s_nbad[s];
            }
            l_x0[t] = x0; l_y0[t] = y0; l_x1[t] = x1; l_y1[t] = y1; l_c1[t] = c1; l_s1[t] = s1;
            l_x2[t] = x2; l_y2[t] = y2; l_c2[t] = c2; l_s2[t] = s2;
        }
        l_live[t] = live ? 1 : 0;
    }
    if (t <= np) l_off[t] = a.poff[p0 + t];
    __syncthreads();
    const int c = l_off[0] + t;
    const bool valid = c < l_off[np];
    int q = 0, s = -1;
    bool flag = false;
    if (valid) {
        while (l_off[q + 1] <= c) ++q;
        const int i = p0 + q;
        s = a.sid[i];
        if (l_live[q]) {
            const int v = a.vstatic[c];
            const bool first = a.cell && c == l_off[q];   // the skin test uses the first vertex
            double dx = a.vpx[v] - l_x0[q], dy = a.vpy[v] - l_y0[q];
            double rx = l_c1[q] * dx - l_s1[q] * dy, ry = l_s1[q] * dx + l_c1[q] * dy;
            const double v1x = l_x1[q] + rx, v1y = l_y1[q] + ry;
            bool f1 = false, f2 = false;
            if (first) {
                const double ddx = v1x - a.lx[i], ddy = v1y - a.ly[i];
                f1 = (ddx * ddx + ddy * ddy) > a.thresh2[s];
            }
            dx = v1x - l_x1[q]; dy = v1y - l_y1[q];
            rx = l_c2[q] * dx - l_s2[q] * dy; ry = l_s2[q] * dx + l_c2[q] * dy;
            const double v2x = l_x2[q] + rx, v2y = l_y2[q] + ry;
            if (first) {
                const double ddx = v2x - a.lx[i], ddy = v2y - a.ly[i];
                f2 = (ddx * ddx + ddy * ddy) > a.thresh2[s];
            }
            a.vpx[v] = v2x;
            a.vpy[v] = v2y;
            flag = f2 || f1;
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// Particles per workgroup of the vertex-parallel kernels (0: use the per-particle kernels);
// at most DPMD_BLOCK - 1, so the G + 1 particle offsets are written by distinct threads.
int rigid_group(dpmd_ctx* c) {
    if (c->nv_max <= 0 || c->nv_max > DPMD_BLOCK) return 0;
    return std::min(DPMD_BLOCK - 1, DPMD_BLOCK / c->nv_max);
}

void launch_vertex_force(dpmd_ctx* c, int mode, Gate g, bool pe = true) {
    RigidArgs a = rigid_args(c);
    const int G = blocks_for(c->Nv);
    ProfScope prof(c);
    auto go = [&](auto pv) {
        constexpr bool PE = decltype(pv)::value;
        switch (mode) {
            case DPMD_FORCE_PBC: k_vertex_force<DPMD_FORCE_PBC, PE><<<G, DPMD_BLOCK, 0, c->st>>>(a, g); break;
            case DPMD_FORCE_WALL: k_vertex_force<DPMD_FORCE_WALL, PE><<<G, DPMD_BLOCK, 0, c->st>>>(a, g); break;
            default: k_vertex_force<DPMD_FORCE_PER_SYSTEM, PE><<<G, DPMD_BLOCK, 0, c->st>>>(a, g); break;
        }
    };
    if (pe) go(std::true_type{}); else go(std::false_type{});
    HIP_OK(hipGetLastError());
}

}  // namespace

void rigid_compute_forces(dpmd_ctx* c, int mode) {
    const Gate g{c->one.p, nullptr, 0};
    launch_vertex_force(c, mode, g);
    if (const int G = rigid_group(c))
        k_particle_v<<<(c->N + G - 1) / G, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), nullptr, nullptr, G, 0, g);
    else
        k_particle_force<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c));
    HIP_OK(hipGetLastError());
}
void rigid_update_positions(dpmd_ctx* c, const double* scale, double scale2) {
    k_update_positions<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), scale, scale2);
    HIP_OK(hipGetLastError());
}
void rigid_update_velocities(dpmd_ctx* c, const double* scale, double scale2) {
    k_update_velocities<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), scale, scale2);
    HIP_OK(hipGetLastError());
}
void rigid_damping(dpmd_ctx* c, const double* scale) {
    k_damping<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), scale);
    HIP_OK(hipGetLastError());
}
void rigid_scale_velocities(dpmd_ctx* c, const double* scale) {
    k_scale_velocities<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), scale);
    HIP_OK(hipGetLastError());
}
void rigid_scale_positions(dpmd_ctx* c, const double* scale) {
    k_scale_positions<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), scale);
    HIP_OK(hipGetLastError());
}
void rigid_mix(dpmd_ctx* c, const double* alpha) {
    k_mix<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), alpha);
    HIP_OK(hipGetLastError());
}
void rigid_ke(dpmd_ctx* c) {
    k_ke<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->ke.p);
    HIP_OK(hipGetLastError());
}
void rigid_stress(dpmd_ctx* c) {
    k_stress<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vpid.p, c->sxx.p, c->sxy.p, c->syx.p,
                                                        c->syy.p);
    HIP_OK(hipGetLastError());
}
void rigid_fire_drift(dpmd_ctx* c, Gate g) {
    if (const int G = rigid_group(c))
        k_fire_drift_v<<<(c->N + G - 1) / G, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), fire_args(c), G, g);
    else
        k_fire_drift<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), fire_args(c), g);
    HIP_OK(hipGetLastError());
}
void rigid_fire_force(dpmd_ctx* c, Gate g, int kick, int do_conv, int pred, double pa, double pb) {
    launch_vertex_force(c, c->fire_mode, g);
    // one thread per particle: the chunk partials need the particles of a 256-chunk in one
    // workgroup, and a vertex-parallel variant plus a separate partials pass measured slower
    // (24.9 + 4.2 us against 21.0 us on 256 x 256 bodies)
    k_fire_particle<<<c->nchunks, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), fire_args(c), red_args(c), kick, g);
    HIP_OK(hipGetLastError());
    fire_finalize(c, g, do_conv, pred, pa, pb);
}
void rigid_vv_drift(dpmd_ctx* c, Gate g) {
    if (const int G = rigid_group(c))
        k_vv_drift_v<<<(c->N + G - 1) / G, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vv_dt.p, c->vv_dth.p, G, g);
    else
        k_vv_drift<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vv_dt.p, c->vv_dth.p, g);
    HIP_OK(hipGetLastError());
}
// Fused velocity Verlet needs the one-thread-per-vertex form.
bool rigid_vv_fusable(dpmd_ctx* c) { return rigid_group(c) != 0; }
void rigid_vv_fused(dpmd_ctx* c, Gate g) {
    const int G = rigid_group(c);
    need(G != 0, "rigid fused velocity Verlet: a particle has more vertices than a workgroup holds");
    launch_vertex_force(c, c->vv_mode, g, false);   // k_vv_fused_v reads no vertex energy
    k_vv_fused_v<<<(c->N + G - 1) / G, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vv_dt.p, c->vv_dth.p, vv_damping(c),
                                                              G, g);
    HIP_OK(hipGetLastError());
}
void rigid_vv_force(dpmd_ctx* c, Gate g) {
    launch_vertex_force(c, c->vv_mode, g);
    if (const int G = rigid_group(c))
        k_particle_v<<<(c->N + G - 1) / G, DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vv_dth.p, vv_damping(c), G, 1, g);
    else
        k_vv_particle<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_args(c), c->vv_dth.p, vv_damping(c), g);
    HIP_OK(hipGetLastError());
}

}  // namespace dpmd
