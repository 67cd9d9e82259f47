// Hand-written adjoints of the disk contact model (no reference counterpart; the pair
// Hessian block is the one the reference assembles in its Hessian observable,
// src/particles/disk.cu:565-661: H_ij = c n n^T + t (I - n n^T) / d with c = k, t = -k delta).
//
//   U   = sum_pairs k delta^2 / 2 + sum_walls k d_w^2 / 2       (delta = r_i + r_j - d)
//   F_i = -dU/dx_i
//   P   = (sum m v^2 + W) / (2 Lx Ly),  W = sum_pairs k delta d over minimum-image
//         contacts even in walled boxes, no wall virial (the reference's stress, disk.cu:702-769)
//
// Kernels: y = H w (k_hvp); the vector-Jacobian products of (U, P, F) with respect to
// positions, radii and box (k_vjp; box derivatives are taken at fixed reduced coordinates
// x / L, which also moves the walls); the batched conjugate-gradient pieces used for
// implicit differentiation at FIRE minima; canonical <-> current order maps.
#include "internal.hpp"

namespace dpmd {
namespace {

struct AdjGeom {
    const double* x; const double* y; const double* rad; const double* mass; const double* vx; const double* vy;
    const int* sid; const int* off;
    const double* Lx; const double* Ly; const double* iLx; const double* iLy; const double* e;
    const int* boundary; const int* nn; const int* nbr; int n;
    const double* live;   // contact mask (1 = the particle touches something); null = every particle
};

AdjGeom geom_current(dpmd_ctx* c) {
    AdjGeom g;
    g.x = c->x.p; g.y = c->y.p; g.rad = c->rad.p; g.mass = c->mass.p; g.vx = c->vx.p; g.vy = c->vy.p;
    g.sid = c->sid.p; g.off = c->off.p;
    g.Lx = c->Lx.p; g.Ly = c->Ly.p; g.iLx = c->iLx.p; g.iLy = c->iLy.p; g.e = c->e.p; g.boundary = c->boundary.p;
    g.nn = c->nn.p; g.nbr = c->nbr.p; g.n = c->N;
    g.live = nullptr;
    return g;
}

template <int MODE>
__device__ __forceinline__ bool walled(const AdjGeom& g, int s) {
    if (MODE == DPMD_FORCE_PBC) return false;
    if (MODE == DPMD_FORCE_WALL) return true;
    return g.boundary[s] != 0;
}

// Visits every contact (delta > 0) of particle i in list order: f(j, qx, qy, nx, ny, d, delta).
// Geometry is evaluated exactly as the force kernels do (same expressions, same order).
template <class Fn>
__device__ __forceinline__ void for_contacts(const AdjGeom& g, int i, int s, bool wall, Fn&& fn) {
    const double xi = g.x[i], yi = g.y[i], ri = g.rad[i];
    const double Lx = g.Lx[s], Ly = g.Ly[s], iLx = g.iLx[s], iLy = g.iLy[s];
    const int cnt = g.nn[i];
    const int* __restrict__ col = g.nbr + i;
    for (int k = 0; k < cnt; ++k) {
        const int j = col[(size_t)k * g.n];
        double dx = xi - g.x[j], dy = yi - g.y[j];
        if (!wall) { dx = min_image(dx, Lx, iLx); dy = min_image(dy, Ly, iLy); }
        const double r2 = dx * dx + dy * dy;
        const double radsum = ri + g.rad[j];
        if (r2 >= radsum * radsum) continue;
        const double d = sqrt(r2);
        const double inv = 1.0 / d;
        fn(j, dx, dy, dx * inv, dy * inv, d, radsum - d);
    }
}

// K v for one contact: k n (n.v) - (k delta / d) (v - n (n.v)).
__device__ __forceinline__ void pair_block(double e, double nx, double ny, double d, double delta, double vx, double vy,
                                           double& ox, double& oy) {
    const double nv = nx * vx + ny * vy;
    const double t = e * delta / d;
    ox = e * nx * nv - t * (vx - nx * nv);
    oy = e * ny * nv - t * (vy - ny * nv);
}

// y = H w (skip: systems whose rows are left untouched, may be null).
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_hvp(AdjGeom g, const double* __restrict__ wx, const double* __restrict__ wy, double* yx, double* yy, double shift) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (g.live && g.live[i] == 0.0) { yx[i] = 0.0; yy[i] = 0.0; return; }
    const int s = g.sid[i];
    const bool wall = walled<MODE>(g, s);
    const double e = g.e[s];
    const double wxi = wx[i], wyi = wy[i];
    double ax = 0.0, ay = 0.0;
    if (wall) {
        const double xi = g.x[i], yi = g.y[i], ri = g.rad[i];
        if (xi < ri) ax += e * wxi;
        if (xi > g.Lx[s] - ri) ax += e * wxi;
        if (yi < ri) ay += e * wyi;
        if (yi > g.Ly[s] - ri) ay += e * wyi;
    }
    for_contacts(g, i, s, wall, [&](int j, double, double, double nx, double ny, double d, double delta) {
        if (g.live && g.live[j] == 0.0) return;
        double kx, ky;
        pair_block(e, nx, ny, d, delta, wxi - wx[j], wyi - wy[j], kx, ky);
        ax += kx;
        ay += ky;
    });
    yx[i] = ax + shift * wxi;
    yy[i] = ay + shift * wyi;
}

struct VjpIn {
    const double* gU; const double* gP;      // per system (null = 0)
    const double* gfx; const double* gfy;    // per particle (null = 0)
};
struct VjpOut {
    double* dx; double* dy; double* dr;      // per particle (null = not wanted)
    double* bx; double* by;                  // per-particle box contributions (summed per system later)
};

template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK) k_vjp(AdjGeom g, VjpIn in, VjpOut out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.n) return;
    const int s = g.sid[m];
    const bool wall = walled<MODE>(g, s);
    const double e = g.e[s];
    const double Lx = g.Lx[s], Ly = g.Ly[s];
    const double gU = in.gU ? in.gU[s] : 0.0;
    const double gPA = in.gP ? in.gP[s] / (2.0 * (Lx * Ly)) : 0.0;   // dP/dW = 1 / (2 A)
    const bool mlive = !g.live || g.live[m] != 0.0;
    const double gfm_x = (in.gfx && mlive) ? in.gfx[m] : 0.0, gfm_y = (in.gfy && mlive) ? in.gfy[m] : 0.0;
    const double xm = g.x[m], ym = g.y[m], rm = g.rad[m];
    double Fx = 0.0, Fy = 0.0;          // force on m (energy term: dU/dx = -F)
    double hx = 0.0, hy = 0.0;          // (H gF)_m
    double gr = 0.0, bx = 0.0, by = 0.0, wx = 0.0, wy = 0.0;   // wx, wy: dW/dx_m
    if (wall) {
        if (xm < rm) {                  // U = k (r - x)^2 / 2, x = s Lx
            const double d = rm - xm;
            Fx += e * d; hx += e * gfm_x;
            gr += gU * (e * d) + gfm_x * e;
            bx += gU * (-(e * d) * xm / Lx) + gfm_x * (-e * xm / Lx);
        }
        if (xm > Lx - rm) {             // gap g = Lx - x, U = k (r - g)^2 / 2
            const double gap = Lx - xm, d = rm - gap;
            Fx -= e * d; hx += e * gfm_x;
            gr += gU * (e * d) - gfm_x * e;
            bx += gU * (-(e * d) * gap / Lx) + gfm_x * (e * gap / Lx);
        }
        if (ym < rm) {
            const double d = rm - ym;
            Fy += e * d; hy += e * gfm_y;
            gr += gU * (e * d) + gfm_y * e;
            by += gU * (-(e * d) * ym / Ly) + gfm_y * (-e * ym / Ly);
        }
        if (ym > Ly - rm) {
            const double gap = Ly - ym, d = rm - gap;
            Fy -= e * d; hy += e * gfm_y;
            gr += gU * (e * d) - gfm_y * e;
            by += gU * (-(e * d) * gap / Ly) + gfm_y * (e * gap / Ly);
        }
    }
    // pressure W = sum k delta d: dW/dx_m = k (delta - d) n, dW/dr_m = k d; box terms are
    // half of each directed pair with dd/dLx = qx^2 / (d Lx)
    auto virial = [&](double qx, double qy, double nx, double ny, double d, double delta) {
        wx += e * (delta - d) * nx; wy += e * (delta - d) * ny;
        gr += gPA * (e * d);
        bx += gPA * (0.5 * e * (delta - d) * (qx * qx / (d * Lx)));
        by += gPA * (0.5 * e * (delta - d) * (qy * qy / (d * Ly)));
    };
    if (wall && gPA != 0.0)   // walled box: the virial still runs over minimum-image contacts
        for_contacts(g, m, s, false, [&](int, double qx, double qy, double nx, double ny, double d, double delta) {
            virial(qx, qy, nx, ny, d, delta);
        });
    for_contacts(g, m, s, wall, [&](int j, double qx, double qy, double nx, double ny, double d, double delta) {
        const double ed = e * delta;
        Fx += ed * nx; Fy += ed * ny;
        if (!wall) virial(qx, qy, nx, ny, d, delta);
        gr += gU * ed;
        const double cx = qx * qx / (d * Lx), cy = qy * qy / (d * Ly);
        bx += gU * (-0.5 * ed * cx);
        by += gU * (-0.5 * ed * cy);
        if (in.gfx && mlive && (!g.live || g.live[j] != 0.0)) {
            const double dgx = gfm_x - in.gfx[j], dgy = gfm_y - in.gfy[j];
            double kx, ky;
            pair_block(e, nx, ny, d, delta, dgx, dgy, kx, ky);
            hx += kx; hy += ky;
            gr += e * (nx * dgx + ny * dgy);
            // dF_m/dLx = -K (qx, 0) / Lx, dF_m/dLy = -K (0, qy) / Ly
            double ax, ay;
            pair_block(e, nx, ny, d, delta, qx, 0.0, ax, ay);
            bx -= (gfm_x * ax + gfm_y * ay) / Lx;
            pair_block(e, nx, ny, d, delta, 0.0, qy, ax, ay);
            by -= (gfm_x * ax + gfm_y * ay) / Ly;
        }
    });
    if (out.dx) {
        out.dx[m] = gU * (-Fx) + gPA * wx - hx;
        out.dy[m] = gU * (-Fy) + gPA * wy - hy;
    }
    if (out.dr) out.dr[m] = gr;
    out.bx[m] = bx;
    out.by[m] = by;
}

// dL_s = sum_i b_i - gP_s P_s / L_s  (P = (sum m v^2 + W) / (2 Lx Ly) at fixed velocities).
__global__ void k_box_final(int S, const double* gP, const double* P, const double* Lx, const double* Ly,
                            const double* sbx, const double* sby, double* dLx, double* dLy) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double g = gP ? gP[s] : 0.0;
    dLx[s] = sbx[s] - g * P[s] / Lx[s];
    dLy[s] = sby[s] - g * P[s] / Ly[s];
}

// ---- order maps (static_index: canonical -> current) -------------------------------------
__global__ void k_to_current(int n, const int* __restrict__ si, const double* __restrict__ canon, double* cur) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) cur[si[c]] = canon[c];
}
// rows of a row-major array of n-vectors from canonical into current order, and back
__global__ void k_rows_to_current(int n, const int* __restrict__ si, const double* __restrict__ src,
                                  double* __restrict__ dst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const size_t r = (size_t)blockIdx.y * n;
    dst[r + si[c]] = src[r + c];
}
__global__ void k_rows_to_canonical(int n, const int* __restrict__ si, const double* __restrict__ src,
                                    double* __restrict__ dst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const size_t r = (size_t)blockIdx.y * n;
    dst[r + c] = src[r + si[c]];
}

__global__ void k_to_canonical(int n, const int* __restrict__ si, const double* __restrict__ cur, double* canon) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) canon[c] = cur[si[c]];
}

// ---- batched MINRES (one independent solve per system) ----------------------------------
// Paige-Saunders MINRES without preconditioner.  The Hessian at a FIRE minimum can be
// slightly indefinite (nearly collinear two-contact particles: compressive contacts have
// negative transverse stiffness) and is singular (translations, rattlers), so CG is not
// safe; MINRES needs only symmetry and converges to the minimum-norm solution of a
// consistent system.  Per-system scalars live in mr[s * MR_STRIDE + field].
enum { MR_BETA1 = 0, MR_BETA, MR_OLDB, MR_EPS, MR_OLDEPS, MR_DBAR, MR_CS, MR_SN, MR_PHIBAR, MR_PHI, MR_GAMMA,
       MR_DELTA, MR_STOP, MR_STRIDE };
struct V2 { double* x; double* y; };

__global__ void k_mr_init(int S, const double* bb, double* mr, int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    double* m = mr + (size_t)s * MR_STRIDE;
    const double b1 = sqrt(bb[s]);
    for (int k = 0; k < MR_STRIDE; ++k) m[k] = 0.0;
    m[MR_BETA1] = b1; m[MR_BETA] = b1; m[MR_PHIBAR] = b1; m[MR_CS] = -1.0;
    active[s] = b1 > 0.0 ? 1 : 0;
}

// v = r2 / beta
__global__ void k_mr_v(int n, const int* sid, const int* active, const double* mr, V2 y, V2 v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sid[i];
    if (!active[s]) return;
    const double b = mr[(size_t)s * MR_STRIDE + MR_BETA];
    v.x[i] = y.x[i] / b; v.y[i] = y.y[i] / b;
}

// y -= (beta / oldb) r1  (from the second iteration on)
__global__ void k_mr_y1(int n, const int* sid, const int* active, const double* mr, V2 y, V2 r1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sid[i];
    if (!active[s]) return;
    const double* m = mr + (size_t)s * MR_STRIDE;
    const double f = m[MR_BETA] / m[MR_OLDB];
    y.x[i] -= f * r1.x[i]; y.y[i] -= f * r1.y[i];
}

// y -= (alpha / beta) r2
__global__ void k_mr_y2(int n, const int* sid, const int* active, const double* mr, const double* alpha, V2 y, V2 r2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sid[i];
    if (!active[s]) return;
    const double f = alpha[s] / mr[(size_t)s * MR_STRIDE + MR_BETA];
    y.x[i] -= f * r2.x[i]; y.y[i] -= f * r2.y[i];
}

// Lanczos / Givens scalar recurrences; stop when the residual estimate phibar <= tol beta1.
__global__ void k_mr_scalars(int S, const int* active, double* mr, const double* alpha, const double* bnew2, double tol) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !active[s]) return;
    double* m = mr + (size_t)s * MR_STRIDE;
    const double a = alpha[s];
    m[MR_OLDB] = m[MR_BETA];
    const double beta = sqrt(bnew2[s]);
    m[MR_BETA] = beta;
    m[MR_OLDEPS] = m[MR_EPS];
    const double cs = m[MR_CS], sn = m[MR_SN], dbar = m[MR_DBAR];
    m[MR_DELTA] = cs * dbar + sn * a;
    const double gbar = sn * dbar - cs * a;
    m[MR_EPS] = sn * beta;
    m[MR_DBAR] = -cs * beta;
    double gamma = sqrt(gbar * gbar + beta * beta);
    gamma = fmax(gamma, 1e-300);
    m[MR_GAMMA] = gamma;
    m[MR_CS] = gbar / gamma;
    m[MR_SN] = beta / gamma;
    m[MR_PHI] = m[MR_CS] * m[MR_PHIBAR];
    m[MR_PHIBAR] = m[MR_SN] * m[MR_PHIBAR];
    if (m[MR_PHIBAR] <= tol * m[MR_BETA1] || beta == 0.0) m[MR_STOP] = 1.0;
}

// w_new = (v - oldeps w2_old - delta w_old) / gamma (into the old w1 buffer); x += phi w_new
__global__ void k_mr_w(int n, const int* sid, const int* active, const double* mr, V2 v, V2 w2, V2 w, V2 wnew, V2 x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = sid[i];
    if (!active[s]) return;
    const double* m = mr + (size_t)s * MR_STRIDE;
    const double oe = m[MR_OLDEPS], de = m[MR_DELTA], g = m[MR_GAMMA], ph = m[MR_PHI];
    const double nx = (v.x[i] - oe * w2.x[i] - de * w.x[i]) / g;
    const double ny = (v.y[i] - oe * w2.y[i] - de * w.y[i]) / g;
    wnew.x[i] = nx; wnew.y[i] = ny;
    x.x[i] += ph * nx; x.y[i] += ph * ny;
}

__global__ void k_mr_commit(int S, const double* mr, int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S && active[s] && mr[(size_t)s * MR_STRIDE + MR_STOP] != 0.0) active[s] = 0;
}

// The contact mask: a particle is kept while it has a contact (a pair contact with a kept
// particle, or a wall contact); the coordinates of a particle without contacts are exact null
// directions of the Hessian.  Every contact stays, also one that carries (almost) no force:
// in the linear response it is a spring of stiffness k, and a particle between two such
// contacts is a strut that the response goes through.  (Dropping particles with fewer than
// three contacts, the rigid backbone, loses those struts: in one of the 1024-disk packings it
// changed entries of dP/dr by up to 40 %, against central differences that agree with the
// full minimum-norm solve.)
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK) k_touching(AdjGeom g, const double* lin, double* lout, int* changed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    if (lin[i] == 0.0) { lout[i] = 0.0; return; }
    const int s = g.sid[i];
    const bool wall = walled<MODE>(g, s);
    int cnt = 0;
    if (wall) {
        const double xi = g.x[i], yi = g.y[i], ri = g.rad[i];
        cnt += (xi < ri) + (xi > g.Lx[s] - ri) + (yi < ri) + (yi > g.Ly[s] - ri);
    }
    for_contacts(g, i, s, wall, [&](int j, double, double, double, double, double, double) { cnt += lin[j] != 0.0; });
    const bool keep = cnt >= 1;
    lout[i] = keep ? 1.0 : 0.0;
    if (!keep) atomicOr(changed, 1);
}

__global__ void k_norm2_masked(int n, const double* ax, const double* ay, const double* live, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = live[i] != 0.0 ? ax[i] * ax[i] + ay[i] * ay[i] : 0.0;
}

__global__ void k_mask_live(int n, const double* live, double* vx, double* vy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && live[i] == 0.0) { vx[i] = 0.0; vy[i] = 0.0; }
}

// subtract the per-system mean over live particles (periodic systems only: a uniform
// translation of the contact network is a null direction)
__global__ void k_sub_mean_live(int n, const int* sid, const int* boundary, int mode, const double* live,
                                const double* cnt, const double* sx, const double* sy, double* vx, double* vy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || live[i] == 0.0) return;
    const int s = sid[i];
    const bool wall = mode == DPMD_FORCE_WALL || (mode == DPMD_FORCE_PER_SYSTEM && boundary[s]);
    if (wall || cnt[s] == 0.0) return;
    vx[i] -= sx[s] / cnt[s];
    vy[i] -= sy[s] / cnt[s];
}

__global__ void k_copy2(int n, const double* ax, const double* ay, double* bx, double* by) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bx[i] = ax[i]; by[i] = ay[i];
}

__global__ void k_norm2(int n, const double* ax, const double* ay, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ax[i] * ax[i] + ay[i] * ay[i];
}

__global__ void k_add(int64_t n, const double* a, const double* b, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

__global__ void k_init_active(int S, const double* bb, int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < S) active[s] = bb[s] > 0.0 ? 1 : 0;
}

// y = a x + y, per particle
__global__ void k_axpy2(int n, double a, const double* xx, const double* xy, double* yx, double* yy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    yx[i] += a * xx[i];
    yy[i] += a * xy[i];
}

// ---- reverse mode through velocity Verlet -------------------------------------------------
// vbar += dt_s xbar; w = vbar / m
__global__ void k_rev_vw(int n, const int* sid, const double* dt, const double* m, const double* xbx, const double* xby,
                         double* vbx, double* vby, double* wx, double* wy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double h = dt[sid[i]];
    const double vx = vbx[i] + h * xbx[i], vy = vby[i] + h * xby[i];
    vbx[i] = vx; vby[i] = vy;
    wx[i] = vx / m[i]; wy[i] = vy / m[i];
}
// w = vbar / m
__global__ void k_rev_w(int n, const double* m, const double* vbx, const double* vby, double* wx, double* wy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    wx[i] = vbx[i] / m[i]; wy[i] = vby[i] / m[i];
}
// xbar -= coef dt_s z
__global__ void k_rev_x(int n, const int* sid, const double* dt, double coef, const double* zx, const double* zy,
                        double* xbx, double* xby) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double h = coef * dt[sid[i]];
    xbx[i] -= h * zx[i];
    xby[i] -= h * zy[i];
}

// One whole reverse velocity-Verlet step in one launch: z = H(x_slot) w (k_hvp's sum, no mask,
// no shift), xbar -= coef dt_s z (k_rev_x), and, when next is set, the following step's
// vbar += dt_s xbar, w' = vbar / m (k_rev_vw) into the other w buffer.  Each expression is the
// one of the three separate kernels, so the results are the same bits.
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_rev_step(AdjGeom g, const double* __restrict__ wx, const double* __restrict__ wy, const double* __restrict__ dt,
           double coef, const double* __restrict__ m, double* xbx, double* xby, double* vbx, double* vby, double* wox,
           double* woy, int next) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const int s = g.sid[i];
    const bool wall = walled<MODE>(g, s);
    const double e = g.e[s];
    const double wxi = wx[i], wyi = wy[i];
    double ax = 0.0, ay = 0.0;
    if (wall) {
        const double xi = g.x[i], yi = g.y[i], ri = g.rad[i];
        if (xi < ri) ax += e * wxi;
        if (xi > g.Lx[s] - ri) ax += e * wxi;
        if (yi < ri) ay += e * wyi;
        if (yi > g.Ly[s] - ri) ay += e * wyi;
    }
    for_contacts(g, i, s, wall, [&](int j, double, double, double nx, double ny, double d, double delta) {
        double kx, ky;
        pair_block(e, nx, ny, d, delta, wxi - wx[j], wyi - wy[j], kx, ky);
        ax += kx;
        ay += ky;
    });
    const double zx = ax + 0.0 * wxi, zy = ay + 0.0 * wyi;
    const double h = coef * dt[s];
    const double xbx_i = xbx[i] - h * zx, xby_i = xby[i] - h * zy;
    xbx[i] = xbx_i; xby[i] = xby_i;
    if (next) {
        const double hh = dt[s];
        const double vx = vbx[i] + hh * xbx_i, vy = vby[i] + hh * xby_i;
        vbx[i] = vx; vby[i] = vy;
        wox[i] = vx / m[i]; woy[i] = vy / m[i];
    }
}

// Canonical-order copy of the current contact list: nn_c[c] = nn[si[c]],
// nbr_c[k n + c] = canonical index of nbr[k n + si[c]] (inv: current -> canonical).
__global__ void k_inv_perm(int n, const int* si, int* inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) inv[si[c]] = c;
}
__global__ void k_list_canonical(int n, int K, const int* si, const int* inv, const int* nn, const int* nbr, int* nn_c,
                                 int* nbr_c) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const int i = si[c];
    const int cnt = nn[i];
    nn_c[c] = cnt;
    for (int k = 0; k < cnt; ++k) nbr_c[(size_t)k * n + c] = inv[nbr[(size_t)k * n + i]];
}

// Largest squared displacement of each particle over stored slots base + 1 .. base + len
// against slot base (layout seg[k * 2n + {0, n} + c]).
__global__ void k_seg_disp2(int n, int base, int len, const double* seg, double* d2) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double x0 = seg[(size_t)base * 2 * n + c], y0 = seg[(size_t)base * 2 * n + n + c];
    double m = 0.0;
    for (int k = base + 1; k <= base + len; ++k) {
        const double dx = seg[(size_t)k * 2 * n + c] - x0, dy = seg[(size_t)k * 2 * n + n + c] - y0;
        m = fmax(m, dx * dx + dy * dy);
    }
    d2[c] = m;
}

// Positions of the step just drifted, in canonical order, into slot run[RUN_REC] + k + 1.
__global__ void k_record_pos(int n, const int* si, const double* x, const double* y, const double2* xy, double* seg,
                             Gate g) {
    if (!gate_open(g)) return;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const size_t slot = (size_t)(g.run[RUN_REC] + g.k + 1);
    const int i = si[c];
    const double2 p = xy ? xy[i] : make_double2(x[i], y[i]);
    seg[slot * 2 * n + c] = p.x;
    seg[slot * 2 * n + n + c] = p.y;
}

template <class F>
void by_mode(int mode, F&& f) {
    switch (mode) {
        case DPMD_FORCE_PBC: f(std::integral_constant<int, DPMD_FORCE_PBC>{}); break;
        case DPMD_FORCE_WALL: f(std::integral_constant<int, DPMD_FORCE_WALL>{}); break;
        default: f(std::integral_constant<int, DPMD_FORCE_PER_SYSTEM>{}); break;
    }
}

void ensure(dpmd::DVec<double>& v, int64_t n, hipStream_t st) {
    if (v.n != n) { v.resize(n); v.zero(st); }
}

}  // namespace

// ---------------------------------------------------------------------------------------
void adj_record_positions(dpmd_ctx* c, Gate g, const double* xy) {
    k_record_pos<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, c->static_index.p, c->x.p, c->y.p,
                                                            (const double2*)xy, c->rec_seg, g);
    HIP_OK(hipGetLastError());
}

void adj_list_canonical(dpmd_ctx* c, int* inv_scratch) {
    const int N = c->N;
    c->ck_nn.resize(N);
    c->ck_nbr.resize((int64_t)N * c->K);
    k_inv_perm<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->static_index.p, inv_scratch);
    k_list_canonical<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->K, c->static_index.p, inv_scratch, c->nn.p,
                                                              c->nbr.p, c->ck_nn.p, c->ck_nbr.p);
    HIP_OK(hipGetLastError());
}

void adj_seg_disp2(dpmd_ctx* c, int base, int len, double* d2) {
    k_seg_disp2<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, base, len, c->ck_seg.p, d2);
    HIP_OK(hipGetLastError());
}

// One reverse step's pieces (canonical order; positions from segment slot `slot`).
void adj_rev_vw(dpmd_ctx* c, const double* m, double* xb, double* vb, double* w) {
    const int N = c->N;
    k_rev_vw<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->sid.p, c->vv_dt.p, m, xb, xb + N, vb, vb + N, w, w + N);
    HIP_OK(hipGetLastError());
}
void adj_rev_w(dpmd_ctx* c, const double* m, const double* vb, double* w) {
    const int N = c->N;
    k_rev_w<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, m, vb, vb + N, w, w + N);
    HIP_OK(hipGetLastError());
}
void adj_rev_x(dpmd_ctx* c, double coef, const double* z, double* xb) {
    const int N = c->N;
    k_rev_x<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->sid.p, c->vv_dt.p, coef, z, z + N, xb, xb + N);
    HIP_OK(hipGetLastError());
}
void adj_hvp_segment(dpmd_ctx* c, int mode, int slot, const double* rad_c, const double* w, double* z) {
    const int N = c->N;
    AdjGeom g = geom_current(c);
    g.x = c->ck_seg.p + (size_t)slot * 2 * N;
    g.y = g.x + N;
    g.rad = rad_c;
    g.nn = c->ck_nn.p;
    g.nbr = c->ck_nbr.p;
    by_mode(mode, [&](auto m) {
        k_hvp<decltype(m)::value><<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(g, w, w + N, z, z + N, 0.0);
    });
    HIP_OK(hipGetLastError());
}

void adj_rev_step(dpmd_ctx* c, int mode, int slot, const double* rad_c, const double* m, const double* w, double coef,
                  double* xb, double* vb, double* w_next, bool next, const double* seg) {
    const int N = c->N;
    AdjGeom g = geom_current(c);
    g.x = seg + (size_t)slot * 2 * N;
    g.y = g.x + N;
    g.rad = rad_c;
    g.nn = c->ck_nn.p;
    g.nbr = c->ck_nbr.p;
    by_mode(mode, [&](auto md) {
        k_rev_step<decltype(md)::value><<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(
            g, w, w + N, c->vv_dt.p, coef, m, xb, xb + N, vb, vb + N, w_next, w_next + N, next ? 1 : 0);
    });
    HIP_OK(hipGetLastError());
}

void adj_add(dpmd_ctx* c, int64_t n, const double* a, const double* b, double* out) {
    k_add<<<blocks_for(n), DPMD_BLOCK, 0, c->st>>>(n, a, b, out);
    HIP_OK(hipGetLastError());
}
void adj_norm2(dpmd_ctx* c, const double* ax, const double* ay, double* out) {
    k_norm2<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, ax, ay, out);
    HIP_OK(hipGetLastError());
}
void adj_axpy2(dpmd_ctx* c, double a, const double* xx, const double* xy, double* yx, double* yy) {
    k_axpy2<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, a, xx, xy, yx, yy);
    HIP_OK(hipGetLastError());
}

void adj_rows_to_current(dpmd_ctx* c, const double* canon, double* cur, int rows) {
    if (rows <= 0) return;
    k_rows_to_current<<<dim3(blocks_for(c->N), rows), DPMD_BLOCK, 0, c->st>>>(c->N, c->static_index.p, canon, cur);
    HIP_OK(hipGetLastError());
}
void adj_rows_to_canonical(dpmd_ctx* c, const double* cur, double* canon, int rows) {
    if (rows <= 0) return;
    k_rows_to_canonical<<<dim3(blocks_for(c->N), rows), DPMD_BLOCK, 0, c->st>>>(c->N, c->static_index.p, cur, canon);
    HIP_OK(hipGetLastError());
}

void adj_to_current(dpmd_ctx* c, const double* canon, double* cur) {
    k_to_current<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, c->static_index.p, canon, cur);
    HIP_OK(hipGetLastError());
}
void adj_to_canonical(dpmd_ctx* c, const double* cur, double* canon) {
    k_to_canonical<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, c->static_index.p, cur, canon);
    HIP_OK(hipGetLastError());
}

// The contact mask of the current contact network, into c->adj_live (1 / 0 per particle).
const double* adj_touching(dpmd_ctx* c, int mode) {
    const int N = c->N;
    hipStream_t st = c->st;
    c->adj_live.resize(N); c->adj_live2.resize(N); c->adj_flag.resize(1);
    fill_f64(st, c->adj_live.p, N, 1.0);
    AdjGeom g = geom_current(c);
    double* a = c->adj_live.p;
    double* b = c->adj_live2.p;
    int changed = 1;
    for (int pass = 0; changed && pass < 100000; ++pass) {
        fill_i32(st, c->adj_flag.p, 1, 0);
        by_mode(mode, [&](auto m) {
            k_touching<decltype(m)::value><<<blocks_for(N), DPMD_BLOCK, 0, st>>>(g, a, b, c->adj_flag.p);
        });
        HIP_OK(hipGetLastError());
        std::swap(a, b);
        HIP_OK(hipMemcpyAsync(&changed, c->adj_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
    }
    if (a != c->adj_live.p) HIP_OK(hipMemcpyAsync(c->adj_live.p, a, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
    return c->adj_live.p;
}

double adj_touching_max_force(dpmd_ctx* c, const double* live) {
    const int N = c->N, S = c->S;
    c->adj_bx.resize(N);
    ensure(c->adj_sa, S, c->st);
    k_norm2_masked<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->fx.p, c->fy.p, live, c->adj_bx.p);
    HIP_OK(hipGetLastError());
    sys_reduce_max(c, c->adj_bx.p, c->adj_sa.p);
    std::vector<double> m(S);
    HIP_OK(hipMemcpyAsync(m.data(), c->adj_sa.p, (size_t)S * 8, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    double w = 0.0;
    for (double v : m) w = std::max(w, v);
    return std::sqrt(w);
}

void adj_hvp_current(dpmd_ctx* c, int mode, const double* wx, const double* wy, double* yx, double* yy,
                     const double* live, double shift) {
    AdjGeom g = geom_current(c);
    g.live = live;
    by_mode(mode, [&](auto m) {
        k_hvp<decltype(m)::value><<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(g, wx, wy, yx, yy, shift);
    });
    HIP_OK(hipGetLastError());
}

void adj_vjp_current(dpmd_ctx* c, int mode, const double* gU, const double* gP, const double* gfx, const double* gfy,
                     double* dx, double* dy, double* dr, double* dLx, double* dLy, const double* live) {
    const int N = c->N, S = c->S;
    ensure(c->adj_bx, N, c->st); ensure(c->adj_by, N, c->st);
    AdjGeom g = geom_current(c);
    g.live = live;
    VjpIn in{gU, gP, gfx, gfy};
    VjpOut out{dx, dy, dr, c->adj_bx.p, c->adj_by.p};
    by_mode(mode, [&](auto m) { k_vjp<decltype(m)::value><<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(g, in, out); });
    HIP_OK(hipGetLastError());
    if (dLx || dLy) {
        ensure(c->adj_sa, S, c->st); ensure(c->adj_sb, S, c->st); ensure(c->adj_sc, S, c->st);
        ensure(c->adj_sd, S, c->st);
        sys_reduce_arr(c, c->adj_bx.p, c->adj_sa.p);
        sys_reduce_arr(c, c->adj_by.p, c->adj_sb.p);
        if (gP) {   // P at the current state (stress tensor, reference definition)
            disk_stress(c);
            per_system_pressure(c);
        }
        k_box_final<<<blocks_for(S), DPMD_BLOCK, 0, c->st>>>(S, gP, c->press.p, c->Lx.p, c->Ly.p, c->adj_sa.p,
                                                            c->adj_sb.p, dLx ? dLx : c->adj_sc.p,
                                                            dLy ? dLy : c->adj_sd.p);
        HIP_OK(hipGetLastError());
    }
}

// Solves (H + shift I) x = b per system by MINRES (current order).  With on_touching the
// Hessian is restricted to the particles with contacts, and b and the result are projected
// onto the complement of its null directions (coordinates of particles without contacts;
// uniform translations in periodic systems); the mask stays in c->adj_live.  Otherwise every particle takes part (use shift > 0).  Returns the
// iterations taken; *rel_resid = max_s |b - H x| / |b| (true residual, recomputed at the end).
int adj_cg_current(dpmd_ctx* c, int mode, const double* bx, const double* by, double* xx, double* xy, double tol,
                   int max_iter, double* rel_resid, bool on_touching, double shift) {
    const int N = c->N, S = c->S;
    hipStream_t st = c->st;
    ensure(c->mr_vec, 12 * (int64_t)N, st);
    dpmd::DVec<double>* ss[] = {&c->cg_rr, &c->cg_rrn, &c->cg_pq, &c->cg_bb, &c->adj_sa, &c->adj_sb};
    for (auto* v : ss) ensure(*v, S, st);
    ensure(c->mr_s, (int64_t)S * MR_STRIDE, st);
    c->cg_active.resize(S);
    const int G = blocks_for(N), GS = blocks_for(S);
    auto vec = [&](int k) { return V2{c->mr_vec.p + (size_t)(2 * k) * N, c->mr_vec.p + (size_t)(2 * k + 1) * N}; };
    // null directions: coordinates of particles without contacts and (periodic systems)
    // uniform translations
    ensure(c->cg_beta, S, st);
    const double* live = nullptr;
    if (on_touching) {
        live = adj_touching(c, mode);
    } else {   // every particle: with shift > 0 the shifted matrix is definite
        c->adj_live.resize(N);
        fill_f64(st, c->adj_live.p, N, 1.0);
    }
    const double* mask = c->adj_live.p;
    double* nlive = c->cg_beta.p;
    sys_reduce_arr(c, mask, nlive);
    auto project = [&](double* vx, double* vy) {
        k_mask_live<<<G, DPMD_BLOCK, 0, st>>>(N, mask, vx, vy);
        sys_reduce_arr(c, vx, c->adj_sa.p);
        sys_reduce_arr(c, vy, c->adj_sb.p);
        k_sub_mean_live<<<G, DPMD_BLOCK, 0, st>>>(N, c->sid.p, c->boundary.p, mode, mask, nlive, c->adj_sa.p,
                                                 c->adj_sb.p, vx, vy);
    };
    V2 r1 = vec(0), r2 = vec(1), y = vec(2), v = vec(3), w1 = vec(4), w2 = vec(5);
    ensure(c->cg_qx, N, st); ensure(c->cg_qy, N, st);
    V2 w{c->cg_qx.p, c->cg_qy.p};
    const V2 X{xx, xy};
    // b (projected) -> r1 = r2; x = w = w1 = w2 = 0
    k_copy2<<<G, DPMD_BLOCK, 0, st>>>(N, bx, by, r1.x, r1.y);
    project(r1.x, r1.y);
    k_copy2<<<G, DPMD_BLOCK, 0, st>>>(N, r1.x, r1.y, r2.x, r2.y);
    for (double* p : {xx, xy, w.x, w.y, w1.x, w1.y, w2.x, w2.y}) fill_f64(st, p, N, 0.0);
    sys_reduce_dot(c, r1.x, r1.x, r1.y, r1.y, c->cg_bb.p);
    k_mr_init<<<GS, DPMD_BLOCK, 0, st>>>(S, c->cg_bb.p, c->mr_s.p, c->cg_active.p);
    HIP_OK(hipGetLastError());
    int it = 0;
    std::vector<int> act(S);
    while (it < max_iter) {
        ++it;
        k_mr_v<<<G, DPMD_BLOCK, 0, st>>>(N, c->sid.p, c->cg_active.p, c->mr_s.p, r2, v);
        adj_hvp_current(c, mode, v.x, v.y, y.x, y.y, live, shift);
        if (it >= 2) k_mr_y1<<<G, DPMD_BLOCK, 0, st>>>(N, c->sid.p, c->cg_active.p, c->mr_s.p, y, r1);
        sys_reduce_dot(c, v.x, y.x, v.y, y.y, c->cg_pq.p);               // alpha
        k_mr_y2<<<G, DPMD_BLOCK, 0, st>>>(N, c->sid.p, c->cg_active.p, c->mr_s.p, c->cg_pq.p, y, r2);
        sys_reduce_dot(c, y.x, y.x, y.y, y.y, c->cg_rrn.p);              // beta_new^2
        k_mr_scalars<<<GS, DPMD_BLOCK, 0, st>>>(S, c->cg_active.p, c->mr_s.p, c->cg_pq.p, c->cg_rrn.p, tol);
        k_mr_w<<<G, DPMD_BLOCK, 0, st>>>(N, c->sid.p, c->cg_active.p, c->mr_s.p, v, w2, w, w1, X);
        k_mr_commit<<<GS, DPMD_BLOCK, 0, st>>>(S, c->mr_s.p, c->cg_active.p);
        HIP_OK(hipGetLastError());
        // r1 <- r2, r2 <- y, y <- the old r1 buffer
        const V2 old_r1 = r1;
        r1 = r2; r2 = y; y = old_r1;
        // w1 <- w2, w2 <- w, w <- the new direction (written into the old w1 buffer)
        const V2 old_w1 = w1;
        w1 = w2; w2 = w; w = old_w1;
        if (it % 16 == 0 || it == max_iter) {
            HIP_OK(hipMemcpyAsync(act.data(), c->cg_active.p, (size_t)S * 4, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            bool any = false;
            for (int s = 0; s < S; ++s) any = any || act[s];
            if (!any) break;
        }
    }
    // project the solution, then the true residual b - H x
    V2 Q = vec(3), R = vec(2);
    project(xx, xy);
    adj_hvp_current(c, mode, xx, xy, Q.x, Q.y, live, shift);
    k_copy2<<<G, DPMD_BLOCK, 0, st>>>(N, bx, by, R.x, R.y);
    project(R.x, R.y);
    k_axpy2<<<G, DPMD_BLOCK, 0, st>>>(N, -1.0, Q.x, Q.y, R.x, R.y);
    sys_reduce_dot(c, R.x, R.x, R.y, R.y, c->cg_rrn.p);
    HIP_OK(hipGetLastError());
    std::vector<double> rr(S), bb(S);
    HIP_OK(hipMemcpyAsync(rr.data(), c->cg_rrn.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(bb.data(), c->cg_bb.p, (size_t)S * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double worst = 0.0;
    for (int s = 0; s < S; ++s)
        if (bb[s] > 0.0) worst = std::max(worst, std::sqrt(rr[s] / bb[s]));
    if (rel_resid) *rel_resid = worst;
    return it;
}

}  // namespace dpmd
