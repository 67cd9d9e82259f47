// =====================================================================================
//  dpmd oracle -- TEST INFRASTRUCTURE ONLY.
//
//  A plain C++17 FP64 host restatement of the reference (mmccraw98/dpmd) routines on the
//  batched 2D granular-MD hot path.  Only tests/, __graft_entry__.smoke() and bench.py's
//  cpu_baseline leg may load this library, and only as the checker / the timed CPU
//  baseline.  The product (dpmd_amd/libdpmd_amd.so) never links, loads or calls it.
//
//  Every routine cites the reference file:line it restates.  Where the reference is
//  nondeterministic (atomic bucket scatter, CUB segmented-sum order) this file fixes one
//  deterministic rule and states it; the HIP library implements the same rule, so the
//  two agree bit for bit:
//    * in-cell order: stable by current index (the reference's Standard sort path,
//      include/particles/base_point_particle.hpp:423-447, with a stable key sort);
//    * neighbor order: 3x3 stencil (dy outer, dx inner) then in-cell order
//      (src/kernels/base_point_particle_kernels.cu:72-94);
//    * per-system sums: system-aligned 256-element chunks, each reduced by a
//      64-lane xor butterfly per wave and ((w0+w1)+(w2+w3)) across 4 waves; chunk
//      partials are folded by 256 strided sequential accumulators and the same tree.
//  Compile with -ffp-contract=off: no a*b+c is ever fused (the HIP side is built the same
//  way).
//
//  Pinning (what fixes this restatement independently of the HIP code): Random123
//  Philox4x32-10 published known-answer vectors; reference unit-test asserts ported as
//  known-answer tests (tests/test_base_particle.cu:52-91, test_base_point_particle.cu,
//  test_disk.cu); closed-form two/three-disk forces and energies; brute-force O(N^2)
//  neighbor sets; Newton's third law; the harmonic-oscillator VV error order; FIRE state
//  machine branch tables; rigid-body closed forms (I = m R^2, torque = r x F).  See
//  tests/test_oracle_*.py.
// =====================================================================================
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <string>
#include <omp.h>

namespace {

// ------------------------------------------------------------------------------------
// Deterministic per-system reduction tree (the HIP kernels implement the same tree).
// ------------------------------------------------------------------------------------
double wave_tree(const double* v) {
    double a[64], b[64];
    for (int l = 0; l < 64; ++l) a[l] = v[l];
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < 64; ++l) b[l] = a[l] + a[l ^ off];
        for (int l = 0; l < 64; ++l) a[l] = b[l];
    }
    return a[0];
}

double block_tree(const double* v256) {
    double w0 = wave_tree(v256), w1 = wave_tree(v256 + 64);
    double w2 = wave_tree(v256 + 128), w3 = wave_tree(v256 + 192);
    return (w0 + w1) + (w2 + w3);
}

// Sum of vals[0..n) with the chunked tree.
double sys_sum(const double* vals, int n) {
    const int nch = (n + 255) / 256;
    std::vector<double> part(nch > 0 ? nch : 1, 0.0);
    double buf[256];
    for (int c = 0; c < nch; ++c) {
        for (int t = 0; t < 256; ++t) {
            const int idx = c * 256 + t;
            buf[t] = idx < n ? vals[idx] : 0.0;
        }
        part[c] = block_tree(buf);
    }
    for (int t = 0; t < 256; ++t) {
        double acc = 0.0;
        for (int c = t; c < nch; c += 256) acc = acc + part[c];
        buf[t] = acc;
    }
    return block_tree(buf);
}

// ------------------------------------------------------------------------------------
// Geometry helpers -- include/kernels/base_particle_kernels.cuh:117-179
// ------------------------------------------------------------------------------------
inline double wrap01(double u) {            // base_particle_kernels.cuh:117-120
    u -= std::floor(u);
    return (u >= 1.0) ? 0.0 : u;
}

inline double disp_pbc_L(double xi, double yi, double xj, double yj, double Lx, double Ly,
                         double ix, double iy, double& dx, double& dy) {   // :169-179
    dx = xi - xj; dy = yi - yj;
    dx = (Lx > 0.0) ? (dx - Lx * std::nearbyint(dx * ix)) : dx;
    dy = (Ly > 0.0) ? (dy - Ly * std::nearbyint(dy * iy)) : dy;
    return dx * dx + dy * dy;
}

// ------------------------------------------------------------------------------------
// Philox4x32-10 -- include/utils/device_fields.cuh:150-207 (Random123 constants)
// ------------------------------------------------------------------------------------
void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

inline double u01d(uint32_t hi, uint32_t lo) {  // device_fields.cuh:203-207
    const uint64_t bits = ((uint64_t)hi << 21) | ((uint64_t)lo >> 11);
    return ((double)bits + 0.5) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------------------------
// Disk system state (SoA, current order) -- include/particles/base_particle.hpp:31-103,
// include/particles/base_point_particle.hpp:86-89
// ------------------------------------------------------------------------------------
struct FireState {   // include/integrators/fire.cuh:116-147
    std::vector<double> dt, dt_max, dt_min, dt_rev, vscale, alpha, pe_prev;
    std::vector<int> n_good, n_bad;
    double alpha_init = 0.1, f_inc = 1.1, f_dec = 0.5, f_alpha = 0.99;
    int n_min = 5, n_bad_max = 10;
    double dt_max_scale = 10.0, dt_min_scale = 0.001;
};

struct Disk {
    int S = 0, N = 0, K = 0;       // K = neighbor capacity per particle
    int method = 1;                 // 0 naive, 1 cell
    std::vector<int> off, nx, ny, cstart, boundary, sid, static_index, cell_id;
    std::vector<double> Lx, Ly, ix, iy, e, cutoff, thresh2;
    std::vector<double> x, y, vx, vy, fx, fy, pe, rad, mass, area, lastx, lasty;
    std::vector<unsigned> flag;
    std::vector<std::vector<int>> nbr;
    int max_count = 0;
    FireState fire;
    std::vector<double> vv_dt, vv_dt_half, vv_damp;
    // saved state (base_particle.hpp:708-736; disk.cu:348-432)
    std::vector<double> sx, sy, srad, smass, sLx, sLy;
    std::vector<int> sstatic;

    void sync_box() {   // base_particle_kernels.cu:50-65
        ix.resize(S); iy.resize(S);
        for (int s = 0; s < S; ++s) {
            ix[s] = Lx[s] > 0.0 ? 1.0 / Lx[s] : 0.0;
            iy[s] = Ly[s] > 0.0 ? 1.0 / Ly[s] : 0.0;
        }
    }
    void sync_cells() {  // base_particle.hpp:552-592 (cell_system_start = exclusive scan)
        cstart.assign(S + 1, 0);
        for (int s = 0; s < S; ++s) cstart[s + 1] = cstart[s] + nx[s] * ny[s];
    }

    int hash_local(int i) const {   // base_particle_kernels.cu:205-233
        const int s = sid[i];
        const double u = wrap01(x[i] * ix[s]);
        const double v = wrap01(y[i] * iy[s]);
        int cx = (int)std::floor(u * nx[s]);
        int cy = (int)std::floor(v * ny[s]);
        cx = cx < nx[s] ? (cx > 0 ? cx : 0) : nx[s] - 1;   // guard: u*nx may round up to nx
        cy = cy < ny[s] ? (cy > 0 ? cy : 0) : ny[s] - 1;
        return cx + cy * nx[s];
    }

    // Cell rebuild of one system: hash, stable counting sort, reorder, static index,
    // reset displacements, neighbor fill.  base_point_particle.hpp:180-201,401-447;
    // disk.cu:1097-1125; base_particle_kernels.cu:172-180.
    void rebuild_system(int s) {
        const int b = off[s], en = off[s + 1], n = en - b;
        const int C = nx[s] * ny[s];
        std::vector<int> cid(n), cnt(C + 1, 0), order(n), inv(n);
        for (int t = 0; t < n; ++t) { cid[t] = hash_local(b + t); cnt[cid[t] + 1]++; }
        for (int c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
        std::vector<int> cs(cnt.begin(), cnt.end());   // cell starts (local)
        std::vector<int> wp(cnt.begin(), cnt.end() - 1);
        for (int t = 0; t < n; ++t) order[wp[cid[t]]++] = t;   // stable by current index
        for (int p = 0; p < n; ++p) inv[order[p]] = p;
        auto gather = [&](std::vector<double>& a) {
            std::vector<double> tmp(n);
            for (int p = 0; p < n; ++p) tmp[p] = a[b + order[p]];
            for (int p = 0; p < n; ++p) a[b + p] = tmp[p];
        };
        gather(x); gather(y); gather(vx); gather(vy); gather(fx); gather(fy);
        gather(rad); gather(mass); gather(pe); gather(area);
        for (int c = b; c < en; ++c) static_index[c] = b + inv[static_index[c] - b];
        for (int p = 0; p < n; ++p) {
            cell_id[b + p] = cstart[s] + cid[order[p]];
            lastx[b + p] = x[b + p]; lasty[b + p] = y[b + p];
        }
        // neighbor fill -- base_point_particle_kernels.cu:98-157 (each object's list is
        // independent: parallel over objects)
        int mc = max_count;
#pragma omp parallel for schedule(static) reduction(max : mc)
        for (int p = 0; p < n; ++p) {
            const int i = b + p;
            std::vector<int>& L = nbr[i];
            L.clear();
            const int loc = cid[order[p]];
            const int cxi = loc % nx[s], cyi = loc / nx[s];
            for (int dy = -1; dy <= 1; ++dy) {
                int yy = cyi + dy; if (yy < 0) yy += ny[s]; else if (yy >= ny[s]) yy -= ny[s];
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = cxi + dx; if (xx < 0) xx += nx[s]; else if (xx >= nx[s]) xx -= nx[s];
                    const int c = yy * nx[s] + xx;
                    for (int q = cs[c]; q < cs[c + 1]; ++q) {
                        const int j = b + q;
                        if (j == i) continue;
                        double ddx, ddy;
                        const double r2 = disp_pbc_L(x[i], y[i], x[j], y[j], Lx[s], Ly[s], ix[s], iy[s], ddx, ddy);
                        const double cut = rad[i] + rad[j] + cutoff[s];
                        if (r2 < cut * cut) L.push_back(j);
                    }
                }
            }
            mc = std::max(mc, (int)L.size());
        }
        max_count = mc;
        flag[s] = 0u;
    }

    void build_naive() {   // base_point_particle_kernels.cu:6-35
        for (int s = 0; s < S; ++s)
            for (int i = off[s]; i < off[s + 1]; ++i) {
                nbr[i].clear();
                for (int j = off[s]; j < off[s + 1]; ++j) if (j != i) nbr[i].push_back(j);
                max_count = std::max(max_count, (int)nbr[i].size());
            }
    }

    void check_neighbors(bool force) {  // base_particle.hpp:283-288,428; disk.cu:1127-1138
        if (method == 0) return;
        for (int s = 0; s < S; ++s) if (force || flag[s]) rebuild_system(s);
    }

    void init_neighbors() {   // base_particle.hpp:265-272
        nbr.assign(N, {});
        cell_id.assign(N, 0);
        lastx = x; lasty = y;
        flag.assign(S, 0u);
        if (method == 0) build_naive(); else check_neighbors(true);
    }

    // Harmonic PBC contact -- src/particles/disk.cu:103-160
    void force_pbc_particle(int i) {
        const int s = sid[i];
        const double ei = e[s];
        const double xi = x[i], yi = y[i], ri = rad[i];
        double fxi = 0.0, fyi = 0.0, pei = 0.0;
        for (int j : nbr[i]) {
            double dx, dy;
            const double r2 = disp_pbc_L(xi, yi, x[j], y[j], Lx[s], Ly[s], ix[s], iy[s], dx, dy);
            const double radsum = ri + rad[j];
            if (r2 >= radsum * radsum) continue;
            const double r = std::sqrt(r2);
            const double inv = 1.0 / r;
            const double nxv = dx * inv, nyv = dy * inv;
            const double delta = radsum - r;
            const double fmag = ei * delta;
            fxi += fmag * nxv;
            fyi += fmag * nyv;
            pei += (0.5 * ei * delta * delta) * 0.5;
        }
        fx[i] = fxi; fy[i] = fyi; pe[i] = pei;
    }

    // Harmonic walls + non-periodic repulsive pair term -- src/particles/disk.cu:163-246.
    // Deviation (documented): the pair displacement is xi - xj (repulsive, consistent with
    // the energy 0.5 k delta^2); the reference writes xj - xi there (:218).
    void force_wall_particle(int i) {
        const int s = sid[i];
        const double ei = e[s];
        const double bx = Lx[s], by = Ly[s];
        const double xi = x[i], yi = y[i], ri = rad[i];
        double fxi = 0.0, fyi = 0.0, pei = 0.0;
        if (xi < ri) { const double d = ri - xi; fxi += ei * d; pei += (0.5 * ei * d * d); }
        if (xi > bx - ri) { const double d = ri - (bx - xi); fxi -= ei * d; pei += (0.5 * ei * d * d); }
        if (yi < ri) { const double d = ri - yi; fyi += ei * d; pei += (0.5 * ei * d * d); }
        if (yi > by - ri) { const double d = ri - (by - yi); fyi -= ei * d; pei += (0.5 * ei * d * d); }
        for (int j : nbr[i]) {
            const double dx = xi - x[j], dy = yi - y[j];
            const double r2 = dx * dx + dy * dy;
            const double radsum = ri + rad[j];
            if (r2 >= radsum * radsum) continue;
            const double r = std::sqrt(r2);
            const double inv = 1.0 / r;
            const double nxv = dx * inv, nyv = dy * inv;
            const double delta = radsum - r;
            const double fmag = ei * delta;
            fxi += fmag * nxv;
            fyi += fmag * nyv;
            pei += (0.5 * ei * delta * delta) * 0.5;
        }
        fx[i] = fxi; fy[i] = fyi; pe[i] = pei;
    }

    // mode 0: PBC kernel for every system (compute_forces), 1: wall kernel for every
    // system (compute_wall_forces), 2: per-system boundary flag.
    void compute_forces(int mode) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            const bool wall = mode == 1 || (mode == 2 && boundary[sid[i]] != 0);
            if (wall) force_wall_particle(i); else force_pbc_particle(i);
        }
    }

    // disk.cu:20-57 (cell) / :60-75 (naive)
    void update_positions(const double* scale, double scale2) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            const int s = sid[i];
            const double sc = scale[s] * scale2;
            x[i] += vx[i] * sc;
            y[i] += vy[i] * sc;
        }
        if (method == 1) {
            for (int i = 0; i < N; ++i) {
                const double dx = x[i] - lastx[i], dy = y[i] - lasty[i];
                const double d2 = dx * dx + dy * dy;
                if (d2 > thresh2[sid[i]]) flag[sid[i]] = 1u;
            }
        }
    }
    void update_velocities(const double* scale, double scale2) {   // disk.cu:78-100
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) {
            const double smi = scale[sid[i]] * scale2 / mass[i];
            vx[i] = vx[i] + fx[i] * smi;
            vy[i] = vy[i] + fy[i] * smi;
        }
    }
    // compute_damping_forces_kernel (disk.cu:249-263)
    void damping(const double* scale) {
        for (int i = 0; i < N; ++i) {
            const double sc = scale[sid[i]];
            fx[i] += -vx[i] * sc;
            fy[i] += -vy[i] * sc;
        }
    }
    void scale_velocities(const double* scale) {   // disk.cu:278-289
        for (int i = 0; i < N; ++i) { vx[i] *= scale[sid[i]]; vy[i] *= scale[sid[i]]; }
    }
    void scale_positions(const double* scale) {    // disk.cu:292-303
        for (int i = 0; i < N; ++i) { x[i] *= scale[sid[i]]; y[i] *= scale[sid[i]]; }
    }
    void mix(const double* alpha) {   // disk.cu:307-335
        for (int i = 0; i < N; ++i) {
            const double a = alpha[sid[i]];
            double vxi = vx[i], vyi = vy[i];
            const double fxi = fx[i], fyi = fy[i];
            const double fn = std::sqrt(fxi * fxi + fyi * fyi);
            const double vn = std::sqrt(vxi * vxi + vyi * vyi);
            double ratio = 0.0;
            if (fn > 1e-16) ratio = vn / fn * a;
            else { vxi = 0.0; vyi = 0.0; ratio = 0.0; }
            vx[i] = vxi * (1 - a) + fxi * ratio;
            vy[i] = vyi * (1 - a) + fyi * ratio;
        }
    }

    template <class F>
    void sys_reduce(F f, double* out) {
        std::vector<double> v(N);
        for (int i = 0; i < N; ++i) v[i] = f(i);
        for (int s = 0; s < S; ++s) out[s] = sys_sum(v.data() + off[s], off[s + 1] - off[s]);
    }
    void pe_total(double* out) { sys_reduce([&](int i) { return pe[i]; }, out); }
    void ke_total(double* out) {   // disk.cu:266-275
        sys_reduce([&](int i) { return 0.5 * (vx[i] * vx[i] + vy[i] * vy[i]) * mass[i]; }, out);
    }
    void fpower_total(double* out) {   // disk.cu:338-346
        sys_reduce([&](int i) { return fx[i] * vx[i] + fy[i] * vy[i]; }, out);
    }
    void max_force2(double* out) {
        for (int s = 0; s < S; ++s) {
            double m = 0.0;
            for (int i = off[s]; i < off[s + 1]; ++i) m = std::max(m, fx[i] * fx[i] + fy[i] * fy[i]);
            out[s] = m;
        }
    }
    void temperature(double* out) {   // base_particle_kernels.cu:115-124, disk.cu:1251-1254
        ke_total(out);
        for (int s = 0; s < S; ++s) out[s] = out[s] * 2.0 / (2 * (off[s + 1] - off[s]));
    }
    void packing_fraction(double* out) {   // base_particle_kernels.cu:102-113
        sys_reduce([&](int i) { return area[i] / (Lx[sid[i]] * Ly[sid[i]]); }, out);
    }
    // Stress tensor (disk.cu:702-769) and pressure 0.5*(sxx+syy) (base_particle_kernels.cu:126-135)
    void stress(std::vector<double>& sxx, std::vector<double>& sxy, std::vector<double>& syx, std::vector<double>& syy) {
        sxx.assign(N, 0); sxy.assign(N, 0); syx.assign(N, 0); syy.assign(N, 0);
        for (int i = 0; i < N; ++i) {
            const int s = sid[i];
            const double ei = e[s], m = mass[i];
            double axx = m * vx[i] * vx[i], axy = m * vx[i] * vy[i];
            double ayx = axy, ayy = m * vy[i] * vy[i];
            const double area_box = Lx[s] * Ly[s];
            for (int j : nbr[i]) {
                double dx, dy;
                const double r2 = disp_pbc_L(x[i], y[i], x[j], y[j], Lx[s], Ly[s], ix[s], iy[s], dx, dy);
                const double radsum = rad[i] + rad[j];
                if (r2 >= radsum * radsum) continue;
                const double r = std::sqrt(r2);
                const double inv = 1.0 / r;
                const double nxv = dx * inv, nyv = dy * inv;
                const double delta = radsum - r;
                const double fmag = ei * delta;
                const double fxv = fmag * nxv, fyv = fmag * nyv;
                axx += dx * fxv / 2.0; axy += dx * fyv / 2.0;
                ayx += dy * fxv / 2.0; ayy += dy * fyv / 2.0;
            }
            sxx[i] = axx / area_box; sxy[i] = axy / area_box;
            syx[i] = ayx / area_box; syy[i] = ayy / area_box;
        }
    }
    void pressure(double* out) {
        std::vector<double> a, b, c, d;
        stress(a, b, c, d);
        std::vector<double> txx(S), tyy(S);
        for (int s = 0; s < S; ++s) {
            txx[s] = sys_sum(a.data() + off[s], off[s + 1] - off[s]);
            tyy[s] = sys_sum(d.data() + off[s], off[s + 1] - off[s]);
            out[s] = 0.5 * (txx[s] + tyy[s]);
        }
    }
    // set_average_velocity (disk.cu:771-811): mean by the chunked tree, then v -= mean.
    void remove_mean_velocity() {
        std::vector<double> mx(S), my(S);
        sys_reduce([&](int i) { return vx[i]; }, mx.data());
        sys_reduce([&](int i) { return vy[i]; }, my.data());
        for (int s = 0; s < S; ++s) {
            const int n = off[s + 1] - off[s];
            const double ax = mx[s] / n, ay = my[s] / n;
            for (int i = off[s]; i < off[s + 1]; ++i) { vx[i] -= ax; vy[i] -= ay; }
        }
    }
    // set_temperature (base_particle.hpp:449-462): scale from the current T, then remove
    // the mean (if asked), then scale.
    void set_temperature(const double* T, int remove_avg) {
        std::vector<double> cur(S), sc(S);
        temperature(cur.data());
        for (int s = 0; s < S; ++s) sc[s] = std::sqrt(T[s] / cur[s]);
        if (remove_avg) remove_mean_velocity();
        scale_velocities(sc.data());
    }

    // ---------------- FIRE (include/integrators/fire.cuh, src/integrators/fire.cu) ------
    void fire_init(const double* dt0, int mode) {   // fire.cuh:44-56,90-96,136-147
        FireState& F = fire;
        F.dt.assign(dt0, dt0 + S);
        F.dt_max.resize(S); F.dt_min.resize(S);
        for (int s = 0; s < S; ++s) { F.dt_max[s] = F.dt[s] * F.dt_max_scale; F.dt_min[s] = F.dt[s] * F.dt_min_scale; }
        F.dt_rev.assign(S, 0.0); F.vscale.assign(S, 0.0); F.alpha.assign(S, F.alpha_init);
        F.n_good.assign(S, 0); F.n_bad.assign(S, 0); F.pe_prev.assign(S, 1e9);
        scale_velocities(F.vscale.data());
        F.vscale.assign(S, 1.0);
        compute_forces(mode);
    }
    void fire_update(const double* power) {   // fire.cu:7-57
        FireState& F = fire;
        for (int s = 0; s < S; ++s) {
            F.dt_rev[s] = 0; F.vscale[s] = 1.0;
            if (F.dt[s] == 0.0) { F.vscale[s] = 0.0; F.alpha[s] = 0.0; continue; }
            if (power[s] > 0) {
                F.n_good[s]++; F.n_bad[s] = 0;
                if (F.n_good[s] > F.n_min) {
                    F.dt[s] = std::fmin(F.dt[s] * F.f_inc, F.dt_max[s]);
                    F.alpha[s] *= F.f_alpha;
                }
            } else {
                F.n_good[s] = 0; F.n_bad[s]++;
                if (F.n_bad[s] > F.n_bad_max) { F.dt[s] = 0.0; continue; }
                F.dt[s] = std::fmax(F.dt[s] * F.f_dec, F.dt_min[s]);
                F.alpha[s] = F.alpha_init;
                F.dt_rev[s] = -F.dt[s] * 0.5;
                F.vscale[s] = 0.0;
            }
        }
    }
    void fire_step(int mode) {   // fire.cuh:59-88
        std::vector<double> P(S);
        fpower_total(P.data());
        fire_update(P.data());
        update_positions(fire.dt_rev.data(), 1.0);
        scale_velocities(fire.vscale.data());
        update_velocities(fire.dt.data(), 0.5);
        mix(fire.alpha.data());
        scale_velocities(fire.vscale.data());
        update_positions(fire.dt.data(), 0.5);
        check_neighbors(false);
        compute_forces(mode);
        update_velocities(fire.dt.data(), 0.5);
    }
    // fire.cuh:98-109 with fire.cu:59-78.  pred 0: the reference's <pe> test (note the
    // double division of pe_prev, kept as written); pred 1: max|F_i| < tol (added for the
    // max|F| stopping rule; pe_prev is still updated the same way); pred 2: max|F_i| < tol
    // or <pe> < pe_target.
    int fire_converged(double diff_target, double pe_target, int pred, double ftol) {
        FireState& F = fire;
        std::vector<double> pt(S), mf(S);
        pe_total(pt.data());
        if (pred != 0) max_force2(mf.data());
        bool all = true;
        for (int s = 0; s < S; ++s) {
            const int dof = off[s + 1] - off[s];
            const double avg = pt[s] / dof;
            const double avg_prev = F.pe_prev[s] / dof;
            const double diff = avg_prev == 0 ? 0 : std::abs(avg / avg_prev - 1);
            if (pred == 0) {
                if (diff < diff_target || avg < pe_target) F.dt[s] = 0.0;
            } else if (pred == 1) {
                if (std::sqrt(mf[s]) < ftol) F.dt[s] = 0.0;
            } else {
                if (std::sqrt(mf[s]) < ftol || avg < pe_target) F.dt[s] = 0.0;
            }
            F.pe_prev[s] = avg;
            if (F.dt[s] != 0.0) all = false;
        }
        return all ? 1 : 0;
    }

    // ---------------- velocity Verlet (include/integrators/velocity_verlet.cuh:34-41) ---
    // velocity_verlet.cuh:16-32; with damping: damped_velocity_verlet.cuh:16-56 (init
    // computes the forces without the drag)
    void vv_init(const double* dt, int mode, const double* damp = nullptr) {
        vv_dt.assign(dt, dt + S);
        vv_dt_half.resize(S);
        for (int s = 0; s < S; ++s) vv_dt_half[s] = dt[s] * 0.5;
        if (damp) vv_damp.assign(damp, damp + S); else vv_damp.clear();
        compute_forces(mode);
    }
    void vv_step(int mode) {
        update_velocities(vv_dt_half.data(), 1.0);
        update_positions(vv_dt.data(), 1.0);
        check_neighbors(false);
        compute_forces(mode);
        if (!vv_damp.empty()) damping(vv_damp.data());   // damped_velocity_verlet.cuh:42-50
        update_velocities(vv_dt_half.data(), 1.0);
    }

    // increment_packing_fraction -- base_particle.hpp:505-529, kernels.cu:148-170.
    // Guard (T7): when the shrunken cell no longer covers 2*r_max + cutoff the cell grid
    // is re-derived with the same rule the library uses.
    void increment_packing_fraction(const double* inc) {
        std::vector<double> phi(S), sc(S);
        packing_fraction(phi.data());
        for (int s = 0; s < S; ++s) {
            const double nphi = phi[s] + inc[s];
            sc[s] = std::sqrt(phi[s] / nphi);
            Lx[s] = Lx[s] * sc[s]; Ly[s] = Ly[s] * sc[s];
        }
        scale_positions(sc.data());
        sync_box();
        regrid_if_needed();
        check_neighbors(true);
    }
    // Cell grid rule shared with the library: a zero dimension is derived (the most cells
    // whose side covers 2*r_max + cutoff, at least 3); after a box change (box_changed) a
    // grid whose cells became smaller than that is re-derived too.  User grids are kept at
    // set-up (the reference's KATs use cells smaller than the cutoff on purpose).
    void regrid_if_needed(bool box_changed = true) {
        if (method != 1) return;
        bool changed = false;
        for (int s = 0; s < S; ++s) {
            double rmax = 0.0;
            for (int i = off[s]; i < off[s + 1]; ++i) rmax = std::max(rmax, rad[i]);
            const double need = 2.0 * rmax + cutoff[s];
            if (nx[s] <= 0 || ny[s] <= 0 || (box_changed && (Lx[s] / nx[s] < need || Ly[s] / ny[s] < need))) {
                nx[s] = std::max(3, (int)std::floor(Lx[s] / need));
                ny[s] = std::max(3, (int)std::floor(Ly[s] / need));
                changed = true;
            }
        }
        if (changed) sync_cells();
    }

    // save/restore state (disk.cu:348-432,1158-1225; base_particle.hpp:708-736)
    void save_state(const int* flg, int tv) {
        if ((int)sx.size() != N) { sx = x; sy = y; srad = rad; smass = mass; sstatic = static_index; sLx = Lx; sLy = Ly; }
        for (int i = 0; i < N; ++i) if (flg[sid[i]] == tv) {
            sx[i] = x[i]; sy[i] = y[i]; srad[i] = rad[i]; smass[i] = mass[i]; sstatic[i] = static_index[i];
        }
        for (int s = 0; s < S; ++s) if (flg[s] == tv) { sLx[s] = Lx[s]; sLy[s] = Ly[s]; }
    }
    void restore_state(const int* flg, int tv) {
        for (int i = 0; i < N; ++i) if (flg[sid[i]] == tv) {
            x[i] = sx[i]; y[i] = sy[i]; rad[i] = srad[i]; mass[i] = smass[i]; static_index[i] = sstatic[i];
        }
        for (int s = 0; s < S; ++s) if (flg[s] == tv) { Lx[s] = sLx[s]; Ly[s] = sLy[s]; }
        sync_box();
        check_neighbors(true);
    }
};

std::string g_err;

// ------------------------------------------------------------------------------------
// sin/cos for the rigid-body rotation (rigid_bumpy.cu:51, sincos of the angle increment).
// Method: Cody-Waite reduction by pi/2 with the fdlibm two-part constant (pio2_1 = first 33
// bits of pi/2, pio2_1t = the remainder), then Taylor series to x^17 / x^18 on
// |r| <= pi/4, Horner form, no contraction.  The HIP library evaluates the identical
// expression sequence, so both produce the same bits; accuracy is checked against libm.
// ------------------------------------------------------------------------------------
void own_sincos(double x, double* sp, double* cp) {
    const double pio2_1 = 1.57079632673412561417e+00;
    const double pio2_1t = 6.07710050650619224932e-11;
    const double two_over_pi = 6.36619772367581382433e-01;
    const double k = std::nearbyint(x * two_over_pi);
    const double r = (x - k * pio2_1) - k * pio2_1t;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;            // 1/17!
    ps = -1.0 / 1307674368000.0 + r2 * ps;          // 1/15!
    ps = 1.0 / 6227020800.0 + r2 * ps;              // 1/13!
    ps = -1.0 / 39916800.0 + r2 * ps;               // 1/11!
    ps = 1.0 / 362880.0 + r2 * ps;                  // 1/9!
    ps = -1.0 / 5040.0 + r2 * ps;                   // 1/7!
    ps = 1.0 / 120.0 + r2 * ps;                     // 1/5!
    ps = -1.0 / 6.0 + r2 * ps;                      // 1/3!
    const double s = r + (r * r2) * ps;
    double pc = -1.0 / 6402373705728000.0;          // 1/18!
    pc = 1.0 / 20922789888000.0 + r2 * pc;          // 1/16!
    pc = -1.0 / 87178291200.0 + r2 * pc;            // 1/14!
    pc = 1.0 / 479001600.0 + r2 * pc;               // 1/12!
    pc = -1.0 / 3628800.0 + r2 * pc;                // 1/10!
    pc = 1.0 / 40320.0 + r2 * pc;                   // 1/8!
    pc = -1.0 / 720.0 + r2 * pc;                    // 1/6!
    pc = 1.0 / 24.0 + r2 * pc;                      // 1/4!
    const double c = (1.0 - 0.5 * r2) + (r2 * r2) * pc;
    const long q = ((long)k) & 3;
    if (q == 0) { *sp = s; *cp = c; }
    else if (q == 1) { *sp = c; *cp = -s; }
    else if (q == 2) { *sp = -s; *cp = -c; }
    else { *sp = -c; *cp = s; }
}

// ------------------------------------------------------------------------------------
// Rigid bumpy polygons -- src/particles/rigid_bumpy.cu, include/particles/base_poly_particle.hpp.
// Particle arrays (COM, angle, velocities, forces) keep canonical order; vertex arrays are
// reordered by the cell rebuild; static_index maps canonical vertex -> current vertex.
// ------------------------------------------------------------------------------------
struct Rigid {
    int S = 0, N = 0, Nv = 0;
    int method = 1;
    std::vector<int> off, voff, poff, nx, ny, cstart, boundary, sid, vsid, vpid, static_index;
    std::vector<double> Lx, Ly, ix, iy, e, cutoff, thresh2;
    std::vector<double> x, y, th, vx, vy, w, fx, fy, tq, pe, mass, inertia, area, lastx, lasty;
    std::vector<double> vpx, vpy, vfx, vfy, vpe, vrad, vmass;
    std::vector<unsigned> flag;
    std::vector<std::vector<int>> nbr;
    int max_count = 0;
    FireState fire;
    std::vector<double> vv_dt, vv_dt_half, vv_damp;

    void sync_box() {
        ix.resize(S); iy.resize(S);
        for (int s = 0; s < S; ++s) { ix[s] = Lx[s] > 0.0 ? 1.0 / Lx[s] : 0.0; iy[s] = Ly[s] > 0.0 ? 1.0 / Ly[s] : 0.0; }
    }
    void sync_cells() {
        cstart.assign(S + 1, 0);
        for (int s = 0; s < S; ++s) cstart[s + 1] = cstart[s] + nx[s] * ny[s];
    }
    void regrid_if_needed(bool box_changed) {
        for (int s = 0; s < S; ++s) {
            double rmax = 0.0;
            for (int v = voff[s]; v < voff[s + 1]; ++v) rmax = std::max(rmax, vrad[v]);
            const double need = 2.0 * rmax + cutoff[s];
            if (nx[s] <= 0 || ny[s] <= 0 || (box_changed && (Lx[s] / nx[s] < need || Ly[s] / ny[s] < need))) {
                nx[s] = std::max(3, (int)std::floor(Lx[s] / need));
                ny[s] = std::max(3, (int)std::floor(Ly[s] / need));
            }
        }
        sync_cells();
    }
    int hash_local(int v) const {
        const int s = vsid[v];
        const double u = wrap01(vpx[v] * ix[s]);
        const double q = wrap01(vpy[v] * iy[s]);
        int cx = (int)std::floor(u * nx[s]);
        int cy = (int)std::floor(q * ny[s]);
        cx = cx < nx[s] ? (cx > 0 ? cx : 0) : nx[s] - 1;
        cy = cy < ny[s] ? (cy > 0 ? cy : 0) : ny[s] - 1;
        return cx + cy * nx[s];
    }
    // vertex pipeline: base_poly_particle.hpp:239-260, base_poly_particle_kernels.cu:78-195,
    // reorder rigid_bumpy.cu:1471-1492, reset_displacements_kernel :1070-1085
    int rebuild_system(int s) {   // returns the system's largest list
        int mcount = 0;
        const int b = voff[s], en = voff[s + 1], n = en - b;
        const int C = nx[s] * ny[s];
        std::vector<int> cid(n), cnt(C + 1, 0), order(n), inv(n);
        for (int t = 0; t < n; ++t) { cid[t] = hash_local(b + t); cnt[cid[t] + 1]++; }
        for (int c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
        std::vector<int> cs(cnt.begin(), cnt.end()), wp(cnt.begin(), cnt.end() - 1);
        for (int t = 0; t < n; ++t) order[wp[cid[t]]++] = t;
        for (int p = 0; p < n; ++p) inv[order[p]] = p;
        auto gather = [&](auto& a) {   // this system's slice only (systems rebuild in parallel)
            std::vector<typename std::decay_t<decltype(a)>::value_type> tmp(n);
            for (int p = 0; p < n; ++p) tmp[p] = a[b + order[p]];
            for (int p = 0; p < n; ++p) a[b + p] = tmp[p];
        };
        gather(vpx); gather(vpy); gather(vfx); gather(vfy); gather(vmass); gather(vrad); gather(vpe); gather(vpid);
        for (int c = b; c < en; ++c) static_index[c] = b + inv[static_index[c] - b];
        std::vector<int> pcid(n);
        for (int p = 0; p < n; ++p) pcid[p] = cid[order[p]];
        for (int p = 0; p < n; ++p) {
            const int i = b + p;
            std::vector<int>& L = nbr[i];
            L.clear();
            const int loc = pcid[p];
            const int cxi = loc % nx[s], cyi = loc / nx[s];
            for (int dy = -1; dy <= 1; ++dy) {
                int yy = cyi + dy; if (yy < 0) yy += ny[s]; else if (yy >= ny[s]) yy -= ny[s];
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = cxi + dx; if (xx < 0) xx += nx[s]; else if (xx >= nx[s]) xx -= nx[s];
                    const int c = yy * nx[s] + xx;
                    for (int q = cs[c]; q < cs[c + 1]; ++q) {
                        const int j = b + q;
                        if (j == i || vpid[j] == vpid[i]) continue;
                        double ddx, ddy;
                        const double r2 = disp_pbc_L(vpx[i], vpy[i], vpx[j], vpy[j], Lx[s], Ly[s], ix[s], iy[s], ddx, ddy);
                        const double cut = vrad[i] + vrad[j] + cutoff[s];
                        if (r2 < cut * cut) L.push_back(j);
                    }
                }
            }
            mcount = std::max(mcount, (int)L.size());
        }
        for (int i = off[s]; i < off[s + 1]; ++i) {
            const int fv = static_index[poff[i]];
            lastx[i] = vpx[fv]; lasty[i] = vpy[fv];
        }
        flag[s] = 0u;
        return mcount;
    }
    void check_neighbors(bool force) {
        // systems are independent; max_count is combined after the parallel loop
        std::vector<int> mc(S, 0);
        const int m0 = max_count;
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < S; ++s) if (force || flag[s]) mc[s] = rebuild_system(s);
        max_count = m0;
        for (int s = 0; s < S; ++s) max_count = std::max(max_count, mc[s]);
    }

    // rigid_bumpy.cu:164-221 (PBC) / :224-307 (walls; pair term repulsive xi - xj)
    void vertex_force(int i, bool wall) {
        const int s = vsid[i];
        const double ei = e[s];
        const double xi = vpx[i], yi = vpy[i], ri = vrad[i];
        double fxi = 0.0, fyi = 0.0, pei = 0.0;
        if (wall) {
            const double bx = Lx[s], by = Ly[s];
            if (xi < ri) { const double d = ri - xi; fxi += ei * d; pei += (0.5 * ei * d * d); }
            if (xi > bx - ri) { const double d = ri - (bx - xi); fxi -= ei * d; pei += (0.5 * ei * d * d); }
            if (yi < ri) { const double d = ri - yi; fyi += ei * d; pei += (0.5 * ei * d * d); }
            if (yi > by - ri) { const double d = ri - (by - yi); fyi -= ei * d; pei += (0.5 * ei * d * d); }
        }
        for (int j : nbr[i]) {
            double dx = xi - vpx[j], dy = yi - vpy[j];
            if (!wall) {
                dx = (Lx[s] > 0.0) ? (dx - Lx[s] * std::nearbyint(dx * ix[s])) : dx;
                dy = (Ly[s] > 0.0) ? (dy - Ly[s] * std::nearbyint(dy * iy[s])) : dy;
            }
            const double r2 = dx * dx + dy * dy;
            const double radsum = ri + vrad[j];
            if (r2 >= radsum * radsum) continue;
            const double r = std::sqrt(r2);
            const double inv = 1.0 / r;
            const double nxv = dx * inv, nyv = dy * inv;
            const double delta = radsum - r;
            const double fmag = ei * delta;
            fxi += fmag * nxv;
            fyi += fmag * nyv;
            pei += (0.5 * ei * delta * delta) * 0.5;
        }
        vfx[i] = fxi; vfy[i] = fyi; vpe[i] = pei;
    }
    // compute_particle_forces_kernel (rigid_bumpy.cu:331-370)
    void particle_force(int i) {
        const double px = x[i], py = y[i];
        double fxi = 0.0, fyi = 0.0, t = 0.0, p = 0.0;
        for (int j = poff[i]; j < poff[i + 1]; ++j) {
            const int v = static_index[j];
            const double fvx = vfx[v], fvy = vfy[v];
            fxi += fvx; fyi += fvy;
            p += vpe[v];
            const double dx = vpx[v] - px, dy = vpy[v] - py;
            t += fvy * dx - fvx * dy;
        }
        fx[i] = fxi; fy[i] = fyi; tq[i] = t; pe[i] = p;
    }
    void compute_forces(int mode) {
#pragma omp parallel for schedule(static)
        for (int v = 0; v < Nv; ++v) vertex_force(v, mode == 1 || (mode == 2 && boundary[vsid[v]] != 0));
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; ++i) particle_force(i);
    }
    // update_positions_kernel_naive/_cell (rigid_bumpy.cu:20-127)
    void update_positions(const double* scale, double scale2) {
        for (int i = 0; i < N; ++i) {
            const int s = sid[i];
            const double sc = scale[s] * scale2;
            const double x0 = x[i], y0 = y[i], a0 = th[i];
            const double nx_ = x0 + vx[i] * sc, ny_ = y0 + vy[i] * sc, na = a0 + w[i] * sc;
            x[i] = nx_; y[i] = ny_; th[i] = na;
            double sn, cs;
            own_sincos(na - a0, &sn, &cs);
            for (int j = poff[i]; j < poff[i + 1]; ++j) {
                const int v = static_index[j];
                const double dx = vpx[v] - x0, dy = vpy[v] - y0;
                const double rx = cs * dx - sn * dy, ry = sn * dx + cs * dy;
                vpx[v] = nx_ + rx; vpy[v] = ny_ + ry;
            }
            if (method == 1) {
                const int fv = static_index[poff[i]];
                const double dx = vpx[fv] - lastx[i], dy = vpy[fv] - lasty[i];
                if (dx * dx + dy * dy > thresh2[s]) flag[s] = 1u;
            }
        }
    }
    // update_velocities_kernel (rigid_bumpy.cu:130-161)
    // compute_damping_forces_kernel (rigid_bumpy.cu:310-327)
    void damping(const double* scale) {
        for (int i = 0; i < N; ++i) {
            const double sc = scale[sid[i]];
            fx[i] += -vx[i] * sc;
            fy[i] += -vy[i] * sc;
            tq[i] += -w[i] * sc;
        }
    }
    void update_velocities(const double* scale, double scale2) {
        for (int i = 0; i < N; ++i) {
            const double sc = scale[sid[i]] * scale2;
            const double smi = sc / mass[i];
            const double sii = inertia[i] != 0.0 ? sc / inertia[i] : 0.0;
            vx[i] = vx[i] + fx[i] * smi; vy[i] = vy[i] + fy[i] * smi; w[i] = w[i] + tq[i] * sii;
        }
    }
    void scale_velocities(const double* sc) {   // rigid_bumpy.cu:540-553
        for (int i = 0; i < N; ++i) { vx[i] *= sc[sid[i]]; vy[i] *= sc[sid[i]]; w[i] *= sc[sid[i]]; }
    }
    void scale_positions(const double* sc) {    // rigid_bumpy.cu:555-583
        for (int i = 0; i < N; ++i) {
            const double px = x[i], py = y[i];
            const double nx_ = px * sc[sid[i]], ny_ = py * sc[sid[i]];
            x[i] = nx_; y[i] = ny_;
            for (int j = poff[i]; j < poff[i + 1]; ++j) {
                const int v = static_index[j];
                vpx[v] += (nx_ - px); vpy[v] += (ny_ - py);
            }
        }
    }
    void mix(const double* alpha) {   // rigid_bumpy.cu:586-629
        for (int i = 0; i < N; ++i) {
            const double a = alpha[sid[i]];
            if (a == 0.0) continue;
            double vxi = vx[i], vyi = vy[i], wi = w[i];
            const double fxi = fx[i], fyi = fy[i], ti = tq[i];
            const double fn = std::sqrt(fxi * fxi + fyi * fyi), vn = std::sqrt(vxi * vxi + vyi * vyi);
            double ratio = 0.0;
            if (fn > 1e-16) ratio = vn / fn * a; else { vxi = 0.0; vyi = 0.0; }
            const double tn = std::fabs(ti), wn = std::fabs(wi);
            double tratio = 0.0;
            if (tn > 1e-16) tratio = wn / tn * a; else { wi = 0.0; tratio = 0.0; }
            vx[i] = vxi * (1 - a) + fxi * ratio;
            vy[i] = vyi * (1 - a) + fyi * ratio;
            w[i] = wi * (1 - a) + ti * tratio;
        }
    }
    template <class Fn>
    void sys_reduce(Fn f, double* out) {
        std::vector<double> v(N);
        for (int i = 0; i < N; ++i) v[i] = f(i);
        for (int s = 0; s < S; ++s) out[s] = sys_sum(v.data() + off[s], off[s + 1] - off[s]);
    }
    void pe_total(double* o) { sys_reduce([&](int i) { return pe[i]; }, o); }
    void ke_total(double* o) {
        sys_reduce([&](int i) { return 0.5 * ((vx[i] * vx[i] + vy[i] * vy[i]) * mass[i] + w[i] * w[i] * inertia[i]); }, o);
    }
    void fpower_total(double* o) {
        sys_reduce([&](int i) { return fx[i] * vx[i] + fy[i] * vy[i] + tq[i] * w[i]; }, o);
    }
    void max_force2(double* o) {
        for (int s = 0; s < S; ++s) {
            double m = 0.0;
            for (int i = off[s]; i < off[s + 1]; ++i) m = std::max(m, fx[i] * fx[i] + fy[i] * fy[i]);
            o[s] = m;
        }
    }
    // set_n_dof_kernel (rigid_bumpy.cu:743-757): 2 per particle plus 1 per particle whose
    // moment of inertia is positive.
    int n_dof(int s) const {
        int d = 0;
        for (int i = off[s]; i < off[s + 1]; ++i) d += 2 + (inertia[i] > 0.0 ? 1 : 0);
        return d;
    }
    // compute_temperature_kernel (base_particle_kernels.cu:115-124): T = 2 KE / n_dof.
    void temperature(double* o) {
        ke_total(o);
        for (int s = 0; s < S; ++s) o[s] = o[s] * 2.0 / (n_dof(s));
    }
    // compute_stress_tensor_virial_kernel + _kinetic_kernel (rigid_bumpy.cu:968-1068): per
    // particle, vertex contributions summed in canonical vertex order (the reference adds
    // them with atomics in arbitrary order).
    void pressure(double* o) {
        std::vector<double> sxx(N), syy(N);
        for (int i = 0; i < N; ++i) {
            const int s = sid[i];
            const double ei = e[s];
            double axx = 0.0, ayy = 0.0;
            for (int jc = poff[i]; jc < poff[i + 1]; ++jc) {
                const int v = static_index[jc];
                double vxx = 0.0, vyy = 0.0;
                for (int u : nbr[v]) {
                    double dx, dy, dxp, dyp;
                    const double r2 = disp_pbc_L(vpx[v], vpy[v], vpx[u], vpy[u], Lx[s], Ly[s], ix[s], iy[s], dx, dy);
                    disp_pbc_L(x[i], y[i], x[vpid[u]], y[vpid[u]], Lx[s], Ly[s], ix[s], iy[s], dxp, dyp);
                    const double radsum = vrad[v] + vrad[u];
                    if (r2 >= radsum * radsum) continue;
                    const double r = std::sqrt(r2);
                    const double inv = 1.0 / r;
                    const double delta = radsum - r;
                    const double fmag = ei * delta;
                    const double fxv = fmag * (dx * inv), fyv = fmag * (dy * inv);
                    vxx += dxp * fxv / 2.0;
                    vyy += dyp * fyv / 2.0;
                }
                axx += vxx; ayy += vyy;
            }
            const double area_box = Lx[s] * Ly[s];
            sxx[i] = (axx + mass[i] * vx[i] * vx[i]) / area_box;
            syy[i] = (ayy + mass[i] * vy[i] * vy[i]) / area_box;
        }
        for (int s = 0; s < S; ++s) {
            const double a = sys_sum(sxx.data() + off[s], off[s + 1] - off[s]);
            const double b = sys_sum(syy.data() + off[s], off[s + 1] - off[s]);
            o[s] = 0.5 * (a + b);
        }
    }
    void fire_init(const double* dt0, int mode) {
        FireState& F = fire;
        F.dt.assign(dt0, dt0 + S);
        F.dt_max.resize(S); F.dt_min.resize(S);
        for (int s = 0; s < S; ++s) { F.dt_max[s] = F.dt[s] * F.dt_max_scale; F.dt_min[s] = F.dt[s] * F.dt_min_scale; }
        F.dt_rev.assign(S, 0.0); F.vscale.assign(S, 0.0); F.alpha.assign(S, F.alpha_init);
        F.n_good.assign(S, 0); F.n_bad.assign(S, 0); F.pe_prev.assign(S, 1e9);
        scale_velocities(F.vscale.data());
        F.vscale.assign(S, 1.0);
        compute_forces(mode);
    }
    void fire_step(int mode) {
        std::vector<double> P(S);
        fpower_total(P.data());
        Disk d; d.S = S; d.fire = fire;
        d.fire_update(P.data());
        fire = d.fire;
        update_positions(fire.dt_rev.data(), 1.0);
        scale_velocities(fire.vscale.data());
        update_velocities(fire.dt.data(), 0.5);
        mix(fire.alpha.data());
        scale_velocities(fire.vscale.data());
        update_positions(fire.dt.data(), 0.5);
        check_neighbors(false);
        compute_forces(mode);
        update_velocities(fire.dt.data(), 0.5);
    }
    int fire_converged(double diff_target, double pe_target, int pred, double ftol) {
        FireState& F = fire;
        std::vector<double> pt(S), mf(S);
        pe_total(pt.data());
        max_force2(mf.data());
        bool all = true;
        for (int s = 0; s < S; ++s) {
            const int dof = off[s + 1] - off[s];
            const double avg = pt[s] / dof, avg_prev = F.pe_prev[s] / dof;
            const double diff = avg_prev == 0 ? 0 : std::abs(avg / avg_prev - 1);
            if (pred == 0) { if (diff < diff_target || avg < pe_target) F.dt[s] = 0.0; }
            else if (pred == 1) { if (std::sqrt(mf[s]) < ftol) F.dt[s] = 0.0; }
            else { if (std::sqrt(mf[s]) < ftol || avg < pe_target) F.dt[s] = 0.0; }
            F.pe_prev[s] = avg;
            if (F.dt[s] != 0.0) all = false;
        }
        return all ? 1 : 0;
    }
    // compute_fractional_packing_fraction_kernel (base_particle_kernels.cu:102-113)
    void packing_fraction(double* out) {
        sys_reduce([&](int i) { return area[i] / (Lx[sid[i]] * Ly[sid[i]]); }, out);
    }
    // increment_packing_fraction (base_particle.hpp:505-529) with the library's grid guard
    void increment_packing_fraction(const double* inc) {
        std::vector<double> phi(S), sc(S);
        packing_fraction(phi.data());
        for (int s = 0; s < S; ++s) {
            const double nphi = phi[s] + inc[s];
            sc[s] = std::sqrt(phi[s] / nphi);
            Lx[s] = Lx[s] * sc[s]; Ly[s] = Ly[s] * sc[s];
        }
        scale_positions(sc.data());
        sync_box();
        regrid_if_needed(true);
        check_neighbors(true);
    }
    // save_state_impl / restore_state_impl (rigid_bumpy.cu:632-741,1763-1854): COM and angle
    // per particle; vertex positions, radii, particle ids and static_index in current order
    // (vertex masses kept with them); box.  restore_state (base_particle.hpp:717-727) then
    // syncs the box and forces a rebuild.
    std::vector<double> sx, sy, sth, svpx, svpy, svrad, svmass, sLx, sLy;
    std::vector<int> svpid, sstatic;
    void save_state(const int* flg, int tv) {
        if ((int)sx.size() != N) {
            sx = x; sy = y; sth = th; svpx = vpx; svpy = vpy; svrad = vrad; svmass = vmass; svpid = vpid;
            sstatic = static_index; sLx = Lx; sLy = Ly;
        }
        for (int i = 0; i < N; ++i) if (flg[sid[i]] == tv) { sx[i] = x[i]; sy[i] = y[i]; sth[i] = th[i]; }
        for (int v = 0; v < Nv; ++v) if (flg[vsid[v]] == tv) {
            svpx[v] = vpx[v]; svpy[v] = vpy[v]; svrad[v] = vrad[v]; svmass[v] = vmass[v]; svpid[v] = vpid[v];
            sstatic[v] = static_index[v];
        }
        for (int s = 0; s < S; ++s) if (flg[s] == tv) { sLx[s] = Lx[s]; sLy[s] = Ly[s]; }
    }
    void restore_state(const int* flg, int tv) {
        for (int i = 0; i < N; ++i) if (flg[sid[i]] == tv) { x[i] = sx[i]; y[i] = sy[i]; th[i] = sth[i]; }
        for (int v = 0; v < Nv; ++v) if (flg[vsid[v]] == tv) {
            vpx[v] = svpx[v]; vpy[v] = svpy[v]; vrad[v] = svrad[v]; vmass[v] = svmass[v]; vpid[v] = svpid[v];
            static_index[v] = sstatic[v];
        }
        for (int s = 0; s < S; ++s) if (flg[s] == tv) { Lx[s] = sLx[s]; Ly[s] = sLy[s]; }
        sync_box();
        check_neighbors(true);
    }

    // velocity_verlet.cuh:16-32; with damping: damped_velocity_verlet.cuh:16-56 (init
    // computes the forces without the drag)
    void vv_init(const double* dt, int mode, const double* damp = nullptr) {
        vv_dt.assign(dt, dt + S);
        vv_dt_half.resize(S);
        for (int s = 0; s < S; ++s) vv_dt_half[s] = dt[s] * 0.5;
        if (damp) vv_damp.assign(damp, damp + S); else vv_damp.clear();
        compute_forces(mode);
    }
    void vv_step(int mode) {
        update_velocities(vv_dt_half.data(), 1.0);
        update_positions(vv_dt.data(), 1.0);
        check_neighbors(false);
        compute_forces(mode);
        if (!vv_damp.empty()) damping(vv_damp.data());   // damped_velocity_verlet.cuh:42-50
        update_velocities(vv_dt_half.data(), 1.0);
    }
};

}  // namespace

extern "C" {
void or_sincos(double x, double* s, double* c) { own_sincos(x, s, c); }

// Rigid constructor: per system: off (S+1 particle offsets); per particle: nvpp, x, y, angle,
// mass, inertia, area; per vertex (canonical order): vpx, vpy, vrad, vmass.
void* or_rigid_create(int S, int N, int Nv, const int* off, const double* Lx, const double* Ly, const double* e,
                      const double* cutoff, const double* thresh2, const int* nx, const int* ny,
                      const int* boundary, const int* nvpp, const double* x, const double* y, const double* th,
                      const double* mass, const double* inertia, const double* area, const double* vpx,
                      const double* vpy, const double* vrad, const double* vmass) {
    Rigid* r = new Rigid();
    r->S = S; r->N = N; r->Nv = Nv;
    r->off.assign(off, off + S + 1);
    r->Lx.assign(Lx, Lx + S); r->Ly.assign(Ly, Ly + S); r->e.assign(e, e + S);
    r->cutoff.assign(cutoff, cutoff + S); r->thresh2.assign(thresh2, thresh2 + S);
    r->nx.assign(nx, nx + S); r->ny.assign(ny, ny + S); r->boundary.assign(boundary, boundary + S);
    r->x.assign(x, x + N); r->y.assign(y, y + N); r->th.assign(th, th + N);
    r->mass.assign(mass, mass + N); r->inertia.assign(inertia, inertia + N); r->area.assign(area, area + N);
    r->vx.assign(N, 0.0); r->vy.assign(N, 0.0); r->w.assign(N, 0.0);
    r->fx.assign(N, 0.0); r->fy.assign(N, 0.0); r->tq.assign(N, 0.0); r->pe.assign(N, 0.0);
    r->lastx.assign(N, 0.0); r->lasty.assign(N, 0.0);
    r->vpx.assign(vpx, vpx + Nv); r->vpy.assign(vpy, vpy + Nv); r->vrad.assign(vrad, vrad + Nv);
    r->vmass.assign(vmass, vmass + Nv);
    r->vfx.assign(Nv, 0.0); r->vfy.assign(Nv, 0.0); r->vpe.assign(Nv, 0.0);
    r->sid.resize(N); r->poff.assign(N + 1, 0); r->vpid.resize(Nv); r->vsid.resize(Nv);
    for (int s = 0; s < S; ++s) for (int i = off[s]; i < off[s + 1]; ++i) r->sid[i] = s;
    for (int i = 0; i < N; ++i) r->poff[i + 1] = r->poff[i] + nvpp[i];
    for (int i = 0; i < N; ++i) for (int v = r->poff[i]; v < r->poff[i + 1]; ++v) { r->vpid[v] = i; r->vsid[v] = r->sid[i]; }
    r->voff.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) r->voff[s + 1] = r->poff[off[s + 1]];
    r->static_index.resize(Nv);
    for (int v = 0; v < Nv; ++v) r->static_index[v] = v;
    r->sync_box();
    r->regrid_if_needed(false);
    r->nbr.assign(Nv, {});
    r->flag.assign(S, 0u);
    r->check_neighbors(true);
    return r;
}
void or_rigid_destroy(void* h) { delete (Rigid*)h; }
// field ids: particle 0 x 1 y 2 angle 3 vx 4 vy 5 w 6 fx 7 fy 8 torque 9 pe; vertex 20 vpx 21 vpy
// 22 vfx 23 vfy 24 vpe 25 vrad
static std::vector<double>* rfield(Rigid* r, int f) {
    switch (f) {
        case 0: return &r->x; case 1: return &r->y; case 2: return &r->th; case 3: return &r->vx;
        case 4: return &r->vy; case 5: return &r->w; case 6: return &r->fx; case 7: return &r->fy;
        case 8: return &r->tq; case 9: return &r->pe;
        case 20: return &r->vpx; case 21: return &r->vpy; case 22: return &r->vfx; case 23: return &r->vfy;
        case 24: return &r->vpe; case 25: return &r->vrad;
    }
    return nullptr;
}
void or_rigid_get(void* h, int f, double* out) { auto* v = rfield((Rigid*)h, f); std::memcpy(out, v->data(), v->size() * 8); }
void or_rigid_set(void* h, int f, const double* in) { auto* v = rfield((Rigid*)h, f); std::memcpy(v->data(), in, v->size() * 8); }
void or_rigid_get_static_index(void* h, int* out) { Rigid* r = (Rigid*)h; std::memcpy(out, r->static_index.data(), r->Nv * 4); }
void or_rigid_flags(void* h, unsigned* out) { Rigid* r = (Rigid*)h; for (int s = 0; s < r->S; ++s) out[s] = r->flag[s]; }
void or_rigid_get_cells(void* h, int* nx, int* ny) { Rigid* r = (Rigid*)h; for (int s = 0; s < r->S; ++s) { nx[s] = r->nx[s]; ny[s] = r->ny[s]; } }
int or_rigid_n_neighbors(void* h) { Rigid* r = (Rigid*)h; long t = 0; for (auto& L : r->nbr) t += L.size(); return (int)t; }
int or_rigid_max_count(void* h) { return ((Rigid*)h)->max_count; }
void or_rigid_get_neighbors(void* h, int* start, int* ids) {
    Rigid* r = (Rigid*)h;
    int w = 0;
    for (int i = 0; i < r->Nv; ++i) { start[i] = w; for (int j : r->nbr[i]) ids[w++] = j; }
    start[r->Nv] = w;
}
void or_rigid_check_neighbors(void* h, int force) { ((Rigid*)h)->check_neighbors(force != 0); }
void or_rigid_compute_forces(void* h, int mode) { ((Rigid*)h)->compute_forces(mode); }
void or_rigid_update_positions(void* h, const double* sc, double s2) { ((Rigid*)h)->update_positions(sc, s2); }
void or_rigid_update_velocities(void* h, const double* sc, double s2) { ((Rigid*)h)->update_velocities(sc, s2); }
void or_rigid_scale_velocities(void* h, const double* sc) { ((Rigid*)h)->scale_velocities(sc); }
void or_rigid_scale_positions(void* h, const double* sc) { ((Rigid*)h)->scale_positions(sc); }
void or_rigid_mix(void* h, const double* a) { ((Rigid*)h)->mix(a); }
void or_rigid_pe_total(void* h, double* o) { ((Rigid*)h)->pe_total(o); }
void or_rigid_ke_total(void* h, double* o) { ((Rigid*)h)->ke_total(o); }
void or_rigid_fpower_total(void* h, double* o) { ((Rigid*)h)->fpower_total(o); }
void or_rigid_pressure(void* h, double* o) { ((Rigid*)h)->pressure(o); }
void or_rigid_temperature(void* h, double* o) { ((Rigid*)h)->temperature(o); }
int or_rigid_n_dof(void* h, int s) { return ((Rigid*)h)->n_dof(s); }
void or_rigid_max_force2(void* h, double* o) { ((Rigid*)h)->max_force2(o); }
void or_rigid_fire_init(void* h, const double* dt, int mode) { ((Rigid*)h)->fire_init(dt, mode); }
void or_rigid_fire_step(void* h, int mode) { ((Rigid*)h)->fire_step(mode); }
int or_rigid_fire_converged(void* h, double d, double t, int pred, double ftol) { return ((Rigid*)h)->fire_converged(d, t, pred, ftol); }
int or_rigid_fire_minimize(void* h, const double* dt, int max_steps, double d, double t, int pred, double ftol, int mode, int* conv) {
    Rigid* r = (Rigid*)h;
    r->fire_init(dt, mode);
    int step = 0, c = 0;
    while (step < max_steps && !c) { r->fire_step(mode); c = r->fire_converged(d, t, pred, ftol); ++step; }
    *conv = c;
    return step;
}
void or_rigid_vv_init(void* h, const double* dt, int mode) { ((Rigid*)h)->vv_init(dt, mode); }
void or_rigid_damped_vv_init(void* h, const double* dt, const double* damp, int mode) {
    ((Rigid*)h)->vv_init(dt, mode, damp);
}
void or_rigid_damping(void* h, const double* scale) { ((Rigid*)h)->damping(scale); }
void or_rigid_vv_run(void* h, int n, int mode) { Rigid* r = (Rigid*)h; for (int k = 0; k < n; ++k) r->vv_step(mode); }
}  // extern "C"

// =====================================================================================
//  extern "C" surface for the Python test harness (ctypes)
// =====================================================================================
extern "C" {

const char* or_last_error() { return g_err.c_str(); }

void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox10(c, key[0], key[1]);
    for (int k = 0; k < 4; ++k) out[k] = c[k];
}

// device_fields.cuh:256-278 (double branch)
void or_rand_uniform_2d(int n, uint64_t seed, uint64_t call, double lox, double hix, double loy,
                        double hiy, double* ox, double* oy) {
    for (int i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)call, 0u, 0u};
        philox10(c, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
        double ux = u01d(c[0], c[1]); if (ux == 1.0) ux = 0.0;
        double uy = u01d(c[2], c[3]); if (uy == 1.0) uy = 0.0;
        ox[i] = lox + (hix - lox) * ux;
        oy[i] = loy + (hiy - loy) * uy;
    }
}
// device_fields.cuh:219-236 (double branch)
void or_rand_uniform_1d(int n, uint64_t seed, uint64_t call, double lo, double hi, double* o) {
    for (int i = 0; i < n; ++i) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)call, 0u, 0u};
        philox10(c, (uint32_t)(seed & 0xFFFFFFFFull), (uint32_t)(seed >> 32));
        double u = u01d(c[0], c[1]); if (u == 1.0) u = 0.0;
        o[i] = lo + (hi - lo) * u;
    }
}

double or_sys_sum(const double* v, int n) { return sys_sum(v, n); }

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void* or_disk_create(int S, int N, int K, int method, const int* off, const double* Lx, const double* Ly,
                     const double* e, const double* cutoff, const double* thresh2, const int* nx,
                     const int* ny, const int* boundary, const double* x, const double* y,
                     const double* rad, const double* mass, const double* area) {
    Disk* d = new Disk();
    d->S = S; d->N = N; d->K = K; d->method = method;
    d->off.assign(off, off + S + 1);
    d->Lx.assign(Lx, Lx + S); d->Ly.assign(Ly, Ly + S);
    d->e.assign(e, e + S); d->cutoff.assign(cutoff, cutoff + S); d->thresh2.assign(thresh2, thresh2 + S);
    d->nx.assign(nx, nx + S); d->ny.assign(ny, ny + S); d->boundary.assign(boundary, boundary + S);
    d->x.assign(x, x + N); d->y.assign(y, y + N); d->rad.assign(rad, rad + N); d->mass.assign(mass, mass + N);
    d->area.assign(area, area + N);
    d->vx.assign(N, 0.0); d->vy.assign(N, 0.0); d->fx.assign(N, 0.0); d->fy.assign(N, 0.0); d->pe.assign(N, 0.0);
    d->sid.resize(N);
    for (int s = 0; s < S; ++s) for (int i = off[s]; i < off[s + 1]; ++i) d->sid[i] = s;
    d->static_index.resize(N);
    for (int i = 0; i < N; ++i) d->static_index[i] = i;
    d->sync_box();
    d->regrid_if_needed(false);
    d->sync_cells();
    d->init_neighbors();
    return d;
}
void or_disk_destroy(void* h) { delete (Disk*)h; }

// field ids: 0 x,1 y,2 vx,3 vy,4 fx,5 fy,6 pe,7 rad,8 mass,9 area,10 lastx,11 lasty
static std::vector<double>* dfield(Disk* d, int f) {
    switch (f) {
        case 0: return &d->x; case 1: return &d->y; case 2: return &d->vx; case 3: return &d->vy;
        case 4: return &d->fx; case 5: return &d->fy; case 6: return &d->pe; case 7: return &d->rad;
        case 8: return &d->mass; case 9: return &d->area; case 10: return &d->lastx; case 11: return &d->lasty;
        case 12: return &d->Lx; case 13: return &d->Ly;
    }
    return nullptr;
}
void or_disk_get(void* h, int f, double* out) { auto* v = dfield((Disk*)h, f); std::memcpy(out, v->data(), v->size() * 8); }
void or_disk_set(void* h, int f, const double* in) { auto* v = dfield((Disk*)h, f); std::memcpy(v->data(), in, v->size() * 8); }
void or_disk_get_static_index(void* h, int* out) { Disk* d = (Disk*)h; std::memcpy(out, d->static_index.data(), d->N * 4); }
void or_disk_get_cells(void* h, int* nx, int* ny) {
    Disk* d = (Disk*)h;
    for (int s = 0; s < d->S; ++s) { nx[s] = d->nx[s]; ny[s] = d->ny[s]; }
}
int or_disk_max_count(void* h) { return ((Disk*)h)->max_count; }
void or_disk_get_cell_ids(void* h, int* out) { Disk* d = (Disk*)h; std::memcpy(out, d->cell_id.data(), d->N * 4); }
void or_disk_get_cell_system_start(void* h, int* out) { Disk* d = (Disk*)h; for (int s = 0; s <= d->S; ++s) out[s] = d->cstart[s]; }

// init_cell_sizes_kernel (base_particle_kernels.cu:68-99) + the exclusive scan of cells per
// system (base_particle.hpp:583-591).
void or_cell_layout(int S, const double* Lx, const double* Ly, const int* nx, const int* ny, double* csx,
                    double* csy, double* invx, double* invy, int* start) {
    start[0] = 0;
    for (int s = 0; s < S; ++s) {
        csx[s] = nx[s] > 0 ? Lx[s] / nx[s] : 0.0;
        csy[s] = ny[s] > 0 ? Ly[s] / ny[s] : 0.0;
        invx[s] = csx[s] > 0.0 ? 1.0 / csx[s] : 0.0;
        invy[s] = csy[s] > 0.0 ? 1.0 / csy[s] : 0.0;
        start[s + 1] = start[s] + nx[s] * ny[s];
    }
}
void or_box_inv(int S, const double* Lx, const double* Ly, double* ix, double* iy) {   // kernels.cu:50-65
    for (int s = 0; s < S; ++s) { ix[s] = Lx[s] > 0.0 ? 1.0 / Lx[s] : 0.0; iy[s] = Ly[s] > 0.0 ? 1.0 / Ly[s] : 0.0; }
}
int or_disk_n_neighbors(void* h) { Disk* d = (Disk*)h; long t = 0; for (auto& L : d->nbr) t += L.size(); return (int)t; }
// CSR export: start[N+1], ids[total]
void or_disk_get_neighbors(void* h, int* start, int* ids) {
    Disk* d = (Disk*)h;
    int w = 0;
    for (int i = 0; i < d->N; ++i) { start[i] = w; for (int j : d->nbr[i]) ids[w++] = j; }
    start[d->N] = w;
}
// compute_contacts_kernel / compute_overlaps_kernel (disk.cu:434-471,663-700): contacts[i],
// summed overlap and contact count over the Verlet list, minimum image in every box.
void or_disk_contacts(void* h, int* contacts, double* overlap, double* count) {
    Disk* d = (Disk*)h;
    for (int i = 0; i < d->N; ++i) {
        const int s = d->sid[i];
        int nc = 0;
        double ov = 0.0, cc = 0.0;
        for (int j : d->nbr[i]) {
            double dx, dy;
            const double r2 = disp_pbc_L(d->x[i], d->y[i], d->x[j], d->y[j], d->Lx[s], d->Ly[s], d->ix[s], d->iy[s], dx, dy);
            const double radsum = d->rad[i] + d->rad[j];
            if (r2 >= radsum * radsum) continue;
            ++nc; ov += radsum - std::sqrt(r2); cc += 1.0;
        }
        contacts[i] = nc; overlap[i] = ov; count[i] = cc;
    }
}
// compute_pair_dist / compute_particle_pair_forces / compute_hessian kernels (disk.cu:473-661)
// per CSR entry of the current lists, canonical ids.  The Hessian loop is the reference's,
// derivative factors and all, accumulated into zeroed storage.
void or_disk_pair_obs(void* h, int* pi, int* pj, double* dist, double* fx, double* fy, double* hs) {
    Disk* d = (Disk*)h;
    std::vector<int> canon(d->N);
    for (int c = 0; c < d->N; ++c) canon[d->static_index[c]] = c;
    long k = 0, tot = 0;
    for (auto& L : d->nbr) tot += (long)L.size();
    for (int i = 0; i < d->N; ++i) {
        const int s = d->sid[i];
        const double e = d->e[s];
        for (int j : d->nbr[i]) {
            double dx, dy;
            const double r2 = disp_pbc_L(d->x[i], d->y[i], d->x[j], d->y[j], d->Lx[s], d->Ly[s], d->ix[s], d->iy[s], dx, dy);
            pi[k] = canon[i]; pj[k] = canon[j];
            dist[k] = std::sqrt(r2);
            fx[k] = 0.0; fy[k] = 0.0;
            for (int m = 0; m < 8; ++m) hs[m * tot + k] = 0.0;
            const double radsum = d->rad[i] + d->rad[j];
            if (r2 < radsum * radsum) {
                const double r = std::sqrt(r2);
                const double inv = 1.0 / r;
                const double nx = dx * inv, ny = dy * inv;
                const double delta = radsum - r;
                const double fmag = e * delta;
                fx[k] += fmag * nx;
                fy[k] += fmag * ny;
                const double t_ij = -e * delta, c_ij = e;
                double hib[4] = {0.0, 0.0, 0.0, 0.0}, hjb[4] = {0.0, 0.0, 0.0, 0.0};
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const double d_ia_dx = 1 * (a == 0), d_ia_dy = 1 * (a == 1);
                        const double d_jb_dx = -1 * (b == 0), d_jb_dy = -1 * (b == 1);
                        const double q_ia = dx * d_ia_dx + dy * d_ia_dy;
                        const double q_jb = dx * d_jb_dx + dy * d_jb_dy;
                        const double d_ia_q_jb = d_ia_dx * d_jb_dx + dx * 0.0 + d_ia_dy * d_jb_dy + dy * 0.0;
                        hjb[a * 2 + b] += c_ij * (q_ia * q_jb) / (r2) + t_ij * (d_ia_q_jb / r - (q_ia * q_jb) / (r2 * r));
                        const double d_ib_dx = 1 * (b == 0), d_ib_dy = 1 * (b == 1);
                        const double q_ib = dx * d_ib_dx + dy * d_ib_dy;
                        const double d_ia_q_ib = d_ia_dx * d_ib_dx + dx * 0.0 + d_ia_dy * d_ib_dy + dy * 0.0;
                        hib[a * 2 + b] += c_ij * (q_ia * q_ib) / (r2) + t_ij * (d_ia_q_ib / r - (q_ia * q_ib) / (r2 * r));
                    }
                // columns: ix_ix, ix_jx, ix_iy, ix_jy, iy_ix, iy_jx, iy_iy, iy_jy
                for (int ab = 0; ab < 4; ++ab) {
                    hs[(2 * ab) * tot + k] += hib[ab];
                    hs[(2 * ab + 1) * tot + k] += hjb[ab];
                }
            }
            ++k;
        }
    }
}

void or_disk_flags(void* h, unsigned* out) { Disk* d = (Disk*)h; for (int s = 0; s < d->S; ++s) out[s] = d->flag[s]; }
void or_disk_check_neighbors(void* h, int force) { ((Disk*)h)->check_neighbors(force != 0); }
void or_disk_compute_forces(void* h, int mode) { ((Disk*)h)->compute_forces(mode); }
void or_disk_update_positions(void* h, const double* sc, double s2) { ((Disk*)h)->update_positions(sc, s2); }
void or_disk_update_velocities(void* h, const double* sc, double s2) { ((Disk*)h)->update_velocities(sc, s2); }
void or_disk_scale_velocities(void* h, const double* sc) { ((Disk*)h)->scale_velocities(sc); }
void or_disk_scale_positions(void* h, const double* sc) { ((Disk*)h)->scale_positions(sc); }
void or_disk_mix(void* h, const double* a) { ((Disk*)h)->mix(a); }
void or_disk_pe_total(void* h, double* o) { ((Disk*)h)->pe_total(o); }
void or_disk_ke_total(void* h, double* o) { ((Disk*)h)->ke_total(o); }
void or_disk_fpower_total(void* h, double* o) { ((Disk*)h)->fpower_total(o); }
void or_disk_temperature(void* h, double* o) { ((Disk*)h)->temperature(o); }
void or_disk_pressure(void* h, double* o) { ((Disk*)h)->pressure(o); }
void or_disk_packing_fraction(void* h, double* o) { ((Disk*)h)->packing_fraction(o); }
void or_disk_max_force2(void* h, double* o) { ((Disk*)h)->max_force2(o); }
void or_disk_stress(void* h, double* xx, double* xy, double* yx, double* yy) {
    Disk* d = (Disk*)h;
    std::vector<double> a, b, c, e;
    d->stress(a, b, c, e);
    std::memcpy(xx, a.data(), d->N * 8); std::memcpy(xy, b.data(), d->N * 8);
    std::memcpy(yx, c.data(), d->N * 8); std::memcpy(yy, e.data(), d->N * 8);
}
void or_disk_set_temperature(void* h, const double* T, int rm) { ((Disk*)h)->set_temperature(T, rm); }
void or_disk_remove_mean_velocity(void* h) { ((Disk*)h)->remove_mean_velocity(); }
void or_disk_increment_packing_fraction(void* h, const double* inc) { ((Disk*)h)->increment_packing_fraction(inc); }
void or_disk_save_state(void* h, const int* f, int tv) { ((Disk*)h)->save_state(f, tv); }
void or_disk_restore_state(void* h, const int* f, int tv) { ((Disk*)h)->restore_state(f, tv); }

void or_disk_fire_init(void* h, const double* dt, int mode) { ((Disk*)h)->fire_init(dt, mode); }
void or_disk_fire_step(void* h, int mode) { ((Disk*)h)->fire_step(mode); }
int or_disk_fire_converged(void* h, double diff, double tgt, int pred, double ftol) {
    return ((Disk*)h)->fire_converged(diff, tgt, pred, ftol);
}
void or_disk_fire_get(void* h, double* dt, double* alpha, int* ngood, int* nbad) {
    Disk* d = (Disk*)h;
    for (int s = 0; s < d->S; ++s) { dt[s] = d->fire.dt[s]; alpha[s] = d->fire.alpha[s]; ngood[s] = d->fire.n_good[s]; nbad[s] = d->fire.n_bad[s]; }
}
// FIRE scalar state machine on its own (fire.cu:7-57), for the branch-table tests.
void or_fire_update_scalar(double* dt, double* alpha, int* ngood, int* nbad, double power, double dt_max,
                           double dt_min, double* dt_rev, double* vscale) {
    Disk d; d.S = 1;
    d.fire.dt = {*dt}; d.fire.alpha = {*alpha}; d.fire.n_good = {*ngood}; d.fire.n_bad = {*nbad};
    d.fire.dt_max = {dt_max}; d.fire.dt_min = {dt_min}; d.fire.dt_rev = {0}; d.fire.vscale = {1};
    double p = power;
    d.fire_update(&p);
    *dt = d.fire.dt[0]; *alpha = d.fire.alpha[0]; *ngood = d.fire.n_good[0]; *nbad = d.fire.n_bad[0];
    *dt_rev = d.fire.dt_rev[0]; *vscale = d.fire.vscale[0];
}
// minimize_fire (include/routines/minimizers.cuh:10-31): returns steps taken; *conv = 1 if
// every system stopped.
int or_disk_fire_minimize(void* h, const double* dt, int max_steps, double diff, double tgt, int pred,
                          double ftol, int mode, int* conv) {
    Disk* d = (Disk*)h;
    d->fire_init(dt, mode);
    int step = 0; int c = 0;
    while (step < max_steps && !c) {
        d->fire_step(mode);
        c = d->fire_converged(diff, tgt, pred, ftol);
        ++step;
    }
    *conv = c;
    return step;
}
// minimize_fire with per-system work counting: a system costs a step until the step after
// the one that stopped it (that step zeroes its velocities; later steps leave it unchanged).
// stopped[s] on return: 1 if system s stopped.
static int fire_minimize_counted(Disk* d, const double* dt, int max_steps, int pred, double a, double b, int mode,
                                 std::vector<int>& nsteps, std::vector<int>& stopped) {
    const int S = d->S;
    d->fire_init(dt, mode);
    std::vector<int> frozen(S), prev(S);
    for (int s = 0; s < S; ++s) { frozen[s] = dt[s] == 0.0; prev[s] = frozen[s]; }
    nsteps.assign(S, 0);
    int step = 0, c = 0;
    bool all = true;
    for (int s = 0; s < S; ++s) all = all && dt[s] == 0.0;
    if (all) c = 1;
    while (step < max_steps && !c) {
        for (int s = 0; s < S; ++s) if (!frozen[s]) nsteps[s]++;
        d->fire_step(mode);
        c = pred == 0 ? d->fire_converged(a, b, 0, 0.0) : d->fire_converged(0.0, b, pred, a);
        for (int s = 0; s < S; ++s) {
            if (prev[s]) frozen[s] = 1;
            prev[s] = d->fire.dt[s] == 0.0;
        }
        ++step;
    }
    stopped.resize(S);
    for (int s = 0; s < S; ++s) stopped[s] = d->fire.dt[s] == 0.0;
    return step;
}

// Pressure-target jamming by compression, the same rounds as the library's
// dpmd_jam_compress: minimise the systems still compressing (dt 0 for the rest), P_s;
// jammed if P_s >= p_target, else phi_s += dphi (increment_packing_fraction,
// base_particle.hpp:505-529), until max_comp compressions.  jam_state: 1 jammed, 0 never
// run or budget used up, -1 the initial minimisation did not stop (dropped, as
// remove_badly_initialized_systems does, jammers.cu:7-17; after a compression the state a
// minimisation reached decides whether it stopped or not, as jammers.cuh:100-110 does).
// retries: a round whose minimisation hit fire_max is run again from where it stopped (FIRE
// restarted, no compression; every system is rebuilt at the round boundary) up to `retries`
// times per system before that rule applies.
long long or_disk_jam_compress(void* h, const double* dt0, double dphi, double p_target, int max_comp, int fire_max,
                               int pred, double a, double b, int mode, int* jam_state, int* n_comp, int* nsteps_out,
                               int retries) {
    Disk* d = (Disk*)h;
    const int S = d->S;
    std::vector<int> state(S), ncomp(S, 0), nsteps(S, 0), ns, stopped, nretry(S, 0);
    std::vector<double> dts(S), inc(S), P(S);
    for (int s = 0; s < S; ++s) state[s] = dt0[s] == 0.0 ? 3 : 0;
    long long work = 0;
    for (;;) {
        bool any = false;
        for (int s = 0; s < S; ++s) { dts[s] = state[s] == 0 ? dt0[s] : 0.0; any = any || dts[s] != 0.0; }
        if (!any) break;
        fire_minimize_counted(d, dts.data(), fire_max, pred, a, b, mode, ns, stopped);
        d->pressure(P.data());
        bool grow = false, retry = false;
        for (int s = 0; s < S; ++s) {
            inc[s] = 0.0;
            if (dts[s] == 0.0) continue;
            nsteps[s] += ns[s];
            work += (long long)(d->off[s + 1] - d->off[s]) * ns[s];
            if (!stopped[s] && nretry[s] < retries) { nretry[s] += 1; retry = true; }
            else if (!stopped[s] && ncomp[s] == 0) state[s] = -1;
            else if (P[s] >= p_target) state[s] = 1;
            else if (ncomp[s] >= max_comp) state[s] = 2;
            else { inc[s] = dphi; ncomp[s] += 1; grow = true; }
        }
        if (!grow && !retry) break;
        d->increment_packing_fraction(inc.data());
    }
    for (int s = 0; s < S; ++s) {
        jam_state[s] = state[s] == 1 ? 1 : state[s] == -1 ? -1 : 0;
        n_comp[s] = ncomp[s];
        nsteps_out[s] = nsteps[s];
    }
    return work;
}

// jam_binary_search_pbc / _wall (include/routines/jammers.cuh:60-256, src/routines/jammers.cu:7-92)
// over the primitives above.  minimize_fire_copy_dt (minimizers.cuh:34-58) passes
// (avg_pe_target, avg_pe_diff_target) into converged(diff, target) (fire.cuh:98), so the
// diff test uses avg_pe_target and the <pe> test avg_pe_diff_target, as written there.
// Returns the number of compression rounds; action/phi as the library's
// dpmd_jam_binary_search.
}  // extern "C"
template <class P>
static int jam_bisect(P* d, const double* dt, int max_comp, int max_min, double pe_t, double pe_diff_t, double dphi,
                      double phi_tol, int mode, int* action_out, double* phi_out) {
    const int S = d->S;
    std::vector<double> dt_copy(dt, dt + S), dtw(S), zero(S, 0.0);
    std::vector<int> action(S, -1);
    auto minimize_copy_dt = [&](const double* dts) {
        d->fire_init(dts, mode);
        int step = 0, conv = 0;
        while (step < max_min && !conv) {
            d->fire_step(mode);
            conv = d->fire_converged(pe_t, pe_diff_t, 0, 0.0);
            ++step;
        }
        std::vector<double> out = d->fire.dt;
        d->scale_velocities(zero.data());
        return out;
    };
    if (mode != 1) d->scale_velocities(zero.data());   // jammers.cuh:85 zero_velocities
    std::vector<double> dtf = minimize_copy_dt(dt_copy.data());
    for (int s = 0; s < S; ++s) if (dtf[s] != 0.0) action[s] = 0;   // jammers.cu:7-17
    d->save_state(action.data(), -1);
    std::vector<double> phi(S);
    d->packing_fraction(phi.data());
    std::vector<double> phi_low = phi, phi_high(S, -1.0), pe(S), A(S), sf(S);
    int step = 0;
    while (step < max_comp) {
        for (int s = 0; s < S; ++s) dtw[s] = action[s] != 0 ? dt_copy[s] : 0.0;   // jammers.cu:19-28
        minimize_copy_dt(dtw.data());
        d->pe_total(pe.data());
        for (int s = 0; s < S; ++s) {   // jammers.cu:30-69
            if (action[s] == 0) continue;
            action[s] = -2;
            const int dof = d->off[s + 1] - d->off[s];
            const double avg_pe = pe[s] / dof;
            if (avg_pe > pe_t) {
                phi_high[s] = phi[s];
                phi[s] = (phi_high[s] + phi_low[s]) / 2.0;
                action[s] = 1;
            } else {
                action[s] = -1;
                phi_low[s] = phi[s];
                if (phi_high[s] > 0) phi[s] = (phi_high[s] + phi_low[s]) / 2.0;
                else phi[s] += dphi;
            }
            if (std::abs(phi_high[s] / phi_low[s] - 1) < phi_tol && phi_high[s] > 0) action[s] = 0;
        }
        d->save_state(action.data(), -1);
        d->restore_state(action.data(), 1);
        d->sys_reduce([&](int i) { return d->area[i]; }, A.data());
        for (int s = 0; s < S; ++s) {   // jammers.cu:71-90
            if (action[s] == 0 || phi[s] <= 0.0) { sf[s] = 1.0; continue; }
            const double L = std::sqrt(A[s] / phi[s]);
            sf[s] = L / d->Lx[s];
            d->Lx[s] = L; d->Ly[s] = L;
        }
        d->sync_box();
        d->regrid_if_needed(true);
        d->check_neighbors(true);
        d->scale_positions(sf.data());
        bool done = true;
        for (int s = 0; s < S; ++s) done = done && action[s] == 0;
        if (done) break;
        ++step;
    }
    d->compute_forces(mode);
    for (int s = 0; s < S; ++s) { action_out[s] = action[s]; phi_out[s] = phi[s]; }
    return step;
}
extern "C" {
int or_disk_jam_bisect(void* h, const double* dt, int max_comp, int max_min, double pe_t, double pe_diff_t,
                       double dphi, double phi_tol, int mode, int* action_out, double* phi_out) {
    return jam_bisect((Disk*)h, dt, max_comp, max_min, pe_t, pe_diff_t, dphi, phi_tol, mode, action_out, phi_out);
}
int or_rigid_jam_bisect(void* h, const double* dt, int max_comp, int max_min, double pe_t, double pe_diff_t,
                        double dphi, double phi_tol, int mode, int* action_out, double* phi_out) {
    return jam_bisect((Rigid*)h, dt, max_comp, max_min, pe_t, pe_diff_t, dphi, phi_tol, mode, action_out, phi_out);
}
// build_particle_neighbors (base_poly_particle.hpp:299-375): per particle, the sorted distinct
// particle ids of its vertices' list entries.  Returns the total; start [N+1], ids may be null.
int or_rigid_particle_neighbors(void* h, int* start, int* ids) {
    Rigid* r = (Rigid*)h;
    int tot = 0;
    for (int i = 0; i < r->N; ++i) {
        std::vector<int> L;
        for (int c = r->poff[i]; c < r->poff[i + 1]; ++c)
            for (int w : r->nbr[r->static_index[c]]) L.push_back(r->vpid[w]);
        std::sort(L.begin(), L.end());
        L.erase(std::unique(L.begin(), L.end()), L.end());
        if (start) start[i] = tot;
        if (ids) for (size_t q = 0; q < L.size(); ++q) ids[tot + q] = L[q];
        tot += (int)L.size();
    }
    if (start) start[r->N] = tot;
    return tot;
}
// count_particle_contacts / compute_pair_dist / compute_friction_coeff / pair forces
// (rigid_bumpy.cu:759-928,1261-1327) per particle-neighbor entry: vertices of i in canonical
// order, each vertex's list in order (the reference adds the pair forces atomically).
void or_rigid_pair_obs(void* h, int* pi, int* pj, double* dist, double* fx, double* fy, int* vci, int* vcj,
                       double* friction, int* contacts) {
    Rigid* r = (Rigid*)h;
    std::vector<int> st(r->N + 1);
    const int tot = or_rigid_particle_neighbors(h, st.data(), nullptr);
    std::vector<int> ids(std::max(1, tot));
    or_rigid_particle_neighbors(h, st.data(), ids.data());
    for (int i = 0; i < r->N; ++i) {
        const int s = r->sid[i];
        const double e = r->e[s];
        int ncontact = 0;
        for (int pk = st[i]; pk < st[i + 1]; ++pk) {
            const int j = ids[pk];
            int nvi = 0;
            double f_x = 0.0, f_y = 0.0;
            for (int c = r->poff[i]; c < r->poff[i + 1]; ++c) {
                const int u = r->static_index[c];
                bool found = false;
                for (int w : r->nbr[u]) {
                    if (r->vpid[w] != j) continue;
                    double dx, dy;
                    const double r2 = disp_pbc_L(r->vpx[u], r->vpy[u], r->vpx[w], r->vpy[w], r->Lx[s], r->Ly[s], r->ix[s],
                                                 r->iy[s], dx, dy);
                    const double radsum = r->vrad[u] + r->vrad[w];
                    if (r2 >= radsum * radsum) continue;
                    found = true;
                    const double rr = std::sqrt(r2);
                    const double inv = 1.0 / rr;
                    const double nx = dx * inv, ny = dy * inv;
                    const double delta = radsum - rr;
                    const double fmag = e * delta;
                    f_x += fmag * nx;
                    f_y += fmag * ny;
                }
                if (found) ++nvi;
            }
            if (nvi > 0) ++ncontact;
            double cx, cy;
            const double rp = std::sqrt(disp_pbc_L(r->x[i], r->y[i], r->x[j], r->y[j], r->Lx[s], r->Ly[s], r->ix[s],
                                                   r->iy[s], cx, cy));
            pi[pk] = i; pj[pk] = j; dist[pk] = rp; fx[pk] = f_x; fy[pk] = f_y; vci[pk] = nvi;
            const double nx = cx / rp, ny = cy / rp;
            const double fn = f_x * nx + f_y * ny;
            double mu = 0.0;
            if (std::fabs(fn) != 0.0) {
                const double tx = f_x - fn * nx, ty = f_y - fn * ny;
                mu = std::sqrt(tx * tx + ty * ty) / std::fabs(fn);
            }
            friction[pk] = mu;
        }
        contacts[i] = ncontact;
    }
    for (int i = 0; i < r->N; ++i)
        for (int pk = st[i]; pk < st[i + 1]; ++pk) {
            const int j = ids[pk];
            vcj[pk] = 0;
            for (int pl = st[j]; pl < st[j + 1]; ++pl) if (ids[pl] == i) vcj[pk] = vci[pl];
        }
}
// compute_hessian_kernel (rigid_bumpy.cu:1129-1259): per particle-neighbor entry (i, j), the
// contributions of i's vertex contacts with j to the (x, y, theta) blocks H_i,i and H_i,j.
// hs[(2 m + side) * tot + k], m = 3 a + b over (x, y, theta), side 0 = (i, i), 1 = (i, j).
// Walls contribute nothing, and vertex offsets from the centres use no minimum image, as in
// the reference.
void or_rigid_hessian(void* h, double* hs) {
    Rigid* r = (Rigid*)h;
    std::vector<int> st(r->N + 1);
    const int tot = or_rigid_particle_neighbors(h, st.data(), nullptr);
    std::vector<int> ids(std::max(1, tot));
    or_rigid_particle_neighbors(h, st.data(), ids.data());
    for (int i = 0; i < r->N; ++i) {
        const int s = r->sid[i];
        const double e = r->e[s];
        const double xi = r->x[i], yi = r->y[i];
        for (int k = st[i]; k < st[i + 1]; ++k) {
            const int j = ids[k];
            const double xj = r->x[j], yj = r->y[j];
            double hib[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, hjb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int c = r->poff[i]; c < r->poff[i + 1]; ++c) {
                const int mu = r->static_index[c];
                const double rxm = r->vpx[mu] - xi, rym = r->vpy[mu] - yi;
                for (int nu : r->nbr[mu]) {
                    if (r->vpid[nu] != j) continue;
                    double dx, dy;
                    const double r2 = disp_pbc_L(r->vpx[mu], r->vpy[mu], r->vpx[nu], r->vpy[nu], r->Lx[s], r->Ly[s],
                                                 r->ix[s], r->iy[s], dx, dy);
                    const double radsum = r->vrad[mu] + r->vrad[nu];
                    if (r2 >= radsum * radsum) continue;
                    const double rr = std::sqrt(r2);
                    const double delta = radsum - rr;
                    const double fmag = e * delta;
                    const double t = -fmag, cc = e;
                    const double rxn = r->vpx[nu] - xj, ryn = r->vpy[nu] - yj;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            const double ia_x = (a == 0) - rym * (a == 2), ia_y = (a == 1) + rxm * (a == 2);
                            const double jb_x = -1.0 * ((b == 0) - ryn * (b == 2));
                            const double jb_y = -1.0 * ((b == 1) + rxn * (b == 2));
                            const double qa = dx * ia_x + dy * ia_y, qjb = dx * jb_x + dy * jb_y;
                            const double dq_j = (ia_x * jb_x + ia_y * jb_y) / rr;
                            hjb[a * 3 + b] += cc * qa * qjb / r2 + t * (dq_j - qa * qjb / (r2 * rr));
                            const double ib_x = (b == 0) - rym * (b == 2), ib_y = (b == 1) + rxm * (b == 2);
                            const double iab_x = -rxm * (a == 2) * (b == 2), iab_y = -rym * (a == 2) * (b == 2);
                            const double qib = dx * ib_x + dy * ib_y;
                            const double dq_i = (ia_x * ib_x + dx * iab_x + ia_y * ib_y + dy * iab_y) / rr;
                            hib[a * 3 + b] += cc * qa * qib / r2 + t * (dq_i - qa * qib / (r2 * rr));
                        }
                }
            }
            for (int m = 0; m < 9; ++m) {
                hs[(2 * m) * (size_t)tot + k] = 0.0 + hib[m];
                hs[(2 * m + 1) * (size_t)tot + k] = 0.0 + hjb[m];
            }
        }
    }
}

// compute_overlaps_kernel (rigid_bumpy.cu:929-966), per vertex in current order
void or_rigid_overlaps(void* h, double* overlap, double* count) {
    Rigid* r = (Rigid*)h;
    for (int u = 0; u < r->Nv; ++u) {
        const int s = r->vsid[u];
        double ov = 0.0, cc = 0.0;
        for (int w : r->nbr[u]) {
            double dx, dy;
            const double r2 = disp_pbc_L(r->vpx[u], r->vpy[u], r->vpx[w], r->vpy[w], r->Lx[s], r->Ly[s], r->ix[s], r->iy[s],
                                         dx, dy);
            const double radsum = r->vrad[u] + r->vrad[w];
            if (r2 >= radsum * radsum) continue;
            ov += radsum - std::sqrt(r2);
            cc += 1.0;
        }
        overlap[u] = ov; count[u] = cc;
    }
}
void or_rigid_packing_fraction(void* h, double* o) { ((Rigid*)h)->packing_fraction(o); }
void or_rigid_increment_packing_fraction(void* h, const double* inc) { ((Rigid*)h)->increment_packing_fraction(inc); }
void or_rigid_save_state(void* h, const int* flg, int tv) { ((Rigid*)h)->save_state(flg, tv); }
void or_rigid_restore_state(void* h, const int* flg, int tv) { ((Rigid*)h)->restore_state(flg, tv); }

// Rescale-thermostat NVT with affine compression (the library's dpmd_nvt_rescale_run; the
// loop of tests/nvt_rescale_disk_pbc_compress.cu:41-68): per global step k = step0 + i one
// velocity-Verlet step, then increment_packing_fraction if k % cf == 0 (cf > 0), then
// set_temperature if k % tf == 0 (tf > 0).
void or_disk_nvt_rescale_run(void* h, int n, int step0, const double* inc, int cf, const double* T, int tf, int rm,
                             int mode) {
    Disk* d = (Disk*)h;
    for (int i = 0; i < n; ++i) {
        const int k = step0 + i;
        d->vv_step(mode);
        if (cf > 0 && k % cf == 0) d->increment_packing_fraction(inc);
        if (tf > 0 && k % tf == 0) d->set_temperature(T, rm);
    }
}

void or_disk_vv_init(void* h, const double* dt, int mode) { ((Disk*)h)->vv_init(dt, mode); }
void or_disk_damped_vv_init(void* h, const double* dt, const double* damp, int mode) {
    ((Disk*)h)->vv_init(dt, mode, damp);
}
void or_disk_damping(void* h, const double* scale) { ((Disk*)h)->damping(scale); }
void or_disk_vv_run(void* h, int n, int mode) { Disk* d = (Disk*)h; for (int k = 0; k < n; ++k) d->vv_step(mode); }

// Brute-force O(N^2) min-image neighbor set of particle i (for pinning the cell list).
int or_disk_brute_neighbors(void* h, int i, int* out) {
    Disk* d = (Disk*)h;
    const int s = d->sid[i];
    int w = 0;
    for (int j = d->off[s]; j < d->off[s + 1]; ++j) {
        if (j == i) continue;
        double dx, dy;
        const double r2 = disp_pbc_L(d->x[i], d->y[i], d->x[j], d->y[j], d->Lx[s], d->Ly[s], d->ix[s], d->iy[s], dx, dy);
        const double cut = d->rad[i] + d->rad[j] + d->cutoff[s];
        if (r2 < cut * cut) out[w++] = j;
    }
    return w;
}

}  // extern "C"
