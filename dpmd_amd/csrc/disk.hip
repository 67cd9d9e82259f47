// Disk kernels: harmonic contact forces (periodic and walled boxes), the velocity-Verlet
// and FIRE particle primitives, and the fused per-step kernels.
//
// Reference: src/particles/disk.cu (kernels :20-335, :702-769), src/integrators/fire.cu,
// include/integrators/{fire,velocity_verlet}.cuh.  The arithmetic of every primitive is the
// reference's expression for expression (operand order kept, no contraction: the library
// is built with -ffp-contract=off), so a fused kernel produces the same bits as the
// reference's sequence of primitives.
#include "internal.hpp"
#include "fire.hpp"
#include "disk_dev.hpp"

namespace dpmd {
namespace {

template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK) k_force(DiskArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double fx, fy, pe;
    disk_force_mode<MODE>(a, i, a.sid[i], fx, fy, pe);
    a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
}

// update_positions_kernel_cell / _naive (disk.cu:20-75)
__global__ void k_update_positions(DiskArgs a, const double* __restrict__ scale, double scale2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid) {
        const double sc = scale[s] * scale2;
        double x = a.x[i], y = a.y[i];
        x += a.vx[i] * sc;
        y += a.vy[i] * sc;
        a.x[i] = x; a.y[i] = y;
        if (a.cell) {
            const double dx = x - a.lx[i], dy = y - a.ly[i];
            flag = (dx * dx + dy * dy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// update_velocities_kernel (disk.cu:78-100)
__global__ void k_update_velocities(DiskArgs a, const double* __restrict__ scale, double scale2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double smi = scale[a.sid[i]] * scale2 / a.mass[i];
    a.vx[i] = a.vx[i] + a.fx[i] * smi;
    a.vy[i] = a.vy[i] + a.fy[i] * smi;
}

// compute_damping_forces_kernel (disk.cu:249-263): F_i += -v_i * scale[s]
__global__ void k_damping(DiskArgs a, const double* __restrict__ scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double sc = scale[a.sid[i]];
    a.fx[i] += -a.vx[i] * sc;
    a.fy[i] += -a.vy[i] * sc;
}

__global__ void k_scale_velocities(DiskArgs a, const double* __restrict__ scale) {   // disk.cu:278-289
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double sc = scale[a.sid[i]];
    a.vx[i] *= sc; a.vy[i] *= sc;
}

__global__ void k_scale_positions(DiskArgs a, const double* __restrict__ scale) {    // disk.cu:292-303
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const double sc = scale[a.sid[i]];
    a.x[i] *= sc; a.y[i] *= sc;
}

__global__ void k_mix(DiskArgs a, const double* __restrict__ alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double vx = a.vx[i], vy = a.vy[i];
    disk_mix_one(vx, vy, a.fx[i], a.fy[i], alpha[a.sid[i]]);
    a.vx[i] = vx; a.vy[i] = vy;
}

// compute_stress_tensor_kernel (disk.cu:702-769)
__global__ void k_stress(DiskArgs a, double* sxx, double* sxy, double* syx, double* syy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double ei = a.e[s], Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xi = a.x[i], yi = a.y[i], vxi = a.vx[i], vyi = a.vy[i], m = a.mass[i], ri = a.rad[i];
    double axx = m * vxi * vxi, axy = m * vxi * vyi;
    double ayx = axy, ayy = m * vyi * vyi;
    const double area_box = Lx * Ly;
    const int cnt = a.nn[i];
    for (int k = 0; k < cnt; ++k) {
        const int j = a.nbr[(size_t)k * a.n + i];
        const double dx = min_image(xi - a.x[j], Lx, iLx);
        const double dy = min_image(yi - a.y[j], Ly, iLy);
        const double r2 = dx * dx + dy * dy;
        const double radsum = ri + a.rad[j];
        if (r2 >= radsum * radsum) continue;
        const double r = sqrt(r2);
        const double inv = 1.0 / r;
        const double nx = dx * inv, ny = dy * inv;
        const double delta = radsum - r;
        const double fmag = ei * delta;
        const double fx = fmag * nx, fy = fmag * ny;
        axx += dx * fx / 2.0; axy += dx * fy / 2.0;
        ayx += dy * fx / 2.0; ayy += dy * fy / 2.0;
    }
    sxx[i] = axx / area_box; sxy[i] = axy / area_box;
    syx[i] = ayx / area_box; syy[i] = ayy / area_box;
}

// ---------------------------------------------------------------------------------------
// FIRE, fused.  One step of BaseFIRE::step_impl (fire.cuh:59-88) is:
//   k_fire_drift   fire_update result -> reverse drift, vscale, half kick, mixing, vscale,
//                  half drift, skin check (all per particle, in the reference's order);
//   rebuild        (device-side conditional, cells.hip);
//   k_fire_force   forces, half kick, per-chunk F.v / pe / max|F|^2 partials;
//   k_fire_finalize one workgroup per system folds the partials (fixed tree), applies the
//                  stopping test (fire.cu:59-78) and evaluates fire_update (fire.cu:7-57)
//                  for the next step into the s_* ("next-step") scalars.
// k_fire_drift commits s_* into the live scalars, so the live dt/alpha/N_good/N_bad are
// always what the reference holds after the same number of steps.
// ---------------------------------------------------------------------------------------

template <int B>
__global__ void __launch_bounds__(B) k_fire_drift(DiskArgs a, FireArgs f, Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid && !fire_frozen(f, s)) flag = fire_drift_one(a, f, i, s);
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// k_fire_drift over the chunks of the systems still running (one workgroup per chunk).
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_drift_act(DiskArgs a, FireArgs f, RedArgs r, const int* __restrict__ act, Gate g) {
    if (!gate_open(g)) return;
    const int c = act[blockIdx.x];
    const int s = r.chunk_sys[c];
    const int i = r.chunk_beg[c] + threadIdx.x;
    const bool valid = i < a.off[s + 1];
    bool flag = false;
    if (valid && !fire_frozen(f, s)) flag = fire_drift_one(a, f, i, s);
    if (a.cell) raise_flags(a.flag, a.any, valid ? s : -1, flag);
}

// The fold of one system's chunk partials in a fixed order: 256 strided sequential
// accumulators (thread t adds chunks t, t + 256, ... in turn), then the block tree.  Every
// thread of the 256-thread workgroup gets the power, energy and max |F|^2 totals (lds: 12
// doubles, not reused by the caller before a barrier).
__device__ __forceinline__ void fold_partials(const RedArgs& r, int s, double* lds, double& ap, double& ape,
                                              double& af) {
    const int cb = r.sys_chunk[s], ce = r.sys_chunk[s + 1];
    ap = 0.0; ape = 0.0; af = 0.0;
    // each thread's chunks in sequence, sixteen loads issued ahead of their additions
    constexpr int U = 16;
    for (int k0 = cb + threadIdx.x; k0 < ce; k0 += U * DPMD_BLOCK) {
        double q0[U], q1[U], q2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * DPMD_BLOCK;
            const int kk = k < ce ? k : k0;
            q0[u] = r.part[kk]; q1[u] = r.part[r.nch + kk]; q2[u] = r.part[2 * r.nch + kk];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k0 + u * DPMD_BLOCK >= ce) break;
            ap = ap + q0[u];
            ape = ape + q1[u];
            af = fmax(af, q2[u]);
        }
    }
    // block_sum256 / block_max256 of the three values with one barrier: wavefront butterflies,
    // then (w0 + w1) + (w2 + w3)
    ap = wave_sum(ap); ape = wave_sum(ape); af = wave_max(af);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { lds[w] = ap; lds[4 + w] = ape; lds[8 + w] = af; }
    __syncthreads();
    ap = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    ape = (lds[4] + lds[5]) + (lds[6] + lds[7]);
    af = fmax(fmax(lds[8], lds[9]), fmax(lds[10], lds[11]));
}

// The fold of one system's partials, the stopping test and the next step's FIRE update, by a
// whole 256-thread workgroup.  A separate launch orders it after every chunk.  (Folding in the
// last chunk workgroup of k_fire_force instead costs more at 2^20 disks: every workgroup then
// waits for an agent-scope atomic's round trip before it retires.)
__device__ __forceinline__ void fire_fold_system(const int* off, const FireArgs& f, const RedArgs& r, int s,
                                                 int do_conv, int do_update, int pred, double pa, double pb,
                                                 const Gate& g, double* lds) {
    // thread 0 fetches the system's FIRE scalars now, overlapping the fold below
    FireSys q{};
    int was_stopped = 0;
    if (threadIdx.x == 0) { q = fire_load(f, s); was_stopped = f.stopped[s]; }
    double ap, ape, af;
    fold_partials(r, s, lds, ap, ape, af);
    if (threadIdx.x == 0) {
        // A system that had stopped before this step has now had its velocities zeroed
        // (vscale = 0 after a stop), so every later step leaves it unchanged: freeze it.
        // The scalars were loaded together and are stored once.
        r.pw_tot[s] = ap; r.pe_tot[s] = ape; r.maxf2[s] = af;
        if (do_conv) {
            fire_convergence_core(q, off[s + 1] - off[s], ape, af, pred, pa, pb);
            f.dt[s] = q.dt;
            f.peprev[s] = q.peprev;
            if (q.dt == 0.0 && !was_stopped) {   // fire_count_stopped
                f.stopped[s] = 1;
                if (atomicSub(&f.run[RUN_RUNNING], 1) == 1) f.run[RUN_CONV_AT] = g.k;
            }
        }
        if (do_update) {
            fire_update_core(q, ap);
            f.s_dt[s] = q.s_dt; f.s_alpha[s] = q.s_alpha; f.s_ngood[s] = q.s_ngood; f.s_nbad[s] = q.s_nbad;
            f.s_dtrev[s] = q.s_dtrev; f.s_vscale[s] = q.s_vscale;
        }
        if (f.frozen && was_stopped) f.frozen[s] = 1;
        if (do_conv) f.nsteps[s] += 1;
    }
}

template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_force(DiskArgs a, FireArgs f, RedArgs r, int kick, const int* __restrict__ act, Gate g) {
    __shared__ double lds[12];
    if (!gate_force(g)) return;
    // act: the chunks of the systems still running, when the launch covers only those
    const int c = act ? act[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
    const int s = r.chunk_sys[c];
    if (fire_frozen(f, s)) return;   // whole workgroup: chunks never straddle systems
    const int beg = r.chunk_beg[c];
    const int end = min(beg + DPMD_BLOCK, a.off[s + 1]);
    const int i = beg + threadIdx.x;
    double vp = 0.0, vpe = 0.0, vf2 = 0.0;
    if (i < end) {
        double fx, fy, pe;
        disk_force_mode<MODE>(a, i, s, fx, fy, pe);
        a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
        double vx = a.vx[i], vy = a.vy[i];
        if (kick) {
            const double smi = f.dt[s] * 0.5 / a.mass[i];
            vx = vx + fx * smi; vy = vy + fy * smi;
            a.vx[i] = vx; a.vy[i] = vy;
        }
        vp = fx * vx + fy * vy;
        vpe = pe;
        vf2 = fx * fx + fy * fy;
    }
    // block_sum256 / block_max256 of the three values with one barrier: wavefront
    // butterflies, then (w0 + w1) + (w2 + w3)
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

// One workgroup per system folds the chunk partials (fire_fold_system).
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_finalize(const int* off, FireArgs f, RedArgs r, int do_conv, int do_update, int pred, double pa, double pb,
                const int* __restrict__ act, Gate g) {
    __shared__ double lds[12];
    if (!gate_open(g)) return;
    const int s = act ? act[blockIdx.x] : blockIdx.x;   // act: the systems still running
    if (fire_frozen(f, s)) return;
    fire_fold_system(off, f, r, s, do_conv, do_update, pred, pa, pb, g, lds);
}

// ---------------------------------------------------------------------------------------
// FIRE across step boundaries (runs of one system).  A step's drift needs the system's power
// F.v, a sum over every particle, so FIRE cannot close a step and open the next in one launch
// the way velocity Verlet does.  It can guess, though: F.v > 0 (the common case) fixes the
// next step's dt, alpha, counters, reverse drift (0) and velocity scale (1) from the current
// scalars alone.  Two launches per step:
//   k_fire_fused   step k's forces, kick and chunk partials, then step k + 1's drift with the
//                  guessed scalars into the other buffer pair (positions and velocities
//                  ping-pong between two interleaved (x, y) and (vx, vy) pairs, so step k's
//                  inputs survive); it also records the scalars it started from (snapshot);
//   k_fire_settle  every workgroup folds the partials in the fixed tree and evaluates the
//                  stopping test and the update from the snapshot (the same bits in each);
//                  workgroup 0 stores the scalars and commits the next step's; when the guess
//                  was wrong, all workgroups redo step k's forces and kick from the kept
//                  inputs and drift with the true scalars, and when the system stopped they
//                  write step k's final state to the context's arrays.
// The bookkeeping that other workgroups of k_fire_settle would read while it is written (the
// running count, the step a stop happened at) is done by the next launch's gate, driven by
// the run word RUN_MIS (k + 1 after a wrong guess at step k, -(k + 1) after a stop, else 0).
// Every value goes through the expressions of k_fire_drift, k_fire_force and k_fire_finalize
// in their order, so a fused run equals the three-launch run bit for bit.  Rebuild tests raise
// only the run words: W(k + 1) for the guessed drift, RUN_FIXW for a redone one; a halted batch
// flags the system on the host before the rebuild.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void raise_word(int* w, bool f) {
    const unsigned long long m = __ballot(f);
    if (m != 0 && (int)(threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) atomicOr(w, 1);
}

// Gate of a fused FIRE launch of step k (k_fire_fused, k_fire_force_pk).
__device__ __forceinline__ bool gate_fire_fused(const Gate& g) {
    if (g.run[RUN_HALT] != 0 || g.run[RUN_RUNNING] <= 0) return false;
    const int v = g.k > 0 ? g.run[RUN_MIS] : 0;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (g.k > 0 && v == -g.k) {   // the system stopped at step k - 1, whose state is final
        if (lead) { g.run[RUN_CONV_AT] = g.k - 1; atomicSub(&g.run[RUN_RUNNING], 1); }
        return false;
    }
    const bool miss = g.k > 0 && v == g.k;   // step k's positions come from a redone drift
    if (g.run[miss ? RUN_FIXW : g.any] != 0) {
        if (lead) { g.run[RUN_HALT_AT] = g.k; g.run[RUN_HALT] = 1; }
        return false;
    }
    if (miss && lead) g.run[g.any] = 0;   // withdraw what the wrong guess raised
    return true;
}

// The first drift of a fused batch (k_fire_drift's body) from the context's arrays into the
// buffer pair (xo, vo).
template <int B>
__global__ void __launch_bounds__(B)
k_fire_drift_to(DiskArgs a, FireArgs f, double2* __restrict__ xo, double2* __restrict__ vo, FusedSys fs, Gate g) {
    if (!gate_open(g)) return;
    const int i = fs.b0 + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < fs.b0 + fs.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid && !fire_frozen(f, s)) {
        const double dt = f.s_dt[s], alpha = f.s_alpha[s], dtrev = f.s_dtrev[s], vscale = f.s_vscale[s];
        double x = a.x[i], y = a.y[i], vx = a.vx[i], vy = a.vy[i];
        const double lx = a.cell ? a.lx[i] : 0.0, ly = a.cell ? a.ly[i] : 0.0, th2 = a.cell ? a.thresh2[s] : 0.0;
        flag = fire_drift_vals(x, y, vx, vy, a.fx[i], a.fy[i], a.mass[i], a.cell, lx, ly, th2, dt, alpha, dtrev,
                               vscale);
        xo[i] = make_double2(x, y); vo[i] = make_double2(vx, vy);
        if (i == a.off[s]) {
            f.dt[s] = dt; f.alpha[s] = alpha; f.ngood[s] = f.s_ngood[s]; f.nbad[s] = f.s_nbad[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// Forces and kick of one particle at the positions xs and velocities vs (k_fire_force's body):
// returns the kicked velocity, the force and the energy.
template <int MODE>
__device__ __forceinline__ void fire_force_kick(const DiskArgs& a, const double2* __restrict__ xs,
                                                const double2* __restrict__ vs, int i, int s, double dt, double& fx,
                                                double& fy, double& pe, double& vx, double& vy) {
    disk_force_mode_pos<MODE>(a, PackedPos{xs}, i, s, fx, fy, pe);
    const double2 v = vs[i];
    vx = v.x; vy = v.y;
    const double smi = dt * 0.5 / a.mass[i];
    vx = vx + fx * smi; vy = vy + fy * smi;
}

// Partials of chunk c (the tree of k_fire_force).
__device__ __forceinline__ void fire_chunk_partials(const RedArgs& r, int c, double vp, double vpe, double vf2,
                                                    double* lds) {
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

// One system (fs): its chunk c covers its particles b0 + [256 c, 256 c + 256).
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_fused(DiskArgs a, FireArgs f, RedArgs r, const double2* __restrict__ xs, const double2* __restrict__ vs,
             double2* __restrict__ xo, double2* __restrict__ vo, double* __restrict__ snap, int* __restrict__ isnap,
             FusedSys fs, Gate g) {
    __shared__ double lds[12];
    if (!gate_fire_fused(g)) return;
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    const int i = fs.b0 + c * DPMD_BLOCK + threadIdx.x;
    const int s = fs.s;
    // this step's dt (kick) and the next step's scalars as F.v > 0 makes them
    FireSys q{};
    q.dt = f.dt[s]; q.alpha = f.alpha[s]; q.ngood = f.ngood[s]; q.nbad = f.nbad[s];
    q.dtmax = f.dtmax[s]; q.dtmin = f.dtmin[s];
    if (c == 0 && threadIdx.x == 0) {
        snap[0] = q.dt; snap[1] = q.alpha; snap[2] = f.peprev[s];
        isnap[0] = q.ngood; isnap[1] = q.nbad;
    }
    const double dt = q.dt;
    fire_update_core(q, 1.0);
    double vp = 0.0, vpe = 0.0, vf2 = 0.0;
    bool flag = false;
    if (i < fs.b0 + fs.n) {
        double fx, fy, pe, vx, vy;
        fire_force_kick<MODE>(a, xs, vs, i, s, dt, fx, fy, pe, vx, vy);
        vp = fx * vx + fy * vy;
        vpe = pe;
        vf2 = fx * fx + fy * fy;
        const double2 p = xs[i];
        double x = p.x, y = p.y;
        const double lx = a.cell ? a.lx[i] : 0.0, ly = a.cell ? a.ly[i] : 0.0, th2 = a.cell ? a.thresh2[s] : 0.0;
        flag = fire_drift_vals(x, y, vx, vy, fx, fy, a.mass[i], a.cell, lx, ly, th2, q.s_dt, q.s_alpha, q.s_dtrev,
                               q.s_vscale);
        xo[i] = make_double2(x, y); vo[i] = make_double2(vx, vy);
    }
    if (a.cell) raise_word(a.any, flag);
    fire_chunk_partials(r, fs.ch0 + c, vp, vpe, vf2, lds);
}

// After k_fire_fused of step k (gate: nothing halted, the system running, as k_fire_fused saw
// it; neither word changes during this launch).  anyw: the run word the guessed drift raised.
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_settle(DiskArgs a, FireArgs f, RedArgs r, const double* __restrict__ snap, const int* __restrict__ isnap,
              int pred, double pa, double pb, const double2* __restrict__ xs, const double2* __restrict__ vs,
              double2* __restrict__ xo, double2* __restrict__ vo, FusedSys fs, Gate g, int anyw) {
    __shared__ double lds[12];
    __shared__ double sh[5];
    __shared__ int shi[2];
    if (g.run[RUN_HALT] != 0 || g.run[RUN_RUNNING] <= 0) return;
    const int s = fs.s;
    double ap, ape, af;
    fold_partials(r, s, lds, ap, ape, af);
    if (threadIdx.x == 0) {
        FireSys q{};
        q.dt = snap[0]; q.alpha = snap[1]; q.peprev = snap[2]; q.ngood = isnap[0]; q.nbad = isnap[1];
        q.dtmax = f.dtmax[s]; q.dtmin = f.dtmin[s];
        const double kick_dt = q.dt;
        FireSys guess = q;
        fire_update_core(guess, 1.0);
        fire_convergence_core(q, fs.n, ape, af, pred, pa, pb);
        const bool stop = q.dt == 0.0;   // fire_count_stopped (the system was running)
        fire_update_core(q, ap);
        const bool miss = stop || q.s_dt != guess.s_dt || q.s_alpha != guess.s_alpha ||
                          q.s_dtrev != guess.s_dtrev || q.s_vscale != guess.s_vscale ||
                          q.s_ngood != guess.s_ngood || q.s_nbad != guess.s_nbad;
        if (blockIdx.x == 0) {
            r.pw_tot[s] = ap; r.pe_tot[s] = ape; r.maxf2[s] = af;
            f.peprev[s] = q.peprev;
            f.s_dt[s] = q.s_dt; f.s_alpha[s] = q.s_alpha; f.s_ngood[s] = q.s_ngood; f.s_nbad[s] = q.s_nbad;
            f.s_dtrev[s] = q.s_dtrev; f.s_vscale[s] = q.s_vscale;
            if (stop) {             // k_fire_finalize's state after a stop; no drift follows
                f.dt[s] = q.dt; f.stopped[s] = 1;
                f.run[anyw] = 0;
            } else {                // what k_fire_drift commits at the start of the next step
                f.dt[s] = q.s_dt; f.alpha[s] = q.s_alpha; f.ngood[s] = q.s_ngood; f.nbad[s] = q.s_nbad;
            }
            f.nsteps[s] += 1;
            f.run[RUN_MIS] = stop ? -(g.k + 1) : (miss ? g.k + 1 : 0);
        }
        shi[0] = miss; shi[1] = stop;
        sh[0] = kick_dt; sh[1] = q.s_dt; sh[2] = q.s_alpha; sh[3] = q.s_dtrev; sh[4] = q.s_vscale;
    }
    __syncthreads();
    if (!shi[0]) return;
    const bool stop = shi[1] != 0;
    const double kick_dt = sh[0];
    for (int c = blockIdx.x; c < fs.nch; c += gridDim.x) {
        const int i = fs.b0 + c * DPMD_BLOCK + threadIdx.x;
        bool flag = false;
        if (i < fs.b0 + fs.n) {
            double fx, fy, pe, vx, vy;
            fire_force_kick<MODE>(a, xs, vs, i, s, kick_dt, fx, fy, pe, vx, vy);
            const double2 p = xs[i];
            if (stop) {   // the step's final state, as k_fire_force leaves it
                a.x[i] = p.x; a.y[i] = p.y; a.vx[i] = vx; a.vy[i] = vy;
                a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
            } else {
                double x = p.x, y = p.y;
                const double lx = a.cell ? a.lx[i] : 0.0, ly = a.cell ? a.ly[i] : 0.0;
                const double th2 = a.cell ? a.thresh2[s] : 0.0;
                flag = fire_drift_vals(x, y, vx, vy, fx, fy, a.mass[i], a.cell, lx, ly, th2, sh[1], sh[2], sh[3], sh[4]);
                xo[i] = make_double2(x, y); vo[i] = make_double2(vx, vy);
            }
        }
        if (a.cell && !stop) raise_word(&g.run[RUN_FIXW], flag);
    }
}

// The last step of a fused batch: k_fire_force at (xs, vs), also writing positions and
// velocities back to the context's arrays; k_fire_finalize follows.
template <int MODE>
__global__ void __launch_bounds__(DPMD_BLOCK)
k_fire_force_pk(DiskArgs a, FireArgs f, RedArgs r, const double2* __restrict__ xs, const double2* __restrict__ vs,
                FusedSys fs, Gate g) {
    __shared__ double lds[12];
    if (!gate_fire_fused(g)) return;
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    const int i = fs.b0 + c * DPMD_BLOCK + threadIdx.x;
    double vp = 0.0, vpe = 0.0, vf2 = 0.0;
    if (i < fs.b0 + fs.n) {
        double fx, fy, pe, vx, vy;
        fire_force_kick<MODE>(a, xs, vs, i, fs.s, f.dt[fs.s], fx, fy, pe, vx, vy);
        const double2 p = xs[i];
        a.x[i] = p.x; a.y[i] = p.y; a.vx[i] = vx; a.vy[i] = vy;
        a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
        vp = fx * vx + fy * vy;
        vpe = pe;
        vf2 = fx * fx + fy * fy;
    }
    fire_chunk_partials(r, fs.ch0 + c, vp, vpe, vf2, lds);
}

// Particles b0 .. b0 + n - 1 from the interleaved pairs to the context's arrays.
__global__ void k_unpack_pv(int b0, int n, const double2* __restrict__ xs, const double2* __restrict__ vs,
                            double* __restrict__ x, double* __restrict__ y, double* __restrict__ vx,
                            double* __restrict__ vy) {
    const int i = b0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b0 + n) return;
    const double2 p = xs[i], v = vs[i];
    x[i] = p.x; y[i] = p.y; vx[i] = v.x; vy[i] = v.y;
}

// Stopping test + next-step update for the step()/converged() call pattern (fire.cu:59-78).
__global__ void k_fire_converge(int S, const int* off, FireArgs f, const double* pe_tot, const double* maxf2,
                                const double* pw_tot, int pred, double pa, double pb) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    fire_convergence_sys(f, off, s, pe_tot[s], maxf2[s], pred, pa, pb);
    fire_count_stopped(f, s, 0);
    fire_update_sys(f, s, pw_tot[s]);
}

__global__ void k_fire_update(int S, FireArgs f, const double* pw_tot) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    fire_update_sys(f, s, pw_tot[s]);
}

// ---------------------------------------------------------------------------------------
// Velocity Verlet, fused (BaseVelocityVerlet::step_impl, velocity_verlet.cuh:34-41):
//   k_vv_drift: update_velocities(dt_half) + update_positions(dt) + skin check;
//   rebuild;  k_vv_force: compute_forces + update_velocities(dt_half).
// ---------------------------------------------------------------------------------------
template <int B>
__global__ void __launch_bounds__(B)
k_vv_drift(DiskArgs a, const double* __restrict__ dt, const double* __restrict__ dth, Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid) {
        const double smi = dth[s] * 1.0 / a.mass[i];
        double vx = a.vx[i] + a.fx[i] * smi;
        double vy = a.vy[i] + a.fy[i] * smi;
        const double sc = dt[s] * 1.0;
        double x = a.x[i], y = a.y[i];
        x += vx * sc; y += vy * sc;
        a.x[i] = x; a.y[i] = y; a.vx[i] = vx; a.vy[i] = vy;
        if (a.cell) {
            const double dx = x - a.lx[i], dy = y - a.ly[i];
            flag = (dx * dx + dy * dy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

template <int MODE, int B>
__global__ void __launch_bounds__(B)
k_vv_force(DiskArgs a, const double* __restrict__ dth, const double* __restrict__ damp, Gate g) {
    if (!gate_force(g)) return;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * B + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    double fx, fy, pe;
    disk_force_mode<MODE>(a, i, s, fx, fy, pe);
    if (damp) {   // compute_damping_forces_kernel (disk.cu:249-263) at the half-step velocity
        const double d = damp[s];
        fx += -a.vx[i] * d;
        fy += -a.vy[i] * d;
    }
    a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
    const double smi = dth[s] * 1.0 / a.mass[i];
    a.vx[i] = a.vx[i] + fx * smi;
    a.vy[i] = a.vy[i] + fy * smi;
}

// The step's first half (update_velocities(dt_half) + update_positions(dt)) writing the new
// positions to (xo, yo): the start of a fused run.
template <int B>
__global__ void __launch_bounds__(B)
k_vv_drift_to(DiskArgs a, double2* __restrict__ xo, const double* __restrict__ dt, const double* __restrict__ dth,
              Gate g) {
    if (!gate_open(g)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? a.sid[i] : -1;
    bool flag = false;
    if (valid) {
        const double smi = dth[s] * 1.0 / a.mass[i];
        const double vx = a.vx[i] + a.fx[i] * smi;
        const double vy = a.vy[i] + a.fy[i] * smi;
        const double sc = dt[s] * 1.0;
        double x = a.x[i], y = a.y[i];
        x += vx * sc; y += vy * sc;
        xo[i] = make_double2(x, y); a.vx[i] = vx; a.vy[i] = vy;
        if (a.cell) {
            const double dx = x - a.lx[i], dy = y - a.ly[i];
            flag = (dx * dx + dy * dy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// Velocity Verlet across a step boundary in one launch: forces at the positions xs and the
// step's closing half kick, then the next step's opening half kick, its drift into xo and
// the skin test of the new positions.  Every force must see its neighbors' old positions,
// hence two buffers; both hold interleaved (x, y) pairs, so each neighbor costs one 16-byte
// gather.  Each expression is the one of k_vv_force / k_vv_drift, in the same order, so a
// fused run equals the two-launch run bit for bit.  ONE: the context holds one system, so no
// system id is loaded per particle.
template <int MODE, int B, bool ONE>
__global__ void __launch_bounds__(B)
k_vv_fused(DiskArgs a, const double2* __restrict__ xs, double2* __restrict__ xo, const double* __restrict__ dt,
           const double* __restrict__ dth, const double* __restrict__ damp, Gate g) {
    if (!gate_force(g)) return;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * B + threadIdx.x;
    const bool valid = i < a.n;
    const int s = valid ? (ONE ? 0 : a.sid[i]) : -1;
    bool flag = false;
    if (valid) {
        // The forces are used here only: every fused batch ends with k_vv_force, and a batch
        // that a rebuild halts has its halted step's forces computed again, so fx, fy and pe
        // are not stored.
        double fx, fy, pe;
        disk_force_mode_pos<MODE>(a, PackedPos{xs}, i, s, fx, fy, pe);
        (void)pe;
        double vx = a.vx[i], vy = a.vy[i];
        if (damp) {   // damped velocity Verlet: the drag at the half-step velocity joins the force
            const double d = damp[s];
            fx += -vx * d;
            fy += -vy * d;
        }
        const double smi = dth[s] * 1.0 / a.mass[i];
        vx = vx + fx * smi;                      // end of this step
        vy = vy + fy * smi;
        vx = vx + fx * smi;                      // next step: update_velocities(dt_half)
        vy = vy + fy * smi;
        const double sc = dt[s] * 1.0;           // update_positions(dt)
        const double2 p = xs[i];
        double x = p.x, y = p.y;
        x += vx * sc; y += vy * sc;
        xo[i] = make_double2(x, y); a.vx[i] = vx; a.vy[i] = vy;
        if (a.cell) {
            const double dx = x - a.lx[i], dy = y - a.ly[i];
            flag = (dx * dx + dy * dy) > a.thresh2[s];
        }
    }
    if (a.cell) raise_flags(a.flag, a.any, s, flag);
}

// The last step of a fused batch: k_vv_force at the interleaved positions xs, which it also
// writes back to the context's (x, y).
template <int MODE, int B>
__global__ void __launch_bounds__(B)
k_vv_force_pk(DiskArgs a, const double2* __restrict__ xs, const double* __restrict__ dth,
              const double* __restrict__ damp, Gate g) {
    if (!gate_force(g)) return;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * B + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    double fx, fy, pe;
    disk_force_mode_pos<MODE>(a, PackedPos{xs}, i, s, fx, fy, pe);
    if (damp) {
        const double d = damp[s];
        fx += -a.vx[i] * d;
        fy += -a.vy[i] * d;
    }
    const double2 p = xs[i];
    a.x[i] = p.x; a.y[i] = p.y;
    a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
    const double smi = dth[s] * 1.0 / a.mass[i];
    a.vx[i] = a.vx[i] + fx * smi;
    a.vy[i] = a.vy[i] + fy * smi;
}

__global__ void k_unpack(int n, const double2* __restrict__ xs, double* __restrict__ x, double* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 p = xs[i];
    x[i] = p.x; y[i] = p.y;
}

// Workgroup size of the per-particle step kernels (the tunable block shape).
template <class F>
void by_block(int block, F&& f) {
    switch (block) {
        case 64: f(std::integral_constant<int, 64>{}); break;
        case 128: f(std::integral_constant<int, 128>{}); break;
        case 512: f(std::integral_constant<int, 512>{}); break;
        default: f(std::integral_constant<int, 256>{}); break;
    }
}

template <class F>
void by_mode(int mode, F&& f) {
    switch (mode) {
        case DPMD_FORCE_PBC: f(std::integral_constant<int, DPMD_FORCE_PBC>{}); break;
        case DPMD_FORCE_WALL: f(std::integral_constant<int, DPMD_FORCE_WALL>{}); break;
        default: f(std::integral_constant<int, DPMD_FORCE_PER_SYSTEM>{}); break;
    }
}

}  // namespace

void disk_compute_forces(dpmd_ctx* c, int mode) {
    DiskArgs a = disk_args(c);
    by_mode(mode, [&](auto m) { k_force<decltype(m)::value><<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(a); });
    HIP_OK(hipGetLastError());
}
void disk_update_positions(dpmd_ctx* c, const double* scale, double scale2) {
    k_update_positions<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), scale, scale2);
    HIP_OK(hipGetLastError());
}
void disk_update_velocities(dpmd_ctx* c, const double* scale, double scale2) {
    k_update_velocities<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), scale, scale2);
    HIP_OK(hipGetLastError());
}
void disk_damping(dpmd_ctx* c, const double* scale) {
    k_damping<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), scale);
    HIP_OK(hipGetLastError());
}
void disk_scale_velocities(dpmd_ctx* c, const double* scale) {
    k_scale_velocities<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), scale);
    HIP_OK(hipGetLastError());
}
void disk_scale_positions(dpmd_ctx* c, const double* scale) {
    k_scale_positions<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), scale);
    HIP_OK(hipGetLastError());
}
void disk_mix(dpmd_ctx* c, const double* alpha) {
    k_mix<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), alpha);
    HIP_OK(hipGetLastError());
}
void disk_stress(dpmd_ctx* c) {
    k_stress<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), c->sxx.p, c->sxy.p, c->syx.p, c->syy.p);
    HIP_OK(hipGetLastError());
}

void disk_fire_drift(dpmd_ctx* c, Gate g) {
    if (c->n_act_chunk > 0) {
        k_fire_drift_act<<<c->n_act_chunk, DPMD_BLOCK, 0, c->st>>>(disk_args(c), fire_args(c), red_args(c),
                                                                     c->act_chunk.p, g);
        HIP_OK(hipGetLastError());
        return;
    }
    by_block(c->block, [&](auto b) {
        constexpr int B = decltype(b)::value;
        k_fire_drift<B><<<blocks_for(c->N, B), B, 0, c->st>>>(disk_args(c), fire_args(c), g);
    });
    HIP_OK(hipGetLastError());
}

void disk_fire_force(dpmd_ctx* c, Gate g, int kick, int do_conv, int pred, double pa, double pb) {
    DiskArgs a = disk_args(c);
    FireArgs f = fire_args(c);
    RedArgs r = red_args(c);
    {
        ProfScope prof(c);
        by_mode(c->fire_mode, [&](auto m) {
            const bool act = c->n_act_chunk > 0;
            k_fire_force<decltype(m)::value><<<act ? c->n_act_chunk : c->nchunks, DPMD_BLOCK, 0, c->st>>>(
                a, f, r, kick, act ? c->act_chunk.p : nullptr, g);
        });
    }
    HIP_OK(hipGetLastError());
    fire_finalize(c, g, do_conv, pred, pa, pb);
}

void fire_finalize(dpmd_ctx* c, Gate g, int do_conv, int pred, double pa, double pb) {
    const bool act = c->n_act_sys > 0;
    k_fire_finalize<<<act ? c->n_act_sys : c->S, DPMD_BLOCK, 0, c->st>>>(c->off.p, fire_args(c), red_args(c), do_conv,
                                                                         1, pred, pa, pb,
                                                                         act ? c->act_sys.p : nullptr, g);
    HIP_OK(hipGetLastError());
}

FusedSys fused_sys(dpmd_ctx* c) {
    if (c->S == 1) return FusedSys{0, 0, c->N, 0, c->nchunks};
    need_state(c->fsys >= 0 && c->fsys < c->S, "fused FIRE: no single running system");
    const int s = c->fsys, b0 = c->h_off[s], n = c->h_off[s + 1] - b0;
    return FusedSys{s, b0, n, c->fsys_ch0, (n + DPMD_BLOCK - 1) / DPMD_BLOCK};
}

void disk_fire_drift_to(dpmd_ctx* c, Gate g, double* xo, double* vo, int any_word) {
    DiskArgs a = disk_args(c);
    a.any = c->run.p + any_word;
    by_block(c->block, [&](auto b) {
        constexpr int B = decltype(b)::value;
        const FusedSys fs = fused_sys(c);
        k_fire_drift_to<B><<<blocks_for(fs.n, B), B, 0, c->st>>>(a, fire_args(c), (double2*)xo, (double2*)vo, fs, g);
    });
    HIP_OK(hipGetLastError());
}

void disk_fire_fused(dpmd_ctx* c, Gate g, const double* xs, const double* vs, double* xo, double* vo, int any_word) {
    DiskArgs a = disk_args(c);
    a.any = c->run.p + any_word;
    by_mode(c->fire_mode, [&](auto m) {
        const FusedSys fs = fused_sys(c);
        k_fire_fused<decltype(m)::value><<<fs.nch, DPMD_BLOCK, 0, c->st>>>(
            a, fire_args(c), red_args(c), (const double2*)xs, (const double2*)vs, (double2*)xo, (double2*)vo,
            c->f_snap.p, c->i_snap.p, fs, g);
    });
    HIP_OK(hipGetLastError());
}

void disk_fire_settle(dpmd_ctx* c, Gate g, int pred, double pa, double pb, const double* xs, const double* vs,
                      double* xo, double* vo, int any_word) {
    const FusedSys fs = fused_sys(c);
    const int grid = std::min(fs.nch, kSettleGrid);
    by_mode(c->fire_mode, [&](auto m) {
        k_fire_settle<decltype(m)::value><<<grid, DPMD_BLOCK, 0, c->st>>>(
            disk_args(c), fire_args(c), red_args(c), c->f_snap.p, c->i_snap.p, pred, pa, pb, (const double2*)xs,
            (const double2*)vs, (double2*)xo, (double2*)vo, fs, g, any_word);
    });
    HIP_OK(hipGetLastError());
}

void disk_fire_force_packed(dpmd_ctx* c, Gate g, const double* xs, const double* vs, int pred, double pa, double pb) {
    by_mode(c->fire_mode, [&](auto m) {
        const FusedSys fs = fused_sys(c);
        k_fire_force_pk<decltype(m)::value><<<fs.nch, DPMD_BLOCK, 0, c->st>>>(
            disk_args(c), fire_args(c), red_args(c), (const double2*)xs, (const double2*)vs, fs, g);
    });
    HIP_OK(hipGetLastError());
    fire_finalize(c, g, 1, pred, pa, pb);
}

void disk_unpack_pv(dpmd_ctx* c, const double* xs, const double* vs) {
    const FusedSys fs = fused_sys(c);
    k_unpack_pv<<<blocks_for(fs.n), DPMD_BLOCK, 0, c->st>>>(fs.b0, fs.n, (const double2*)xs, (const double2*)vs,
                                                            c->x.p, c->y.p, c->vx.p, c->vy.p);
    HIP_OK(hipGetLastError());
}

void disk_vv_drift(dpmd_ctx* c, Gate g) {
    by_block(c->block, [&](auto b) {
        constexpr int B = decltype(b)::value;
        k_vv_drift<B><<<blocks_for(c->N, B), B, 0, c->st>>>(disk_args(c), c->vv_dt.p, c->vv_dth.p, g);
    });
    HIP_OK(hipGetLastError());
}

void disk_vv_drift_to(dpmd_ctx* c, Gate g, double* xo, int any_word) {
    DiskArgs a = disk_args(c);
    a.any = c->run.p + any_word;
    by_block(c->block, [&](auto b) {
        constexpr int B = decltype(b)::value;
        k_vv_drift_to<B><<<blocks_for(c->N, B), B, 0, c->st>>>(a, (double2*)xo, c->vv_dt.p, c->vv_dth.p, g);
    });
    HIP_OK(hipGetLastError());
}

void disk_vv_fused(dpmd_ctx* c, Gate g, const double* xs, double* xo, int any_word) {
    DiskArgs a = disk_args(c);
    a.any = c->run.p + any_word;
    by_mode(c->vv_mode, [&](auto m) {
        by_block(c->block, [&](auto b) {
            constexpr int B = decltype(b)::value;
            if (c->S == 1)
                k_vv_fused<decltype(m)::value, B, true><<<blocks_for(c->N, B), B, 0, c->st>>>(
                    a, (const double2*)xs, (double2*)xo, c->vv_dt.p, c->vv_dth.p, vv_damping(c), g);
            else
                k_vv_fused<decltype(m)::value, B, false><<<blocks_for(c->N, B), B, 0, c->st>>>(
                    a, (const double2*)xs, (double2*)xo, c->vv_dt.p, c->vv_dth.p, vv_damping(c), g);
        });
    });
    HIP_OK(hipGetLastError());
}

void disk_vv_force_packed(dpmd_ctx* c, Gate g, const double* xs) {
    DiskArgs a = disk_args(c);
    ProfScope prof(c);
    by_mode(c->vv_mode, [&](auto m) {
        by_block(c->block, [&](auto b) {
            constexpr int B = decltype(b)::value;
            k_vv_force_pk<decltype(m)::value, B><<<blocks_for(c->N, B), B, 0, c->st>>>(a, (const double2*)xs,
                                                                                        c->vv_dth.p, vv_damping(c), g);
        });
    });
    HIP_OK(hipGetLastError());
}

void disk_unpack_positions(dpmd_ctx* c, const double* xs) {
    k_unpack<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, (const double2*)xs, c->x.p, c->y.p);
    HIP_OK(hipGetLastError());
}

void disk_vv_force(dpmd_ctx* c, Gate g) {
    DiskArgs a = disk_args(c);
    ProfScope prof(c);
    by_mode(c->vv_mode, [&](auto m) {
        by_block(c->block, [&](auto b) {
            constexpr int B = decltype(b)::value;
            k_vv_force<decltype(m)::value, B><<<blocks_for(c->N, B), B, 0, c->st>>>(a, c->vv_dth.p, vv_damping(c),
                                                                                     g);
        });
    });
    HIP_OK(hipGetLastError());
}

__global__ void k_ke(DiskArgs a, double* ke) {   // compute_ke_kernel (disk.cu:266-275)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    ke[i] = 0.5 * (a.vx[i] * a.vx[i] + a.vy[i] * a.vy[i]) * a.mass[i];
}

void disk_ke(dpmd_ctx* c) {
    k_ke<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(disk_args(c), c->ke.p);
    HIP_OK(hipGetLastError());
}

void fire_apply_convergence(dpmd_ctx* c, int pred, double a, double b) {
    k_fire_converge<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->off.p, fire_args(c), c->pe_tot.p,
                                                               c->maxf2.p, c->pw_tot.p, pred, a, b);
    HIP_OK(hipGetLastError());
}

void fire_recompute_update(dpmd_ctx* c) {
    k_fire_update<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, fire_args(c), c->pw_tot.p);
    HIP_OK(hipGetLastError());
}

}  // namespace dpmd
