// Persistent per-system FIRE for disk batches: one workgroup runs all FIRE steps of one
// system, with workgroup barriers between the phases of a step and no grid-wide
// synchronisation, so systems of a batch advance independently and a stopped system costs
// nothing.  This replaces, for batches of systems up to kPersistMaxN particles, the
// device-driven graph batches of drift / force / finalize launches, whose per-step launch
// latency dominates once few systems are still running.
//
// Each phase is the per-element body the multi-launch path uses (fire_drift_core, the
// rebuild phase bodies, disk_force_mode, fire_convergence_core, fire_update_core), and every
// per-system sum is formed in the same tree (256-particle chunks, 64-lane butterfly,
// (w0 + w1) + (w2 + w3), chunk partials folded by 256 strided accumulators), so a system's
// trajectory is bit-identical to the one the multi-launch path produces.
//
// Reference: BaseFIRE::step_impl (include/integrators/fire.cuh:59-88), fire.cu:7-78,
// minimize_fire (include/routines/minimizers.cuh:10-31), the rebuild of
// base_point_particle.hpp:180-201.
#include "internal.hpp"
#include "fire.hpp"
#include "disk_dev.hpp"
#include "cells_dev.hpp"
#include <climits>

namespace dpmd {
namespace {

// The LDS-resident engine: systems of up to kLdsMaxN particles, one 512-thread workgroup
// each; 11 per-particle arrays of doubles take 132 KB of the CU's 160 KB.
constexpr int kLdsMaxN = 1536;
constexpr int kLdsBlock = 512;
// That state needs gfx950's LDS; device code built for a GPU with less per workgroup gets an
// empty LDS-resident kernel, which the host never selects there (lds_engine_fits).
#if defined(__HIP_DEVICE_COMPILE__) && !defined(__gfx950__)
constexpr int kLdsArchN = 1;
#else
constexpr int kLdsArchN = kLdsMaxN;
#endif

struct PersistArgs {
    const int* order;   // workgroup b runs system order[b]
    const int* todo;    // per system: 0 nothing, 1 run until it stops or max_steps, 2 one step
    int* stop_step;     // per system: the step whose stopping test stopped it (-1: none)
    int max_steps;
    int pred; double pa, pb;
};

// Exclusive scan of one int per thread over a workgroup of W threads.
template <int W>
__device__ __forceinline__ int wg_excl_scan_int(int v, int* lds, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) lds[w] = inc;
    __syncthreads();
    int pre = 0;
    total = 0;
    for (int k = 0; k < W / 64; ++k) {
        if (k < w) pre += lds[k];
        total += lds[k];
    }
    __syncthreads();
    return pre + inc - v;
}

// The rebuild of one system by its workgroup: the phases of the multi-launch chain in the
// same order (cell ids and counts, exclusive scan of the system's cells, atomic scatter,
// per-cell insertion sort by current index, gather, copy back, neighbor lists).
template <int W>
__device__ void rebuild_system(const RebuildArgs& r, int s, int beg, int end, int* ilds) {
    const int t = threadIdx.x;
    const int cs = r.ca.cstart[s], ce = r.ca.cstart[s + 1];
    if (t == 0) r.ca.flag[s] = 1u;
    __syncthreads();
    for (int c = cs + t; c < ce; c += W) r.ca.cell_count[c] = 0;
    __syncthreads();
    for (int i = beg + t; i < end; i += W) ph_hash(r, i);
    __syncthreads();
    // cell starts: the system's particles begin at beg, its cells are contiguous
    const int nc = ce - cs;
    const int per = (nc + W - 1) / W;
    const int c0 = cs + t * per, c1 = min(c0 + per, ce);
    int sum = 0;
    for (int c = c0; c < c1; ++c) sum += r.ca.cell_count[c];
    int total;
    int run = beg + wg_excl_scan_int<W>(sum, ilds, total);
    for (int c = c0; c < c1; ++c) {
        r.ca.cell_start[c] = run;
        r.ca.cell_wptr[c] = run;
        run += r.ca.cell_count[c];
    }
    __syncthreads();
    for (int i = beg + t; i < end; i += W) ph_scatter(r, i);
    __syncthreads();
    for (int c = cs + t; c < ce; c += W) ph_sort(r, c);
    __syncthreads();
    for (int p = beg + t; p < end; p += W) ph_gather(r, p);
    __syncthreads();
    for (int p = beg + t; p < end; p += W) ph_copy_back(r, p);
    __syncthreads();
    for (int i = beg + t; i < end; i += W) ph_neighbors(r, i);
    __syncthreads();
    if (t == 0) r.ca.flag[s] = 0u;
}

// A system's per-particle state held in LDS by the LDS-resident engine (NL > 0): positions,
// velocities, forces, energies, radii, masses and the displacement reference, 88 bytes per
// particle.  The neighbor lists stay in global memory (read-only between rebuilds).
template <int NL>
struct LdsState {
    double x[NL], y[NL], vx[NL], vy[NL], fx[NL], fy[NL], pe[NL], r[NL], m[NL], lx[NL], ly[NL];
};

template <int W, int NL>
__device__ void lds_load(LdsState<NL>& L, const DiskArgs& a, int beg, int end) {
    for (int i = beg + (int)threadIdx.x; i < end; i += W) {
        const int k = i - beg;
        L.x[k] = a.x[i]; L.y[k] = a.y[i]; L.vx[k] = a.vx[i]; L.vy[k] = a.vy[i];
        L.fx[k] = a.fx[i]; L.fy[k] = a.fy[i]; L.pe[k] = a.pe[i]; L.r[k] = a.rad[i]; L.m[k] = a.mass[i];
        L.lx[k] = a.cell ? a.lx[i] : 0.0; L.ly[k] = a.cell ? a.ly[i] : 0.0;
    }
}

// What the step changes: positions, velocities, forces and energies.
template <int W, int NL>
__device__ void lds_store(const LdsState<NL>& L, const DiskArgs& a, int beg, int end) {
    for (int i = beg + (int)threadIdx.x; i < end; i += W) {
        const int k = i - beg;
        a.x[i] = L.x[k]; a.y[i] = L.y[k]; a.vx[i] = L.vx[k]; a.vy[i] = L.vy[k];
        a.fx[i] = L.fx[k]; a.fy[i] = L.fy[k]; a.pe[i] = L.pe[k];
    }
}

// NL = 0: the system's state lives in global memory (L2-resident for these sizes).
// NL > 0: it is loaded into LDS once per launch, every phase but the rebuild works there,
// and it goes back to global memory before a rebuild (which permutes it), is reloaded after,
// and is stored at the end.  The arithmetic and the order of every sum are the same in both.
template <int MODE, int W, int NL>
__device__ __forceinline__ void fire_persist_body(DiskArgs a, FireArgs f, RedArgs rd, RebuildArgs r, PersistArgs p) {
    constexpr int NW = W / 64;
    // one partial per wavefront of particles (a 256-particle chunk is 4 of them), so the
    // force pass needs a single barrier before the chunk sums are formed
    __shared__ double red[3][(NL ? NL : kPersistMaxN) / 64];
    __shared__ LdsState<NL ? NL : 1> L;
    __shared__ double fold[3][4];
    __shared__ int ilds[NW];
    __shared__ double sh_dt, sh_alpha, sh_dtrev, sh_vscale, sh_live_dt;
    __shared__ int sh_flag, sh_stop, sh_frozen;
    const int s = p.order[blockIdx.x];
    const int todo = p.todo[s];
    if (todo == 0) return;   // the whole workgroup
    const int beg = a.off[s], end = a.off[s + 1];
    const int nch = (end - beg + 255) / 256;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int steps = todo == 2 ? 1 : p.max_steps;
    // thread 0 owns the system's FIRE scalars for the run (stored back at the end)
    FireSys q{};
    int stopped = 0, frozen = 0, nsteps = 0;
    if (t == 0) {
        q = fire_load(f, s);
        stopped = f.stopped[s];
        frozen = f.frozen[s];
        nsteps = f.nsteps[s];
        sh_frozen = frozen;
    }
    if constexpr (NL > 0) lds_load<W, NL>(L, a, beg, end);
    __syncthreads();
    for (int k = 0; k < steps; ++k) {
        if (sh_frozen) break;
        if (t == 0) {
            // k_fire_drift: the step uses the next-step scalars, which become the live ones
            sh_dt = q.s_dt; sh_alpha = q.s_alpha; sh_dtrev = q.s_dtrev; sh_vscale = q.s_vscale;
            q.dt = q.s_dt; q.alpha = q.s_alpha; q.ngood = q.s_ngood; q.nbad = q.s_nbad;
            sh_live_dt = q.dt;
            sh_flag = 0;
            sh_stop = 0;
        }
        __syncthreads();
        bool flag = false;
        if constexpr (NL > 0) {
            const double dt = sh_dt, alpha = sh_alpha, dtrev = sh_dtrev, vscale = sh_vscale;
            const double th2 = a.thresh2[s];
            for (int q0 = t; q0 < end - beg; q0 += W) {
                double x = L.x[q0], y = L.y[q0], vx = L.vx[q0], vy = L.vy[q0];
                flag = fire_drift_vals(x, y, vx, vy, L.fx[q0], L.fy[q0], L.m[q0], a.cell, L.lx[q0], L.ly[q0], th2, dt,
                                       alpha, dtrev, vscale) || flag;
                L.x[q0] = x; L.y[q0] = y; L.vx[q0] = vx; L.vy[q0] = vy;
            }
        } else {
            // DU particles per thread per pass, every load issued before the first store
            constexpr int DU = 4;
            const double dt = sh_dt, alpha = sh_alpha, dtrev = sh_dtrev, vscale = sh_vscale;
            const double th2 = a.thresh2[s];
            for (int i0 = beg + t; i0 < end; i0 += DU * W) {
                double x[DU], y[DU], vx[DU], vy[DU], fx[DU], fy[DU], m[DU], lx[DU], ly[DU];
#pragma unroll
                for (int u = 0; u < DU; ++u) {
                    const int i = i0 + u * W;
                    if (i < end) {
                        x[u] = a.x[i]; y[u] = a.y[i]; vx[u] = a.vx[i]; vy[u] = a.vy[i];
                        fx[u] = a.fx[i]; fy[u] = a.fy[i]; m[u] = a.mass[i];
                        lx[u] = a.cell ? a.lx[i] : 0.0; ly[u] = a.cell ? a.ly[i] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < DU; ++u) {
                    const int i = i0 + u * W;
                    if (i < end) {
                        flag = fire_drift_vals(x[u], y[u], vx[u], vy[u], fx[u], fy[u], m[u], a.cell, lx[u], ly[u], th2,
                                               dt, alpha, dtrev, vscale) || flag;
                        a.x[i] = x[u]; a.y[i] = y[u]; a.vx[i] = vx[u]; a.vy[i] = vy[u];
                    }
                }
            }
        }
        if (flag) sh_flag = 1;
        __syncthreads();
        if (sh_flag) {
            if constexpr (NL > 0) {
                lds_store<W, NL>(L, a, beg, end);
                __syncthreads();
            }
            rebuild_system<W>(r, s, beg, end, ilds);
            if constexpr (NL > 0) {
                lds_load<W, NL>(L, a, beg, end);
                __syncthreads();
            }
        }
        // forces, kick and wavefront partials (k_fire_force)
        const double dt = sh_live_dt;
        for (int b0 = beg; b0 < end; b0 += W) {
            const int i = b0 + t;
            double vp = 0.0, vpe = 0.0, vf2 = 0.0;
            if (i < end && NL > 0) {
                const int q0 = i - beg;
                double fx, fy, pe;
                disk_force_mode_pos<MODE>(a, LdsPos{L.x, L.y, L.r, beg}, i, s, fx, fy, pe);
                L.fx[q0] = fx; L.fy[q0] = fy; L.pe[q0] = pe;
                double vx = L.vx[q0], vy = L.vy[q0];
                const double smi = dt * 0.5 / L.m[q0];
                vx = vx + fx * smi; vy = vy + fy * smi;
                L.vx[q0] = vx; L.vy[q0] = vy;
                vp = fx * vx + fy * vy;
                vpe = pe;
                vf2 = fx * fx + fy * fy;
            } else if (i < end) {
                double fx, fy, pe;
                disk_force_mode<MODE>(a, i, s, fx, fy, pe);
                a.fx[i] = fx; a.fy[i] = fy; a.pe[i] = pe;
                double vx = a.vx[i], vy = a.vy[i];
                const double smi = dt * 0.5 / a.mass[i];
                vx = vx + fx * smi; vy = vy + fy * smi;
                a.vx[i] = vx; a.vy[i] = vy;
                vp = fx * vx + fy * vy;
                vpe = pe;
                vf2 = fx * fx + fy * fy;
            }
            vp = wave_sum(vp); vpe = wave_sum(vpe); vf2 = wave_max(vf2);
            const int gw = (b0 - beg) / 64 + w;
            if (lane == 0) { red[0][gw] = vp; red[1][gw] = vpe; red[2][gw] = vf2; }
        }
        __syncthreads();
        // chunk sums ((w0 + w1) + (w2 + w3) of its wavefronts), folded as k_fire_finalize
        // does: 256 strided accumulators (one chunk each here), then the 256-thread tree
        double ap = 0.0, ape = 0.0, af = 0.0;
        if (t < nch) {
            const int q4 = 4 * t;
            const double cp = (red[0][q4] + red[0][q4 + 1]) + (red[0][q4 + 2] + red[0][q4 + 3]);
            const double ce = (red[1][q4] + red[1][q4 + 1]) + (red[1][q4 + 2] + red[1][q4 + 3]);
            const double cf = fmax(fmax(red[2][q4], red[2][q4 + 1]), fmax(red[2][q4 + 2], red[2][q4 + 3]));
            ap = ap + cp; ape = ape + ce; af = fmax(af, cf);
        }
        if (w < 4) {
            ap = wave_sum(ap); ape = wave_sum(ape); af = wave_max(af);
            if (lane == 0) { fold[0][w] = ap; fold[1][w] = ape; fold[2][w] = af; }
        }
        __syncthreads();
        if (t == 0) {
            ap = (fold[0][0] + fold[0][1]) + (fold[0][2] + fold[0][3]);
            ape = (fold[1][0] + fold[1][1]) + (fold[1][2] + fold[1][3]);
            af = fmax(fmax(fold[2][0], fold[2][1]), fmax(fold[2][2], fold[2][3]));
            const int was_stopped = stopped;
            rd.pw_tot[s] = ap; rd.pe_tot[s] = ape; rd.maxf2[s] = af;
            fire_convergence_core(q, end - beg, ape, af, p.pred, p.pa, p.pb);
            if (q.dt == 0.0 && !was_stopped) {
                stopped = 1;
                if (todo == 1) p.stop_step[s] = k;
                sh_stop = 1;
            }
            fire_update_core(q, ap);
            if (was_stopped) { frozen = 1; sh_frozen = 1; }
            nsteps += 1;
        }
        __syncthreads();
        if (sh_stop) break;
    }
    if constexpr (NL > 0) lds_store<W, NL>(L, a, beg, end);
    if (t == 0) {
        fire_store(f, s, q);
        f.stopped[s] = stopped;
        f.frozen[s] = frozen;
        f.nsteps[s] = nsteps;
    }
}

template <int MODE, int W, int NL>
__global__ void __launch_bounds__(W) k_fire_persist(DiskArgs a, FireArgs f, RedArgs rd, RebuildArgs r, PersistArgs p) {
    if constexpr (NL <= kLdsArchN) fire_persist_body<MODE, W, NL>(a, f, rd, r, p);
}

// Whether this device gives one workgroup the LDS-resident kernel's static LDS (checked once
// per device).
bool lds_engine_fits() {
    static int fits[64];   // 0 unknown, 1 yes, 2 no
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return false;
    if (fits[dev] == 0) {
        int maxl = 0;
        HIP_OK(hipDeviceGetAttribute(&maxl, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        hipFuncAttributes fa{};
        HIP_OK(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(
                                             &k_fire_persist<DPMD_FORCE_PBC, kLdsBlock, kLdsMaxN>)));
        fits[dev] = fa.sharedSizeBytes > 1024 && fa.sharedSizeBytes <= (size_t)maxl ? 1 : 2;
    }
    return fits[dev] == 1;
}

template <class F>
void persist_by_mode(int mode, F&& f) {
    switch (mode) {
        case DPMD_FORCE_PBC: f(std::integral_constant<int, DPMD_FORCE_PBC>{}); break;
        case DPMD_FORCE_WALL: f(std::integral_constant<int, DPMD_FORCE_WALL>{}); break;
        default: f(std::integral_constant<int, DPMD_FORCE_PER_SYSTEM>{}); break;
    }
}

// nmax: the largest system of the launch; systems of up to kLdsMaxN particles run in LDS.
void launch(dpmd_ctx* c, const PersistArgs& p, int nblocks, int nmax) {
    if (nblocks <= 0) return;
    DiskArgs a = disk_args(c);
    FireArgs f = fire_args(c);
    RedArgs rd = red_args(c);
    RebuildArgs r = rebuild_args(c);
    const bool lds = c->persist_lds && nmax <= kLdsMaxN && lds_engine_fits();
    persist_by_mode(c->fire_mode, [&](auto m) {
        constexpr int M = decltype(m)::value;
        if (lds) {
            k_fire_persist<M, kLdsBlock, kLdsMaxN><<<nblocks, kLdsBlock, 0, c->st>>>(a, f, rd, r, p);
            return;
        }
        switch (c->persist_block) {
            case 256: k_fire_persist<M, 256, 0><<<nblocks, 256, 0, c->st>>>(a, f, rd, r, p); break;
            case 1024: k_fire_persist<M, 1024, 0><<<nblocks, 1024, 0, c->st>>>(a, f, rd, r, p); break;
            default: k_fire_persist<M, 512, 0><<<nblocks, 512, 0, c->st>>>(a, f, rd, r, p); break;
        }
    });
    HIP_OK(hipGetLastError());
}

}  // namespace

bool fire_persist_ok(dpmd_ctx* c) {
    if (!c->persist_fire || c->kind == DPMD_RIGID_BUMPY || c->method != DPMD_NEIGHBOR_CELL) return false;
    if (!c->mask_frozen || c->capturing || c->rec_seg) return false;
    for (int s = 0; s < c->S; ++s)
        if (c->h_off[s + 1] - c->h_off[s] > kPersistMaxN) return false;
    return true;
}

// Per-step cost models (microseconds) of the two engines on the MI355X, from timed runs of
// FIRE on single systems and on the 512-system batch: a persistent workgroup is bound by one
// CU's memory and FP64 throughput (about 9 ns per particle; two systems share a CU when more
// than 256 run).  A step of per-step launches over the chunks of the running systems costs a
// few us of launch latency plus a share of their particles (5 us + 0.08 ns per particle, set
// by the time of the 512-system jamming workload: 47 s, against 61 s with 18 us + 0.18 ns).
// Against the LDS-resident engine the per-step launches also pay the rebuild halts that the
// engine's in-workgroup rebuilds avoid, so while the running systems fit that engine the
// launches are charged 18 us + 0.18 ns per particle (a 128 x 1024 FIRE run to 1e-12 took
// 1.15 s with the lower charge, 0.86 s with this one).
double persist_step_us(int nmax, int nrun) {
    return (3.2 + 0.009 * nmax) * std::max(1, (nrun + 255) / 256);
}
double launch_step_us(int64_t ntotal, bool vs_lds) {
    return vs_lds ? 18.0 + 0.00018 * (double)ntotal : 5.0 + 0.00008 * (double)ntotal;
}

// The device-driven FIRE run of lean_run with the same results.  Running systems advance in
// persistent phases (persist_phase steps), each system in its own workgroup until its
// stopping test fires; at the end of a phase every system still running is at the same step
// index, so when the per-step launches have become the cheaper engine for what is left
// (few systems, large ones) the run continues with them in lockstep.  The run's length is
// the lean run's (one past the last stop when every system stops, else max_steps), and each
// system that stopped before the end takes the one further step the lean run gives it (the
// step that zeroes its velocities and freezes it).
int fire_persistent_run(dpmd_ctx* c, int max_steps, int pred, double a, double b, bool use_graph, int* conv) {
    const int S = c->S;
    std::vector<int> frozen(S), todo(S), stop(S);
    HIP_OK(hipMemcpyAsync(frozen.data(), c->f_frozen.p, S * 4, hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    std::vector<int> running;
    for (int s = 0; s < S; ++s) if (!frozen[s]) running.push_back(s);
    if (running.empty() || max_steps <= 0) {
        if (conv) *conv = running.empty() ? 1 : 0;
        return 0;
    }
    // largest systems first: the hardware dispatches workgroups in order
    auto size = [&](int s) { return c->h_off[s + 1] - c->h_off[s]; };
    std::stable_sort(running.begin(), running.end(), [&](int x, int y) { return size(x) > size(y); });
    int done = 0;
    for (;;) {
        int nmax = 0;
        int64_t ntot = 0;
        for (int s : running) { nmax = std::max(nmax, size(s)); ntot += size(s); }
        const bool lds_run = c->persist_lds && nmax <= kLdsMaxN && lds_engine_fits();
        if (c->persist_fire == 1 && launch_step_us(ntot, lds_run) < persist_step_us(nmax, (int)running.size())) {
            // hand the rest of the run to the per-step launches, in lockstep from step `done`;
            // their launches cover only the chunks of the systems still running (the others
            // would return at once, but a launch over every chunk of a large batch costs its
            // dispatch), the set shrinks as systems freeze, and the last system left takes the
            // fused FIRE step
            const int nrun = (int)running.size();
            int run[RUN_WORDS] = {0, 0, nrun, 1, 0, -1, -1, 0, 0, 0};
            c->run.upload(run, RUN_WORDS, c->st);
            std::vector<int> first(S + 1, 0), chunks;
            for (int s = 0; s < S; ++s) first[s + 1] = first[s] + (size(s) + DPMD_BLOCK - 1) / DPMD_BLOCK;
            std::vector<int> sys(running);
            std::sort(sys.begin(), sys.end());
            for (int s : sys)
                for (int k = first[s]; k < first[s + 1]; ++k) chunks.push_back(k);
            c->act_chunk.upload(chunks.data(), (int64_t)chunks.size(), c->st);
            c->act_sys.upload(sys.data(), (int64_t)sys.size(), c->st);
            struct ActScope {
                dpmd_ctx* c;
                ~ActScope() { c->n_act_chunk = 0; c->n_act_sys = 0; c->fsys = -1; }
            } scope{c};
            c->n_act_chunk = (int)chunks.size();
            c->n_act_sys = (int)sys.size();
            if (sys.size() == 1 && c->fuse_fire) {
                c->fsys = sys[0];
                c->fsys_ch0 = first[sys[0]];
            }
            int lconv = 0;
            const int steps = lean_run_batches(c, max_steps - done, true, pred, a, b, use_graph, &lconv);
            if (conv) *conv = lconv;
            return done + steps;
        }
        const int phase = std::min(c->persist_phase, max_steps - done);
        std::fill(todo.begin(), todo.end(), 0);
        for (int s : running) todo[s] = 1;
        std::fill(stop.begin(), stop.end(), -1);
        c->persist_order.upload(running.data(), (int64_t)running.size(), c->st);
        c->persist_todo.upload(todo.data(), S, c->st);
        c->persist_stop.upload(stop.data(), S, c->st);
        PersistArgs p{c->persist_order.p, c->persist_todo.p, c->persist_stop.p, phase, pred, a, b};
        launch(c, p, (int)running.size(), nmax);
        HIP_OK(hipMemcpyAsync(stop.data(), c->persist_stop.p, S * 4, hipMemcpyDeviceToHost, c->st));
        HIP_OK(hipStreamSynchronize(c->st));
        std::vector<int> still;
        int last = -1;
        for (int s : running) {
            if (stop[s] < 0) still.push_back(s);
            else last = std::max(last, done + stop[s]);
        }
        const int end = done + phase;
        // the run's length if it ends now (INT_MAX while it goes on)
        const int total = still.empty() ? last + 1 : (end == max_steps ? max_steps : INT_MAX);
        std::vector<int> extra;
        std::fill(todo.begin(), todo.end(), 0);
        int emax = 0;
        for (int s : running)
            if (stop[s] >= 0 && done + stop[s] + 1 < total) {
                todo[s] = 2;
                extra.push_back(s);
                emax = std::max(emax, size(s));
            }
        if (!extra.empty()) {
            c->persist_order.upload(extra.data(), (int64_t)extra.size(), c->st);
            c->persist_todo.upload(todo.data(), S, c->st);
            launch(c, p, (int)extra.size(), emax);
        }
        if (total != INT_MAX) {
            int nrun = (int)still.size();   // run words as the lean run leaves them
            HIP_OK(hipMemcpyAsync(c->run.p + RUN_RUNNING, &nrun, sizeof(int), hipMemcpyHostToDevice, c->st));
            HIP_OK(hipStreamSynchronize(c->st));
            if (conv) *conv = still.empty() ? 1 : 0;
            return total;
        }
        running.swap(still);
        done = end;
    }
}

}  // namespace dpmd
