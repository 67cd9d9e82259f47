// Cell hash, deterministic cell sort, SoA reorder and Verlet (ELL) neighbor build.
//
// Reference pipeline: base_point_particle.hpp:180-201 (update_cell_neighbors_impl),
// base_poly_particle.hpp:239-260 (vertex variant), kernels in
// src/kernels/base_particle_kernels.cu:205-259 and base_point_particle_kernels.cu:37-157.
//
// Differences by design:
//  * only systems whose rebuild flag is set are rebuilt (the reference rebuilds every
//    system when any flag is set, disk.cu:1127-1138); the decision is taken on the device,
//    so no host round trip happens per step;
//  * in-cell order is deterministic (ascending current index), by an atomic scatter
//    followed by a per-cell insertion sort; the reference's atomic scatter order is
//    nondeterministic (base_particle_kernels.cu:246-259);
//  * the neighbor list is a fixed-capacity column-major ELL array nbr[k*n+i] (coalesced
//    for one-thread-per-object force kernels), with the reference's stencil order.
#include "internal.hpp"
#include "cells_dev.hpp"
namespace dpmd {
namespace {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = DPMD_BLOCK * SCAN_ITEMS;

__device__ __forceinline__ int block_excl_scan_int(int v, int* lds4, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) lds4[w] = inc;
    __syncthreads();
    int pre = 0;
    for (int k = 0; k < w; ++k) pre += lds4[k];
    total = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return pre + inc - v;
}

// ---- per-element phase bodies -------------------------------------------------------

// Block-level: sum of one tile of cell counts.
__device__ __forceinline__ void ph_tile_sum(const RebuildArgs& r, int t, int* lds) {
    const int C = r.ca.C;
    const int base = t * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) s += (base + k < C) ? r.ca.cell_count[base + k] : 0;
    int total;
    (void)block_excl_scan_int(s, lds, total);
    if (threadIdx.x == 0) r.tile_sum[t] = total;
}

// Block-level: exclusive scan of the tile sums (one block).
__device__ __forceinline__ void ph_top_scan(const RebuildArgs& r, int* lds) {
    int carry = 0;
    for (int b = 0; b < r.nt; b += DPMD_BLOCK) {
        const int t = b + threadIdx.x;
        const int v = t < r.nt ? r.tile_sum[t] : 0;
        int total;
        const int ex = block_excl_scan_int(v, lds, total);
        if (t < r.nt) r.tile_sum[t] = carry + ex;
        carry += total;
    }
}

// Block-level: cell_start for one tile; write pointers for flagged systems' cells.  Counts
// of systems that are not rebuilt are unchanged since their last rebuild, so their starts
// are recomputed to the same values.
__device__ __forceinline__ void ph_tile_down(const RebuildArgs& r, int t, int* lds) {
    const int C = r.ca.C;
    const int base = t * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = (base + k < C) ? r.ca.cell_count[base + k] : 0; s += v[k]; }
    int total;
    int run = block_excl_scan_int(s, lds, total) + r.tile_sum[t];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int c = base + k;
        if (c < C) {
            r.ca.cell_start[c] = run;
            if (r.ca.flag[r.ca.cell_sys[c]]) r.ca.cell_wptr[c] = run;
            if (c == C - 1) r.ca.cell_start[C] = run + v[k];
        }
        run += v[k];
    }
}

// reset_displacements_kernel for rigid bodies (rigid_bumpy.cu:1070-1085): the skin
// reference point is the first vertex of each particle.
__device__ __forceinline__ void ph_rigid_reset(const RebuildArgs& r, int i) {
    if (!r.ca.flag[r.psid[i]]) return;
    const int fv = r.static_index[r.poff[i]];
    r.plx[i] = r.o.x[fv]; r.ply[i] = r.o.y[fv];
}

// ---- multi-launch form ----------------------------------------------------------------
#define DPMD_EACH(name, count, body)                                                          \
    __global__ void name(RebuildArgs r, const int* active) {                                  \
        const int i = blockIdx.x * blockDim.x + threadIdx.x;                                  \
        if (i >= (count) || !*active) return;                                                 \
        body(r, i);                                                                           \
    }
DPMD_EACH(k_zero, r.ca.C, ph_zero)
DPMD_EACH(k_hash, r.o.n, ph_hash)
DPMD_EACH(k_scatter, r.o.n, ph_scatter)
DPMD_EACH(k_sort, r.ca.C, ph_sort)
DPMD_EACH(k_gather, r.o.n, ph_gather)
DPMD_EACH(k_copy_back, r.o.n, ph_copy_back)
DPMD_EACH(k_rigid_reset, r.N, ph_rigid_reset)
#undef DPMD_EACH

// ---- neighbor build, output staged in LDS -----------------------------------------------
// Each object appends its neighbors to its own column of an LDS table nbr_l[k][thread]; the
// workgroup then writes rows 0 .. (its largest count - 1) of the column-major ELL table as
// whole 1 KB rows.  Written directly, each lane stores its k-th neighbor at a different time
// of its divergent stencil loop, and rows become scattered partial cache lines (about 4x the
// payload in HBM writes).  Objects of systems that are not rebuilt keep their lists.
__global__ void __launch_bounds__(DPMD_BLOCK) k_neighbors(RebuildArgs r, const int* active) {
    extern __shared__ int nbr_l[];   // K rows x DPMD_BLOCK
    __shared__ int wmax[DPMD_BLOCK / 64];
    if (!*active) return;
    const ObjArgs& o = r.o; const CellArgs& ca = r.ca; const NbrArgs& nb = r.nb;
    const int i = blockIdx.x * DPMD_BLOCK + threadIdx.x;
    const int K = nb.K;
    int cnt = 0;
    bool rebuilt = false;
    if (i < o.n) {
        const int s = o.sid[i];
        rebuilt = ca.flag[s] != 0;
        if (rebuilt) {
            cnt = neighbor_scan(r, i, s, [&](int k, int j) {
                if (k < K) nbr_l[k * DPMD_BLOCK + threadIdx.x] = j;
            });
            if (cnt > K) atomicMax(nb.overflow, cnt);
            nb.nn[i] = cnt < K ? cnt : K;
        }
    }
    // largest stored count of the workgroup: the rows to write.  (No global maximum is kept
    // here: one same-address atomic per wave serialises the whole launch.)
    int m = cnt;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    int rows = 0;
#pragma unroll
    for (int w = 0; w < DPMD_BLOCK / 64; ++w) rows = max(rows, wmax[w]);
    rows = min(rows, K);
    const int mine = min(cnt, K);
    if (rebuilt)
        for (int k = 0; k < rows; ++k)
            nb.nbr[(size_t)k * o.n + i] = k < mine ? nbr_l[k * DPMD_BLOCK + threadIdx.x] : -1;
}

// Capacities above kLdsCapacity: write the ELL table directly.
constexpr int kLdsCapacity = 64;
__global__ void __launch_bounds__(DPMD_BLOCK) k_neighbors_direct(RebuildArgs r, const int* active) {
    if (!*active) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r.o.n) ph_neighbors(r, i);
}

__global__ void k_tile_sum(RebuildArgs r, const int* active) {
    __shared__ int lds[4];
    if (!*active) return;
    ph_tile_sum(r, blockIdx.x, lds);
}
__global__ void k_top_scan(RebuildArgs r, const int* active) {
    __shared__ int lds[4];
    if (!*active) return;
    ph_top_scan(r, lds);
}
__global__ void k_tile_down(RebuildArgs r, const int* active) {
    __shared__ int lds[4];
    if (!*active) return;
    ph_tile_down(r, blockIdx.x, lds);
}

// Gate of a device-side rebuild: copy the "some flag is set" word (the rebuild clears it)
// and count the rebuilds run.  One thread.
__global__ void k_rebuild_gate(int* run) {
    const int a = run[RUN_ANY] != 0 ? 1 : 0;
    run[RUN_GATE] = a;
    if (a) run[RUN_REBUILDS] += 1;
}

__global__ void k_set_flags(int S, unsigned* flag, const int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !*active) return;
    flag[s] = 1u;
}

__global__ void k_clear_flags(int S, unsigned* flag, int* any, const int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !*active) return;
    flag[s] = 0u;
    if (s == 0) *any = 0;
}

// set_naive_neighbor_count + fill_naive_neighbor_list_kernel (base_point_particle_kernels.cu:
// 6-35); vertex form excludes same-particle vertices (base_poly_particle_kernels.cu:18-55).
__global__ void k_naive(ObjArgs o, NbrArgs nb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= o.n) return;
    const int s = o.sid[i];
    const int pi = o.pid ? o.pid[i] : -1;
    int cnt = 0;
    for (int j = o.off[s]; j < o.off[s + 1]; ++j) {
        if (j == i) continue;
        if (o.pid && o.pid[j] == pi) continue;
        if (cnt < nb.K) nb.nbr[(size_t)cnt * o.n + i] = j;
        ++cnt;
    }
    if (cnt > nb.K) atomicMax(nb.overflow, cnt);
    nb.nn[i] = cnt < nb.K ? cnt : nb.K;
}

ObjArgs obj_args(dpmd_ctx* c) {
    ObjArgs o;
    if (c->kind == DPMD_RIGID_BUMPY) {
        o.x = c->vpx.p; o.y = c->vpy.p; o.rad = c->vrad.p; o.sid = c->vsid.p; o.off = c->voff.p;
        o.pid = c->vpid.p; o.n = c->Nv;
    } else {
        o.x = c->x.p; o.y = c->y.p; o.rad = c->rad.p; o.sid = c->sid.p; o.off = c->off.p;
        o.pid = nullptr; o.n = c->N;
    }
    return o;
}

CellArgs cell_args(dpmd_ctx* c) {
    CellArgs a;
    a.cdx = c->cdx.p; a.cdy = c->cdy.p; a.cstart = c->cstart.p;
    a.Lx = c->Lx.p; a.Ly = c->Ly.p; a.iLx = c->iLx.p; a.iLy = c->iLy.p; a.cutoff = c->cutoff.p;
    a.cell_id = c->cell_id.p; a.cell_count = c->cell_count.p; a.cell_start = c->cell_start.p;
    a.cell_wptr = c->cell_wptr.p; a.cell_sys = c->cell_sys.p; a.C = c->C; a.flag = c->flag.p;
    return a;
}

NbrArgs nbr_args(dpmd_ctx* c) {
    NbrArgs nb;
    nb.nn = c->nn.p; nb.nbr = c->nbr.p; nb.K = c->K; nb.overflow = c->overflow.p;
    return nb;
}

PermArgs perm_args(dpmd_ctx* c) {
    PermArgs pa{};
    if (c->kind == DPMD_RIGID_BUMPY) {
        double* d[] = {c->vpx.p, c->vpy.p, c->vfx.p, c->vfy.p, c->vmass.p, c->vrad.p, c->vpe.p};
        pa.nd = 7;
        for (int k = 0; k < pa.nd; ++k) pa.d[k] = d[k];
        pa.i[0] = c->vpid.p; pa.i[1] = c->cell_id.p; pa.ni = 2;
    } else {
        double* d[] = {c->x.p, c->y.p, c->vx.p, c->vy.p, c->fx.p, c->fy.p, c->rad.p, c->mass.p, c->pe.p, c->area.p};
        pa.nd = 10;
        for (int k = 0; k < pa.nd; ++k) pa.d[k] = d[k];
        pa.i[0] = c->cell_id.p; pa.ni = 1;
    }
    pa.tmp_d = c->tmp_d.p; pa.tmp_i = c->tmp_i.p;
    return pa;
}

RebuildArgs make_rebuild_args(dpmd_ctx* c) {
    RebuildArgs r;
    r.o = obj_args(c); r.ca = cell_args(c); r.pa = perm_args(c); r.nb = nbr_args(c);
    r.order = c->order.p; r.order_inv = c->order_inv.p; r.static_index = c->static_index.p;
    const bool rigid = c->kind == DPMD_RIGID_BUMPY;
    r.lx = rigid ? nullptr : c->lx.p; r.ly = rigid ? nullptr : c->ly.p;
    r.N = c->N; r.psid = c->sid.p; r.poff = c->poff.p;
    r.plx = rigid ? c->lx.p : nullptr; r.ply = rigid ? c->ly.p : nullptr;
    r.tile_sum = c->scan_part.p; r.nt = (c->C + SCAN_TILE - 1) / SCAN_TILE; r.S = c->S;
    r.any = c->run.p + RUN_ANY;
    return r;
}

void rebuild_multi(dpmd_ctx* c, const RebuildArgs& r, const int* active) {
    hipStream_t st = c->st;
    const int n = r.o.n, C = r.ca.C;
    k_zero<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(r, active);
    k_hash<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_tile_sum<<<r.nt, DPMD_BLOCK, 0, st>>>(r, active);
    k_top_scan<<<1, DPMD_BLOCK, 0, st>>>(r, active);
    k_tile_down<<<r.nt, DPMD_BLOCK, 0, st>>>(r, active);
    k_scatter<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_sort<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(r, active);
    k_gather<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_copy_back<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    if (c->K <= kLdsCapacity)
        k_neighbors<<<blocks_for(n), DPMD_BLOCK, (size_t)c->K * DPMD_BLOCK * sizeof(int), st>>>(r, active);
    else
        k_neighbors_direct<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    if (r.plx) k_rigid_reset<<<blocks_for(r.N), DPMD_BLOCK, 0, st>>>(r, active);
    k_clear_flags<<<blocks_for(r.S), DPMD_BLOCK, 0, st>>>(r.S, r.ca.flag, r.any, active);
}

}  // namespace

RebuildArgs rebuild_args(dpmd_ctx* c) { return make_rebuild_args(c); }

void build_chunks(dpmd_ctx* c) {
    std::vector<int> cs, cb, sc(c->S + 1, 0);
    for (int s = 0; s < c->S; ++s) {
        sc[s] = (int)cs.size();
        for (int b = c->h_off[s]; b < c->h_off[s + 1]; b += DPMD_BLOCK) { cs.push_back(s); cb.push_back(b); }
    }
    sc[c->S] = (int)cs.size();
    c->nchunks = (int)cs.size();
    c->chunk_sys.upload(cs.data(), cs.size(), c->st);
    c->chunk_beg.upload(cb.data(), cb.size(), c->st);
    c->sys_chunk.upload(sc.data(), sc.size(), c->st);
    c->sys_counter.resize(c->S); c->sys_counter.zero(c->st);
    c->part.resize(4 * (int64_t)std::max(1, c->nchunks));
}

// Moves every particle to a given current position: order[p] = the current index of the
// particle that goes to p (a permutation within each system).  Uses the rebuild's gather and
// copy-back phases, so every per-particle array, static_index and the displacement reference
// follow; cells and neighbor lists are stale afterwards (the caller rebuilds).
void permute_particles(dpmd_ctx* c, const int* order_host) {
    const int S = c->S, n = c->n_obj();
    hipStream_t st = c->st;
    c->order.upload(order_host, n, st);
    const int* active = c->one.p;
    k_set_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, c->flag.p, active);
    RebuildArgs r = make_rebuild_args(c);
    k_gather<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_copy_back<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_clear_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, r.ca.flag, r.any, active);
    HIP_OK(hipGetLastError());
}

void rebuild_on_device(dpmd_ctx* c) {
    k_rebuild_gate<<<1, 1, 0, c->st>>>(c->run.p);
    HIP_OK(hipGetLastError());
    rebuild_neighbors(c, c->run.p + RUN_GATE, false);
}

void rebuild_neighbors(dpmd_ctx* c, const int* active, bool force_all) {
    const int S = c->S;
    hipStream_t st = c->st;
    if (force_all) k_set_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, c->flag.p, active);
    RebuildArgs r = make_rebuild_args(c);
    rebuild_multi(c, r, active);
    HIP_OK(hipGetLastError());
}

void clear_flags(dpmd_ctx* c, const int* active) {
    k_clear_flags<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->flag.p, c->run.p + RUN_ANY, active);
    HIP_OK(hipGetLastError());
}

void naive_neighbors(dpmd_ctx* c) {
    k_naive<<<blocks_for(c->n_obj()), DPMD_BLOCK, 0, c->st>>>(obj_args(c), nbr_args(c));
    HIP_OK(hipGetLastError());
}

void check_overflow(dpmd_ctx* c) {
    int ov = 0;
    HIP_OK(hipMemcpyAsync(&ov, c->overflow.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    if (ov > 0)
        throw OverflowError("neighbor list overflow: an object has " + std::to_string(ov) +
                            " neighbors, capacity is " + std::to_string(c->K) +
                            " (raise it with dpmd_set_neighbor_capacity)");
}

}  // namespace dpmd
