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

namespace dpmd {
namespace {

__global__ void k_set_flags(int S, unsigned* flag, const int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !*active) return;
    flag[s] = 1u;
}

__global__ void k_clear_flags(int S, unsigned* flag, const int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !*active) return;
    flag[s] = 0u;
}

__global__ void k_cell_zero(int C, const int* __restrict__ cell_sys, const unsigned* __restrict__ flag,
                            int* __restrict__ count, const int* active) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C || !*active) return;
    if (flag[cell_sys[c]]) count[c] = 0;
}

// assign_cell_ids_kernel (base_particle_kernels.cu:205-233) + count_cells_kernel (:235-244)
__global__ void k_cell_hash_count(ObjArgs o, CellArgs ca, const int* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= o.n || !*active) return;
    const int s = o.sid[i];
    if (!ca.flag[s]) return;
    const double u = wrap01(o.x[i] * ca.iLx[s]);
    const double v = wrap01(o.y[i] * ca.iLy[s]);
    const int nx = ca.cdx[s], ny = ca.cdy[s];
    int cx = (int)floor(u * nx);
    int cy = (int)floor(v * ny);
    cx = cx < nx ? (cx > 0 ? cx : 0) : nx - 1;   // guard: u*nx may round up to nx
    cy = cy < ny ? (cy > 0 ? cy : 0) : ny - 1;
    const int cell = ca.cstart[s] + cx + cy * nx;
    ca.cell_id[i] = cell;
    atomicAdd(&ca.cell_count[cell], 1);
}

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

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = DPMD_BLOCK * SCAN_ITEMS;

__global__ void k_scan_tiles(int C, const int* __restrict__ count, int* __restrict__ tile_sum, const int* active) {
    if (!*active) return;
    __shared__ int lds[4];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) s += (base + k < C) ? count[base + k] : 0;
    int total;
    (void)block_excl_scan_int(s, lds, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ void k_scan_top(int nt, int* __restrict__ tile_sum, const int* active) {
    if (!*active) return;
    __shared__ int lds[4];
    int carry = 0;
    for (int b = 0; b < nt; b += DPMD_BLOCK) {
        const int t = b + threadIdx.x;
        const int v = t < nt ? tile_sum[t] : 0;
        int total;
        const int ex = block_excl_scan_int(v, lds, total);
        if (t < nt) tile_sum[t] = carry + ex;
        carry += total;
    }
}

__global__ void k_scan_down(int C, const int* __restrict__ count, const int* __restrict__ tile_sum,
                            int* __restrict__ cell_start, int* __restrict__ wptr,
                            const int* __restrict__ cell_sys, const unsigned* __restrict__ flag,
                            const int* active) {
    if (!*active) return;
    __shared__ int lds[4];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = (base + k < C) ? count[base + k] : 0; s += v[k]; }
    int total;
    int run = block_excl_scan_int(s, lds, total) + tile_sum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int c = base + k;
        if (c < C) {
            cell_start[c] = run;
            if (flag[cell_sys[c]]) wptr[c] = run;
            if (c == C - 1) cell_start[C] = run + v[k];
        }
        run += v[k];
    }
}

__global__ void k_scatter(ObjArgs o, CellArgs ca, int* __restrict__ order, const int* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= o.n || !*active) return;
    if (!ca.flag[o.sid[i]]) return;
    const int slot = atomicAdd(&ca.cell_wptr[ca.cell_id[i]], 1);
    order[slot] = i;
}

// Fix the in-cell order: ascending current index (stable sort by cell id).
__global__ void k_cell_sort(CellArgs ca, int* __restrict__ order, const int* active) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ca.C || !*active) return;
    if (!ca.flag[ca.cell_sys[c]]) return;
    const int b = ca.cell_start[c], e = ca.cell_start[c + 1];
    for (int k = b + 1; k < e; ++k) {
        const int key = order[k];
        int m = k - 1;
        while (m >= b && order[m] > key) { order[m + 1] = order[m]; --m; }
        order[m + 1] = key;
    }
}

// reorder_particles_impl (disk.cu:1102-1125): gather into scratch by the new order.
__global__ void k_gather(int n, const int* __restrict__ sid, const unsigned* __restrict__ flag,
                         const int* __restrict__ order, int* __restrict__ order_inv, PermArgs pa,
                         const int* active) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !*active) return;
    if (!flag[sid[p]]) return;
    const int o = order[p];
    for (int k = 0; k < pa.nd; ++k) pa.tmp_d[(size_t)k * n + p] = pa.d[k][o];
    for (int k = 0; k < pa.ni; ++k) pa.tmp_i[(size_t)k * n + p] = pa.i[k][o];
    order_inv[o] = p;
}

// Copy back, update static_index (update_static_index_kernel, base_particle_kernels.cu:
// 172-180) and reset the displacement reference (reset_displacements_impl, disk.cu:1097).
__global__ void k_copy_back(int n, const int* __restrict__ sid, const unsigned* __restrict__ flag,
                            const int* __restrict__ order_inv, int* __restrict__ static_index, PermArgs pa,
                            double* __restrict__ lx, double* __restrict__ ly, const int* active) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !*active) return;
    if (!flag[sid[p]]) return;
    for (int k = 0; k < pa.nd; ++k) pa.d[k][p] = pa.tmp_d[(size_t)k * n + p];
    for (int k = 0; k < pa.ni; ++k) pa.i[k][p] = pa.tmp_i[(size_t)k * n + p];
    static_index[p] = order_inv[static_index[p]];
    if (lx) { lx[p] = pa.tmp_d[p]; ly[p] = pa.tmp_d[(size_t)n + p]; }
}

// count_cell_neighbors_kernel + fill_neighbors_cell_kernel (base_point_particle_kernels.cu:
// 37-157) in one pass, writing a fixed-capacity ELL list.
__global__ void __launch_bounds__(DPMD_BLOCK)
k_neighbors(ObjArgs o, CellArgs ca, NbrArgs nb, const int* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= o.n || !*active) return;
    const int s = o.sid[i];
    if (!ca.flag[s]) return;
    const int nx = ca.cdx[s], ny = ca.cdy[s], cs = ca.cstart[s];
    const int loc = ca.cell_id[i] - cs;
    const int cxi = loc % nx, cyi = loc / nx;
    const double Lx = ca.Lx[s], Ly = ca.Ly[s], iLx = ca.iLx[s], iLy = ca.iLy[s], cut0 = ca.cutoff[s];
    const double xi = o.x[i], yi = o.y[i], ri = o.rad[i];
    const int pi = o.pid ? o.pid[i] : -1;
    int cnt = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        int yy = cyi + dy; if (yy < 0) yy += ny; else if (yy >= ny) yy -= ny;
        for (int dx = -1; dx <= 1; ++dx) {
            int xx = cxi + dx; if (xx < 0) xx += nx; else if (xx >= nx) xx -= nx;
            const int nc = cs + yy * nx + xx;
            const int b = ca.cell_start[nc], e = ca.cell_start[nc + 1];
            for (int j = b; j < e; ++j) {
                if (j == i) continue;
                if (o.pid && o.pid[j] == pi) continue;
                const double ddx = min_image(xi - o.x[j], Lx, iLx);
                const double ddy = min_image(yi - o.y[j], Ly, iLy);
                const double r2 = ddx * ddx + ddy * ddy;
                const double cut = ri + o.rad[j] + cut0;
                if (r2 < cut * cut) {
                    if (cnt < nb.K) nb.nbr[(size_t)cnt * o.n + i] = j;
                    ++cnt;
                }
            }
        }
    }
    if (cnt > nb.K) atomicMax(nb.overflow, cnt);
    atomicMax(nb.max_count, cnt);
    nb.nn[i] = cnt < nb.K ? cnt : nb.K;
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
    atomicMax(nb.max_count, cnt);
    nb.nn[i] = cnt < nb.K ? cnt : nb.K;
}

// reset_displacements_kernel for rigid bodies (rigid_bumpy.cu:1070-1085): the skin
// reference point is the first vertex of each particle.
__global__ void k_rigid_reset_disp(int N, const int* __restrict__ sid, const unsigned* __restrict__ flag,
                                   const int* __restrict__ poff, const int* __restrict__ vstatic,
                                   const double* __restrict__ vpx, const double* __restrict__ vpy,
                                   double* __restrict__ lx, double* __restrict__ ly, const int* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !*active) return;
    if (!flag[sid[i]]) return;
    const int fv = vstatic[poff[i]];
    lx[i] = vpx[fv]; ly[i] = vpy[fv];
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
    nb.nn = c->nn.p; nb.nbr = c->nbr.p; nb.K = c->K; nb.overflow = c->overflow.p; nb.max_count = c->max_count.p;
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

}  // namespace

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

void rebuild_neighbors(dpmd_ctx* c, const int* active, bool force_all) {
    const int n = c->n_obj();
    const int S = c->S;
    hipStream_t st = c->st;
    if (force_all) k_set_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, c->flag.p, active);
    ObjArgs o = obj_args(c);
    CellArgs ca = cell_args(c);
    NbrArgs nb = nbr_args(c);
    PermArgs pa = perm_args(c);
    const int C = c->C;
    const int nt = (C + SCAN_TILE - 1) / SCAN_TILE;
    k_cell_zero<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(C, c->cell_sys.p, c->flag.p, c->cell_count.p, active);
    k_cell_hash_count<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(o, ca, active);
    k_scan_tiles<<<nt, DPMD_BLOCK, 0, st>>>(C, c->cell_count.p, c->scan_part.p, active);
    k_scan_top<<<1, DPMD_BLOCK, 0, st>>>(nt, c->scan_part.p, active);
    k_scan_down<<<nt, DPMD_BLOCK, 0, st>>>(C, c->cell_count.p, c->scan_part.p, c->cell_start.p,
                                          c->cell_wptr.p, c->cell_sys.p, c->flag.p, active);
    k_scatter<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(o, ca, c->order.p, active);
    k_cell_sort<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(ca, c->order.p, active);
    k_gather<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(n, o.sid, c->flag.p, c->order.p, c->order_inv.p, pa, active);
    const bool rigid = c->kind == DPMD_RIGID_BUMPY;
    k_copy_back<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(n, o.sid, c->flag.p, c->order_inv.p, c->static_index.p, pa,
                                                     rigid ? nullptr : c->lx.p, rigid ? nullptr : c->ly.p, active);
    k_neighbors<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(obj_args(c), ca, nb, active);
    if (rigid)
        k_rigid_reset_disp<<<blocks_for(c->N), DPMD_BLOCK, 0, st>>>(c->N, c->sid.p, c->flag.p, c->poff.p,
                                                                    c->static_index.p, c->vpx.p, c->vpy.p,
                                                                    c->lx.p, c->ly.p, active);
    k_clear_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, c->flag.p, active);
    HIP_OK(hipGetLastError());
}

void clear_flags(dpmd_ctx* c, const int* active) {
    k_clear_flags<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->flag.p, active);
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
