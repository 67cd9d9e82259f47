// Rebuild phase bodies shared by the multi-launch rebuild (cells.hip) and the persistent
// per-system FIRE kernel (fire_persist.hip).  Each body handles one element (object or
// cell) and does nothing for objects of systems whose rebuild flag is clear.
//
// Reference: src/kernels/base_particle_kernels.cu:205-259, base_point_particle_kernels.cu:
// 37-157, src/particles/disk.cu:1097-1125.
#pragma once
#include "internal.hpp"

namespace dpmd {

constexpr int kTileRec = 8;

struct RebuildArgs {
    ObjArgs o; CellArgs ca; PermArgs pa; NbrArgs nb;
    int* order; int* order_inv; int* static_index;
    double* lx; double* ly;           // disks: displacement reference, reset per object
    // rigid bodies: per-particle reference = first vertex of each particle
    int N; const int* psid; const int* poff; double* plx; double* ply;
    const int* psys_off;              // per-system particle offsets (S+1)
    int* tile_sum; int nt; int S;
    int* any;                          // the run word "some rebuild flag is set"
    // Positions within the permuted arrays: pa.d[0], pa.d[1] are x, y; pa.d[ird] is the
    // radius; pa.i[0] is cell_id and pa.i[1] (rigid bodies) the owning particle.
    int ird;
    // Row tiles of the cell grids for the tiled neighbor build: kTileRec ints per tile
    // (system, cell row, first cell, one past the last cell of the row segment, the system's
    // cell grid nx, ny and first global cell, padding), so a tile's workgroup needs one load.
    const int* tiles; int ntiles;
    const int* sys_tile;               // tiles of system s: [sys_tile[s], sys_tile[s + 1])
    int* list; int* list_count;        // tiles of the systems the per-system engine rebuilt
};

__device__ __forceinline__ void ph_zero(const RebuildArgs& r, int c) {
    if (r.ca.flag[r.ca.cell_sys[c]]) r.ca.cell_count[c] = 0;
}

// assign_cell_ids_kernel (base_particle_kernels.cu:205-233): the cell of object i within its
// system's grid (cx + cy * nx).
__device__ __forceinline__ int hash_local(const ObjArgs& o, const CellArgs& ca, int i, int s) {
    const double u = wrap01(o.x[i] * ca.iLx[s]);
    const double v = wrap01(o.y[i] * ca.iLy[s]);
    const int nx = ca.cdx[s], ny = ca.cdy[s];
    int cx = (int)floor(u * nx);
    int cy = (int)floor(v * ny);
    cx = cx < nx ? (cx > 0 ? cx : 0) : nx - 1;   // guard: u*nx may round up to nx
    cy = cy < ny ? (cy > 0 ? cy : 0) : ny - 1;
    return cx + cy * nx;
}

// assign_cell_ids_kernel + count_cells_kernel (base_particle_kernels.cu:205-244)
__device__ __forceinline__ void ph_hash(const RebuildArgs& r, int i) {
    const ObjArgs& o = r.o; const CellArgs& ca = r.ca;
    const int s = o.sid[i];
    if (!ca.flag[s]) return;
    const int cell = ca.cstart[s] + hash_local(o, ca, i, s);
    ca.cell_id[i] = cell;
    atomicAdd(&ca.cell_count[cell], 1);
}

__device__ __forceinline__ void ph_scatter(const RebuildArgs& r, int i) {
    if (!r.ca.flag[r.o.sid[i]]) return;
    const int slot = atomicAdd(&r.ca.cell_wptr[r.ca.cell_id[i]], 1);
    r.order[slot] = i;
}

// In-cell order: ascending current index (a stable sort by cell id).
__device__ __forceinline__ void ph_sort(const RebuildArgs& r, int c) {
    if (!r.ca.flag[r.ca.cell_sys[c]]) return;
    const int b = r.ca.cell_start[c], e = r.ca.cell_start[c + 1];
    int* ord = r.order;
    for (int k = b + 1; k < e; ++k) {
        const int key = ord[k];
        int m = k - 1;
        while (m >= b && ord[m] > key) { ord[m + 1] = ord[m]; --m; }
        ord[m + 1] = key;
    }
}

// reorder_particles_impl (disk.cu:1102-1125): gather into scratch by the new order.  Every
// load is issued before the first store (the compiler cannot rule out aliasing, which would
// serialise the arrays).
__device__ __forceinline__ void ph_gather(const RebuildArgs& r, int p) {
    if (!r.ca.flag[r.o.sid[p]]) return;
    const int n = r.o.n;
    const int o = r.order[p];
    constexpr int MD = 10, MI = 2;
    double v[MD];
    int w[MI];
#pragma unroll
    for (int k = 0; k < MD; ++k) v[k] = k < r.pa.nd ? r.pa.d[k][o] : 0.0;
#pragma unroll
    for (int k = 0; k < MI; ++k) w[k] = k < r.pa.ni ? r.pa.i[k][o] : 0;
#pragma unroll
    for (int k = 0; k < MD; ++k) if (k < r.pa.nd) r.pa.tmp_d[(size_t)k * n + p] = v[k];
#pragma unroll
    for (int k = 0; k < MI; ++k) if (k < r.pa.ni) r.pa.tmp_i[(size_t)k * n + p] = w[k];
    r.order_inv[o] = p;
}

// Copy back, update static_index (update_static_index_kernel, base_particle_kernels.cu:
// 172-180) and reset the displacement reference (reset_displacements_impl, disk.cu:1097).
__device__ __forceinline__ void ph_copy_back(const RebuildArgs& r, int p) {
    if (!r.ca.flag[r.o.sid[p]]) return;
    const int n = r.o.n;
    constexpr int MD = 10, MI = 2;
    double v[MD];
    int w[MI];
#pragma unroll
    for (int k = 0; k < MD; ++k) v[k] = k < r.pa.nd ? r.pa.tmp_d[(size_t)k * n + p] : 0.0;
#pragma unroll
    for (int k = 0; k < MI; ++k) w[k] = k < r.pa.ni ? r.pa.tmp_i[(size_t)k * n + p] : 0;
    const int si = r.order_inv[r.static_index[p]];
#pragma unroll
    for (int k = 0; k < MD; ++k) if (k < r.pa.nd) r.pa.d[k][p] = v[k];
#pragma unroll
    for (int k = 0; k < MI; ++k) if (k < r.pa.ni) r.pa.i[k][p] = w[k];
    r.static_index[p] = si;
    if (r.lx) { r.lx[p] = v[0]; r.ly[p] = v[1]; }
}

// count_cell_neighbors_kernel + fill_neighbors_cell_kernel (base_point_particle_kernels.cu:
// 37-157): object i's neighbors over the periodic 3x3 stencil of its system's cell grid, in
// the reference's order (stencil rows dy = -1, 0, 1; in a row cells dx = -1, 0, 1; objects
// of a cell in cell order).  emit(k, j) receives the k-th neighbor j; returns the count.
// A stencil row is one contiguous object range unless it wraps in x (then two), and the
// candidates of a range are loaded four at a time before they are tested in order.
template <class Emit>
__device__ __forceinline__ int neighbor_scan(const RebuildArgs& r, int i, int s, Emit&& emit) {
    const ObjArgs& o = r.o; const CellArgs& ca = r.ca;
    const int nx = ca.cdx[s], ny = ca.cdy[s], cs = ca.cstart[s];
    const int loc = ca.cell_id[i] - cs;
    const int cxi = loc % nx, cyi = loc / nx;
    const double Lx = ca.Lx[s], Ly = ca.Ly[s], iLx = ca.iLx[s], iLy = ca.iLy[s], cut0 = ca.cutoff[s];
    const double xi = o.x[i], yi = o.y[i], ri = o.rad[i];
    const int pi = o.pid ? o.pid[i] : -1;
    // cell ranges [c0, c1) of the stencil, in order
    int c0[6], c1[6];
    int nr = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        int yy = cyi + dy; if (yy < 0) yy += ny; else if (yy >= ny) yy -= ny;
        const int row = cs + yy * nx;
        if (cxi == 0) {               // cells nx-1 | 0, 1
            c0[nr] = row + nx - 1; c1[nr] = row + nx; ++nr;
            c0[nr] = row; c1[nr] = row + 2; ++nr;
        } else if (cxi == nx - 1) {   // cells nx-2, nx-1 | 0
            c0[nr] = row + nx - 2; c1[nr] = row + nx; ++nr;
            c0[nr] = row; c1[nr] = row + 1; ++nr;
        } else {                      // cells cx-1 .. cx+1
            c0[nr] = row + cxi - 1; c1[nr] = row + cxi + 2; ++nr;
        }
    }
    int ob[6], oe[6];
    for (int q = 0; q < nr; ++q) { ob[q] = ca.cell_start[c0[q]]; oe[q] = ca.cell_start[c1[q]]; }
    int cnt = 0;
    constexpr int G = 4;
    for (int q = 0; q < nr; ++q) {
        const int b = ob[q], e = oe[q];
        for (int j0 = b; j0 < e; j0 += G) {
            int jj[G];
            double xj[G], yj[G], rj[G];
            int pj[G];
#pragma unroll
            for (int u = 0; u < G; ++u) jj[u] = j0 + u < e ? j0 + u : j0;
#pragma unroll
            for (int u = 0; u < G; ++u) {
                xj[u] = o.x[jj[u]]; yj[u] = o.y[jj[u]]; rj[u] = o.rad[jj[u]];
                pj[u] = o.pid ? o.pid[jj[u]] : -2;
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
                const int j = j0 + u;
                if (j >= e) break;
                if (j == i) continue;
                if (o.pid && pj[u] == pi) continue;
                const double ddx = min_image(xi - xj[u], Lx, iLx);
                const double ddy = min_image(yi - yj[u], Ly, iLy);
                const double r2 = ddx * ddx + ddy * ddy;
                const double cut = ri + rj[u] + cut0;
                if (r2 < cut * cut) {
                    emit(cnt, j);
                    ++cnt;
                }
            }
        }
    }
    return cnt;
}

// The ELL list of object i written directly (objects of systems not rebuilt keep theirs).
__device__ __forceinline__ void ph_neighbors(const RebuildArgs& r, int i) {
    const NbrArgs& nb = r.nb;
    const int s = r.o.sid[i];
    if (!r.ca.flag[s]) return;
    const int n = r.o.n, K = nb.K;
    int* col = nb.nbr + i;
    const int cnt = neighbor_scan(r, i, s, [&](int k, int j) {
        if (k < K) col[(size_t)k * n] = j;
    });
    if (cnt > K) atomicMax(nb.overflow, cnt);
    nb.nn[i] = cnt < K ? cnt : K;
}

RebuildArgs rebuild_args(dpmd_ctx* c);

}  // namespace dpmd
