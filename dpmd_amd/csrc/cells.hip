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

// Block-level: cell_start for one tile, from the sum of the counts before it (carry).  Counts
// of systems that are not rebuilt are unchanged since their last rebuild, so their starts
// are recomputed to the same values.
__device__ __forceinline__ void ph_tile_down(const RebuildArgs& r, int t, int* lds, int carry) {
    const int C = r.ca.C;
    const int base = t * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) { v[k] = (base + k < C) ? r.ca.cell_count[base + k] : 0; s += v[k]; }
    int total;
    int run = block_excl_scan_int(s, lds, total) + carry;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int c = base + k;
        if (c < C) {
            r.ca.cell_start[c] = run;
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

// ---- device-wide chain ---------------------------------------------------------------------
#define DPMD_EACH(name, count, body)                                                          \
    __global__ void name(RebuildArgs r, const int* active) {                                  \
        const int i = blockIdx.x * blockDim.x + threadIdx.x;                                  \
        if (i >= (count) || !*active) return;                                                 \
        body(r, i);                                                                           \
    }
DPMD_EACH(k_zero, r.ca.C, ph_zero)
DPMD_EACH(k_sort, r.ca.C, ph_sort)
DPMD_EACH(k_gather, r.o.n, ph_gather)
DPMD_EACH(k_copy_back, r.o.n, ph_copy_back)
DPMD_EACH(k_rigid_reset, r.N, ph_rigid_reset)
#undef DPMD_EACH

// Cell id and the object's rank in its cell.  The objects are still in the cell order of the
// last rebuild, so consecutive lanes mostly share a cell: the first lane of each run of equal
// cells in a wavefront takes the run's ranks with one atomic (its count's old value is the
// run's first rank).  The ranks of a cell are distinct and dense, in no set order; k_sort puts
// each cell in ascending current index.  The rank goes to order_inv, which the gather
// overwrites later.
__global__ void k_hash_rank(RebuildArgs r, const int* active) {
    if (!*active) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int cell = -1;
    if (i < r.o.n) {
        const int s = r.o.sid[i];
        if (r.ca.flag[s]) cell = r.ca.cstart[s] + hash_local(r.o, r.ca, i, s);
    }
    const int prev = __shfl_up(cell, 1, 64);
    const bool head = lane == 0 || prev != cell;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (2ull << lane) - 1ull;       // lanes 0 .. lane (all at lane 63)
    const int first = 63 - __builtin_clzll(heads & upto);       // head of this lane's run
    const unsigned long long after = heads & ~upto;
    const int end = after ? __builtin_ffsll((long long)after) - 1 : 64;
    int base = 0;
    if (head && cell >= 0) base = atomicAdd(&r.ca.cell_count[cell], end - lane);
    base = __shfl(base, first, 64);
    if (cell >= 0) {
        r.ca.cell_id[i] = cell;
        r.order_inv[i] = base + (lane - first);
    }
}

// Scatter by start + rank: no second atomic.  The rank order is arbitrary; k_sort then puts
// each cell in ascending current index.
__global__ void k_scatter_rank(RebuildArgs r, const int* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.o.n || !*active) return;
    if (!r.ca.flag[r.o.sid[i]]) return;
    r.order[r.ca.cell_start[r.ca.cell_id[i]] + r.order_inv[i]] = i;
}

// The per-cell sort (one thread per cell: ph_sort), then the gather of the workgroup's cells'
// objects.  The objects of cells [c0, c0 + 256) are one range of new positions and this
// workgroup alone sorted them, so a workgroup barrier orders the sort before the gather.  Two
// objects per thread and round, every load issued before the first store (ph_gather's order).
__global__ void __launch_bounds__(DPMD_BLOCK) k_sort_gather(RebuildArgs r, const int* active) {
    if (!*active) return;
    const int C = r.ca.C, n = r.o.n;
    const int c0 = blockIdx.x * DPMD_BLOCK;
    if (c0 + (int)threadIdx.x < C) ph_sort(r, c0 + threadIdx.x);
    __syncthreads();
    const int pb = r.ca.cell_start[c0], pe = r.ca.cell_start[min(c0 + DPMD_BLOCK, C)];
    constexpr int MD = 10, MI = 2;
    for (int p0 = pb + threadIdx.x; p0 < pe; p0 += 2 * DPMD_BLOCK) {
        const int p1 = p0 + DPMD_BLOCK;
        const bool a0 = r.ca.flag[r.o.sid[p0]] != 0u;
        const bool a1 = p1 < pe && r.ca.flag[r.o.sid[p1]] != 0u;
        const int o0 = a0 ? r.order[p0] : 0, o1 = a1 ? r.order[p1] : 0;
        double v0[MD], v1[MD];
        int w0[MI], w1[MI];
#pragma unroll
        for (int k = 0; k < MD; ++k) {
            v0[k] = a0 && k < r.pa.nd ? r.pa.d[k][o0] : 0.0;
            v1[k] = a1 && k < r.pa.nd ? r.pa.d[k][o1] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < MI; ++k) {
            w0[k] = a0 && k < r.pa.ni ? r.pa.i[k][o0] : 0;
            w1[k] = a1 && k < r.pa.ni ? r.pa.i[k][o1] : 0;
        }
        if (a0) {
#pragma unroll
            for (int k = 0; k < MD; ++k) if (k < r.pa.nd) r.pa.tmp_d[(size_t)k * n + p0] = v0[k];
#pragma unroll
            for (int k = 0; k < MI; ++k) if (k < r.pa.ni) r.pa.tmp_i[(size_t)k * n + p0] = w0[k];
            r.order_inv[o0] = p0;
        }
        if (a1) {
#pragma unroll
            for (int k = 0; k < MD; ++k) if (k < r.pa.nd) r.pa.tmp_d[(size_t)k * n + p1] = v1[k];
#pragma unroll
            for (int k = 0; k < MI; ++k) if (k < r.pa.ni) r.pa.tmp_i[(size_t)k * n + p1] = w1[k];
            r.order_inv[o1] = p1;
        }
    }
}

// ---- tiled neighbor build (device-wide chain) ----------------------------------------------
// One workgroup per row tile: the cells [cx0, cx1) of one cell row of one system.  The tile's
// objects and the objects of the cells around it (rows cy - 1, cy, cy + 1, cells cx0 - 1 ..
// cx1, periodic) are contiguous object ranges in cell order (at most three per row), read
// from the gathered arrays with coalesced loads into LDS; every stencil candidate is then an
// LDS read.  Objects are visited in the reference's stencil order (rows dy = -1, 0, 1, cells
// dx = -1, 0, 1 within a row, objects of a cell in cell order), as neighbor_scan does.  The
// workgroup also copies its own objects' gathered state back (the gather wrote it to the
// scratch arrays), updates static_index over the same index range and resets the
// displacement reference, so no separate copy-back pass runs.  A tile whose surroundings
// hold more than kTileCap objects reads its candidates from global memory instead.
constexpr int kTileThreads = 256;
// LDS per workgroup is about 30 KB, so five workgroups share a CU.
constexpr int kTileCap = 640;               // staged candidates, at most three per thread (slots fit 16 bits)
constexpr int kTileMaxCells = 96;           // cells of one row segment
constexpr int kHalo = kTileMaxCells + 2;    // cells of one halo row
constexpr int kStageRows = 12;              // list entries staged in LDS per object (the rest go out directly)
constexpr int kTileObjects = 170;           // objects per tile the tile widths aim at
static_assert(kTileCap <= 3 * kTileThreads, "three staging loads per thread");

template <int W>
__device__ __forceinline__ int wg_scan_excl(int v, int* lds, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(inc, d, 64);
        if (lane >= d) inc += u;
    }
    if (lane == 63) lds[wv] = inc;
    __syncthreads();
    int pre = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < W / 64; ++k) {
        if (k < wv) pre += lds[k];
        total += lds[k];
    }
    __syncthreads();
    return pre + inc - v;
}

// The copy-back of one own object, in two halves: own_load issues every load (the cell id and,
// for a gathered tile, the gathered state and static_index), own_store writes the state back,
// updates static_index at the same index and resets the displacement reference.  The loads
// of a workgroup's first round go out together with its candidate loads.
struct OwnLoad {
    double v[10];
    int w[2];
    int si, cell;
};

template <bool INPLACE>
__device__ __forceinline__ void own_load(const RebuildArgs& r, int i, int n, const int* gcell, OwnLoad& q,
                                         bool staged) {
    constexpr int MD = 10, MI = 2;
    q.cell = gcell[i];
    if (INPLACE) return;
    // staged: x, y and the radius come from the object's own LDS record instead
#pragma unroll
    for (int k = 0; k < MD; ++k)
        q.v[k] = k < r.pa.nd && !(staged && (k == 0 || k == 1 || k == r.ird)) ? r.pa.tmp_d[(size_t)k * n + i] : 0.0;
#pragma unroll
    for (int k = 0; k < MI; ++k) q.w[k] = k < r.pa.ni ? r.pa.tmp_i[(size_t)k * n + i] : 0;
    q.si = r.static_index[i];
}

__device__ __forceinline__ void own_store(const RebuildArgs& r, int i, const OwnLoad& q, int si_new) {
    constexpr int MD = 10, MI = 2;
#pragma unroll
    for (int k = 0; k < MD; ++k) if (k < r.pa.nd) r.pa.d[k][i] = q.v[k];
#pragma unroll
    for (int k = 0; k < MI; ++k) if (k < r.pa.ni) r.pa.i[k][i] = q.w[k];
    r.static_index[i] = si_new;
    if (r.lx) { r.lx[i] = q.v[0]; r.ly[i] = q.v[1]; }
}

// A staged candidate: position, radius, object index and owning particle in one 32-byte
// record (two 16-byte LDS reads per candidate).
struct __align__(16) TileCand {
    double x, y, r;
    int j, p;
};

struct TileLds {
    TileCand hc[kTileCap];                             // staged candidates, in halo-cell order
    int hg0[3 * kHalo], hbs[3 * kHalo], hcnt[3 * kHalo];   // per halo cell: first object, LDS slot, count
    unsigned short nbr_l[kStageRows * kTileThreads];   // the first list entries, as slot numbers
    int scan_w[kTileThreads / 64];
};

// The neighbor lists of one tile.  INPLACE: the objects are already in their new order in the
// live arrays (the per-system engine permuted them); otherwise they are in the gather's
// scratch arrays, and the workgroup copies its objects back.
//
// Halo cells are numbered h = q * w + u (stencil row q = 0, 1, 2 for dy = -1, 0, 1; u = 0 ..
// w - 1 for cells cx0 - 1 .. cx1, periodic); each thread owns a contiguous run of them, so the
// slots follow from one workgroup scan of the cells' counts.  Two global round trips (cell
// starts, then the candidates and the own objects' state, all in one round) precede the
// candidate tests, which read only LDS.  The first kStageRows entries of each object's list are
// staged in its own LDS column as 16-bit slot numbers, then written out k-th entry by k-th
// entry, so the k-th ids of a wave's objects are one coalesced store; later entries are
// written as they are found (entries past an object's count are not written: every reader
// stops at the count).
template <bool RIGID, bool INPLACE>
__device__ void tile_neighbors(const RebuildArgs& r, int tile, TileLds& L) {
    const int4* tl = (const int4*)(r.tiles + kTileRec * tile);
    const int4 ta = tl[0], tb = tl[1];
    const int s = ta.x, cy = ta.y, cx0 = ta.z, cx1 = ta.w;
    const int nx = tb.x, ny = tb.y, cs = tb.z;
    const int t = threadIdx.x;
    const int n = r.o.n, K = r.nb.K;
    // the box and cutoff are needed only by the tests; their loads go out now
    const double Lx = r.ca.Lx[s], Ly = r.ca.Ly[s], iLx = r.ca.iLx[s], iLy = r.ca.iLy[s], cut0 = r.ca.cutoff[s];
    const int* cst = r.ca.cell_start;
    const double* gx = INPLACE ? r.o.x : r.pa.tmp_d;
    const double* gy = INPLACE ? r.o.y : r.pa.tmp_d + (size_t)n;
    const double* gr = INPLACE ? r.o.rad : r.pa.tmp_d + (size_t)r.ird * n;
    const int* gcell = INPLACE ? r.ca.cell_id : r.pa.tmp_i;
    const int* gpid = RIGID ? (INPLACE ? r.o.pid : r.pa.tmp_i + (size_t)n) : nullptr;
    const int w = cx1 - cx0 + 2;                        // halo cells per row
    const int H = 3 * w;
    const int per = (H + kTileThreads - 1) / kTileThreads;
    const int h0 = min(t * per, H), h1 = min(h0 + per, H);
    int sum = 0;
    for (int h = h0; h < h1; ++h) {
        const int q = h / w, u = h - q * w;
        int yy = cy + q - 1; if (yy < 0) yy += ny; else if (yy >= ny) yy -= ny;
        int x = cx0 - 1 + u; if (x < 0) x += nx; else if (x >= nx) x -= nx;
        const int cell = cs + yy * nx + x;
        const int g0 = cst[cell], c = cst[cell + 1] - g0;
        L.hg0[h] = g0; L.hcnt[h] = c;
        sum += c;
    }
    int total;
    int base = wg_scan_excl<kTileThreads>(sum, L.scan_w, total);
    const bool staged = total <= kTileCap;
    for (int h = h0; h < h1; ++h) {
        L.hbs[h] = staged ? base : L.hg0[h];            // unstaged: candidates read from global memory
        base += L.hcnt[h];
    }
    __syncthreads();
    // own objects: the cells cx0 .. cx1 - 1 of row cy are the halo cells w + 1 .. 2w - 2
    const int o0 = L.hg0[w + 1], o1 = L.hg0[2 * w - 2] + L.hcnt[2 * w - 2];
    const int own_slot0 = L.hbs[w + 1];
    OwnLoad q;
    if (o0 + t < o1) own_load<INPLACE>(r, o0 + t, n, gcell, q, staged);
    if (staged) {
        // The halo cells u = [0, b1), [b1, b2), [b2, w) of a row are consecutive cells of the
        // grid (the row wraps in x at most twice), so each of these pieces is one object range
        // whose slots are consecutive too: slot m of the piece starting at slot sl holds object
        // g + (m - sl).  Three slots per thread, all loads issued before the first LDS store.
        const int b1 = min(cx0 == 0 ? 1 : nx - cx0 + 1, w), b2 = min(b1 + nx, w);
        int sl[9], gd[9];
#pragma unroll
        for (int p = 0; p < 9; ++p) {
            const int rq = p / 3, k = p % 3;
            const int ua = k == 0 ? 0 : (k == 1 ? b1 : b2), ub = k == 0 ? b1 : (k == 1 ? b2 : w);
            sl[p] = ua < ub ? L.hbs[rq * w + ua] : INT_MAX;
            gd[p] = ua < ub ? L.hg0[rq * w + ua] - sl[p] : 0;
        }
        TileCand cq[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int m = t + k * kTileThreads;
            int d = gd[0];
#pragma unroll
            for (int p = 1; p < 9; ++p) d = sl[p] <= m ? gd[p] : d;
            const int j = m + d;
            if (m < total) {
                cq[k].x = gx[j]; cq[k].y = gy[j]; cq[k].r = gr[j]; cq[k].j = j;
                cq[k].p = RIGID ? gpid[j] : -2;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (t + k * kTileThreads < total) L.hc[t + k * kTileThreads] = cq[k];
    }
    __syncthreads();
    for (int ob = o0; ob < o1; ob += kTileThreads) {
        const int i = ob + t;
        const bool mine = i < o1;
        int cnt = 0;
        if (mine) {
            const int si_new = INPLACE ? 0 : r.order_inv[q.si];   // used at the end: latency hidden
            double xi, yi, ri;
            int pi = -1;
            if (staged) {
                const TileCand& me = L.hc[own_slot0 + (i - o0)];
                xi = me.x; yi = me.y; ri = me.r;
                if (RIGID) pi = me.p;
                q.v[0] = xi; q.v[1] = yi;
#pragma unroll
                for (int k = 2; k < 10; ++k) if (k == r.ird) q.v[k] = ri;
            } else {
                xi = gx[i]; yi = gy[i]; ri = gr[i];
                if (RIGID) pi = gpid[i];
            }
            const int ui = q.cell - cs - cy * nx - cx0 + 1;   // halo position of the own cell
            int* col = r.nb.nbr + i;
            auto test = [&](int m, int j, double xj, double yj, double rj, int pj) {
                if (j == i) return;
                if (RIGID && pj == pi) return;
                const double ddx = min_image(xi - xj, Lx, iLx);
                const double ddy = min_image(yi - yj, Ly, iLy);
                const double r2 = ddx * ddx + ddy * ddy;
                const double cut = ri + rj + cut0;
                if (r2 < cut * cut) {
                    if (cnt < K) {
                        if (staged && cnt < kStageRows) L.nbr_l[cnt * kTileThreads + t] = (unsigned short)m;
                        else col[(size_t)cnt * n] = j;
                    }
                    ++cnt;
                }
            };
            if (staged) {
                // the three cells of a stencil row are consecutive halo cells, so their slots form
                // one range in visiting order; candidates are read four at a time, every read of a
                // group issued before the first test
                constexpr int G = 4;
                for (int rq = 0; rq < 3; ++rq) {
                    const int h = rq * w + ui;
                    const int lo = L.hbs[h - 1], hi = L.hbs[h + 1] + L.hcnt[h + 1];
                    for (int m0 = lo; m0 < hi; m0 += G) {
                        TileCand cg[G];
#pragma unroll
                        for (int u = 0; u < G; ++u) cg[u] = L.hc[min(m0 + u, hi - 1)];
#pragma unroll
                        for (int u = 0; u < G; ++u)
                            if (m0 + u < hi) test(m0 + u, cg[u].j, cg[u].x, cg[u].y, cg[u].r, cg[u].p);
                    }
                }
            } else {
                for (int rq = 0; rq < 3; ++rq)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int h = rq * w + ui + dx;
                        const int lo = L.hbs[h], hi = lo + L.hcnt[h];
                        for (int m = lo; m < hi; ++m) test(m, m, gx[m], gy[m], gr[m], RIGID ? gpid[m] : -2);
                    }
            }
            if (cnt > K) atomicMax(r.nb.overflow, cnt);
            r.nb.nn[i] = cnt < K ? cnt : K;
            if (!INPLACE) own_store(r, i, q, si_new);
        }
        if (staged && mine) {   // the k-th ids of a wave's objects go out together
            const int have = min(min(cnt, K), kStageRows);
            for (int k = 0; k < have; ++k) r.nb.nbr[(size_t)k * n + i] = L.hc[L.nbr_l[k * kTileThreads + t]].j;
        }
        if (i + kTileThreads < o1) own_load<INPLACE>(r, i + kTileThreads, n, gcell, q, staged);   // next round
    }
}

// Device-wide chain: one workgroup per tile; tiles of systems not rebuilt leave at once.
template <bool RIGID>
__global__ void __launch_bounds__(kTileThreads) k_neighbors_tiled(RebuildArgs r, const int* active) {
    __shared__ TileLds L;
    if (!*active) return;
    // each XCD takes one contiguous range of tiles, so the rows above and below a tile are
    // mostly staged by the same XCD, from its L2
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    if (!r.ca.flag[r.tiles[kTileRec * tile]]) return;
    tile_neighbors<RIGID, false>(r, tile, L);
}

// Per-system engine: the tiles that the sort kernel listed (every tile of each system it
// rebuilt), spread over a fixed grid: workgroup b takes list entries b, b + gridDim.x, ...
constexpr int kListGrid = 1024;  // workgroups of the listed neighbor build

template <bool RIGID>
__global__ void __launch_bounds__(kTileThreads) k_neighbors_listed(RebuildArgs r, const int* active) {
    __shared__ TileLds L;
    if (!*active) return;
    const int T = *r.list_count;
    int tile = r.list[blockIdx.x];   // in bounds (the grid is at most the tile count); used if < T
    for (int v = blockIdx.x; v < T; v += gridDim.x) {
        tile_neighbors<RIGID, true>(r, tile, L);
        __syncthreads();
        if (v + (int)gridDim.x < T) tile = r.list[v + gridDim.x];
    }
}

// ---- per-system workgroup rebuild ------------------------------------------------------------
// One workgroup per system (grid = S); systems whose flag is clear leave at once, so a step
// with nothing to rebuild costs near-empty launches and a rebuild costs work on the flagged
// systems only.  The workgroup sorts its system through LDS: cell counts with LDS atomics,
// the scan of the system's cells, an LDS scatter followed by a per-cell insertion sort
// (ascending current index: the same order as the device-wide chain), the permutation of
// every live array (through registers, or LDS for the largest systems), static_index and the
// displacement reference; it then lists its system's cell tiles for k_neighbors_listed, which
// builds their lists on many workgroups (one workgroup per system spent most of its time there).
constexpr int kSysLdsBytes = 156 * 1024;
constexpr int kSysThreads = 1024;
constexpr int kSysRegArrays = 4;   // arrays permuted per round through registers
constexpr int kSysRegPos = 5;      // positions per thread of that path (systems up to 5 * 1024 objects)

template <bool RIGID>
__global__ void __launch_bounds__(kSysThreads) k_rebuild_sys(RebuildArgs r, const int* active) {
    constexpr int W = kSysThreads;
    __shared__ __align__(16) unsigned char lds_raw[kSysLdsBytes];
    if (!*active) return;
    const int s = blockIdx.x;
    if (!r.ca.flag[s]) return;
    const int t = threadIdx.x;
    const int b = r.o.off[s], e = r.o.off[s + 1], n = e - b;
    const int cs = r.ca.cstart[s], nc = r.ca.cstart[s + 1] - cs;
    int* scratch = (int*)lds_raw;                      // 16 words
    int* ord = scratch + 16;                           // [n] new local position -> old local index
    int* cid = ord + n;                                // [n] old object's local cell, then the inverse
    unsigned char* rest = lds_raw + 64 + 8 * (size_t)n;
    int* cnt = (int*)rest;                             // [nc]
    double* stage = (double*)rest;                     // [n], after the sort
    for (int c = t; c < nc; c += W) cnt[c] = 0;
    __syncthreads();
    for (int q = t; q < n; q += W) {
        const int c = hash_local(r.o, r.ca, b + q, s);
        cid[q] = c;
        atomicAdd(&cnt[c], 1);
    }
    __syncthreads();
    {
        const int per = (nc + W - 1) / W;
        const int c0 = min(t * per, nc), c1 = min(c0 + per, nc);
        int sum = 0;
        for (int c = c0; c < c1; ++c) sum += cnt[c];
        int total;
        int run = wg_scan_excl<W>(sum, scratch, total);
        for (int c = c0; c < c1; ++c) {
            const int k = cnt[c];
            r.ca.cell_start[cs + c] = b + run;
            r.ca.cell_count[cs + c] = k;
            cnt[c] = run;
            run += k;
        }
        if (s == r.S - 1 && t == 0) r.ca.cell_start[cs + nc] = e;
    }
    __syncthreads();
    for (int q = t; q < n; q += W) ord[atomicAdd(&cnt[cid[q]], 1)] = q;
    __syncthreads();
    for (int c = t; c < nc; c += W) {                  // cnt[c] is now the end of cell c
        const int lo = c == 0 ? 0 : cnt[c - 1], hi = cnt[c];
        for (int k = lo + 1; k < hi; ++k) {
            const int key = ord[k];
            int m = k - 1;
            while (m >= lo && ord[m] > key) { ord[m + 1] = ord[m]; --m; }
            ord[m + 1] = key;
        }
    }
    __syncthreads();
    int* cell_id = r.pa.i[0];
    for (int p = t; p < n; p += W) cell_id[b + p] = cs + cid[ord[p]];
    __syncthreads();
    for (int p = t; p < n; p += W) cid[ord[p]] = p;   // inverse permutation
    __syncthreads();
    for (int q = t; q < n; q += W) r.static_index[b + q] = b + cid[r.static_index[b + q] - b];
    if (n <= kSysRegPos * W) {
        // kSysRegArrays arrays at a time: every thread gathers its positions' values of the
        // group into registers, a barrier, then the stores (the permutation is in place)
        for (int k0 = 0; k0 < r.pa.nd; k0 += kSysRegArrays) {
            const bool last = k0 + kSysRegArrays >= r.pa.nd;
            double v[kSysRegArrays][kSysRegPos];
            int w[kSysRegPos];
#pragma unroll
            for (int u = 0; u < kSysRegPos; ++u) {
                const int p = t + u * W;
                const int o = p < n ? b + ord[p] : b;
#pragma unroll
                for (int k = 0; k < kSysRegArrays; ++k)
                    v[k][u] = (p < n && k0 + k < r.pa.nd) ? r.pa.d[k0 + k][o] : 0.0;
                w[u] = (last && p < n && r.pa.ni > 1) ? r.pa.i[1][o] : 0;
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kSysRegPos; ++u) {
                const int p = t + u * W;
                if (p >= n) continue;
#pragma unroll
                for (int k = 0; k < kSysRegArrays; ++k) {
                    if (k0 + k >= r.pa.nd) continue;
                    r.pa.d[k0 + k][b + p] = v[k][u];
                    if (k0 + k == 0 && r.lx) r.lx[b + p] = v[k][u];
                    if (k0 + k == 1 && r.ly) r.ly[b + p] = v[k][u];
                }
                if (last && r.pa.ni > 1) r.pa.i[1][b + p] = w[u];
            }
            __syncthreads();
        }
    } else {
        for (int k = 0; k < r.pa.nd; ++k) {
            double* d = r.pa.d[k];
            for (int p = t; p < n; p += W) stage[p] = d[b + ord[p]];
            __syncthreads();
            double* l = (k == 0) ? r.lx : (k == 1 ? r.ly : nullptr);
            for (int p = t; p < n; p += W) {
                d[b + p] = stage[p];
                if (l) l[b + p] = stage[p];
            }
            __syncthreads();
        }
        for (int k = 1; k < r.pa.ni; ++k) {
            int* d = r.pa.i[k];
            int* st = (int*)stage;
            for (int p = t; p < n; p += W) st[p] = d[b + ord[p]];
            __syncthreads();
            for (int p = t; p < n; p += W) d[b + p] = st[p];
            __syncthreads();
        }
    }
    if (RIGID) {
        __syncthreads();
        for (int i = r.psys_off[s] + t; i < r.psys_off[s + 1]; i += W) ph_rigid_reset(r, i);
    }
    // the neighbor kernel takes this system's tiles from the list
    __shared__ int list_base;
    const int tb = r.sys_tile[s], nts = r.sys_tile[s + 1] - tb;
    if (t == 0) list_base = atomicAdd(r.list_count, nts);
    __syncthreads();
    for (int k = t; k < nts; k += W) r.list[list_base + k] = tb + k;
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
__global__ void k_tile_down(RebuildArgs r, const int* active) {   // after k_top_scan
    __shared__ int lds[4];
    if (!*active) return;
    ph_tile_down(r, blockIdx.x, lds, r.tile_sum[blockIdx.x]);
}
// Without k_top_scan: each workgroup adds up the tile sums before its own (at most
// kSelfScanTiles of them).
constexpr int kSelfScanTiles = 1024;
__global__ void k_tile_down_self(RebuildArgs r, const int* active) {
    __shared__ int lds[4];
    if (!*active) return;
    int v = 0;
    for (int u = threadIdx.x; u < (int)blockIdx.x; u += DPMD_BLOCK) v += r.tile_sum[u];
    int carry;
    (void)block_excl_scan_int(v, lds, carry);
    ph_tile_down(r, blockIdx.x, lds, carry);
}

__global__ void k_set_flags(int S, unsigned* flag, const int* active) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S || !*active) return;
    flag[s] = 1u;
}

// End of a rebuild: clear every flag and the any-flag word (and the per-system engine's list).
// Unconditional, because the device-gated chain's gate is the any-flag word itself, which this
// kernel clears; a flag is only ever set together with that word, or for a host-decided
// rebuild, so clearing a flag that was not rebuilt cannot happen.  count: a device-gated
// rebuild, counted in RUN_REBUILDS when it ran.
__global__ void k_clear_flags(int S, unsigned* flag, int* run, int* list_count, int count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    flag[s] = 0u;
    if (s == 0) {
        if (count && run[RUN_ANY]) run[RUN_REBUILDS] += 1;
        run[RUN_ANY] = 0;
        if (list_count) *list_count = 0;
    }
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

// Arrays a rebuild permutes.  The live ones come first; inside a step (between the drift and
// the force evaluation) the forces and energies are about to be overwritten, so only the
// live arrays move: disks x, y, vx, vy, rad, mass, area (then fx, fy, pe); vertices vpx, vpy,
// vmass, vrad (then vfx, vfy, vpe).  pa.i: cell_id, then the owning particle of a vertex.
PermArgs perm_args(dpmd_ctx* c, bool in_step) {
    PermArgs pa{};
    if (c->kind == DPMD_RIGID_BUMPY) {
        double* d[] = {c->vpx.p, c->vpy.p, c->vmass.p, c->vrad.p, c->vfx.p, c->vfy.p, c->vpe.p};
        pa.nd = in_step ? 4 : 7;
        for (int k = 0; k < 7; ++k) pa.d[k] = d[k];
        pa.i[0] = c->cell_id.p; pa.i[1] = c->vpid.p; pa.ni = 2;
    } else {
        double* d[] = {c->x.p, c->y.p, c->vx.p, c->vy.p, c->rad.p, c->mass.p, c->area.p, c->fx.p, c->fy.p, c->pe.p};
        pa.nd = in_step ? 7 : 10;
        for (int k = 0; k < 10; ++k) pa.d[k] = d[k];
        pa.i[0] = c->cell_id.p; pa.ni = 1;
    }
    pa.tmp_d = c->tmp_d.p; pa.tmp_i = c->tmp_i.p;
    return pa;
}

RebuildArgs make_rebuild_args(dpmd_ctx* c, bool in_step) {
    RebuildArgs r;
    r.o = obj_args(c); r.ca = cell_args(c); r.pa = perm_args(c, in_step); r.nb = nbr_args(c);
    r.order = c->order.p; r.order_inv = c->order_inv.p; r.static_index = c->static_index.p;
    const bool rigid = c->kind == DPMD_RIGID_BUMPY;
    r.lx = rigid ? nullptr : c->lx.p; r.ly = rigid ? nullptr : c->ly.p;
    r.N = c->N; r.psid = c->sid.p; r.poff = c->poff.p;
    r.plx = rigid ? c->lx.p : nullptr; r.ply = rigid ? c->ly.p : nullptr;
    r.psys_off = c->off.p;
    r.tile_sum = c->scan_part.p; r.nt = (c->C + SCAN_TILE - 1) / SCAN_TILE; r.S = c->S;
    r.any = c->run.p + RUN_ANY;
    r.ird = rigid ? 3 : 4;
    r.tiles = c->tiles.p; r.ntiles = c->ntiles; r.sys_tile = c->sys_tile.p;
    r.list = c->sys_list.p; r.list_count = c->sys_list_n.p;
    return r;
}

// The device-wide chain: cell ids and ranks, the scan of every cell, the rank scatter, the
// per-cell sort with the gather of the live arrays, then the tiled neighbor build (which also
// copies the gathered state back).
void rebuild_multi(dpmd_ctx* c, const RebuildArgs& r, const int* active, int count) {
    hipStream_t st = c->st;
    const int n = r.o.n, C = r.ca.C;
    k_zero<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(r, active);
    k_hash_rank<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_tile_sum<<<r.nt, DPMD_BLOCK, 0, st>>>(r, active);
    if (r.nt <= kSelfScanTiles) {
        k_tile_down_self<<<r.nt, DPMD_BLOCK, 0, st>>>(r, active);
    } else {
        k_top_scan<<<1, DPMD_BLOCK, 0, st>>>(r, active);
        k_tile_down<<<r.nt, DPMD_BLOCK, 0, st>>>(r, active);
    }
    k_scatter_rank<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_sort_gather<<<blocks_for(C), DPMD_BLOCK, 0, st>>>(r, active);
    need(r.ntiles > 0, "rebuild: no cell tiles (cells not set up)");
    if (c->kind == DPMD_RIGID_BUMPY) k_neighbors_tiled<true><<<r.ntiles, kTileThreads, 0, st>>>(r, active);
    else k_neighbors_tiled<false><<<r.ntiles, kTileThreads, 0, st>>>(r, active);
    if (r.plx) k_rigid_reset<<<blocks_for(r.N), DPMD_BLOCK, 0, st>>>(r, active);
    k_clear_flags<<<blocks_for(r.S), DPMD_BLOCK, 0, st>>>(r.S, r.ca.flag, c->run.p, nullptr, count);
}

// Whether every system fits the per-system workgroup rebuild's LDS: the cell counts and
// the permutation stage after the two index arrays.
bool sys_rebuild_fits(dpmd_ctx* c) {
    const int64_t need = 64 + 8 * (int64_t)c->max_sys_obj + std::max<int64_t>(4 * (int64_t)c->max_sys_cells,
                                                                                   8 * (int64_t)c->max_sys_obj);
    return c->max_sys_obj > 0 && need <= kSysLdsBytes;
}

bool use_sys_rebuild(dpmd_ctx* c) {
    if (c->rebuild_engine == 2) return false;
    const bool fits = sys_rebuild_fits(c);
    if (c->rebuild_engine == 1)
        need(fits, "DPMD_OPT_REBUILD_ENGINE = 1: a system is too large for the per-system workgroup rebuild");
    return fits;
}

}  // namespace

RebuildArgs rebuild_args(dpmd_ctx* c) { return make_rebuild_args(c, false); }

void build_tiles(dpmd_ctx* c, const std::vector<int>& cdx, const std::vector<int>& cdy) {
    const std::vector<int>& off = c->kind == DPMD_RIGID_BUMPY ? c->h_voff : c->h_off;
    std::vector<int> tl;
    int mo = 0, mc = 0, cs = 0;
    for (int s = 0; s < c->S; cs += std::max(0, cdx[s]) * std::max(0, cdy[s]), ++s) {   // cs: first cell
        const int n = off[s + 1] - off[s], nx = cdx[s], ny = cdy[s];
        mo = std::max(mo, n);
        mc = std::max(mc, nx * ny);
        if (nx <= 0 || ny <= 0) continue;
        // row segments of about kTileObjects objects (at most one per thread of a tile workgroup)
        const double rho = (double)n / ((double)nx * ny);
        int tw = rho > 0.0 ? (int)(kTileObjects / rho) : nx;
        tw = std::max(1, std::min(tw, std::min(nx, kTileMaxCells)));
        const int nseg = (nx + tw - 1) / tw;
        for (int cy = 0; cy < ny; ++cy)
            for (int k = 0; k < nseg; ++k) {
                const int rec[kTileRec] = {s, cy, (int)((int64_t)k * nx / nseg), (int)((int64_t)(k + 1) * nx / nseg),
                                           nx, ny, cs, 0};
                tl.insert(tl.end(), rec, rec + kTileRec);
            }
    }
    c->ntiles = (int)(tl.size() / kTileRec);
    std::vector<int> st(c->S + 1, 0);                   // tiles of system s: [st[s], st[s + 1])
    for (int k = 0; k < c->ntiles; ++k) st[tl[kTileRec * k] + 1]++;
    for (int s = 0; s < c->S; ++s) st[s + 1] += st[s];
    if (tl.empty()) tl.assign(kTileRec, 0);
    c->tiles.upload(tl.data(), (int64_t)tl.size(), c->st);
    c->sys_tile.upload(st.data(), c->S + 1, c->st);
    c->sys_list.resize(std::max(1, c->ntiles));          // the per-system engine's tile list
    if (c->sys_list_n.n == 0) { c->sys_list_n.resize(1); c->sys_list_n.zero(c->st); }
    c->max_sys_obj = mo;
    c->max_sys_cells = mc;
}

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
    RebuildArgs r = make_rebuild_args(c, false);
    k_gather<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_copy_back<<<blocks_for(n), DPMD_BLOCK, 0, st>>>(r, active);
    k_clear_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, r.ca.flag, c->run.p, nullptr, 0);
    HIP_OK(hipGetLastError());
}

namespace {
void rebuild_impl(dpmd_ctx* c, const int* active, bool force_all, bool in_step, int count);
}

// The device-gated chain: its gate is the any-flag word the drift kernels raise, which the
// chain's last kernel clears; a step without flags costs only near-empty launches.
void rebuild_on_device(dpmd_ctx* c) { rebuild_impl(c, c->run.p + RUN_ANY, false, true, 1); }

void rebuild_neighbors(dpmd_ctx* c, const int* active, bool force_all, bool in_step) {
    rebuild_impl(c, active, force_all, in_step, 0);
}

namespace {
void rebuild_impl(dpmd_ctx* c, const int* active, bool force_all, bool in_step, int count) {
    const int S = c->S;
    hipStream_t st = c->st;
    if (force_all) k_set_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, c->flag.p, active);
    RebuildArgs r = make_rebuild_args(c, in_step);
    if (use_sys_rebuild(c)) {
        const int grid = std::max(1, std::min(r.ntiles, kListGrid));
        if (c->kind == DPMD_RIGID_BUMPY) {
            k_rebuild_sys<true><<<S, kSysThreads, 0, st>>>(r, active);
            k_neighbors_listed<true><<<grid, kTileThreads, 0, st>>>(r, active);
        } else {
            k_rebuild_sys<false><<<S, kSysThreads, 0, st>>>(r, active);
            k_neighbors_listed<false><<<grid, kTileThreads, 0, st>>>(r, active);
        }
        k_clear_flags<<<blocks_for(S), DPMD_BLOCK, 0, st>>>(S, r.ca.flag, c->run.p, r.list_count, count);
    } else {
        rebuild_multi(c, r, active, count);
    }
    HIP_OK(hipGetLastError());
}
}  // namespace

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
