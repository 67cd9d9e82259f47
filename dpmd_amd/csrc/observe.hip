// Disk contact observables over the Verlet lists: contact counts and overlaps per particle,
// contacts per system, and the per-neighbor-entry pair distance, pair force and Hessian
// blocks.
//
// Reference: src/particles/disk.cu:434-700 (compute_contacts_kernel, compute_pair_dist_kernel,
// compute_particle_pair_forces_kernel, compute_hessian_kernel, compute_overlaps_kernel) and
// :962-1074 (their host sides); base_particle.hpp:649-702.  As there, the pair geometry uses
// the minimum image in every box (walled boxes included) and the walls contribute nothing.
// Outputs indexed by neighbor entry follow the CSR order of dpmd_get_neighbors_csr (particles
// in current order, each particle's neighbors in list order); particle ids in pair outputs are
// canonical.  (The reference's pair-distance and Hessian host code hands static_index where
// its kernels expect the inverse map, which yields canonical ids only while the two coincide;
// here every pair output carries canonical ids.)
#include "internal.hpp"

namespace dpmd {
namespace {

struct ObsArgs {
    const double* x; const double* y; const double* rad; const int* sid;
    const double* Lx; const double* Ly; const double* iLx; const double* iLy; const double* e;
    const int* nn; const int* nbr; int n;
};

ObsArgs obs_args(dpmd_ctx* c) {
    ObsArgs a;
    a.x = c->x.p; a.y = c->y.p; a.rad = c->rad.p; a.sid = c->sid.p;
    a.Lx = c->Lx.p; a.Ly = c->Ly.p; a.iLx = c->iLx.p; a.iLy = c->iLy.p; a.e = c->e.p;
    a.nn = c->nn.p; a.nbr = c->nbr.p; a.n = c->N;
    return a;
}

// compute_contacts_kernel + compute_overlaps_kernel (disk.cu:434-471,663-700)
__global__ void k_contacts(ObsArgs a, int* contacts, double* overlap, double* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xi = a.x[i], yi = a.y[i], ri = a.rad[i];
    const int cnt = a.nn[i];
    int nc = 0;
    double ov = 0.0, cc = 0.0;
    for (int k = 0; k < cnt; ++k) {
        const int j = a.nbr[(size_t)k * a.n + i];
        const double dx = min_image(xi - a.x[j], Lx, iLx), dy = min_image(yi - a.y[j], Ly, iLy);
        const double r2 = dx * dx + dy * dy;
        const double radsum = ri + a.rad[j];
        if (r2 >= radsum * radsum) continue;
        ++nc;
        ov += radsum - sqrt(r2);
        cc += 1.0;
    }
    contacts[i] = nc;
    overlap[i] = ov;
    count[i] = cc;
}

// Per-system integer sum (compute_n_contacts_total, base_particle.hpp:653-659): one block
// per system; integer addition makes the order irrelevant.
__global__ void k_int_sys_sum(const int* off, const int* v, int* out) {
    __shared__ int lds[DPMD_BLOCK];
    const int s = blockIdx.x;
    int acc = 0;
    for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x) acc += v[i];
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = lds[0];
}

__global__ void k_invert(int n, const int* perm, int* inv) {   // inv[perm[c]] = c
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) inv[perm[c]] = c;
}

struct PairOut {
    const int* start; const int* canon;
    int* pi; int* pj; double* dist; double* fx; double* fy; double* h[8];
};

// compute_pair_dist_kernel, compute_particle_pair_forces_kernel and compute_hessian_kernel
// (disk.cu:473-661) in one pass.  The Hessian terms are the reference's expressions with its
// unit derivative factors substituted (each substitution is exact): for contact k of i with
// j, d = (dx, dy) = x_i - x_j (minimum image), c = e, t = -e delta,
//   H_ia,ib += c d_a d_b / r^2 + t (delta_ab / r - d_a d_b / (r^2 r)),
//   H_ia,jb += c (-d_a d_b) / r^2 + t (-delta_ab / r - (-d_a d_b) / (r^2 r)).
__global__ void k_pair_obs(ObsArgs a, PairOut o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double e = a.e[s];
    const double Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xi = a.x[i], yi = a.y[i], ri = a.rad[i];
    const int cnt = a.nn[i];
    const int k0 = o.start[i];
    const int ci = o.canon[i];
    for (int q = 0; q < cnt; ++q) {
        const int k = k0 + q;
        const int j = a.nbr[(size_t)q * a.n + i];
        const double dx = min_image(xi - a.x[j], Lx, iLx), dy = min_image(yi - a.y[j], Ly, iLy);
        const double r2 = dx * dx + dy * dy;
        if (o.pi) { o.pi[k] = ci; o.pj[k] = o.canon[j]; }
        if (o.dist) o.dist[k] = sqrt(r2);
        const double radsum = ri + a.rad[j];
        double fxk = 0.0, fyk = 0.0;
        double h[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (r2 < radsum * radsum) {
            const double r = sqrt(r2);
            const double inv = 1.0 / r;
            const double nx = dx * inv, ny = dy * inv;
            const double delta = radsum - r;
            const double fmag = e * delta;
            // the reference adds each term into zero-filled storage: 0.0 + v
            fxk = 0.0 + fmag * nx;
            fyk = 0.0 + fmag * ny;
            if (o.h[0]) {
                const double t = -e * delta, cc = e;
                const double dd[2] = {dx, dy};
                for (int p = 0; p < 2; ++p)
                    for (int b = 0; b < 2; ++b) {
                        const double dib = p == b ? 1.0 : 0.0;    // d(q_ib)/d(x_ia)
                        const double djb = p == b ? -1.0 : 0.0;   // d(q_jb)/d(x_ia)
                        const double qq = dd[p] * dd[b];
                        const double qj = dd[p] * -dd[b];
                        h[(p * 2 + b) * 2 + 0] = 0.0 + (0.0 + (cc * qq / r2 + t * (dib / r - qq / (r2 * r))));
                        h[(p * 2 + b) * 2 + 1] = 0.0 + (0.0 + (cc * qj / r2 + t * (djb / r - qj / (r2 * r))));
                    }
            }
        }
        if (o.fx) { o.fx[k] = fxk; o.fy[k] = fyk; }
        if (o.h[0])
            for (int m = 0; m < 8; ++m) o.h[m][k] = h[m];
    }
}

// ---------------------------------------------------------------------------------------
// Rigid bumpy polygons.  Particles keep canonical order; vertices are in current order and
// particle i's vertices are static_index[particle_offset[i] .. particle_offset[i+1]).
// ---------------------------------------------------------------------------------------
struct RigidObs {
    const double* x; const double* y;                 // centres (canonical)
    const double* vx; const double* vy; const double* vr; const int* vpid;   // vertices (current)
    const int* vstatic; const int* poff; const int* sid;
    const double* Lx; const double* Ly; const double* iLx; const double* iLy; const double* e;
    const int* nn; const int* nbr; int n; int nv;
};

RigidObs rigid_obs(dpmd_ctx* c) {
    RigidObs a;
    a.x = c->x.p; a.y = c->y.p;
    a.vx = c->vpx.p; a.vy = c->vpy.p; a.vr = c->vrad.p; a.vpid = c->vpid.p;
    a.vstatic = c->static_index.p; a.poff = c->poff.p; a.sid = c->sid.p;
    a.Lx = c->Lx.p; a.Ly = c->Ly.p; a.iLx = c->iLx.p; a.iLy = c->iLy.p; a.e = c->e.p;
    a.nn = c->nn.p; a.nbr = c->nbr.p; a.n = c->N; a.nv = c->Nv;
    return a;
}

// build_particle_neighbors (base_poly_particle.hpp:299-375): the particles whose vertices
// appear in the vertex lists of particle i's vertices, ascending and without repeats (what
// sorting and deduplicating the (i, j) keys yields).  FILL = false counts, FILL = true
// writes at start[i].  More than kMaxParticleNeighbors sets *overflow.
constexpr int kMaxParticleNeighbors = 64;
template <bool FILL>
__global__ void k_rigid_pn(RigidObs a, int* count, const int* start, int* ids, int* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    int list[kMaxParticleNeighbors];
    int m = 0;
    bool full = false;
    for (int c = a.poff[i]; c < a.poff[i + 1]; ++c) {
        const int u = a.vstatic[c];
        const int cnt = a.nn[u];
        for (int k = 0; k < cnt; ++k) {
            const int p = a.vpid[a.nbr[(size_t)k * a.nv + u]];
            int pos = 0;
            while (pos < m && list[pos] < p) ++pos;
            if (pos < m && list[pos] == p) continue;
            if (m == kMaxParticleNeighbors) { full = true; continue; }
            for (int q = m; q > pos; --q) list[q] = list[q - 1];
            list[pos] = p;
            ++m;
        }
    }
    if (full) atomicOr(overflow, 1);
    if (FILL) {
        const int b = start[i];
        for (int q = 0; q < m; ++q) ids[b + q] = list[q];
    } else {
        count[i] = m;
    }
}

struct RigidPairOut {
    const int* start; const int* ids;
    int* pi; int* pj; double* dist; double* fx; double* fy; int* vci; int* vcj; double* friction;
    int* contacts;
};

// count_particle_contacts_kernel, compute_pair_dist_kernel, compute_friction_coeff_kernel and
// the pair sums of compute_particle_pair_forces_kernel (rigid_bumpy.cu:759-928,1261-1327),
// one thread per particle, its pair entries in order.  Per entry (i, j): the number of i's
// vertices touching some vertex of j, the centre distance (minimum image), the summed
// contact force on i's vertices from j's vertices (i's vertices in canonical order, each
// vertex's list in order) and the friction ratio |F_t| / |F_n| about the centre line.
__global__ void k_rigid_pairs(RigidObs a, RigidPairOut o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double e = a.e[s];
    const double Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    int ncontact = 0;
    for (int pk = o.start[i]; pk < o.start[i + 1]; ++pk) {
        const int j = o.ids[pk];
        int nvi = 0;
        double fx = 0.0, fy = 0.0;
        for (int c = a.poff[i]; c < a.poff[i + 1]; ++c) {
            const int u = a.vstatic[c];
            const double xu = a.vx[u], yu = a.vy[u], ru = a.vr[u];
            const int cnt = a.nn[u];
            bool found = false;
            for (int k = 0; k < cnt; ++k) {
                const int w = a.nbr[(size_t)k * a.nv + u];
                if (a.vpid[w] != j) continue;
                const double dx = min_image(xu - a.vx[w], Lx, iLx), dy = min_image(yu - a.vy[w], Ly, iLy);
                const double r2 = dx * dx + dy * dy;
                const double radsum = ru + a.vr[w];
                if (r2 >= radsum * radsum) continue;
                found = true;
                const double r = sqrt(r2);
                const double inv = 1.0 / r;
                const double nx = dx * inv, ny = dy * inv;
                const double delta = radsum - r;
                const double fmag = e * delta;
                fx += fmag * nx;
                fy += fmag * ny;
            }
            if (found) ++nvi;
        }
        if (nvi > 0) ++ncontact;
        const double cx = min_image(a.x[i] - a.x[j], Lx, iLx), cy = min_image(a.y[i] - a.y[j], Ly, iLy);
        const double rp = sqrt(cx * cx + cy * cy);
        if (o.pi) { o.pi[pk] = i; o.pj[pk] = j; }
        if (o.dist) o.dist[pk] = rp;
        if (o.fx) { o.fx[pk] = fx; o.fy[pk] = fy; }
        if (o.vci) o.vci[pk] = nvi;
        if (o.friction) {
            const double nx = cx / rp, ny = cy / rp;
            const double fn = fx * nx + fy * ny;
            const double fabs_n = fabs(fn);
            double mu = 0.0;
            if (fabs_n != 0.0) {
                const double tx = fx - fn * nx, ty = fy - fn * ny;
                mu = sqrt(tx * tx + ty * ty) / fabs_n;
            }
            o.friction[pk] = mu;
        }
    }
    if (o.contacts) o.contacts[i] = ncontact;
}

// compute_hessian_kernel (rigid_bumpy.cu:1129-1259): per particle-neighbor entry (i, j), the
// contributions of i's vertex contacts with j to the (x, y, theta) blocks H_i,i and H_i,j,
// expression for expression (unit derivative factors written as the reference writes them).
// Column 2m holds (i, i) and 2m + 1 holds (i, j) for m = 3a + b over (x, y, theta).  Walls
// contribute nothing, and vertex offsets from the centres take no minimum image, as there.
__global__ void k_rigid_hessian(RigidObs a, const int* start, const int* ids, double* const* hs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int s = a.sid[i];
    const double e = a.e[s];
    const double Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xi = a.x[i], yi = a.y[i];
    for (int k = start[i]; k < start[i + 1]; ++k) {
        const int j = ids[k];
        const double xj = a.x[j], yj = a.y[j];
        double hib[9], hjb[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) { hib[m] = 0.0; hjb[m] = 0.0; }
        for (int c = a.poff[i]; c < a.poff[i + 1]; ++c) {
            const int mu = a.vstatic[c];
            const double xm = a.vx[mu], ym = a.vy[mu];
            const double rxm = xm - xi, rym = ym - yi;
            const int cnt = a.nn[mu];
            for (int q = 0; q < cnt; ++q) {
                const int nu = a.nbr[(size_t)q * a.nv + mu];
                if (a.vpid[nu] != j) continue;
                const double xn = a.vx[nu], yn = a.vy[nu];
                const double dx = min_image(xm - xn, Lx, iLx), dy = min_image(ym - yn, Ly, iLy);
                const double r2 = dx * dx + dy * dy;
                const double radsum = a.vr[mu] + a.vr[nu];
                if (r2 >= radsum * radsum) continue;
                const double rr = sqrt(r2);
                const double delta = radsum - rr;
                const double fmag = e * delta;
                const double t = -fmag, cc = e;
                const double rxn = xn - xj, ryn = yn - yj;
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        const double ia_x = (p == 0) - rym * (p == 2), ia_y = (p == 1) + rxm * (p == 2);
                        const double jb_x = -1.0 * ((b == 0) - ryn * (b == 2));
                        const double jb_y = -1.0 * ((b == 1) + rxn * (b == 2));
                        const double qa = dx * ia_x + dy * ia_y, qjb = dx * jb_x + dy * jb_y;
                        const double dq_j = (ia_x * jb_x + ia_y * jb_y) / rr;
                        hjb[p * 3 + b] += cc * qa * qjb / r2 + t * (dq_j - qa * qjb / (r2 * rr));
                        const double ib_x = (b == 0) - rym * (b == 2), ib_y = (b == 1) + rxm * (b == 2);
                        const double iab_x = -rxm * (p == 2) * (b == 2), iab_y = -rym * (p == 2) * (b == 2);
                        const double qib = dx * ib_x + dy * ib_y;
                        const double dq_i = (ia_x * ib_x + dx * iab_x + ia_y * ib_y + dy * iab_y) / rr;
                        hib[p * 3 + b] += cc * qa * qib / r2 + t * (dq_i - qa * qib / (r2 * rr));
                    }
            }
        }
#pragma unroll
        for (int m = 0; m < 9; ++m) {
            hs[2 * m][k] = 0.0 + hib[m];
            hs[2 * m + 1][k] = 0.0 + hjb[m];
        }
    }
}

// vertex contacts of the reverse entry: entry (i, j) gets the count entry (j, i) holds
__global__ void k_rigid_reverse(int n, const int* start, const int* ids, const int* vci, int* vcj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int pk = start[i]; pk < start[i + 1]; ++pk) {
        const int j = ids[pk];
        int lo = start[j], hi = start[j + 1];
        while (lo < hi) {   // ids of a particle are ascending
            const int mid = (lo + hi) >> 1;
            if (ids[mid] < i) lo = mid + 1; else hi = mid;
        }
        vcj[pk] = (lo < start[j + 1] && ids[lo] == i) ? vci[lo] : 0;
    }
}

// compute_overlaps_kernel (rigid_bumpy.cu:929-966): per vertex, over its list
__global__ void k_rigid_overlaps(RigidObs a, const int* vsid, double* overlap, double* count) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.nv) return;
    const int s = vsid[u];
    const double Lx = a.Lx[s], Ly = a.Ly[s], iLx = a.iLx[s], iLy = a.iLy[s];
    const double xu = a.vx[u], yu = a.vy[u], ru = a.vr[u];
    double ov = 0.0, cc = 0.0;
    for (int k = 0; k < a.nn[u]; ++k) {
        const int w = a.nbr[(size_t)k * a.nv + u];
        const double dx = min_image(xu - a.vx[w], Lx, iLx), dy = min_image(yu - a.vy[w], Ly, iLy);
        const double r2 = dx * dx + dy * dy;
        const double radsum = ru + a.vr[w];
        if (r2 >= radsum * radsum) continue;
        ov += radsum - sqrt(r2);
        cc += 1.0;
    }
    overlap[u] = ov;
    count[u] = cc;
}

}  // namespace

// Particle-level neighbor lists of rigid bodies into c->pn_start / c->pn_ids; returns the
// number of entries.
int64_t rigid_particle_neighbors(dpmd_ctx* c) {
    const int N = c->N;
    RigidObs a = rigid_obs(c);
    c->pn_start.resize(N + 1);
    c->pn_count.resize(N);
    c->pn_flag.resize(1);
    c->pn_flag.zero(c->st);
    k_rigid_pn<false><<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(a, c->pn_count.p, nullptr, nullptr, c->pn_flag.p);
    HIP_OK(hipGetLastError());
    std::vector<int> cnt = dl(c, c->pn_count);
    int ov = 0;
    HIP_OK(hipMemcpyAsync(&ov, c->pn_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    need(ov == 0, "build_particle_neighbors: a particle has more than " + std::to_string(kMaxParticleNeighbors) +
                      " particle neighbors");
    std::vector<int> st(N + 1, 0);
    int64_t tot = 0;
    for (int i = 0; i < N; ++i) { st[i] = (int)tot; tot += cnt[i]; }
    need(tot <= 0x7fffffff, "build_particle_neighbors: more than 2^31 entries");
    st[N] = (int)tot;
    c->pn_start.upload(st.data(), N + 1, c->st);
    c->pn_ids.resize(std::max<int64_t>(1, tot));
    k_rigid_pn<true><<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(a, nullptr, c->pn_start.p, c->pn_ids.p, c->pn_flag.p);
    HIP_OK(hipGetLastError());
    c->pn_total = tot;
    return tot;
}

void rigid_pair_obs(dpmd_ctx* c, int* pi, int* pj, double* dist, double* fx, double* fy, int* vci, int* vcj,
                    double* friction) {
    const int N = c->N;
    c->pn_vc.resize(std::max<int64_t>(1, c->pn_total));
    RigidPairOut o;
    o.start = c->pn_start.p; o.ids = c->pn_ids.p;
    o.pi = pi; o.pj = pj; o.dist = dist; o.fx = fx; o.fy = fy; o.friction = friction;
    o.vci = vci ? vci : c->pn_vc.p;
    o.vcj = vcj;
    o.contacts = c->contacts.p;
    k_rigid_pairs<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(rigid_obs(c), o);
    HIP_OK(hipGetLastError());
    if (vcj) {
        k_rigid_reverse<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->pn_start.p, c->pn_ids.p, o.vci, vcj);
        HIP_OK(hipGetLastError());
    }
}

void rigid_hessian(dpmd_ctx* c, double* const* h18) {
    dpmd::DVec<double*> cols;
    std::vector<double*> hp(h18, h18 + 18);
    cols.upload(hp.data(), 18, c->st);
    k_rigid_hessian<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(rigid_obs(c), c->pn_start.p, c->pn_ids.p, cols.p);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->st));   // cols is freed on return
}

void rigid_overlaps(dpmd_ctx* c) {
    c->overlap.resize(c->Nv);
    c->ccount.resize(c->Nv);
    k_rigid_overlaps<<<blocks_for(c->Nv), DPMD_BLOCK, 0, c->st>>>(rigid_obs(c), c->vsid.p, c->overlap.p, c->ccount.p);
    HIP_OK(hipGetLastError());
}

void disk_contacts(dpmd_ctx* c) {
    k_contacts<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(obs_args(c), c->contacts.p, c->overlap.p, c->ccount.p);
    HIP_OK(hipGetLastError());
}

void contacts_total(dpmd_ctx* c) {
    k_int_sys_sum<<<c->S, DPMD_BLOCK, 0, c->st>>>(c->off.p, c->contacts.p, c->ncontacts.p);
    HIP_OK(hipGetLastError());
}

// CSR starts of the current lists (host scan of the counts; the observables are off the
// step path) and the current -> canonical map.
int64_t pair_setup(dpmd_ctx* c) {
    const int N = c->N;
    std::vector<int> cnt = dl(c, c->nn);
    std::vector<int> st(N + 1, 0);
    int64_t tot = 0;
    for (int i = 0; i < N; ++i) { st[i] = (int)tot; tot += cnt[i]; }
    need(tot <= 0x7fffffff, "pair observables: more than 2^31 neighbor entries");
    st[N] = (int)tot;
    c->pair_start.upload(st.data(), N + 1, c->st);
    c->canon.resize(N);
    k_invert<<<blocks_for(N), DPMD_BLOCK, 0, c->st>>>(N, c->static_index.p, c->canon.p);
    HIP_OK(hipGetLastError());
    return tot;
}

void disk_pair_obs(dpmd_ctx* c, int* pi, int* pj, double* dist, double* fx, double* fy, double* const* h) {
    PairOut o;
    o.start = c->pair_start.p; o.canon = c->canon.p;
    o.pi = pi; o.pj = pj; o.dist = dist; o.fx = fx; o.fy = fy;
    for (int m = 0; m < 8; ++m) o.h[m] = h ? h[m] : nullptr;
    k_pair_obs<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(obs_args(c), o);
    HIP_OK(hipGetLastError());
}

}  // namespace dpmd
