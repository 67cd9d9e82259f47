// Deterministic per-system reductions and per-system scalar kernels.
//
// Reference: segmented_sum over system_offset with CUB (base_particle.hpp:987-1039) and the
// per-system kernels of src/kernels/base_particle_kernels.cu:102-170.  Here every sum uses
// a fixed tree (system-aligned 256-wide chunks, xor-butterfly waves, strided fold of chunk
// partials), so results are bit-reproducible run to run.
#include "internal.hpp"

namespace dpmd {
namespace {

struct RedIn {
    const double* a; const double* b; const double* c; const double* d;   // kind-dependent
    const double* e; const double* f;
    const int* sid; const double* Lx; const double* Ly;
    int rigid;
};

__device__ __forceinline__ double red_value(int kind, const RedIn& in, int i) {
    switch (kind) {
        case RED_PE: return in.a[i];
        case RED_KE: {   // disk.cu:266-275 / rigid_bumpy.cu:507-524
            const double vx = in.a[i], vy = in.b[i], m = in.c[i];
            if (!in.rigid) return 0.5 * (vx * vx + vy * vy) * m;
            const double w = in.d[i], I = in.e[i];
            return 0.5 * ((vx * vx + vy * vy) * m + w * w * I);
        }
        case RED_POWER: {   // disk.cu:338-346 / rigid_bumpy.cu:527-537
            const double p = in.a[i] * in.b[i] + in.c[i] * in.d[i];
            if (!in.rigid) return p;
            return p + in.e[i] * in.f[i];
        }
        case RED_VX: case RED_VY: case RED_SXX: case RED_SXY: case RED_SYX: case RED_SYY:
            return in.a[i];
        case RED_PHI: {   // compute_fractional_packing_fraction_kernel (kernels.cu:102-113)
            const int s = in.sid[i];
            const double box_area = in.Lx[s] * in.Ly[s];
            return in.a[i] / box_area;
        }
        case RED_MAXF2: { const double fx = in.a[i], fy = in.b[i]; return fx * fx + fy * fy; }
        case RED_ARR: case RED_MAXARR: return in.a[i];
        case RED_DOT: return in.a[i] * in.b[i] + in.c[i] * in.d[i];
    }
    return 0.0;
}

__global__ void __launch_bounds__(DPMD_BLOCK)
k_red_chunks(int kind, RedIn in, const int* chunk_sys, const int* chunk_beg, const int* off, double* part) {
    __shared__ double lds[4];
    const int c = blockIdx.x;
    const int s = chunk_sys[c];
    const int beg = chunk_beg[c];
    const int end = min(beg + DPMD_BLOCK, off[s + 1]);
    const int i = beg + threadIdx.x;
    double v = 0.0;
    if (i < end) v = red_value(kind, in, i);
    v = (kind == RED_MAXF2 || kind == RED_MAXARR) ? block_max256(v, lds) : block_sum256(v, lds);
    if (threadIdx.x == 0) part[c] = v;
}

__global__ void __launch_bounds__(DPMD_BLOCK)
k_red_final(int kind, const int* sys_chunk, const double* part, double* out) {
    __shared__ double lds[4];
    const int s = blockIdx.x;
    const int cb = sys_chunk[s], ce = sys_chunk[s + 1];
    double acc = 0.0;
    if (kind == RED_MAXF2 || kind == RED_MAXARR) {
        for (int k = cb + threadIdx.x; k < ce; k += DPMD_BLOCK) acc = fmax(acc, part[k]);
        acc = block_max256(acc, lds);
    } else {
        for (int k = cb + threadIdx.x; k < ce; k += DPMD_BLOCK) acc = acc + part[k];
        acc = block_sum256(acc, lds);
    }
    if (threadIdx.x == 0) out[s] = acc;
}

__global__ void k_temperature(int S, const double* ke, const int* ndof, double* T) {   // kernels.cu:115-124
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    T[s] = ke[s] * 2.0 / (ndof[s]);
}

__global__ void k_pressure(int S, const double* sxx, const double* syy, double* P) {  // kernels.cu:126-135
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    P[s] = 0.5 * (sxx[s] + syy[s]);
}

__global__ void k_temp_scale(int S, const double* T, const double* target, double* scale) {  // kernels.cu:137-146
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    scale[s] = sqrt(target[s] / T[s]);
}

// set_average_velocity_kernel (disk.cu:791-811); mean = tree sum / n.
__global__ void k_sub_mean(int N, const int* sid, const int* off, double* vx, double* vy, const double* sumx,
                           const double* sumy, const double* wantx, const double* wanty) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = sid[i];
    const int n = off[s + 1] - off[s];
    const double ax = sumx[s] / n - (wantx ? wantx[s] : 0.0);
    const double ay = sumy[s] / n - (wanty ? wanty[s] : 0.0);
    vx[i] -= ax; vy[i] -= ay;
}

__global__ void k_box_inv(int S, const double* Lx, const double* Ly, double* iLx, double* iLy) {  // kernels.cu:50-65
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double lx = Lx[s], ly = Ly[s];
    iLx[s] = (lx > 0.0) ? (1.0 / lx) : 0.0;
    iLy[s] = (ly > 0.0) ? (1.0 / ly) : 0.0;
}

// scale_box_by_increment_kernel (kernels.cu:148-170)
__global__ void k_scale_box_inc(int S, double* Lx, double* Ly, const double* phi, const double* inc, double* scale) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double nphi = phi[s] + inc[s];
    const double sc = sqrt(phi[s] / nphi);
    scale[s] = sc;
    Lx[s] = Lx[s] * sc;
    Ly[s] = Ly[s] * sc;
}

__global__ void k_run_begin(int* run) {
    const int active = (run[0] < run[1] && run[2] > 0) ? 1 : 0;
    run[3] = active;
    if (active) run[0] += 1;
}

// A halted step batch, before the host finishes the halted step: clears the halt and the
// batch words of fused batches, and flags system 0 (fused FIRE batches of one system raise
// only the run words).
__global__ void k_halt_reset(int* run, unsigned* flag, int flag_sys) {
    run[RUN_HALT] = 0; run[RUN_HALT_AT] = -1;
    run[RUN_ANY2] = 0; run[RUN_MIS] = 0; run[RUN_FIXW] = 0;
    if (flag_sys >= 0) flag[flag_sys] = 1u;
}

// save_/restore_particle_state_kernel, save_/restore_system_state_kernel (disk.cu:348-432).
// (The reference's system kernels index g_sys.id with a system index; here the flag is
// read per system.)
__global__ void k_save_restore_p(int N, const int* sid, const int* flag, int tv, int restore, double* x, double* y,
                                 double* rad, double* mass, int* stat, double* ang,
                                 double* sx, double* sy, double* srad, double* smass, int* sstat, double* sang) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (flag[sid[i]] != tv) return;
    if (restore) {
        x[i] = sx[i]; y[i] = sy[i];
        if (rad) rad[i] = srad[i];
        if (mass) mass[i] = smass[i];
        if (stat) stat[i] = sstat[i];
        if (ang) ang[i] = sang[i];
    } else {
        sx[i] = x[i]; sy[i] = y[i];
        if (rad) srad[i] = rad[i];
        if (mass) smass[i] = mass[i];
        if (stat) sstat[i] = stat[i];
        if (ang) sang[i] = ang[i];
    }
}

// save_/restore_vertex_state_kernel (rigid_bumpy.cu:655-674,711-730) over vertices in current
// order, with the vertex masses kept alongside (they are reordered with the vertices).
__global__ void k_save_restore_v(int Nv, const int* vsid, const int* flag, int tv, int restore, double* vx, double* vy,
                                 double* vr, double* vm, int* vpid, int* stat, double* sx, double* sy, double* sr,
                                 double* sm, int* spid, int* sstat) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv) return;
    if (flag[vsid[v]] != tv) return;
    if (restore) {
        vx[v] = sx[v]; vy[v] = sy[v]; vr[v] = sr[v]; vm[v] = sm[v]; vpid[v] = spid[v]; stat[v] = sstat[v];
    } else {
        sx[v] = vx[v]; sy[v] = vy[v]; sr[v] = vr[v]; sm[v] = vm[v]; spid[v] = vpid[v]; sstat[v] = stat[v];
    }
}

__global__ void k_save_restore_s(int S, const int* flag, int tv, int restore, double* Lx, double* Ly, double* sLx,
                                 double* sLy) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    if (flag[s] != tv) return;
    if (restore) { Lx[s] = sLx[s]; Ly[s] = sLy[s]; }
    else { sLx[s] = Lx[s]; sLy[s] = Ly[s]; }
}

}  // namespace

void sys_reduce(dpmd_ctx* c, int kind, double* out) {
    RedIn in{};
    const bool rigid = c->kind == DPMD_RIGID_BUMPY;
    in.rigid = rigid ? 1 : 0;
    in.sid = c->sid.p; in.Lx = c->Lx.p; in.Ly = c->Ly.p;
    switch (kind) {
        case RED_PE: in.a = c->pe.p; break;
        case RED_KE: in.a = c->vx.p; in.b = c->vy.p; in.c = c->mass.p; in.d = c->avel.p; in.e = c->inertia.p; break;
        case RED_POWER: in.a = c->fx.p; in.b = c->vx.p; in.c = c->fy.p; in.d = c->vy.p; in.e = c->torque.p; in.f = c->avel.p; break;
        case RED_VX: in.a = c->vx.p; break;
        case RED_VY: in.a = c->vy.p; break;
        case RED_PHI: in.a = c->area.p; break;
        case RED_SXX: in.a = c->sxx.p; break;
        case RED_SXY: in.a = c->sxy.p; break;
        case RED_SYX: in.a = c->syx.p; break;
        case RED_SYY: in.a = c->syy.p; break;
        case RED_MAXF2: in.a = c->fx.p; in.b = c->fy.p; break;
    }
    k_red_chunks<<<c->nchunks, DPMD_BLOCK, 0, c->st>>>(kind, in, c->chunk_sys.p, c->chunk_beg.p, c->off.p, c->part.p);
    k_red_final<<<c->S, DPMD_BLOCK, 0, c->st>>>(kind, c->sys_chunk.p, c->part.p, out);
    HIP_OK(hipGetLastError());
}

// Per-system sum of a per-particle array, and of a.b + c.d (same tree as every other sum).
void sys_reduce_arr(dpmd_ctx* c, const double* a, double* out) {
    RedIn in{};
    in.a = a;
    k_red_chunks<<<c->nchunks, DPMD_BLOCK, 0, c->st>>>(RED_ARR, in, c->chunk_sys.p, c->chunk_beg.p, c->off.p, c->part.p);
    k_red_final<<<c->S, DPMD_BLOCK, 0, c->st>>>(RED_ARR, c->sys_chunk.p, c->part.p, out);
    HIP_OK(hipGetLastError());
}
void sys_reduce_max(dpmd_ctx* c, const double* a, double* out) {   // a >= 0
    RedIn in{};
    in.a = a;
    k_red_chunks<<<c->nchunks, DPMD_BLOCK, 0, c->st>>>(RED_MAXARR, in, c->chunk_sys.p, c->chunk_beg.p, c->off.p,
                                                       c->part.p);
    k_red_final<<<c->S, DPMD_BLOCK, 0, c->st>>>(RED_MAXARR, c->sys_chunk.p, c->part.p, out);
    HIP_OK(hipGetLastError());
}
void sys_reduce_dot(dpmd_ctx* c, const double* a, const double* b, const double* cc, const double* d, double* out) {
    RedIn in{};
    in.a = a; in.b = b; in.c = cc; in.d = d;
    k_red_chunks<<<c->nchunks, DPMD_BLOCK, 0, c->st>>>(RED_DOT, in, c->chunk_sys.p, c->chunk_beg.p, c->off.p, c->part.p);
    k_red_final<<<c->S, DPMD_BLOCK, 0, c->st>>>(RED_DOT, c->sys_chunk.p, c->part.p, out);
    HIP_OK(hipGetLastError());
}

void per_system_temperature(dpmd_ctx* c) {
    sys_reduce(c, RED_KE, c->ke_tot.p);
    k_temperature<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->ke_tot.p, c->ndof.p, c->temp.p);
    HIP_OK(hipGetLastError());
}

void per_system_pressure(dpmd_ctx* c) {
    sys_reduce(c, RED_SXX, c->st_xx.p);
    sys_reduce(c, RED_SYY, c->st_yy.p);
    sys_reduce(c, RED_SXY, c->st_xy.p);
    sys_reduce(c, RED_SYX, c->st_yx.p);
    k_pressure<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->st_xx.p, c->st_yy.p, c->press.p);
    HIP_OK(hipGetLastError());
}

void temperature_scale(dpmd_ctx* c, const double* target, double* scale_out) {
    k_temp_scale<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->temp.p, target, scale_out);
    HIP_OK(hipGetLastError());
}

void subtract_mean_velocity(dpmd_ctx* c, const double* cur_x, const double* cur_y, const double* want_x,
                            const double* want_y) {
    k_sub_mean<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(c->N, c->sid.p, c->off.p, c->vx.p, c->vy.p, cur_x, cur_y,
                                                          want_x, want_y);
    HIP_OK(hipGetLastError());
}

void box_inv(dpmd_ctx* c) {
    k_box_inv<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->Lx.p, c->Ly.p, c->iLx.p, c->iLy.p);
    HIP_OK(hipGetLastError());
}

void scale_box_increment(dpmd_ctx* c, const double* inc, double* scale_out) {
    k_scale_box_inc<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, c->Lx.p, c->Ly.p, c->phi.p, inc, scale_out);
    HIP_OK(hipGetLastError());
}

void halt_reset(dpmd_ctx* c, bool flag_one) {
    k_halt_reset<<<1, 1, 0, c->st>>>(c->run.p, c->flag.p, flag_one ? fused_sys(c).s : -1);
    HIP_OK(hipGetLastError());
}

void run_begin_step(dpmd_ctx* c) {
    k_run_begin<<<1, 1, 0, c->st>>>(c->run.p);
    HIP_OK(hipGetLastError());
}

void save_restore(dpmd_ctx* c, const int* flag, int tv, bool restore) {
    const bool rigid = c->kind == DPMD_RIGID_BUMPY;
    k_save_restore_p<<<blocks_for(c->N), DPMD_BLOCK, 0, c->st>>>(
        c->N, c->sid.p, flag, tv, restore ? 1 : 0, c->x.p, c->y.p, rigid ? nullptr : c->rad.p,
        rigid ? nullptr : c->mass.p, rigid ? nullptr : c->static_index.p, rigid ? c->angle.p : nullptr,
        c->sv_x.p, c->sv_y.p, c->sv_rad.p, c->sv_mass.p, c->sv_static.p, c->sv_angle.p);
    if (rigid)
        k_save_restore_v<<<blocks_for(c->Nv), DPMD_BLOCK, 0, c->st>>>(
            c->Nv, c->vsid.p, flag, tv, restore ? 1 : 0, c->vpx.p, c->vpy.p, c->vrad.p, c->vmass.p, c->vpid.p,
            c->static_index.p, c->sv_vpx.p, c->sv_vpy.p, c->sv_vrad.p, c->sv_vmass.p, c->sv_vpid.p, c->sv_vstatic.p);
    k_save_restore_s<<<blocks_for(c->S), DPMD_BLOCK, 0, c->st>>>(c->S, flag, tv, restore ? 1 : 0, c->Lx.p, c->Ly.p,
                                                                c->sv_Lx.p, c->sv_Ly.p);
    HIP_OK(hipGetLastError());
}

}  // namespace dpmd
