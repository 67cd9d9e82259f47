// FIRE per-system state and helpers shared by the disk and rigid-body kernels.
// Reference: src/integrators/fire.cu:7-78, include/integrators/fire.cuh:116-147.
#pragma once
#include "internal.hpp"

namespace dpmd {

struct FireArgs {
    double* dt; double* dtmax; double* dtmin; double* alpha; double* peprev;
    int* ngood; int* nbad; int* stopped;
    int* frozen;    // null unless the run skips frozen systems
    int* nsteps;    // fused steps that did work, per system
    double* s_dt; double* s_alpha; double* s_dtrev; double* s_vscale; int* s_ngood; int* s_nbad;
    int* run;
};

struct RedArgs {
    const int* chunk_sys; const int* chunk_beg; const int* sys_chunk;
    double* part; int nch;
    double* pw_tot; double* pe_tot; double* maxf2;
};


inline FireArgs fire_args(dpmd_ctx* c) {
    FireArgs f;
    f.dt = c->f_dt.p; f.dtmax = c->f_dtmax.p; f.dtmin = c->f_dtmin.p; f.alpha = c->f_alpha.p;
    f.peprev = c->f_peprev.p; f.ngood = c->f_ngood.p; f.nbad = c->f_nbad.p; f.stopped = c->f_stopped.p;
    f.frozen = c->mask_frozen ? c->f_frozen.p : nullptr;
    f.nsteps = c->f_nsteps.p;
    f.s_dt = c->s_dt.p; f.s_alpha = c->s_alpha.p; f.s_dtrev = c->s_dtrev.p; f.s_vscale = c->s_vscale.p;
    f.s_ngood = c->s_ngood.p; f.s_nbad = c->s_nbad.p; f.run = c->run.p;
    return f;
}

inline RedArgs red_args(dpmd_ctx* c) {
    RedArgs r;
    r.chunk_sys = c->chunk_sys.p; r.chunk_beg = c->chunk_beg.p; r.sys_chunk = c->sys_chunk.p;
    r.part = c->part.p; r.nch = c->nchunks;
    r.pw_tot = c->pw_tot.p; r.pe_tot = c->pe_tot.p; r.maxf2 = c->maxf2.p;
    return r;
}

// The FIRE scalars of one system: committed values and the next step's ("s_") values.
// The array forms below and the persistent kernel (which keeps them in registers for a
// whole run) apply the same two functions to them.
struct FireSys {
    double dt, dtmax, dtmin, alpha, peprev;
    int ngood, nbad;
    double s_dt, s_alpha, s_dtrev, s_vscale;
    int s_ngood, s_nbad;
};

// fire_update (fire.cu:7-57): next-step dt, alpha, counters, reverse drift and v scale.
__device__ __forceinline__ void fire_update_core(FireSys& q, double power) {
    double dt = q.dt, alpha = q.alpha;
    int ng = q.ngood, nb = q.nbad;
    double dtrev = 0.0, vscale = 1.0;
    if (dt == 0.0) {
        vscale = 0.0; alpha = 0.0;
    } else if (power > 0) {
        ng++; nb = 0;
        if (ng > 5) { dt = fmin(dt * 1.1, q.dtmax); alpha *= 0.99; }
    } else {
        ng = 0; nb++;
        if (nb > 10) {
            dt = 0.0;
        } else {
            dt = fmax(dt * 0.5, q.dtmin);
            alpha = 0.1;
            dtrev = -dt * 0.5;
            vscale = 0.0;
        }
    }
    q.s_dt = dt; q.s_alpha = alpha; q.s_ngood = ng; q.s_nbad = nb;
    q.s_dtrev = dtrev; q.s_vscale = vscale;
}

// fire_convergence_kernel (fire.cu:59-78) and the added predicates: dt = 0 stops.
__device__ __forceinline__ void fire_convergence_core(FireSys& q, int dof, double pe_total, double maxf2, int pred,
                                                      double a, double b) {
    const double avg = pe_total / dof;
    const double avg_prev = q.peprev / dof;
    const double diff = avg_prev == 0 ? 0 : fabs(avg / avg_prev - 1);
    if (pred == DPMD_FIRE_PRED_REFERENCE) {
        if (diff < a || avg < b) q.dt = 0.0;
    } else if (pred == DPMD_FIRE_PRED_MAX_FORCE) {
        if (sqrt(maxf2) < a) q.dt = 0.0;
    } else {   // DPMD_FIRE_PRED_FORCE_OR_PE
        if (sqrt(maxf2) < a || avg < b) q.dt = 0.0;
    }
    q.peprev = avg;
}

__device__ __forceinline__ FireSys fire_load(const FireArgs& f, int s) {
    FireSys q;
    q.dt = f.dt[s]; q.dtmax = f.dtmax[s]; q.dtmin = f.dtmin[s]; q.alpha = f.alpha[s]; q.peprev = f.peprev[s];
    q.ngood = f.ngood[s]; q.nbad = f.nbad[s];
    q.s_dt = f.s_dt[s]; q.s_alpha = f.s_alpha[s]; q.s_dtrev = f.s_dtrev[s]; q.s_vscale = f.s_vscale[s];
    q.s_ngood = f.s_ngood[s]; q.s_nbad = f.s_nbad[s];
    return q;
}

__device__ __forceinline__ void fire_store(const FireArgs& f, int s, const FireSys& q) {
    f.dt[s] = q.dt; f.alpha[s] = q.alpha; f.peprev[s] = q.peprev; f.ngood[s] = q.ngood; f.nbad[s] = q.nbad;
    f.s_dt[s] = q.s_dt; f.s_alpha[s] = q.s_alpha; f.s_dtrev[s] = q.s_dtrev; f.s_vscale[s] = q.s_vscale;
    f.s_ngood[s] = q.s_ngood; f.s_nbad[s] = q.s_nbad;
}

__device__ __forceinline__ void fire_update_sys(const FireArgs& f, int s, double power) {
    FireSys q;
    q.dt = f.dt[s]; q.alpha = f.alpha[s]; q.ngood = f.ngood[s]; q.nbad = f.nbad[s];
    q.dtmax = f.dtmax[s]; q.dtmin = f.dtmin[s];
    fire_update_core(q, power);
    f.s_dt[s] = q.s_dt; f.s_alpha[s] = q.s_alpha; f.s_ngood[s] = q.s_ngood; f.s_nbad[s] = q.s_nbad;
    f.s_dtrev[s] = q.s_dtrev; f.s_vscale[s] = q.s_vscale;
}

__device__ __forceinline__ void fire_convergence_sys(const FireArgs& f, const int* off, int s, double pe_total,
                                                     double maxf2, int pred, double a, double b) {
    FireSys q;
    q.dt = f.dt[s]; q.peprev = f.peprev[s];
    fire_convergence_core(q, off[s + 1] - off[s], pe_total, maxf2, pred, a, b);
    f.dt[s] = q.dt;
    f.peprev[s] = q.peprev;
}

__device__ __forceinline__ bool fire_frozen(const FireArgs& f, int s) { return f.frozen && f.frozen[s]; }

__device__ __forceinline__ void fire_count_stopped(const FireArgs& f, int s, int k) {
    if (f.dt[s] == 0.0 && !f.stopped[s]) {
        f.stopped[s] = 1;
        if (atomicSub(&f.run[RUN_RUNNING], 1) == 1) f.run[RUN_CONV_AT] = k;
    }
}


// Fold chunk partials, apply the stopping test, evaluate the next FIRE update (disk.hip).
void fire_finalize(dpmd_ctx* c, Gate g, int do_conv, int pred, double pa, double pb);

}  // namespace dpmd
