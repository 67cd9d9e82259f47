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
    unsigned* counter; double* part; int nch;
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
    r.counter = c->sys_counter.p; r.part = c->part.p; r.nch = c->nchunks;
    r.pw_tot = c->pw_tot.p; r.pe_tot = c->pe_tot.p; r.maxf2 = c->maxf2.p;
    return r;
}

__device__ __forceinline__ void fire_update_sys(const FireArgs& f, int s, double power) {
    double dt = f.dt[s], alpha = f.alpha[s];
    int ng = f.ngood[s], nb = f.nbad[s];
    double dtrev = 0.0, vscale = 1.0;
    if (dt == 0.0) {
        vscale = 0.0; alpha = 0.0;
    } else if (power > 0) {
        ng++; nb = 0;
        if (ng > 5) { dt = fmin(dt * 1.1, f.dtmax[s]); alpha *= 0.99; }
    } else {
        ng = 0; nb++;
        if (nb > 10) {
            dt = 0.0;
        } else {
            dt = fmax(dt * 0.5, f.dtmin[s]);
            alpha = 0.1;
            dtrev = -dt * 0.5;
            vscale = 0.0;
        }
    }
    f.s_dt[s] = dt; f.s_alpha[s] = alpha; f.s_ngood[s] = ng; f.s_nbad[s] = nb;
    f.s_dtrev[s] = dtrev; f.s_vscale[s] = vscale;
}

__device__ __forceinline__ void fire_convergence_sys(const FireArgs& f, const int* off, int s, double pe_total,
                                                     double maxf2, int pred, double a, double b) {
    const int dof = off[s + 1] - off[s];
    const double avg = pe_total / dof;
    const double avg_prev = f.peprev[s] / dof;
    const double diff = avg_prev == 0 ? 0 : fabs(avg / avg_prev - 1);
    if (pred == DPMD_FIRE_PRED_REFERENCE) {
        if (diff < a || avg < b) f.dt[s] = 0.0;
    } else if (pred == DPMD_FIRE_PRED_MAX_FORCE) {
        if (sqrt(maxf2) < a) f.dt[s] = 0.0;
    } else {   // DPMD_FIRE_PRED_FORCE_OR_PE
        if (sqrt(maxf2) < a || avg < b) f.dt[s] = 0.0;
    }
    f.peprev[s] = avg;
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
