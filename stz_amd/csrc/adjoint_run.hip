// adjoint_run.hip -- the C ABI of stz_adjoint (include/stz_amd.h): a forward
// run of the FP64 quasi-static step that keeps checkpoints, and the reverse
// sweep that recomputes each segment of steps from its checkpoint and runs
// the adjoint kernels (adjoint.hip) backwards through it.
//
// State kept per checkpoint and per recomputed step: the ten field planes
// (u, v, p, q, s, tau, chi, X, Y, ddev; m*(n+2) each), vel, the time and the
// V-cycle predictor's limit, so that a recomputed step repeats the original
// one exactly (same warm start, same number of V-cycles).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "sim_internal.h"

using namespace stz;

namespace {

constexpr int kPlanes = 10;

void planes_of(const FieldPtrs &f, double *p[kPlanes]) {
    p[0] = f.u; p[1] = f.v; p[2] = f.p; p[3] = f.q; p[4] = f.s;
    p[5] = f.tau; p[6] = f.chi; p[7] = f.X; p[8] = f.Y; p[9] = f.ddev;
}

FieldPtrs fields_at(double *base, size_t pe) {
    FieldPtrs f;
    f.u = base; f.v = base + pe; f.p = base + 2 * pe; f.q = base + 3 * pe; f.s = base + 4 * pe;
    f.tau = base + 5 * pe; f.chi = base + 6 * pe; f.X = base + 7 * pe; f.Y = base + 8 * pe;
    f.ddev = base + 9 * pe;
    return f;
}

}  // namespace

struct stz_adjoint {
    stz_sim *s = nullptr;
    int every = 25;
    double dt = 0, rel = 1e-10;
    int nsteps = 0;
    size_t pe = 0, nv = 0, state = 0;  // plane, vel and state sizes in doubles
    double *ckpt = nullptr;            // checkpoints, then the final state
    int nckpt = 0;
    std::vector<double> ck_time;
    std::vector<int> ck_lim;
    double *seg = nullptr;             // pre-step states of one segment
    double *lam = nullptr;             // three adjoint plane sets: L, A, G
    double *coef = nullptr, *g = nullptr, *w = nullptr, *keep = nullptr, *grad = nullptr;
    double *adv = nullptr;             // the recomputed advection output (chi diffusion only)
    int adj_lim = 0;
    double stats[6] = {0, 0, 0, 0, 0, 0};

    void free_all() {
        for (double **p : {&ckpt, &seg, &lam, &coef, &g, &w, &keep, &grad, &adv}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
};

namespace {

int save_state(stz_adjoint *a, double *dst) {
    stz_sim *s = a->s;
    double *p[kPlanes];
    planes_of(s->cur, p);
    for (int k = 0; k < kPlanes; k++)
        STZ_HIP(hipMemcpyAsync(dst + k * a->pe, p[k], a->pe * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    STZ_HIP(hipMemcpyAsync(dst + kPlanes * a->pe, s->vel, a->nv * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    return STZ_OK;
}

int load_state(stz_adjoint *a, const double *src) {
    stz_sim *s = a->s;
    double *p[kPlanes];
    planes_of(s->cur, p);
    for (int k = 0; k < kPlanes; k++)
        STZ_HIP(hipMemcpyAsync(p[k], src + k * a->pe, a->pe * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    STZ_HIP(hipMemcpyAsync(s->vel, src + kPlanes * a->pe, a->nv * sizeof(double), hipMemcpyDeviceToDevice,
                           s->stream));
    return STZ_OK;
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// w = A^{-1} g to a relative mean-square residual rel^2, warm-started from
// the previous w; the simulation's vel, src and predictor are kept aside
int adjoint_solve(stz_adjoint *a) {
    stz_sim *s = a->s;
    const size_t bytes = a->nv * sizeof(double);
    STZ_HIP(hipMemcpyAsync(a->keep, s->vel, bytes, hipMemcpyDeviceToDevice, s->stream));
    STZ_HIP(hipMemcpyAsync(a->keep + a->nv, s->src, bytes, hipMemcpyDeviceToDevice, s->stream));
    STZ_HIP(hipMemcpyAsync(s->src, a->g, bytes, hipMemcpyDeviceToDevice, s->stream));
    STZ_HIP(hipMemsetAsync(s->vel, 0, bytes, s->stream));
    int e;
    double ms = 0;
    if ((e = stz_mg_l2_error(s, &ms))) return e;  // mean square of g
    int rc = STZ_OK;
    if (ms > 0) {
        STZ_HIP(hipMemcpyAsync(s->vel, a->w, bytes, hipMemcpyDeviceToDevice, s->stream));
        const stz_mg_config cfg = s->cfg;
        const int lim = s->lim, solves = s->solves, vcount = s->vcount;
        const stz_solve_info last = s->last;
        s->cfg.tol_mult = a->rel * a->rel * ms / s->acc_base;
        s->lim = a->adj_lim;
        stz_solve_info info;
        rc = stz_mg_solve(s, &info);
        a->stats[3] += 1;
        a->stats[4] += info.total_cycles;
        a->adj_lim = s->lim;
        s->cfg = cfg;
        s->lim = lim;
        s->solves = solves;
        s->vcount = vcount;
        s->last = last;
        if (hipMemcpyAsync(a->w, s->vel, bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess && !rc)
            rc = STZ_ERR_HIP;
    } else {
        if (hipMemsetAsync(a->w, 0, bytes, s->stream) != hipSuccess) rc = STZ_ERR_HIP;
    }
    if (hipMemcpyAsync(s->vel, a->keep, bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(s->src, a->keep + a->nv, bytes, hipMemcpyDeviceToDevice, s->stream) != hipSuccess) {
        set_error("restoring the solver state after the adjoint solve failed");
        return STZ_ERR_HIP;
    }
    if (rc == STZ_ERR_NOT_CONVERGED) {
        char msg[300];
        std::snprintf(msg, sizeof msg, "adjoint solve: %s", stz_last_error());
        set_error("%s", msg);
    }
    return rc;
}

}  // namespace

extern "C" {

int stz_sim_set_model(stz_sim *s, const stz_model_params *mp) {
    if (!s || !mp) { set_error("null argument"); return STZ_ERR_ARG; }
    if (!(mp->theta > 0) || !(mp->tau0 > 0) || !(mp->c0 > 0)) {
        set_error("model needs theta, tau0, c0 > 0");
        return STZ_ERR_ARG;
    }
    s->c.stz = *mp;
    s->c.thetainv = 1 / mp->theta;
    s->p.stz = *mp;
    return STZ_OK;
}

int stz_adjoint_create(stz_sim *s, int every, stz_adjoint **out) {
    if (!s || !out) { set_error("null argument"); return STZ_ERR_ARG; }
    *out = nullptr;
    if (every < 1) { set_error("checkpoint interval must be >= 1"); return STZ_ERR_ARG; }
    if (s->precision != STZ_PREC_FP64) {
        set_error("the adjoint differentiates the FP64 form; create the simulation with STZ_PREC_FP64");
        return STZ_ERR_ARG;
    }
    stz_adjoint *a = new stz_adjoint();
    a->s = s;
    a->every = every;
    a->pe = (size_t)s->c.m * (s->c.n + 2);
    a->nv = 2 * (size_t)s->c.m * (s->c.n + 1);
    a->state = kPlanes * a->pe + a->nv;
    auto fail = [&](const char *what) {
        set_error("out of device memory for the adjoint (%s)", what);
        a->free_all();
        delete a;
        return STZ_ERR_HIP;
    };
    if (hipMalloc(&a->seg, (size_t)every * a->state * sizeof(double)) != hipSuccess) return fail("segment");
    if (hipMalloc(&a->lam, 3 * kPlanes * a->pe * sizeof(double)) != hipSuccess) return fail("adjoint planes");
    if (hipMalloc(&a->coef, 4 * (size_t)s->c.m * s->c.n * sizeof(double)) != hipSuccess) return fail("scratch");
    if (hipMalloc(&a->g, a->nv * sizeof(double)) != hipSuccess || hipMalloc(&a->w, a->nv * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->keep, 2 * a->nv * sizeof(double)) != hipSuccess || hipMalloc(&a->grad, 3 * sizeof(double)) != hipSuccess)
        return fail("vectors");
    if (hipMalloc(&a->adv, kPlanes * a->pe * sizeof(double)) != hipSuccess) return fail("advection output");
    a->adj_lim = s->cfg.predict_mult * s->cfg.predict_init;
    *out = a;
    return STZ_OK;
}

int stz_adjoint_destroy(stz_adjoint *a) {
    if (!a) return STZ_OK;
    if (a->s && a->s->stream) (void)hipStreamSynchronize(a->s->stream);
    a->free_all();
    delete a;
    return STZ_OK;
}

int stz_adjoint_set_solve_accuracy(stz_adjoint *a, double rel) {
    if (!a || !(rel > 0) || !(rel < 1)) { set_error("need 0 < rel < 1"); return STZ_ERR_ARG; }
    a->rel = rel;
    return STZ_OK;
}

int stz_adjoint_stats(stz_adjoint *a, double *out6) {
    if (!a || !out6) { set_error("null argument"); return STZ_ERR_ARG; }
    for (int k = 0; k < 6; k++) out6[k] = a->stats[k];
    return STZ_OK;
}

int stz_adjoint_forward(stz_adjoint *a, double dt, int nsteps) {
    if (!a) { set_error("null adjoint"); return STZ_ERR_ARG; }
    if (nsteps < 1 || !(dt > 0)) { set_error("need nsteps >= 1 and dt > 0"); return STZ_ERR_ARG; }
    stz_sim *s = a->s;
    STZ_HIP(hipSetDevice(s->device));
    const SimConst &c = s->c;
    if (s->cfg.kernels != 0 && s->cfg.kernels != 1) { set_error("bad kernel setting"); return STZ_ERR_ARG; }
    if (s->rep) {
        // the adjoint's own forward and recomputation take their decisions
        // from the state, as the steps it differentiates did
        set_error("a decision replay is active on this simulation; end it first");
        return STZ_ERR_STATE;
    }
    const int nck = (nsteps + a->every - 1) / a->every;
    if (a->ckpt) (void)hipFree(a->ckpt);
    a->ckpt = nullptr;
    if (hipMalloc(&a->ckpt, (size_t)(nck + 1) * a->state * sizeof(double)) != hipSuccess) {
        set_error("out of device memory for %d checkpoints", nck + 1);
        return STZ_ERR_HIP;
    }
    a->nckpt = nck;
    a->ck_time.assign(nck + 1, 0);
    a->ck_lim.assign(nck + 1, 0);
    a->dt = dt;
    a->nsteps = 0;
    const double t0 = now();
    int e;
    for (int t = 0; t < nsteps; t++) {
        if (t % a->every == 0) {
            const int k = t / a->every;
            if ((e = save_state(a, a->ckpt + (size_t)k * a->state))) return e;
            a->ck_time[k] = s->time;
            a->ck_lim[k] = s->lim;
        }
        if ((e = stz_sim_step_quasistatic(s, dt, 1))) return e;
        a->stats[5] += s->last.total_cycles;
        a->nsteps = t + 1;
    }
    if ((e = save_state(a, a->ckpt + (size_t)nck * a->state))) return e;
    a->ck_time[nck] = s->time;
    a->ck_lim[nck] = s->lim;
    STZ_HIP(hipStreamSynchronize(s->stream));
    a->stats[0] += now() - t0;
    return STZ_OK;
}

int stz_adjoint_backward(stz_adjoint *a, const double *lam_final, double *lam_init, double *grad3) {
    if (!a || !lam_final || !lam_init || !grad3) { set_error("null argument"); return STZ_ERR_ARG; }
    if (a->nsteps < 1 || !a->ckpt) { set_error("stz_adjoint_forward has not run"); return STZ_ERR_STATE; }
    stz_sim *s = a->s;
    if (s->rep) { set_error("a decision replay is active on this simulation; end it first"); return STZ_ERR_STATE; }
    STZ_HIP(hipSetDevice(s->device));
    hipStream_t st = s->stream;
    const double t0 = now();
    int e;
    const size_t PB = kPlanes * a->pe;
    double *lL = a->lam, *lA = a->lam + PB, *lG = a->lam + 2 * PB;
    STZ_HIP(hipMemcpyAsync(lL, lam_final, PB * sizeof(double), hipMemcpyHostToDevice, st));
    STZ_HIP(hipMemsetAsync(a->grad, 0, 3 * sizeof(double), st));
    STZ_HIP(hipMemsetAsync(a->w, 0, a->nv * sizeof(double), st));
    const int fwd_lim = s->lim;
    const double fwd_time = s->time;
    // the recomputed steps' solves are not the forward run's: the solver
    // statistics (avg_iters, last_solve) are restored on every exit
    struct StatsGuard {
        stz_sim *s;
        int solves, vcount;
        stz_solve_info last;
        ~StatsGuard() {
            s->solves = solves;
            s->vcount = vcount;
            s->last = last;
        }
    } guard{s, s->solves, s->vcount, s->last};
    for (int seg = a->nckpt - 1; seg >= 0; seg--) {
        const int t0s = seg * a->every, t1s = std::min(t0s + a->every, a->nsteps);
        if ((e = load_state(a, a->ckpt + (size_t)seg * a->state))) return e;
        s->time = a->ck_time[seg];
        s->lim = a->ck_lim[seg];
        for (int t = t0s; t < t1s; t++) {
            if ((e = save_state(a, a->seg + (size_t)(t - t0s) * a->state))) return e;
            if ((e = stz_sim_step_quasistatic(s, a->dt, 1))) return e;
            a->stats[2] += 1;
        }
        for (int t = t1s - 1; t >= t0s; t--) {
            const FieldPtrs L = fields_at(lL, a->pe), A = fields_at(lA, a->pe), G = fields_at(lG, a->pe);
            const FieldPtrs o = fields_at(a->seg + (size_t)(t - t0s) * a->state, a->pe);
            if ((e = adj_launch_correct(s->c, L, A, a->coef, a->g, a->dt, st))) return e;
            if ((e = adjoint_solve(a))) return e;
            if ((e = adj_launch_source(s->c, A, a->w, a->dt, st))) return e;
            if ((e = adj_zero(lG, PB, st))) return e;
            if (s->c.chi_diff != 0 || s->c.chi_len != 0) {
                // the advection's ddev output, recomputed from the pre-step
                // state (its ghost rows, which the advection leaves, first)
                const FieldPtrs W = fields_at(a->adv, a->pe);
                STZ_HIP(hipMemcpyAsync(W.ddev, o.ddev, a->pe * sizeof(double), hipMemcpyDeviceToDevice, st));
                if ((e = launch_advect(s->c, o, W, a->dt, nullptr, st))) return e;
                if ((e = adj_launch_chi_diffusion(s->c, o, W.ddev, A, G, a->dt, st))) return e;
            }
            if ((e = adj_launch_advect(s->c, o, A, G, a->dt, a->grad, st))) return e;
            std::swap(lL, lG);
        }
    }
    STZ_HIP(hipMemcpyAsync(lam_init, lL, PB * sizeof(double), hipMemcpyDeviceToHost, st));
    STZ_HIP(hipMemcpyAsync(grad3, a->grad, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    // leave the simulation in its state after the forward run
    if ((e = load_state(a, a->ckpt + (size_t)a->nckpt * a->state))) return e;
    s->time = fwd_time;
    s->lim = fwd_lim;
    STZ_HIP(hipStreamSynchronize(st));
    a->stats[1] += now() - t0;
    return STZ_OK;
}

}  // extern "C"
