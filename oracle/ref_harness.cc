// ref_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// A C-callable driver around the reference's own compiled classes
// (sb/main/shear_sim.{hh,cc}, qs_multi.{hh,cc}, tgmg/tgmg.in.cc through the
// reference's generator, shared/mat.cc), built by oracle/Makefile from the
// sources where they lie under the read-only reference tree into
// oracle/_ref/.  It is used to write the golden fixtures under
// tests/golden/, to pin the FP64 restatement (oracle/stz_oracle.c), and as
// the "reference" CPU baseline.  The reference code itself is not modified;
// the harness reaches its private members (qsm, mg levels) to read the
// multigrid hierarchy, and may tighten the solve threshold qs_multi::acc
// (qs_multi.cc:14) so that two different smoothers converge to the same
// solution.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <complex>
#include <limits>
#include <new>
#include <omp.h>

#define private public
#define protected public
#include "shear_sim.hh"
// template bodies for members the reference's own instantiation may have
// inlined away (tgmg_base<qs_multi,...>::mds behind l2_error)
#include "tgmg.cc"
#undef private
#undef protected

#include "../include/stz_amd.h"

struct ref_sim {
    stz_dynamics *stz;
    shear_sim *sim;
    double acc_base;
};

stz_dynamics *ref_make_stz(const stz_model_params &p) {
    switch (p.kind) {
    case STZ_MODEL_LINEAR:
        return new stz_dynamics_linear(p.TZ, p.c0, p.tau0, p.kappa, p.Delta, p.Omegaeps0, p.chi_inf, p.theta, p.rho,
                                       p.t_model);
    case STZ_MODEL_NONLINEAR:
        return new stz_dynamics_nonlinear(p.TZ, p.c0, p.tau0, p.kappa, p.Delta, p.Omegaeps0, p.chi_inf, p.theta, p.rho);
    case STZ_MODEL_NONLINEAR_ATHERMAL:
        return new stz_dynamics_nonlinear_athermal(p.TZ, p.c0, p.tau0, p.kappa, p.Delta, p.Omegaeps0, p.chi_inf,
                                                   p.theta, p.rho);
    default: {
        stz_dynamics_linear_athermal *a = new stz_dynamics_linear_athermal(
            p.TZ, p.c0, p.tau0, p.kappa, p.Delta, p.Omegaeps0, p.chi_inf, p.theta, p.rho, p.t_model);
        a->t_icpt = p.t_icpt;
        a->t_grad = p.t_grad;
        return a;
    }
    }
}

extern "C" {

void *ref_create(const stz_sim_params *p) {
    ref_sim *r = new ref_sim;
    r->stz = ref_make_stz(p->stz);
    r->sim = new shear_sim(p->m, p->n, p->ax, p->bx, p->ay, p->by, p->mu, p->K, p->viscosity, p->chi_diff,
                           p->chi_len, p->t_scale, p->tmult, p->adapt_fac, p->u_bdry, r->stz, ".");
    shear_sim &s = *r->sim;
    s.qsm.mg.verbose = 0;
    // the reference leaves fm/vel/src uninitialised; start from zeros
    std::memset(static_cast<void *>(s.fm - s.m), 0, sizeof(c_field) * (s.mn + 2 * s.m));
    std::memset(static_cast<void *>(s.vel), 0, sizeof(vec) * (s.mn + s.m));
    std::memset(static_cast<void *>(s.src), 0, sizeof(vec) * (s.mn + s.m));
    r->acc_base = s.qsm.acc;
    return r;
}

void ref_destroy(void *h) {
    ref_sim *r = static_cast<ref_sim *>(h);
    delete r->sim;
    delete r->stz;
    delete r;
}

// plane k (c_field member k, fields.hh:7-26) with rows -1..n
void ref_set_field(void *h, int k, const double *plane) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    for (int j = -1; j <= s.n; j++)
        for (int i = 0; i < s.m; i++)
            reinterpret_cast<double *>(&s.fm[i + s.m * j])[k] = plane[i + s.m * (j + 1)];
}

void ref_get_field(void *h, int k, double *plane) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    for (int j = -1; j <= s.n; j++)
        for (int i = 0; i < s.m; i++)
            plane[i + s.m * (j + 1)] = reinterpret_cast<double *>(&s.fm[i + s.m * j])[k];
}

void ref_set_vel(void *h, const double *v) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    std::memcpy(static_cast<void *>(s.vel), v, sizeof(vec) * (s.mn + s.m));
}
void ref_get_vel(void *h, double *v) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    std::memcpy(v, static_cast<void *>(s.vel), sizeof(vec) * (s.mn + s.m));
}
void ref_set_src(void *h, const double *v) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    std::memcpy(static_cast<void *>(s.src), v, sizeof(vec) * (s.mn + s.m));
}
void ref_get_src(void *h, double *v) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    std::memcpy(v, static_cast<void *>(s.src), sizeof(vec) * (s.mn + s.m));
}

void ref_set_time(void *h, double t) { static_cast<ref_sim *>(h)->sim->time = t; }
double ref_get_time(void *h) { return static_cast<ref_sim *>(h)->sim->time; }
void ref_set_chi_inf(void *h, double c) { static_cast<ref_sim *>(h)->sim->chi_inf = c; }

// mean-square residual threshold = mult * tgmg_accuracy(...) (qs_multi.cc:14)
void ref_set_tol_mult(void *h, double mult) {
    ref_sim *r = static_cast<ref_sim *>(h);
    const_cast<double &>(r->sim->qsm.acc) = r->acc_base * mult;
}
double ref_tolerance(void *h) { return static_cast<ref_sim *>(h)->sim->qsm.acc; }

// stz_dynamics::Dplastic of a freshly constructed law (stz_model.hh:28)
double ref_dplastic(const stz_model_params *p, double sbar, double chi, double *dchi1, double *dchi2) {
    stz_dynamics *s = ref_make_stz(*p);
    const double r = s->Dplastic(sbar, chi, *dchi1, *dchi2);
    delete s;
    return r;
}

void ref_init_fields(void *h, int c, double t0, double dt) { static_cast<ref_sim *>(h)->sim->init_fields(c, t0, dt); }
void ref_initialize_random(void *h, double l, double avg, double sd, unsigned seed) {
    srand(seed);
    static_cast<ref_sim *>(h)->sim->initialize_random(l, avg, sd);
}

void ref_step(void *h, double dt, int nsteps) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    for (int k = 0; k < nsteps; k++) s.step_forward_quasistatic(dt);
}
void ref_advection_step(void *h, double dt) { static_cast<ref_sim *>(h)->sim->advection_step(dt); }
void ref_projection_step(void *h, double dt) { static_cast<ref_sim *>(h)->sim->projection_step(dt); }
double ref_avg_iters(void *h) { return static_cast<ref_sim *>(h)->sim->avg_iters(); }

// multigrid access (tgmg.hh:286-495)
void ref_mg_setup(void *h, double viscdt) { static_cast<ref_sim *>(h)->sim->qsm.init(viscdt); }
int ref_mg_solve(void *h) {
    qs_multi &q = static_cast<ref_sim *>(h)->sim->qsm;
    return q.mg.solve_v_cycle(q.tp) ? 0 : 1;
}
void ref_mg_v_cycle(void *h) { static_cast<ref_sim *>(h)->sim->qsm.mg.v_cycle(); }
double ref_mg_l2_error(void *h) { return static_cast<ref_sim *>(h)->sim->qsm.mg.l2_error(); }
int ref_mg_lim(void *h) { return static_cast<ref_sim *>(h)->sim->qsm.tp.lim; }
int ref_mg_levels(void *h, int *dims, int max_levels) {
    tgmg<qs_multi, vec, mat> &mg = static_cast<ref_sim *>(h)->sim->qsm.mg;
    int k = 0;
    if (max_levels > 0) { dims[0] = mg.m; dims[1] = mg.n; k = 1; }
    for (int l = 0; l < mg.ml && k < max_levels; l++, k++) { dims[2 * k] = mg.mg[l]->m; dims[2 * k + 1] = mg.mg[l]->n; }
    return mg.ml + 1;
}
// coarse stencils of level (1 = mg[0]): 10 mats (a,b,c,d) per node
int ref_mg_stencil(void *h, int level, double *out) {
    tgmg<qs_multi, vec, mat> &mg = static_cast<ref_sim *>(h)->sim->qsm.mg;
    if (level < 1 || level > mg.ml) return -1;
    tgmg_level<vec, mat> *L = mg.mg[level - 1];
    std::memcpy(out, static_cast<void *>(L->s), sizeof(mat) * 10 * L->mn);
    return 0;
}

void ref_tractions(void *h, double *o) { static_cast<ref_sim *>(h)->sim->tractions(o[0], o[1], o[2], o[3]); }
void ref_qs_measure(void *h, int staggered, double *Q, double *mq) {
    static_cast<ref_sim *>(h)->sim->qs_measure(*Q, *mq, staggered != 0);
}

// the c_field::cu slot (rows -1..n), the diagnostic output of qs_measure,
// compute_strain and the direct step
void ref_get_cu(void *h, double *plane) { ref_get_field(h, 10, plane); }
void ref_step_forward(void *h, double dt) { static_cast<ref_sim *>(h)->sim->step_forward(dt); }
// the reference has only its own form (corrected is not available)
void ref_compute_strain(void *h, int corrected) {
    (void)corrected;
    static_cast<ref_sim *>(h)->sim->compute_strain();
}
void ref_l2_comparison(void *a, void *b, double *l2) {
    static_cast<ref_sim *>(a)->sim->l2_comparison(*static_cast<ref_sim *>(b)->sim, l2);
}
void ref_initialize_chi_bicubic(void *h, int mm, int nn, const double *arr, double lo, double sca) {
    static_cast<ref_sim *>(h)->sim->initialize_chi_bicubic(mm, nn, const_cast<double *>(arr), lo, sca);
}
// shear_sim::output (shear_sim.cc:688-726) of one field into directory dir
void ref_output(void *h, const char *dir, const char *prefix, int mode, int sn) {
    shear_sim &s = *static_cast<ref_sim *>(h)->sim;
    const char *keep = s.filename;
    s.filename = dir;
    s.output(prefix, mode, sn);
    s.filename = keep;
}

void ref_set_threads(int n) { omp_set_num_threads(n); }
}
