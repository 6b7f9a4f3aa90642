// ref_bd_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// A C-callable driver around the reference's free-boundary classes
// (sb/main/bd_sim.{hh,cc}, qs_multi_bd.{hh,cc}, extra_force.{hh,cc}, the
// Level++ narrow-band level set levelset/*.cc, and bicubic_vel.cc written by
// the reference's own generator sb/main/bicubic_vel.pl), compiled by
// oracle/Makefile from the sources where they lie under the read-only
// reference tree into oracle/_ref/.  It writes the golden fixtures of the
// free-boundary step under tests/golden/ and is the CPU side of the
// component comparisons.  The reference code itself is not modified; the
// harness reaches private members (qsm, ls internals) to read the operator
// rows and to evaluate the interface velocity on its own.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <vector>

#define private public
#define protected public
#include "bd_sim.hh"
#include "extra_force.hh"
#undef private
#undef protected

#include "../include/stz_amd.h"

stz_dynamics *ref_make_stz(const stz_model_params &p);

namespace {

// extra_force whose force(i,j) is a per-node table set by the caller
// (extra_force.hh:5-28); its centre (x,y) stays at the origin
struct table_force : public extra_force {
    std::vector<double> F;
    explicit table_force(bd_sim &b) : extra_force(b) { x = y = 0; }
    virtual vec force(int i, int j) {
        const size_t k = 2 * (size_t)(i + bds.me * j);
        return vec(F[k], F[k + 1]);
    }
};

struct ref_bd {
    stz_dynamics *stz;
    bd_sim *sim;
    table_force *force;
    extra_force *builtin;
};

double *cell_member(c_field &f, int k) {
    switch (k) {
    case STZ_BD_P: return &f.p;
    case STZ_BD_Q: return &f.q;
    case STZ_BD_S: return &f.s;
    case STZ_BD_TAU: return &f.tau;
    case STZ_BD_CHI: return &f.chi;
    default: return &f.ddev;
    }
}
double *node_member(c_field &f, int k) {
    switch (k) {
    case STZ_BD_U: return &f.u;
    case STZ_BD_V: return &f.v;
    case STZ_BD_X: return &f.X;
    default: return &f.Y;
    }
}

}  // namespace

extern "C" {

void *ref_bd_create(const stz_bd_params *p, const char *outdir) {
    ref_bd *r = new ref_bd;
    r->stz = ref_make_stz(p->stz);
    r->sim = new bd_sim(p->m, p->n, p->ax, p->bx, p->ay, p->by, p->mu, p->K, p->viscosity, p->chi_len,
                        p->filter_stress, p->t_scale, p->tmult, p->adapt_fac, r->stz, outdir, 0);
    r->force = 0;
    r->builtin = 0;
    bd_sim &s = *r->sim;
    s.mg.verbose = 0;
    // the reference leaves its arrays uninitialised; start from zeros
    std::memset(static_cast<void *>(s.fm), 0, sizeof(c_field) * s.mne);
    std::memset(static_cast<void *>(s.vel), 0, sizeof(vec) * s.mne);
    std::memset(static_cast<void *>(s.src), 0, sizeof(vec) * s.mne);
    std::memset(s.phi, 0, sizeof(double) * s.mn);
    std::memset(s.ls.vp, 0, sizeof(double) * s.mn);
    return r;
}

void ref_bd_destroy(void *h) {
    ref_bd *r = static_cast<ref_bd *>(h);
    delete r->force;
    delete r->builtin;
    delete r->sim;
    delete r->stz;
    delete r;
}

void ref_bd_init_bar(void *h, int ca) { static_cast<ref_bd *>(h)->sim->init_bar(ca); }

// levelset::build_band on the given signed function, then the node status
void ref_bd_set_phi_build(void *h, const double *phi) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    std::memcpy(s.phi, phi, sizeof(double) * s.mn);
    s.ls.build_band();
    s.set_staggered_status();
}

void ref_bd_get_phi(void *h, double *out) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    std::memcpy(out, s.phi, sizeof(double) * s.mn);
}

void ref_bd_get_status(void *h, int *c, int *cc) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    if (c) std::memcpy(c, s.c, sizeof(int) * s.mn);
    if (cc) std::memcpy(cc, s.cc, sizeof(int) * s.mne);
}

void ref_bd_set_cells(void *h, int k, const double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    for (int j = 0; j < s.n; j++)
        for (int i = 0; i < s.m; i++) *cell_member(s.fm[i + s.me * j], k) = v[i + s.m * j];
}
void ref_bd_get_cells(void *h, int k, double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    for (int j = 0; j < s.n; j++)
        for (int i = 0; i < s.m; i++) v[i + s.m * j] = *cell_member(s.fm[i + s.me * j], k);
}
void ref_bd_set_nodes(void *h, int k, const double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    for (int ij = 0; ij < s.mne; ij++) *node_member(s.fm[ij], k) = v[ij];
}
void ref_bd_get_nodes(void *h, int k, double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    for (int ij = 0; ij < s.mne; ij++) v[ij] = *node_member(s.fm[ij], k);
}
void ref_bd_get_vel(void *h, double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    std::memcpy(v, s.vel, sizeof(vec) * s.mne);
}
void ref_bd_get_src(void *h, double *v) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    std::memcpy(v, s.src, sizeof(vec) * s.mne);
}

void ref_bd_set_time(void *h, double t) { static_cast<ref_bd *>(h)->sim->time = t; }

void ref_bd_set_force(void *h, const double *F) {
    ref_bd *r = static_cast<ref_bd *>(h);
    if (!r->force) r->force = new table_force(*r->sim);
    r->force->F.assign(F, F + 2 * (size_t)r->sim->mne);
}

// the reference's own forcing classes (extra_force.cc:22-42), k = 1, 2, 3
void ref_bd_use_builtin_force(void *h, int k) {
    ref_bd *r = static_cast<ref_bd *>(h);
    if (k == 1) r->builtin = new extra_force1(*r->sim);
    else if (k == 2) r->builtin = new extra_force2(*r->sim);
    else r->builtin = new extra_force3(*r->sim);
}

void ref_bd_advection_step(void *h, double dt) { static_cast<ref_bd *>(h)->sim->advection_step(dt); }
void ref_bd_projection_step(void *h, double dt) { static_cast<ref_bd *>(h)->sim->projection_step(dt); }
void ref_bd_post_process(void *h, int all) { static_cast<ref_bd *>(h)->sim->post_process(all != 0); }
void ref_bd_step(void *h, double dt, int nsteps) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    for (int k = 0; k < nsteps; k++) s.step_forward_quasistatic(dt);
}

// operator rows after qs_multi_bd::init: row type d (-1 interior, -2 box
// edge, -3 outside, >= 0 free boundary), the nine 2x2 blocks a_dl..a_ur
// (qs_multi_bd.hh:31-39, 36 doubles per node) and the source
void ref_bd_operator(void *h, double dt, int *d, double *mats, double *b) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    qs_multi_bd &q = s.qsm;
    q.init(s.viscosity, dt);
    for (int j = 0; j < q.n; j++)
        for (int i = 0; i < q.m; i++) {
            const int ij = i + q.m * j;
            if (d) d[ij] = q.d[ij] >= 0 ? 0 : q.d[ij];
            if (mats) {
                mat a[9] = {q.a_dl(i, ij), q.a_dc(i, ij), q.a_dr(i, ij), q.a_cl(i, ij), q.a_cc(i, ij),
                            q.a_cr(i, ij), q.a_ul(i, ij), q.a_uc(i, ij), q.a_ur(i, ij)};
                double *o = mats + 36 * (size_t)ij;
                for (int k = 0; k < 9; k++) {
                    o[4 * k] = a[k].a;
                    o[4 * k + 1] = a[k].b;
                    o[4 * k + 2] = a[k].c;
                    o[4 * k + 3] = a[k].d;
                }
            }
            if (b) {
                b[2 * ij] = q.b[ij].x;
                b[2 * ij + 1] = q.b[ij].y;
            }
        }
}

// the velocity extension of levelset::move_and_smooth (levelset.hh:159-205)
// for every interface cell (status 3 or 4), without moving the level set;
// other cells get 0
void ref_bd_interface_velocity(void *h, double *vp) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    levelset &ls = s.ls;
    double x0, y0, x, y;
    for (int ij = 0; ij < ls.mn; ij++) {
        vp[ij] = 0;
        if (ls.s[ij] != 3 && ls.s[ij] != 4) continue;
        const int i = ij % ls.m;
        vp[ij] = ls.newton(i, ij, x0, y0, x, y) ? s.velocity(x, y) : s.velocity(x0, y0);
        vp[ij] += s.bc_smooth * ls.laplacian(i, ij);
    }
}

// the band rebuilt around an unmoved interface: levelset::move_and_smooth
// with dt = 0 (levelset.hh:159-205, post_move_update levelset.cc:701-772),
// then the node statuses
void ref_bd_remarch(void *h) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    s.ls.move_and_smooth(&s, 0.0, 0.0);
    s.set_staggered_status();
}

// levelset::newton at one cell: closest point (x,y) and success flag
int ref_bd_newton(void *h, int ij, double *xy) {
    levelset &ls = static_cast<ref_bd *>(h)->sim->ls;
    double x0, y0;
    const bool ok = ls.newton(ij % ls.m, ij, x0, y0, xy[0], xy[1]);
    return ok ? 1 : 0;
}

// levelset::bicubic of phi and bd_sim::bicubic_velocity at a point
double ref_bd_bicubic(void *h, double x, double y, double *grad, double *uv) {
    bd_sim &s = *static_cast<ref_bd *>(h)->sim;
    const double v = s.ls.bicubic(x, y, grad[0], grad[1]);
    s.bicubic_velocity(x, y, uv[0], uv[1]);
    return v;
}

void ref_bd_set_bc_smooth(void *h, double v) { static_cast<ref_bd *>(h)->sim->bc_smooth = v; }

}  // extern "C"
