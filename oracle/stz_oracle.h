/*
 * stz_oracle.h -- CPU FP64 restatement of the reference's quasi-static STZ
 * step, used ONLY as the oracle by the tests, smoke() and bench.py's CPU
 * baseline.  Nothing in the product links or calls this.
 *
 * Parity status of each part (what pins it):
 *  - explicit step, STZ law, boundaries, source, correction: bit-compared
 *    against the compiled reference (oracle/_ref) and its committed golden
 *    fixtures (tests/golden/), plus closed-form tests of eno2 / Dplastic.
 *  - multigrid (hierarchy, Galerkin operators, zebra Gauss-Seidel, transfers,
 *    adaptive V-cycle count): compared against the compiled reference's own
 *    coarse stencils and V-cycle iterates, and the converged solution against
 *    a dense LU solve on tiny grids.
 *  - initialize_random: same glibc rand() stream as the reference; compared
 *    against the compiled reference.
 */
#ifndef STZ_ORACLE_H
#define STZ_ORACLE_H

#include "../include/stz_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const stz_sim_params *p);
void orc_destroy(orc_sim *s);

/* fields: plane layout of stz_amd.h (m*(n+2), row r = j+1) */
void orc_set_field(orc_sim *s, int field, const double *plane);
void orc_get_field(orc_sim *s, int field, double *plane);
void orc_set_vel(orc_sim *s, const double *vec2);
void orc_get_vel(orc_sim *s, double *vec2);
void orc_get_src(orc_sim *s, double *vec2);
void orc_set_src(orc_sim *s, const double *vec2);
void orc_set_time(orc_sim *s, double t);
double orc_get_time(orc_sim *s);
void orc_set_chi_inf(orc_sim *s, double chi_inf);

/* solve controls: tolerance multiplier on the reference threshold */
void orc_set_tol_mult(orc_sim *s, double mult);
double orc_tolerance(orc_sim *s);
void orc_set_decisions(orc_sim *s, int enable);
void orc_get_decisions(orc_sim *s, uint32_t *cells);

void orc_init_fields(orc_sim *s, int chi_case, double tem_base, double tem_delta);
void orc_initialize_random(orc_sim *s, double l, double tem_avg, double tem_stdev, unsigned seed);

int orc_step_quasistatic(orc_sim *s, double dt, int nsteps);
void orc_advection_step(orc_sim *s, double dt);
int orc_projection_step(orc_sim *s, double dt);
void orc_set_boundaries(orc_sim *s);

/* multigrid pieces */
void orc_mg_setup(orc_sim *s, double viscdt);
int orc_mg_solve(orc_sim *s, int *cycles, double *residual);
void orc_mg_v_cycle(orc_sim *s);
double orc_mg_l2_error(orc_sim *s);
int orc_mg_levels(orc_sim *s, int *dims, int max_levels);
/* coarse stencil of level l>=1 (l-1 in the reference's mg[] numbering):
 * 10 2x2 matrices (a,b,c,d) per node, node-major */
int orc_mg_stencil(orc_sim *s, int level, double *out);
double orc_avg_iters(orc_sim *s);
int orc_last_cycles(orc_sim *s);

/* scalar pieces for closed-form tests */
double orc_eno2(double p0, double p1, double p2, double p3);
double orc_dplastic(const stz_model_params *mp, double sbar, double chi,
                    double *dchi1, double *dchi2);
double orc_plastic_term(orc_sim *s, double sbar, double *chiv, double *ddev,
                        double dt, int *trips);

/* diagnostics (shear_sim.cc:733-775, 778-824, 1003-1028) */
void orc_tractions(orc_sim *s, double *out4);
void orc_qs_measure(orc_sim *s, int staggered, double *Q, double *max_qs);
/* the c_field::cu slot (rows -1..n) that qs_measure and compute_strain fill */
void orc_get_cu(orc_sim *s, double *plane);
void orc_step_forward(orc_sim *s, double dt);
void orc_compute_strain(orc_sim *s, int corrected);
void orc_l2_comparison(orc_sim *a, orc_sim *b, double *l2);
void orc_initialize_chi_bicubic(orc_sim *s, int mm, int nn, const double *arr, double lo, double sca);

/* free-boundary band rebuild as the GPU runs it (bd_oracle.c): the
 * interface cells (status 3/4) stay, the rest of the band is marched again
 * in rounds of width delta; phi and st (m*n, row-major) updated in place;
 * returns the number of rounds */
int orc_bd_remarch(int m, int n, double dx, double dy, double delta, double *phi, int *st);

#ifdef __cplusplus
}
#endif
#endif
