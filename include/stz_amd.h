/*
 * stz_amd.h -- C ABI of the MI355X (gfx950) quasi-static STZ timestep.
 *
 * The reference exposes this path only as C++ classes (no FFI): the
 * simulation class `shear_sim` (sb/main/shear_sim.hh:20-151), the multigrid
 * template `tgmg<S,V,M>` (tgmg/tgmg.hh:286-495) with its problem class
 * `qs_multi` (sb/main/qs_multi.hh:11-57), and the plasticity law
 * `stz_dynamics` (sb/main/stz_model.hh:9-170).  Every entry point below is
 * what a thin C++ facade keeping those signatures forwards to; the facade
 * lives in stz_amd/host/ and the per-entry-point mapping is given in each
 * comment.  Plain C types only: pointers, sizes, doubles, ints.
 *
 * Conventions
 *  - All functions return an int status (STZ_OK on success).  The reference
 *    has no status codes: its failures call exit(1) (qs_multi.hh:52,
 *    tgmg.hh:29-32); the C++ facade maps a non-zero status back to that.
 *  - Field planes are FP64, row-major, m columns, n+2 rows: plane row r holds
 *    grid row j=r-1 for j in [-1,n].  This is the reference's `fm` array
 *    (shear_sim.cc:37, c_field[m*(n+2)] offset by m) split into one plane per
 *    c_field member (fields.hh:7-52).  Node fields (u,v,X,Y) use rows 0..n,
 *    cell fields (p,q,s,tau,chi,ddev) use rows -1..n (ghost rows -1 and n).
 *  - The multigrid solution/velocity `vel` and source `src` are vec2 arrays of
 *    m*(n+1) nodes, interleaved (x,y) (shear_sim.cc:37-38, vec.hh:6-8).
 *  - Host pointers are copied synchronously; device pointers returned by the
 *    library stay valid until stz_sim_destroy.
 */
#ifndef STZ_AMD_H
#define STZ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    STZ_OK = 0,
    STZ_ERR_ARG = 1,            /* bad argument (size, null, enum) */
    STZ_ERR_HIP = 2,            /* a HIP runtime call failed */
    STZ_ERR_NOT_CONVERGED = 3,  /* V-cycle budget exceeded (tgmg.in.cc:285-288) */
    STZ_ERR_NAN = 4,            /* residual is NaN (tgmg.in.cc:300) */
    STZ_ERR_NO_DEVICE = 5,      /* no gfx950 device visible */
    STZ_ERR_STATE = 6           /* call out of order (e.g. solve before setup) */
};

/* ---- STZ plasticity law (stz_model.hh) ---------------------------------- */
enum {
    STZ_MODEL_LINEAR_ATHERMAL = 0,    /* stz_model.hh:90-170 (used by drivers) */
    STZ_MODEL_LINEAR = 1,             /* stz_model.hh:50-68 */
    STZ_MODEL_NONLINEAR = 2,          /* stz_model.hh:70-88 */
    STZ_MODEL_NONLINEAR_ATHERMAL = 3  /* stz_model.hh:172-207 */
};

/* Constructor arguments of stz_dynamics (stz_model.hh:22-26, 98-101).
 * theta is passed as given; the law uses 1/theta.  t_model, t_icpt and
 * t_grad are the truncation settings of the linear athermal law
 * (stz_model.hh:93-97); set t_grad = 0.5*Omegaeps0/Delta, t_icpt = 0 for the
 * reference defaults. */
typedef struct stz_model_params {
    int32_t kind;
    int32_t t_model;
    double TZ, c0, tau0, kappa, Delta, Omegaeps0, chi_inf, theta, rho;
    double t_icpt, t_grad;
} stz_model_params;

/* stz_dynamics::Dplastic (stz_model.hh:28; the laws of :50-207): the
 * plastic rate at deviatoric stress sbar and effective temperature chi, and
 * the chi-rate coefficients dchi1, dchi2.  Evaluated on the host with the
 * law code the step's kernels run (C library transcendentals, FMA
 * contraction off), for callers that tabulate the law; never part of a step. */
int stz_model_dplastic(const stz_model_params *mp, double sbar, double chi, double *dpl, double *dchi1,
                       double *dchi2);

/* Constructor arguments of shear_sim (shear_sim.hh:95-99, shear_sim.cc:26-44). */
typedef struct stz_sim_params {
    int32_t m, n;                 /* cells in x (periodic) and y (walls) */
    double ax, bx, ay, by;        /* domain bounds */
    double mu, K, viscosity;      /* shear/bulk modulus, viscous damping */
    double chi_diff, chi_len;     /* chi diffusion (constant, Dpl-mediated) */
    double t_scale, tmult, adapt_fac, u_bdry;
    stz_model_params stz;
} stz_sim_params;

/* ---- field planes (fields.hh:7-52 member order) ------------------------- */
enum {
    STZ_F_U = 0, STZ_F_V, STZ_F_P, STZ_F_Q, STZ_F_S, STZ_F_TAU, STZ_F_CHI,
    STZ_F_X, STZ_F_Y, STZ_F_DDEV, STZ_NFIELDS
};
/* The c_field::cu slot (fields.hh:28), which the reference's diagnostics
 * use as their output plane (qs_measure, compute_strain).  Same plane
 * layout; not part of the state a step reads. */
enum { STZ_F_CU = STZ_NFIELDS };

/* ---- precision of the projection solve --------------------------------- */
enum {
    STZ_PREC_FP64 = 0,   /* all multigrid work in FP64 */
    STZ_PREC_MIXED = 1   /* FP32 V-cycles inside FP64 iterative refinement */
};

/* Multigrid settings.  Defaults follow the reference's solve_v_cycle
 * (tgmg.hh:439): one pre/post sweep, the adaptive V-cycle predictor of
 * tgmg_predict.hh:6-33 and the tolerance tgmg_accuracy(2*(tmufac1+tmufac2),
 * 1e4) of qs_multi.cc:14 scaled by tol_mult (mean-square residual threshold
 * = tol_mult * that value).  The coarse end of the V-cycle is solved exactly
 * (see direct_nodes), where the reference smooths its coarsest level with
 * 2 bottom_sweeps Gauss-Seidel sweeps: bottom_sweeps is kept for the C++
 * facade's tgmg signatures and does not change the solve. */
typedef struct stz_mg_config {
    int32_t nu1, nu2;            /* pre/post smoothing sweeps per level */
    int32_t bottom_sweeps;       /* half the sweeps on the coarsest level */
    int32_t extra_cycles;        /* cycles after convergence (reference: 1) */
    int32_t predict_mult, predict_decay, predict_init, predict_max_iters;
    double tol_mult;             /* multiplier on the mean-square tolerance */
    int32_t max_refine;          /* mixed: FP64 refinement iterations one solve
                                    may take (one per V-cycle) before it fails
                                    with STZ_ERR_NOT_CONVERGED */
    int32_t kernels;             /* 0: LDS-tiled multi-sweep kernels where a
                                    level allows them (default); 1: one
                                    kernel per colour pass everywhere */
    int32_t direct_nodes;        /* the finest coarse level with at most this
                                    many nodes whose operator is invariant
                                    along x is solved exactly (a DFT along x
                                    and one inverse per Fourier mode), and
                                    the V-cycle stops there (default 4200:
                                    the 64x65 level of a 1024^2 grid; levels
                                    of even width up to 128 and at most 72
                                    rows qualify); 0: the V-cycle runs to the
                                    coarsest level, solved exactly */
    int32_t predict_unit;        /* 0: after a failed test run k more cycles
                                    (k = 1, 2, 3, ...) before the next, as
                                    tgmg.in.cc:282-292 (default); 1: one more
                                    cycle per test, for solvers whose test is
                                    cheap next to a cycle (the predicted
                                    count still adapts as tgmg_predict's) */
    int32_t reserved[4];
} stz_mg_config;

/* Result of the last projection solve. */
typedef struct stz_solve_info {
    int32_t status;              /* STZ_OK / STZ_ERR_NOT_CONVERGED / STZ_ERR_NAN */
    int32_t cycles;              /* V-cycles counted as tgmg_predict does */
    int32_t total_cycles;        /* V-cycles actually run (incl. extra) */
    int32_t refinements;         /* FP64 residual evaluations (mixed mode) */
    double residual;             /* mean-square residual (tgmg.hh:146) */
    double tolerance;            /* threshold it was compared against */
} stz_solve_info;

typedef struct stz_sim stz_sim;   /* opaque */

/* Last error message of the calling thread ("" if none). */
const char *stz_last_error(void);

/* Number of visible devices (0 without a GPU; never fails). */
int stz_device_count(int *count);

/* shear_sim::shear_sim (shear_sim.cc:26-44) + qs_multi/tgmg construction
 * (qs_multi.cc:9-15, tgmg.in.cc:10-30).  Fields start zeroed. */
int stz_sim_create(const stz_sim_params *p, int precision, stz_sim **out);
/* shear_sim::~shear_sim (shear_sim.cc:47-52). */
int stz_sim_destroy(stz_sim *sim);

/* Run all following work on this HIP stream (hipStream_t passed as void*;
 * NULL = the library's own stream). */
int stz_sim_set_stream(stz_sim *sim, void *hip_stream);
int stz_sim_synchronize(stz_sim *sim);

/* Field access (public `fm`, shear_sim.hh:79).  Plane = m*(n+2) doubles;
 * field is one of STZ_F_U .. STZ_F_DDEV, or STZ_F_CU.  The device pointer is
 * the field's current plane: p, q, s, tau, chi, X and Y alternate between two
 * planes at every step, so a caller that keeps device pointers queries them
 * again after stepping. */
int stz_sim_upload_field(stz_sim *sim, int field, const double *host_plane);
int stz_sim_download_field(stz_sim *sim, int field, double *host_plane);
int stz_sim_field_device_ptr(stz_sim *sim, int field, double **dev_plane);
/* Public `vel` / `src` (shear_sim.hh:82-85): m*(n+1) interleaved vec2. */
int stz_sim_upload_vel(stz_sim *sim, const double *host_vec2);
int stz_sim_download_vel(stz_sim *sim, double *host_vec2);
int stz_sim_download_src(stz_sim *sim, double *host_vec2);
int stz_sim_vel_device_ptr(stz_sim *sim, double **dev_vec2);

/* Public `time` (shear_sim.hh:87) and `chi_inf` (shear_sim.hh:92). */
int stz_sim_set_time(stz_sim *sim, double t);
int stz_sim_get_time(stz_sim *sim, double *t);
int stz_sim_set_chi_inf(stz_sim *sim, double chi_inf);

/* shear_sim::init_fields (shear_sim.cc:60-114), computed on the device. */
int stz_sim_init_fields(stz_sim *sim, int chi_case, double tem_base, double tem_delta);

/* shear_sim::step_forward_quasistatic (shear_sim.cc:280-291), nsteps times.
 * Stops at the first failed solve and returns its status. */
int stz_sim_step_quasistatic(stz_sim *sim, double dt, int nsteps);
/* shear_sim::advection_step (shear_sim.cc:296-361), chi_diffusion included. */
int stz_sim_advection_step(stz_sim *sim, double dt);
/* stz_sim_advection_step `count` times, asynchronously on the simulation's
 * stream (no counterpart in the reference: for timing the explicit kernels
 * between a caller's own events). */
int stz_sim_advection_launch(stz_sim *sim, double dt, int count);
/* shear_sim::projection_step (shear_sim.cc:391-438). */
int stz_sim_projection_step(stz_sim *sim, double dt);
/* shear_sim::set_boundaries (shear_sim.cc:552-581). */
int stz_sim_set_boundaries(stz_sim *sim);

/* Multigrid controls (tgmg.hh:331-441, tgmg_predict.hh). */
int stz_sim_get_mg_config(stz_sim *sim, stz_mg_config *cfg);
int stz_sim_set_mg_config(stz_sim *sim, const stz_mg_config *cfg);

/* The adaptive V-cycle count's state, tgmg_predict::lim (tgmg_predict.hh:
 * 12-14): the predicted cycles of the next solve times predict_mult.  A
 * caller that restores a simulation's state (a tuner probing settings from
 * one state) restores this with it. */
int stz_sim_get_predictor(stz_sim *sim, int32_t *lim);
int stz_sim_set_predictor(stz_sim *sim, int32_t lim);

int stz_sim_last_solve(stz_sim *sim, stz_solve_info *info);
/* shear_sim::avg_iters -> tgmg_predict::avg_iters (tgmg_predict.hh:28-33):
 * mean V-cycles per solve since the last call; resets the counters. */
int stz_sim_avg_iters(stz_sim *sim, double *avg);
/* Hierarchy: number of levels (top included) and their node dimensions
 * (tgmg.in.cc:10-30); dims holds 2*max_levels ints. */
int stz_sim_mg_hierarchy(stz_sim *sim, int max_levels, int *nlevels, int *dims);

/* qs_multi::init -> tgmg::setup (qs_multi.cc:31-54, tgmg.hh:331-334):
 * rebuild the operator hierarchy for viscosity/dt = viscdt if it changed. */
int stz_mg_setup(stz_sim *sim, double viscdt);
/* tgmg::solve_v_cycle(tgmg_predict&) on the current src/vel
 * (tgmg.in.cc:266-301): adaptive V-cycles until the tolerance is met. */
int stz_mg_solve(stz_sim *sim, stz_solve_info *info);
/* One V-cycle (tgmg.in.cc:313-328) on the current src/vel. */
int stz_mg_v_cycle(stz_sim *sim);
/* tgmg_base::l2_error (tgmg.hh:146): mean-square residual of src - A vel. */
int stz_mg_l2_error(stz_sim *sim, double *err);
/* tgmg_base::gauss_seidel (tgmg.in.cc:127-165): `sweeps` smoothing sweeps
 * on hierarchy level `level` (0 = top) of the hierarchy the solve uses (the
 * FP32 one in the mixed form).  Asynchronous on the simulation's stream, so
 * a caller can bracket it with its own events. */
int stz_mg_smooth(stz_sim *sim, int level, int sweeps);
/* tgmg::fmg_pre (tgmg.in.cc:369-399) on the current src/vel: the
 * full-multigrid pass (smoothing on the way down, the coarse levels solved,
 * prolongation, smoothing and a V-cycle from each level on the way up,
 * ending with the prolongation into vel).  Runs on the FP64 hierarchy in
 * both precision forms, with the per-colour kernels. */
int stz_mg_fmg_pre(stz_sim *sim);
/* tgmg_base::clear_z (tgmg.in.cc:792-797): vel = 0. */
int stz_mg_clear_z(stz_sim *sim);
/* One smoothing launch of the V-cycle on hierarchy level `level`, alone,
 * `count` times (no counterpart in the reference: for timing the kernels a
 * solve runs).  mode 1: the pre-smoothing (nu1 sweeps, on the correction
 * equation from zero in the mixed form, with the restriction of the residual
 * into level+1); mode 2: the post-smoothing (prolongation of level+1, nu2
 * sweeps, and on the mixed form's top level the FP64 correction vel += e).
 * Needs the tiled kernels (stz_mg_config.kernels = 0) on that level;
 * asynchronous on the simulation's stream. */
int stz_mg_vcycle_launch(stz_sim *sim, int level, int mode, int count);
/* Upload a source array (m*(n+1) vec2) for a standalone solve. */
int stz_mg_upload_src(stz_sim *sim, const double *host_vec2);

/* The mean-square residual of stz_mg_l2_error, allocating nothing.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  vel and src are device arrays of
 * m*(n+1) interleaved vec2 nodes; the operator is qs_multi's
 * (qs_multi.cc:31-54) for the grid and moduli of p at viscosity/dt =
 * viscdt; the result, bit-identical to stz_mg_l2_error's on the same arrays,
 * goes to the device double *out.  Asynchronous on hip_stream (NULL: the
 * null stream).  Partial sums live in a fixed device array of the library
 * (one per GPU, held until the process ends), so m*(n+1) may not exceed
 * 16,777,216 (STZ_ERR_ARG beyond), and calls that overlap on one GPU race.
 * The first call on each device looks the array up (hipGetSymbolAddress);
 * make that call outside any graph capture. */
int stz_qs_l2_error_noalloc(const stz_sim_params *p, double viscdt, const double *vel, const double *src,
                            double *out, void *hip_stream);

/* ---- direct step, diagnostics, initial conditions ------------------------ */

/* shear_sim::step_forward (shear_sim.cc:185-276), the direct (dynamic)
 * explicit step, nsteps times. */
int stz_sim_step_forward(stz_sim *sim, double dt, int nsteps);
/* shear_sim::tractions (shear_sim.cc:1003-1028): out4 = (bx, by, tx, ty). */
int stz_sim_tractions(stz_sim *sim, double *out4);
/* shear_sim::qs_measure (shear_sim.cc:733-775); the integrand goes to the
 * STZ_F_CU plane (rows 1..n-2). */
int stz_sim_qs_measure(stz_sim *sim, int staggered, double *Q, double *max_qs);
/* shear_sim::compute_strain (shear_sim.cc:778-824): deviatoric modulus of the
 * Hencky strain into the STZ_F_CU plane (rows 0..n-1).  corrected = 0 gives
 * the reference's result, with its two slips: the y difference on the
 * interior rows takes both ends from row j+1 (shear_sim.cc:804), which yields
 * 0 there, and the seam column i = m-1 subtracts the periodic shift of its
 * right neighbour (:793) instead of adding it.  corrected = 1 uses rows j+1
 * and j-1 on the interior rows and adds the shift. */
int stz_sim_compute_strain(stz_sim *sim, int corrected);
/* shear_sim::l2_comparison (shear_sim.cc:460-502) of a against b (same size,
 * same device): out4 = velocity, stress, chi, reference-map squared norms. */
int stz_sim_l2_comparison(stz_sim *a, stz_sim *b, double *out4);
/* shear_sim::initialize_random (shear_sim.cc:932-985): the noise is drawn
 * from the C library's rand() in the reference's order (seed it with srand
 * first, as the reference's drivers do); the Gaussian smoothing runs on the
 * device.  Writes u, v, X, Y on rows 0..n and chi, p, q, s, tau on rows
 * 0..n-1. */
int stz_sim_initialize_random(stz_sim *sim, double l, double tem_avg, double tem_stdev);
/* shear_sim::initialize_chi_bicubic (shear_sim.cc:832-913): chi on rows
 * 0..n-1 from the mm x nn row-major array arr (lo + sca * interpolant); the
 * other fields reset on rows -1..n. */
int stz_sim_initialize_chi_bicubic(stz_sim *sim, int mm, int nn, const double *arr, double lo, double sca);

/* Per-cell decision bits of the last advection step (shear_sim.cc:296-349,
 * 513-548), one uint32 per cell (i,j), j in [0,n), i fastest.  "ENO" bits are
 * set when eno2 (shear_sim.cc:653-655) took the one-sided second-order
 * branch 3p1-4p2+p3 rather than the centred p0-p2. */
enum {
    STZ_D_NODE_UPX = 1u << 0,    /* j>0: uc=f.u > 0 (x upwind side) */
    STZ_D_NODE_UPY = 1u << 1,    /* j>0: vc=f.v > 0 */
    STZ_D_NODE_ENO_XX = 1u << 2, /* eno2 branch, X along x */
    STZ_D_NODE_ENO_YX = 1u << 3, /* eno2 branch, Y along x */
    STZ_D_NODE_ENO_XY = 1u << 4, /* eno2 branch, X along y */
    STZ_D_NODE_ENO_YY = 1u << 5, /* eno2 branch, Y along y */
    STZ_D_NODE_ONESIDED = 1u << 6, /* y derivative first-order one-sided (wall) */
    STZ_D_CELL_UPX = 1u << 7,    /* staggered uc > 0 */
    STZ_D_CELL_UPY = 1u << 8,    /* staggered vc > 0 */
    STZ_D_CELL_ENO_X0 = 9,       /* bit index: eno2 branch along x, p,q,s,tau,chi -> bits 9..13 */
    STZ_D_CELL_ENO_Y0 = 14,      /* along y -> bits 14..18 */
    STZ_D_CELL_ONESIDED = 1u << 19, /* staggered y derivative one-sided (wall) */
    STZ_D_YIELD = 1u << 20,      /* Dplastic took the sbar>=1 branch on entry */
    STZ_D_SMALLDEV = 1u << 21,   /* sbar < 1e-12 cut-off (shear_sim.cc:522) */
    STZ_D_YIELD_LAST = 1u << 22, /* the last plastic loop trip took the sbar>=1 branch */
    STZ_D_SUBSTEP_SHIFT = 24     /* bits 24..31: plastic loop trips, saturating */
};

/* Recording is off by default (it costs one extra store per cell). */
int stz_sim_record_decisions(stz_sim *sim, int enable);
int stz_sim_download_decisions(stz_sim *sim, uint32_t *host_cells);

/* Frozen-decision forward: the next nsteps advection steps of this
 * simulation (stz_sim_advection_step, stz_sim_step_quasistatic, the
 * launch forms) take every decision listed above from host_cells instead of
 * evaluating it, step k from host_cells[k*m*n ..] in the layout of
 * stz_sim_download_decisions; a cell whose substep count is 255 (saturated)
 * evaluates its plastic loop as usual.  Replaying the record of a run from
 * the same state repeats that run bit for bit; from a nearby state it gives
 * the smooth continuation of the recorded branches, the function whose
 * derivative the adjoint (stz_adjoint_*) computes.  A step after the last
 * recorded one fails with STZ_ERR_STATE; nsteps = 0 ends the replay.  The
 * record is copied to the device before the call returns.  No counterpart in
 * the reference (which evaluates every decision, shear_sim.cc:296-349,
 * 513-548, 653-655). */
int stz_sim_replay_decisions(stz_sim *sim, int nsteps, const uint32_t *host_cells);

/* ======================================================================== *
 * Free-boundary simulation: a body bounded by the zero contour of a level
 * set inside a rectangular box, the reference's `bd_sim` (sb/main/
 * bd_sim.hh:27-190) with its problem class `qs_multi_bd`
 * (sb/main/qs_multi_bd.hh:11-72) and the narrow-band level set `levelset`
 * (levelset/levelset.hh:19-114).
 *
 * Grids: m x n cells (cell (i,j) centred at (ax+(i+1/2)dx, ay+(j+1/2)dy))
 * and (m+1) x (n+1) nodes ((ax+i dx, ay+j dy)).  Cell planes are m*n FP64
 * row-major (p, q, s, tau, chi, ddev and the level set phi), node planes
 * (m+1)*(n+1) (u, v, X, Y); the multigrid velocity/source are interleaved
 * vec2 over the nodes.  phi < 0 inside the body; the band is |phi| < 5 dx
 * and cells outside it hold phi = +-5 dx, as in the reference
 * (bd_sim.cc:40, levelset.cc:412-418).
 *
 * Cell status (levelset.hh s[], one int per cell): 0 far inside, 1 inside
 * and next to the band, 2 inside band, 3 inside and next to the interface,
 * 4 outside next to the interface, 5 outside band, 6 outside next to the
 * band, 7 far outside (the codes of levelset.cc:458-520).  Node status
 * (bd_sim::cc, bd_sim.cc:125-172): 0 inside, 7 outside.
 * ======================================================================== */
typedef struct stz_bd stz_bd;       /* opaque */

/* Constructor arguments of bd_sim (bd_sim.hh:147-151, bd_sim.cc:25-53). */
typedef struct stz_bd_params {
    int32_t m, n;                   /* cells in x and y (both bounded) */
    double ax, bx, ay, by;          /* box */
    double mu, K, viscosity;        /* shear/bulk modulus, viscous damping */
    double chi_len;                 /* Dpl-mediated chi diffusion length (0: off) */
    double filter_stress;           /* stress filtering factor (advection_step) */
    double t_scale, tmult, adapt_fac;
    stz_model_params stz;
} stz_bd_params;

enum { STZ_BD_P = 0, STZ_BD_Q, STZ_BD_S, STZ_BD_TAU, STZ_BD_CHI, STZ_BD_DDEV, STZ_BD_PHI, STZ_BD_NCELL };
enum { STZ_BD_U = 0, STZ_BD_V, STZ_BD_X, STZ_BD_Y, STZ_BD_NNODE };

int stz_bd_create(const stz_bd_params *p, stz_bd **out);
int stz_bd_destroy(stz_bd *b);
int stz_bd_set_stream(stz_bd *b, void *hip_stream);
int stz_bd_synchronize(stz_bd *b);

/* Level set from the signed function phi (m*n cell values, < 0 inside):
 * levelset::build_band (levelset.cc:209-345) -- the interface cells get their
 * distance to the zero contour by the Newton closest-point iteration, the
 * rest of the band by the eikonal update of the reference's fast march --
 * then bd_sim::set_staggered_status (bd_sim.cc:125-172). */
int stz_bd_set_level_set(stz_bd *b, const double *phi);
/* A level set that is already a narrow band, with its cell statuses (codes
 * above), taken as they are -- a saved state or another implementation's;
 * node statuses are derived (bd_sim.cc:125-172). */
int stz_bd_upload_level_set(stz_bd *b, const double *phi, const int32_t *cell_status);
/* bd_sim::init_bar (bd_sim.cc:63-123): geometry case ca (1 bar, 2 four
 * rounded notches, other: ellipse), fields inside the body, extrapolation. */
int stz_bd_init_bar(stz_bd *b, int ca);

int stz_bd_upload_cells(stz_bd *b, int k, const double *host);    /* k < STZ_BD_PHI */
int stz_bd_download_cells(stz_bd *b, int k, double *host);        /* k <= STZ_BD_PHI */
int stz_bd_upload_nodes(stz_bd *b, int k, const double *host);
int stz_bd_download_nodes(stz_bd *b, int k, double *host);
int stz_bd_download_vel(stz_bd *b, double *host);                 /* multigrid solution */
int stz_bd_download_src(stz_bd *b, double *host);                 /* multigrid source */
/* status arrays (int32, either pointer may be NULL) */
int stz_bd_download_status(stz_bd *b, int32_t *cell_status, int32_t *node_status);
int stz_bd_set_time(stz_bd *b, double t);
int stz_bd_get_time(stz_bd *b, double *t);

/* bd_sim::advection_step (bd_sim.cc:374-469): level set moved
 * (levelset::move_and_smooth, levelset.hh:159-205) then the explicit
 * reference-map / stress / chi update. */
int stz_bd_advection_step(stz_bd *b, double dt);
/* bd_sim::projection_step (bd_sim.cc:510-545): qs_multi_bd::init
 * (qs_multi_bd.cc:26-156), the multigrid solve to the mean-square residual
 * 1e-20 (qs_multi_bd.cc:16, tgmg.in.cc:242-255), velocity and stress
 * update. */
int stz_bd_projection_step(stz_bd *b, double dt);
/* bd_sim::post_process (bd_sim.cc:256-259): extrapolate the cell fields
 * (all_fields: also the node fields) across the interface, then the box
 * boundary conditions. */
int stz_bd_post_process(stz_bd *b, int all_fields);
/* bd_sim::step_forward_quasistatic (bd_sim.cc:240-254), nsteps times. */
int stz_bd_step_quasistatic(stz_bd *b, double dt, int nsteps);
int stz_bd_last_solve(stz_bd *b, stz_solve_info *info);

/* Diagnostics of single pieces of the step, on the current state:
 * the interface speed of levelset::move_and_smooth (levelset.hh:159-205) at
 * every interface cell (0 elsewhere), without moving anything; and the
 * operator of qs_multi_bd::init (qs_multi_bd.cc:26-156) for time step dt:
 * row type per node (-1 interior, -2 box edge, -3 outside, 0 free surface),
 * the nine 2x2 blocks a_dl..a_ur per node (36 doubles, block order dl dc dr
 * cl cc cr ul uc ur, each a b c d row-major) and the source.  Any pointer
 * may be NULL. */
int stz_bd_interface_velocity(stz_bd *b, double *vp);
/* The band rebuilt around the unmoved interface (levelset::move_and_smooth
 * with dt = 0, levelset.hh:159-205): interface cells keep phi, the rest of
 * the band is marched again; statuses and node statuses follow. */
int stz_bd_remarch(stz_bd *b);
/* Solver settings: pre/post Gauss-Seidel sweeps per level of the V-cycle
 * preconditioner (default 1, 1 as tgmg.hh:439), sweeps on its coarsest
 * level (20), the
 * BiCGStab iteration cap (500) and the mean-square residual threshold
 * (1e-20, qs_multi_bd.cc:16). */
int stz_bd_set_solver(stz_bd *b, int nu1, int nu2, int bottom_sweeps, int max_iter, double tol);
/* Precision of the projection solve's preconditioner: STZ_PREC_FP64 (the
 * default) or STZ_PREC_MIXED (each V-cycle in FP32 on FP32 copies of the
 * Galerkin levels; BiCGStab, its residuals and the convergence test stay
 * FP64, so the solve stops at the same threshold).  No counterpart in the
 * reference, which solves in FP64 throughout (qs_multi_bd.cc:12-196). */
int stz_bd_set_precision(stz_bd *b, int precision);
/* Work counts of the last step: out4 = {march rounds (both sides), cell
 * extrapolation waves, node extrapolation waves, BiCGStab iterations}. */
int stz_bd_step_stats(stz_bd *b, int32_t *out4);
int stz_bd_operator(stz_bd *b, double dt, int32_t *row_type, double *blocks36, double *src);

/* Per-node body force (interleaved vec2, (m+1)*(n+1) nodes; NULL removes
 * it): the value that extra_force::force(i,j) returns (extra_force.hh:5-28),
 * added as idt*force to the interior rows' source (qs_multi_bd.cc:71). */
int stz_bd_set_body_force(stz_bd *b, const double *force);
/* Edge velocity: u = u_left on the nodes of column 0 and u_right on column
 * m, the reference maps there following the edge (X = ax + integral of
 * u_left, likewise at bx).  Default 0, which is the reference's
 * set_boundaries (bd_sim.cc:732-761); no counterpart in the reference, it
 * loads a body held between two moving grips. */
int stz_bd_set_edge_velocity(stz_bd *b, double u_left, double u_right);

/* ---- ensembles of independent simulations --------------------------------
 * No counterpart in the reference, which runs one shear_sim per process
 * (drivers such as shear_random.cc:106-121 loop over seeds in separate
 * runs).  An ensemble owns `count` simulations of one precision, each with
 * its own parameters (grid size included), and steps them together: every
 * kernel of stz_sim_step_quasistatic runs as ONE launch over all members,
 * each member's adaptive V-cycle count (tgmg_predict) is decided on the
 * device, and a step replays from a HIP graph with one host
 * synchronisation.
 * Every member computes exactly what stz_sim_step_quasistatic computes for
 * it alone.  Requirements: all members have the same stz_mg_config, with
 * kernels = 0, nu1 >= 1 and ceil(nu1/2) + ceil(nu2/2) even (nu1 = nu2 is
 * always fine), and every multigrid level a member's V-cycle smooths (the
 * levels above its direct level, or above its workgroup solver) is covered
 * by the tiled kernels (widths a multiple of 32: true for every side in
 * {64, 128, 256, 384, 512}); stz_ensemble_step_quasistatic reports anything
 * else as STZ_ERR_ARG. */
typedef struct stz_ensemble stz_ensemble;   /* opaque */

/* count simulations, member k from params[k] (as stz_sim_create). */
int stz_ensemble_create(const stz_sim_params *params, int count, int precision, stz_ensemble **out);
int stz_ensemble_destroy(stz_ensemble *e);
int stz_ensemble_size(stz_ensemble *e, int *count);
/* Borrowed handle of member k for every stz_sim_* call (fields, time,
 * initial conditions, diagnostics, single steps); never destroy it. */
int stz_ensemble_member(stz_ensemble *e, int k, stz_sim **member);
/* Multigrid settings of every member (stz_sim_set_mg_config on each). */
int stz_ensemble_set_mg_config(stz_ensemble *e, const stz_mg_config *cfg);
/* Replay steps from HIP graphs (1, default) or launch them directly (0);
 * the launches and results are the same. */
int stz_ensemble_use_graphs(stz_ensemble *e, int on);
/* nsteps quasi-static steps of every member.  Stops after the first step in
 * which a member's solve fails and returns that status; the failing member
 * is left as stz_sim_step_quasistatic leaves it, the others have completed
 * the step.  If instead a launch or copy fails inside a step (STZ_ERR_HIP),
 * every member's fields, velocity and time are undefined until the caller
 * sets them again; the next call rebuilds the ensemble's tables. */
int stz_ensemble_step_quasistatic(stz_ensemble *e, double dt, int nsteps);
/* Counters since creation: out4 = {steps, cycle slots launched, slots
 * launched after the planned ones, graph captures}. */
int stz_ensemble_stats(stz_ensemble *e, int64_t *out4);
/* The tile-shape tuner: when the tables are built (first step, new dt or
 * settings), each batched smoother launch is timed with 32x16 tiles and,
 * where the level widths allow, 64x32 tiles, and the faster kept (the shape
 * changes speed only, never results).  out2 = {launches kept at 32x16,
 * launches kept at 64x32} of the current tables. */
int stz_ensemble_tile_choice(stz_ensemble *e, int32_t *out2);

/* ---- adjoint of the FP64 step ------------------------------------------
 * No counterpart in the reference.  Reverse-mode gradients of a function of
 * the final state with respect to the initial state and the STZ parameters,
 * through nsteps of stz_sim_step_quasistatic.  Each discrete choice of the
 * forward step (upwind side, ENO branch, one-sided wall rule, yield branch,
 * plastic substep count) is the forward's own, taken again on recomputation;
 * the multigrid solve is differentiated as an exact solve (the adjoint
 * solve reuses the forward operator, whose interior block is symmetric);
 * states are kept every `checkpoint_every` steps and each segment is
 * recomputed during the reverse sweep.  FP64 form; every STZ law and
 * truncation model (the plastic loop is differentiated in dual numbers
 * through the step's own law code) and chi diffusion. */
typedef struct stz_adjoint stz_adjoint;   /* opaque */

/* Replace the STZ law's parameters (c0, Omegaeps0, ...; stz_model.hh:22-26)
 * of a simulation; chi_inf of the plastic term is stz_sim_set_chi_inf. */
int stz_sim_set_model(stz_sim *sim, const stz_model_params *mp);
int stz_adjoint_create(stz_sim *sim, int checkpoint_every, stz_adjoint **out);
int stz_adjoint_destroy(stz_adjoint *a);
/* Step the simulation nsteps times, keeping checkpoints. */
int stz_adjoint_forward(stz_adjoint *a, double dt, int nsteps);
/* lam_final: dL/d(final state), ten planes in field order (STZ_F_U ..
 * STZ_F_DDEV, m*(n+2) doubles each).  Writes dL/d(initial state) in the
 * same layout and grad3 = dL/d(c0, Omegaeps0, chi_inf).  The simulation is
 * left in its state after the forward run. */
int stz_adjoint_backward(stz_adjoint *a, const double *lam_final, double *lam_init, double *grad3);
/* Relative residual of each adjoint solve (default 1e-10). */
int stz_adjoint_set_solve_accuracy(stz_adjoint *a, double rel);
/* out6 = {forward seconds, backward seconds, recomputed steps, adjoint
 * solves, adjoint V-cycles, forward V-cycles}, accumulated. */
int stz_adjoint_stats(stz_adjoint *a, double *out6);

#ifdef __cplusplus
}
#endif
#endif /* STZ_AMD_H */
