// bd.h -- device-side structures and launchers of the free-boundary
// (level-set) quasi-static step: bd_sim (sb/main/bd_sim.{hh,cc}),
// qs_multi_bd (sb/main/qs_multi_bd.{hh,cc}) and the narrow-band level set
// (levelset/levelset.{hh,cc}, heap.cc, traverse.hh).
//
// Layout in HBM (structure of arrays, FP64 unless noted):
//   cells  m x n,        index c = i + m*j : p q s tau chi ddev (two buffers,
//                        the advection writes the second), phi, vp, status
//                        (int8, the reference's codes 0..7)
//   nodes (m+1)x(n+1),   index k = i + me*j: u v X Y (two buffers), status
//                        (int8, 0 inside / 7 outside), vel and src (vec2),
//                        row type (int32), body force (vec2, optional)
// The reference keeps both grids in one c_field array indexed on the node
// grid (bd_sim.cc:37-41); splitting them gives each plane unit stride.
#pragma once

#include "device.h"

namespace stz {

struct BdConst {
    int m, n, me, ne;
    double ax, ay, bx, by, dx, dy, xsp, ysp;
    double mu, mu_inv, K, viscosity, chi_len, filter_stress, t_scale, tmult, adapt_fac;
    double chi_inf, TZ, thetainv, bc_smooth;
    stz_model_params stz;
    // level set on the cell centres (bd_sim.cc:40): origin, band limit
    double lax, lay, phi_lim;
    // edge grips (extension; 0 = the reference's fixed box)
    double u_left, u_right, x_left, x_right;
};

struct BdCells {
    double *p, *q, *s, *tau, *chi, *ddev;
};
struct BdNodes {
    double *u, *v, *X, *Y;
};

// level set (bd_levelset.hip)
int bd_ls_scan(const BdConst &c, const double *phi, int8_t *st, hipStream_t s);
int bd_ls_newton_distance(const BdConst &c, const double *phi, const int8_t *st, double *out, int *fail,
                          hipStream_t s);
int bd_ls_interface_velocity(const BdConst &c, const double *phi, const int8_t *st, const BdNodes &nd,
                             double *vp, hipStream_t s);
int bd_ls_advance(const BdConst &c, double *phi, const double *vp, int8_t *st, double dt, hipStream_t s);
// reinitialise the band outward from the interface cells, positive side
// first; scratch: one double and ints for the round loop (device), flag on
// the host (coherent pinned); returns rounds in *rounds
int bd_ls_march(const BdConst &c, double *phi, int8_t *st, double delta, unsigned long long *dmin,
                int *counter, volatile int *hflag, int *rounds, hipStream_t s);
int bd_ls_finish(const BdConst &c, double *phi, double *vp, int8_t *st, int build, hipStream_t s);
int bd_ls_node_status(const BdConst &c, const double *phi, const int8_t *st, int8_t *nst, hipStream_t s);
// extrapolation waves (traverse.hh:176-240, 62-174)
// counter: 16 ints of device scratch; list: 3 max(cells, nodes) ints; hflag:
// 16 ints of host-visible memory
int bd_extrap_cells(const BdConst &c, const double *phi, const int8_t *st, const BdCells &f, int *done,
                    int *counter, int *list, volatile int *hflag, int *waves, hipStream_t s);
int bd_extrap_nodes(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdNodes &f,
                    int *done, double *key, int *counter, int *list, volatile int *hflag, int *waves,
                    hipStream_t s);
// both, their waves sharing launches; each walk its own counters and lists
int bd_extrap_both(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdCells &fc,
                   const BdNodes &fn, int *done_c, int *done_n, double *key, int *cnt_c, int *cnt_n, int *list_c,
                   int *list_n, volatile int *hflag, int *waves_c, int *waves_n, hipStream_t s);

// explicit step, operator, correction (bd_step.hip)
int bd_advect(const BdConst &c, const int8_t *st, const int8_t *nst, const BdCells &o, const BdNodes &on,
              const BdCells &w, const BdNodes &wn, double dt, hipStream_t s);
int bd_chi_diffusion(const BdConst &c, const int8_t *st, const BdCells &o, const BdCells &w, double dt,
                     hipStream_t s);
int bd_boundaries(const BdConst &c, const BdNodes &f, double edge_shift, hipStream_t s);
// qs_multi_bd::init: row types, sources and the free-boundary rows (8
// preconditioned 2x2 blocks each, appended through *nfree)
int bd_operator(const BdConst &c, const double *phi, const int8_t *st, const int8_t *nst, const BdCells &f,
                const BdNodes &nd, const double *force, double dt, int *rowt, double *src, double *fmats,
                int fcap, int *nfree, hipStream_t s);
int bd_correct(const BdConst &c, const int8_t *st, const BdCells &f, const BdNodes &nd, const double *vel,
               double dt, hipStream_t s);
int bd_init_bar_phi(const BdConst &c, int ca, double *phi, hipStream_t s);
int bd_init_bar_fields(const BdConst &c, int ca, const int8_t *st, const int8_t *nst, const BdCells &f,
                       const BdNodes &nd, hipStream_t s);

// operator constants of qs_multi_bd::init (qs_multi_bd.cc:26-58)
struct BdOp {
    double Kmx, Kmy, mx, my, qK, cxx, cyy;
    double bcs, ocs, ncs;
};
BdOp bd_op_constants(const BdConst &c, double dt);

}  // namespace stz
