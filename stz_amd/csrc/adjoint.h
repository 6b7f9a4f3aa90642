// adjoint.h -- launchers of the adjoint kernels (adjoint.hip).
#pragma once

#include "device.h"

namespace stz {

int adj_zero(double *p, size_t count, hipStream_t st);
// lambda after the step (L) -> lambda after the advection (A, all planes
// written) and g = dL/dz on the node grid (walls zero); coef: 4 doubles per
// cell of scratch
int adj_launch_correct(const SimConst &c, const FieldPtrs &L, const FieldPtrs &A, double *coef, double *g,
                       double dt, hipStream_t st);
// A.{p,q,s,tau} += source^T w
int adj_launch_source(const SimConst &c, const FieldPtrs &A, const double *w, double dt, hipStream_t st);
// G (zeroed) += chi_diffusion^T A.chi with respect to the pre-step chi (o)
// and A.ddev += the same with respect to the advection's new ddev (dnew)
int adj_launch_chi_diffusion(const SimConst &c, const FieldPtrs &o, const double *dnew, const FieldPtrs &A,
                             const FieldPtrs &G, double dt, hipStream_t st);
// G += advect'(o)^T A; grad[0..2] += d/d(c0, Omegaeps0, chi_inf)
int adj_launch_advect(const SimConst &c, const FieldPtrs &o, const FieldPtrs &A, const FieldPtrs &G, double dt,
                      double *grad, hipStream_t st);

}  // namespace stz
