// diag.h -- launchers of the diagnostics and initial-condition kernels
// (diag.hip).  Outputs of reductions go to `out`, which may be coherent
// pinned host memory.
#pragma once

#include "device.h"

namespace stz {

// (bx, by, tx, ty) of shear_sim::tractions
int diag_tractions(const SimConst &c, const FieldPtrs &f, double *out, hipStream_t st);
// number of partials diag_l2 / diag_qs may write (x4 doubles)
int diag_partials_needed(const SimConst &c);
// qs_measure integrand into cu; out = (sum, max) of it
int diag_qs(const SimConst &c, const FieldPtrs &f, const double *vel, int staggered, double *cu, double *partial,
            double *out, hipStream_t st);
int diag_strain(const SimConst &c, const FieldPtrs &f, int corrected, double *cu, hipStream_t st);
// out = the four unscaled sums of l2_comparison
int diag_l2(const SimConst &c, const FieldPtrs &a, const FieldPtrs &b, double *partial, double *out, hipStream_t st);
int diag_random_fields(const SimConst &c, const FieldPtrs &f, int cut, const double *w, const double *noise,
                       double *tmp, double chi0, double chi1, hipStream_t st);
// d holds arr, xd, yd, xyd (mm*nn each)
int diag_bicubic(const SimConst &c, const FieldPtrs &f, int mm, int nn, const double *d, double lo, double sca,
                 hipStream_t st);

}  // namespace stz
