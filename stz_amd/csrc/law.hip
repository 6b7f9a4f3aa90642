// law.hip -- point evaluation of the STZ plasticity law on the host:
// stz_dynamics::Dplastic (sb/main/stz_model.hh:28, the four laws of
// :50-207) for callers that evaluate the law outside a step, such as the
// reference's dplas_test.cc:17 table or the C++ facade's Dplastic.  It runs
// the same source the step's kernels run on the device (stz_law.h, FMA
// contraction off), compiled for the host, where exp, cosh, sinh, tanh and
// sqrt are the C library's, as in the reference.
#include <cmath>

#include "stz_law.h"

extern "C" int stz_model_dplastic(const stz_model_params *mp, double sbar, double chi, double *dpl, double *dchi1,
                                  double *dchi2) {
    if (!mp || !dpl || !dchi1 || !dchi2) {
        stz::set_error("null argument");
        return STZ_ERR_ARG;
    }
    if (mp->kind < STZ_MODEL_LINEAR_ATHERMAL || mp->kind > STZ_MODEL_NONLINEAR_ATHERMAL) {
        stz::set_error("unknown STZ law %d", mp->kind);
        return STZ_ERR_ARG;
    }
    double d1 = 0, d2 = 0;
    *dpl = stz::dplastic(*mp, 1 / mp->theta, sbar, chi, d1, d2);
    *dchi1 = d1;
    *dchi2 = d2;
    return STZ_OK;
}
