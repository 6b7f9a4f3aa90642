// stz_law.h -- the STZ plasticity law (sb/main/stz_model.hh:9-207) and the
// adaptive plastic substepping (shear_sim.cc:513-548, bd_sim.cc:693-730),
// shared by the periodic-shear and free-boundary kernels (and, for the law
// alone, by the host point evaluation stz_model_dplastic).  Expressions keep
// the reference's evaluation order; FMA contraction is off.
#pragma once

#include "device.h"

#pragma clang fp contract(off)

namespace stz {

// ---------------------------------------------------------------------------
// STZ law (stz_model.hh:9-207)
// ---------------------------------------------------------------------------

// The law is written once over the scalar type S: double for the step, or a
// Dual (dual.h) for the adjoint's derivatives with respect to s-bar, chi,
// Omega*eps0 and c0 (passed separately as Oe and c0 so they can carry
// derivatives; for S = double they are mp.Omegaeps0 and mp.c0).  Every
// expression is the same for both, so the values agree bit for bit.
template <class S>
__host__ __device__ static S dpl_common_t(const stz_model_params &mp, const S &sbar, const S &chi, S &dchi1,
                                          S &dchi2, const S &sC1, const S &sC2, const S &sC, const S &sT,
                                          const S &c0) {
    // stz_model.hh:30-47
    const double s0 = 1;
    S sbart = sbar > 1e-12 ? sbar : S(1e-12), bfac = exp(-1 / chi),
      pM = 1 + sbart * sT / s0 + mp.rho / (2 * sC2),
      M = s0 / (2 * sbart) * (pM - sqrt(pM * pM - 4 * sbart * sT / s0)),
      Dpl = bfac * sC * (sT - M) / mp.tau0,
      Gamma = 2 * mp.tau0 * Dpl * sbar / (s0 * bfac);
    bfac /= mp.tau0 * c0;
    dchi1 = bfac * Gamma;
    dchi2 = bfac * mp.kappa * sC1 * mp.rho;
    return Dpl;
}

// yield: -1 evaluates the athermal laws' yield test sbar < 1; 0 or 1 takes
// the recorded branch (below or at/above yield) instead, whatever sbar is
template <class S>
__host__ __device__ static S dplastic_t(const stz_model_params &mp, double thetainv, const S &sbar, const S &chi,
                                        const S &Oe, const S &c0, S &dchi1, S &dchi2, int yield = -1) {
    const double s0 = 1;
    const bool below = yield < 0 ? sbar < 1 : yield == 0;
    switch (mp.kind) {
    case STZ_MODEL_LINEAR: {  // stz_model.hh:58-67
        S sC1 = exp(-mp.Delta * thetainv);
        S sC2 = cosh(Oe * sbar * thetainv), sC = sC1 * sC2, sT = tanh(Oe * sbar * thetainv);
        return dpl_common_t(mp, sbar, chi, dchi1, dchi2, sC1, sC2, sC, sT, c0);
    }
    case STZ_MODEL_NONLINEAR: {  // stz_model.hh:76-87
        S x = Oe * sbar / mp.Delta, sC1 = exp(-mp.Delta * thetainv),
          sC = exp(-mp.Delta * cosh(x) * thetainv) * cosh(mp.Delta * sinh(x) * thetainv), sC2 = sC / sC1,
          sT = tanh(mp.Delta * sinh(x) * thetainv);
        return dpl_common_t(mp, sbar, chi, dchi1, dchi2, sC1, sC2, sC, sT, c0);
    }
    case STZ_MODEL_NONLINEAR_ATHERMAL: {  // stz_model.hh:179-206
        dchi2 = S(0);
        if (below) { dchi1 = S(0); return S(0); }
        S x = Oe * sbar / mp.Delta,
          sC = 0.5 * (exp(-mp.Delta * thetainv * exp(x)) + exp(-mp.Delta * thetainv * exp(-x))),
          sT = tanh(mp.Delta * sinh(x) * thetainv);
        S bfac = exp(-1 / chi), Dpl = bfac * sC * (sT - s0 / sbar) / mp.tau0,
          Gamma = 2 * mp.tau0 * Dpl * sbar / (s0 * bfac);
        dchi1 = bfac * Gamma / (mp.tau0 * c0);
        return Dpl;
    }
    default: {  // linear athermal, stz_model.hh:103-151
        S tmp, sC, sT;
        dchi2 = S(0);
        if (below) { dchi1 = S(0); return S(0); }
        if (mp.t_model == -1 || mp.Delta >= Oe * sbar)
            sC = 0.5 * (exp((Oe * sbar - mp.Delta) * thetainv) + exp((-Oe * sbar - mp.Delta) * thetainv));
        else switch (mp.t_model) {
            case 0: sC = 0.5 * (exp((-Oe * sbar - mp.Delta) * thetainv) + 1); break;
            case 1:
                tmp = Oe * sbar / mp.Delta;
                sC = 0.5 * (exp((-Oe * sbar - mp.Delta) * thetainv) + sqrt(0.5 * (1 + tmp * tmp)));
                break;
            case 2:
                tmp = Oe * sbar / mp.Delta;
                sC = 0.5 * (exp((-Oe * sbar - mp.Delta) * thetainv) + (0.5 * (1 + tmp * tmp)));
                break;
            case 3: sC = mp.t_icpt + sbar * mp.t_grad; break;
            case 4:
                sC = 0.5 * (exp((-Oe * sbar - mp.Delta) * thetainv) + (Oe * sbar - mp.Delta) * thetainv + 1);
                break;
            default: sC = S(0);
        }
        sT = mp.t_model != 3 ? tanh(Oe * sbar * thetainv) : S(1);
        S bfac = exp(-1 / chi), Dpl = bfac * sC * (sT - s0 / sbar) / mp.tau0,
          Gamma = 2 * mp.tau0 * Dpl * sbar / (s0 * bfac);
        dchi1 = bfac * Gamma / (mp.tau0 * c0);
        return Dpl;
    }
    }
}

__host__ __device__ static inline double dplastic(const stz_model_params &mp, double thetainv, double sbar,
                                                  double chi, double &dchi1, double &dchi2, int yield = -1) {
    return dplastic_t<double>(mp, thetainv, sbar, chi, mp.Omegaeps0, mp.c0, dchi1, dchi2, yield);
}

// shear_sim.cc:513-548 (adaptive_plastic_term); trips counts loop iterations
// and last_yield is the athermal laws' yield test (sbar >= 1) on the last one
template <class C>
__device__ static double plastic_term(const C &c, double sbar, double &chiv, double &ddev, double dt,
                                      int &trips, bool &last_yield) {
    const double small_dev_cutoff = 1e-12;
    const double theta = 0;
    bool adapt = true;
    int t = 0;
    double dtl = dt, osbar = sbar, dplas, dchi1, dchi2, adt;
    last_yield = false;
    if (sbar < small_dev_cutoff) { ddev = 0; trips = 0; return 0; }
    do {
        last_yield = !(sbar < 1);
        dplas = 2 * c.t_scale * c.mu * dplastic(c.stz, c.thetainv, sbar, chiv, dchi1, dchi2);
        if (fabs(dtl * dplas) > c.adapt_fac) {
            adt = c.adapt_fac / fabs(dplas);
            dtl -= adt;
        } else {
            adapt = false;
            adt = dtl;
        }
        sbar -= adt * dplas;
        chiv += (dchi1 * (c.chi_inf - chiv) + dchi2 * (theta - chiv)) * c.t_scale * adt;
        t++;
    } while (adapt);
    trips = t;
    ddev = osbar - sbar;
    return ddev / osbar;
}

// plastic_term with its decisions taken from a record of an earlier
// evaluation: the small-deviator cut-off, the number of loop trips (every
// trip but the last a substep of adapt_fac / |dplas|, which a nonzero dplas
// needs, so at or above yield) and the yield branch of the last trip.
// Where the record matches what plastic_term would decide, the result is
// plastic_term's bit for bit; nearby, it is the smooth continuation of the
// recorded branch.
template <class C>
__device__ static double plastic_term_replay(const C &c, double sbar, double &chiv, double &ddev, double dt,
                                             bool small, int trips, bool last_yield) {
    const double theta = 0;
    double dtl = dt, osbar = sbar, dplas, dchi1, dchi2, adt;
    if (small) { ddev = 0; return 0; }
    for (int t = 0; t < trips; t++) {
        const int yield = t + 1 < trips ? 1 : (last_yield ? 1 : 0);
        dplas = 2 * c.t_scale * c.mu * dplastic(c.stz, c.thetainv, sbar, chiv, dchi1, dchi2, yield);
        if (t + 1 < trips) {
            adt = c.adapt_fac / fabs(dplas);
            dtl -= adt;
        } else {
            adt = dtl;
        }
        sbar -= adt * dplas;
        chiv += (dchi1 * (c.chi_inf - chiv) + dchi2 * (theta - chiv)) * c.t_scale * adt;
    }
    ddev = osbar - sbar;
    return ddev / osbar;
}

}  // namespace stz
