// stz_model.hh -- the STZ plasticity laws of the reference (sb/main/
// stz_model.hh:9-207) for the gfx950 step.  The constructors, the public
// constants and the virtual Dplastic (stz_model.hh:28) are the reference's;
// Dplastic evaluates the law through the C ABI (stz_model_dplastic), which
// runs on the host the law code the device kernels run, so a driver that
// tabulates the law (dplas_test.cc) gets the numbers the step uses.
// to_params() gives the C ABI struct the simulation is created with.
#ifndef STZ_AMD_STZ_MODEL_HH
#define STZ_AMD_STZ_MODEL_HH

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../include/stz_amd.h"

class stz_dynamics {
  public:
    const double rt3;
    const double TZ, c0, tau0, kappa, Delta, Omegaeps0, chi_inf;
    const double thetainv;
    const double rho;
    const double s0;
    stz_dynamics(double TZ_, double c0_, double tau0_, double kappa_, double Delta_, double Omegaeps0_,
                 double chi_inf_, double theta_, double rho_)
        : rt3(std::sqrt(3.0)), TZ(TZ_), c0(c0_), tau0(tau0_), kappa(kappa_), Delta(Delta_), Omegaeps0(Omegaeps0_),
          chi_inf(chi_inf_), thetainv(1 / theta_), rho(rho_), s0(1), theta(theta_) {}
    virtual ~stz_dynamics() {}
    /** the plastic rate at deviatoric stress sbar and effective temperature
     * chi, with the chi-rate coefficients (stz_model.hh:28) */
    virtual double Dplastic(double sbar, double chi, double &dchi1, double &dchi2) {
        const stz_model_params p = to_params();
        double d = 0;
        if (stz_model_dplastic(&p, sbar, chi, &d, &dchi1, &dchi2) != STZ_OK) {
            std::fprintf(stderr, "stz_amd: Dplastic failed: %s\n", stz_last_error());
            std::exit(1);
        }
        return d;
    }
    virtual stz_model_params to_params() const = 0;

  protected:
    /** theta as the constructor received it (the law uses 1/theta) */
    const double theta;
    stz_model_params base(int kind, int t_model, double t_icpt, double t_grad) const {
        stz_model_params p;
        p.kind = kind;
        p.t_model = t_model;
        p.TZ = TZ;
        p.c0 = c0;
        p.tau0 = tau0;
        p.kappa = kappa;
        p.Delta = Delta;
        p.Omegaeps0 = Omegaeps0;
        p.chi_inf = chi_inf;
        p.theta = theta;
        p.rho = rho;
        p.t_icpt = t_icpt;
        p.t_grad = t_grad;
        return p;
    }
};

// stz_model.hh:50-68
class stz_dynamics_linear : public stz_dynamics {
  public:
    int t_model;
    stz_dynamics_linear(double TZ_, double c0_, double tau0_, double kappa_, double Delta_, double Omegaeps0_,
                        double chi_inf_, double theta, double rho_, int t_model_ = -1)
        : stz_dynamics(TZ_, c0_, tau0_, kappa_, Delta_, Omegaeps0_, chi_inf_, theta, rho_), t_model(t_model_) {}
    stz_model_params to_params() const { return base(STZ_MODEL_LINEAR, t_model, 0, 0); }
};

// stz_model.hh:70-88
class stz_dynamics_nonlinear : public stz_dynamics {
  public:
    stz_dynamics_nonlinear(double TZ_, double c0_, double tau0_, double kappa_, double Delta_, double Omegaeps0_,
                           double chi_inf_, double theta, double rho_)
        : stz_dynamics(TZ_, c0_, tau0_, kappa_, Delta_, Omegaeps0_, chi_inf_, theta, rho_) {}
    stz_model_params to_params() const { return base(STZ_MODEL_NONLINEAR, -1, 0, 0); }
};

// stz_model.hh:90-170, including the truncation-model controls
class stz_dynamics_linear_athermal : public stz_dynamics {
  public:
    /** truncation model (-1: none; 0..4, stz_model.hh:120-139) */
    int t_model;
    /** intercept and gradient of the linear truncation (model 3) */
    double t_icpt, t_grad;
    stz_dynamics_linear_athermal(double TZ_, double c0_, double tau0_, double kappa_, double Delta_,
                                 double Omegaeps0_, double chi_inf_, double theta, double rho_, int t_model_ = -1)
        : stz_dynamics(TZ_, c0_, tau0_, kappa_, Delta_, Omegaeps0_, chi_inf_, theta, rho_), t_model(t_model_),
          t_icpt(0), t_grad(0.5 * Omegaeps0_ / Delta_) {}
    stz_model_params to_params() const { return base(STZ_MODEL_LINEAR_ATHERMAL, t_model, t_icpt, t_grad); }
    /** truncation model 3 with its slope scaled by p (stz_model.hh:156-158) */
    void trunc_modify_mult(double p) {
        t_grad = 0.5 * Omegaeps0 / Delta * p;
        refit_intercept();
    }
    /** truncation model 3 with its slope moved a fraction p from the
     * standard one towards the slope that makes sC differentiable where the
     * truncation starts (stz_model.hh:163-166) */
    void trunc_modify_scale(double p) {
        t_grad = (1 - p) * (0.5 * Omegaeps0 / Delta) +
                 p * (0.5 * (1 - std::exp(-2 * Delta * thetainv)) * Omegaeps0 * thetainv);
        refit_intercept();
    }

  private:
    /** the intercept that makes the line meet the untruncated sC at
     * Omegaeps0 sbar = Delta (stz_model.hh:168-170) */
    void refit_intercept() { t_icpt = 0.5 * (1 + std::exp(-2 * Delta * thetainv)) - Delta / Omegaeps0 * t_grad; }
};

// stz_model.hh:172-207
class stz_dynamics_nonlinear_athermal : public stz_dynamics {
  public:
    stz_dynamics_nonlinear_athermal(double TZ_, double c0_, double tau0_, double kappa_, double Delta_,
                                    double Omegaeps0_, double chi_inf_, double theta, double rho_)
        : stz_dynamics(TZ_, c0_, tau0_, kappa_, Delta_, Omegaeps0_, chi_inf_, theta, rho_) {}
    stz_model_params to_params() const { return base(STZ_MODEL_NONLINEAR_ATHERMAL, -1, 0, 0); }
};

#endif
