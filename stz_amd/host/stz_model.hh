// stz_model.hh -- the STZ plasticity laws of the reference (sb/main/
// stz_model.hh:9-207) as parameter objects for the gfx950 step.  The laws are
// evaluated inside the device kernels; the constructors take the same
// arguments as the reference's, and to_params() gives the C ABI struct.
#ifndef STZ_AMD_STZ_MODEL_HH
#define STZ_AMD_STZ_MODEL_HH

#include "../../include/stz_amd.h"

class stz_dynamics {
  public:
    const double TZ, c0, tau0, kappa, Delta, Omegaeps0, chi_inf, theta, rho;
    stz_dynamics(double TZ_, double c0_, double tau0_, double kappa_, double Delta_, double Omegaeps0_,
                 double chi_inf_, double theta_, double rho_)
        : TZ(TZ_), c0(c0_), tau0(tau0_), kappa(kappa_), Delta(Delta_), Omegaeps0(Omegaeps0_),
          chi_inf(chi_inf_), theta(theta_), rho(rho_) {}
    virtual ~stz_dynamics() {}
    virtual stz_model_params to_params() const = 0;

  protected:
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
    int t_model;
    double t_icpt, t_grad;
    stz_dynamics_linear_athermal(double TZ_, double c0_, double tau0_, double kappa_, double Delta_,
                                 double Omegaeps0_, double chi_inf_, double theta, double rho_, int t_model_ = -1)
        : stz_dynamics(TZ_, c0_, tau0_, kappa_, Delta_, Omegaeps0_, chi_inf_, theta, rho_), t_model(t_model_),
          t_icpt(0), t_grad(0.5 * Omegaeps0_ / Delta_) {}
    stz_model_params to_params() const { return base(STZ_MODEL_LINEAR_ATHERMAL, t_model, t_icpt, t_grad); }
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
