"""Parameter sets for the quasi-static simple-shear workloads.

The physical set follows the reference drivers: the linear athermal STZ law
(stz_model.hh:90-170) constructed as in sb/main/shear_small.cc:30-40 and
shear_random.cc:41-82 (9 constructor arguments, so rho = 0 and
t_model = -1), moduli rescaled by the yield stress, and the walls at y = +/-1
moving at +/-U (shear_sim.hh:120-124).

Choices the reference leaves to the driver, recorded here:
  * Omega = 300 A^3 and Omega*eps0 = Omega*1e-30*eps0*s_y/(T_Z k_B)
    (shear_small.cc:40 form with s_y = 0.85 GPa);
  * viscosity 0.02, adapt_fac 0.002, tmult 0.5, chi_diff = chi_len = 0;
  * U = 1e-4: the wall-to-wall velocity difference 2U over height 2 gives
    the nondimensional shear rate 1e-4;
  * t_scale = 1 (one time unit is one second of STZ time), so the steady
    flow stress sits near 1.2 s_y;
  * the run lasts T = 1000 time units (shear strain 0.1, about twice the
    yield strain mu^-1), dt = T / steps.
"""
import math

import numpy as np

TZ = 21000.0
KB = 1.3806503e-23          # shear_small.cc:14
S_Y = 0.85e9                # Pa
MU = 20e9 / S_Y             # 20 GPa in units of s_y
K_BULK = 143.2e9 / S_Y      # 143.2 GPa in units of s_y
OMEGA = 300.0               # cubic Angstrom
EPS0 = 0.3


def stz_linear_athermal(c0=0.4, eps0=EPS0, chi_inf_K=900.0, tau0=1e-13,
                        Delta_K=8000.0, theta_K=400.0, omega=OMEGA):
    Omegaeps0 = omega * 1e-30 * eps0 * S_Y / (TZ * KB)
    Delta = Delta_K / TZ
    return {
        "kind": 0, "t_model": -1,
        "TZ": TZ, "c0": c0, "tau0": tau0, "kappa": 4.2,
        "Delta": Delta, "Omegaeps0": Omegaeps0, "chi_inf": chi_inf_K / TZ,
        "theta": theta_K / TZ, "rho": 0.0,
        "t_icpt": 0.0, "t_grad": 0.5 * Omegaeps0 / Delta,
    }


def simple_shear(m, n=None, u_bdry=1e-4, t_scale=1.0, viscosity=0.02,
                 adapt_fac=0.002, chi_diff=0.0, chi_len=0.0, **stz_kw):
    """shear_sim constructor arguments for the periodic-x simple shear cell."""
    n = m if n is None else n
    return {
        "m": m, "n": n, "ax": -1.0, "bx": 1.0, "ay": -1.0, "by": 1.0,
        "mu": MU, "K": K_BULK, "viscosity": viscosity,
        "chi_diff": chi_diff, "chi_len": chi_len, "t_scale": t_scale,
        "tmult": 0.5, "adapt_fac": adapt_fac, "u_bdry": u_bdry,
        "stz": stz_linear_athermal(**stz_kw),
    }


# Simple-shear workloads: (cells per side, steps, chi mean K, chi std K,
# correlation length in cells)
WORKLOADS = {
    "shear128": dict(N=128, steps=2000, chi_mean=600.0, chi_std=20.0, corr=2.0),
    "shear1024": dict(N=1024, steps=5000, chi_mean=600.0, chi_std=25.0, corr=3.0),
    "calibration256": dict(N=256, steps=400, chi_mean=620.0, chi_std=25.0, corr=3.0),
}
T_END = 1000.0


def workload(name):
    w = dict(WORKLOADS[name])
    w["params"] = simple_shear(w["N"])
    w["dt"] = T_END / w["steps"]
    return w


def describe(params):
    return ("simple shear %dx%d, mu=%.4g K=%.4g eta=%.3g U=%.3g t_scale=%.3g" %
            (params["m"], params["n"], params["mu"], params["K"],
             params["viscosity"], params["u_bdry"], params["t_scale"]))


assert math.isclose(simple_shear(8)["stz"]["Omegaeps0"], 0.2638, rel_tol=1e-3)


# ---------------------------------------------------------------------------
# Free-boundary workloads (bd_sim).  The body is where phi < 0.
# ---------------------------------------------------------------------------

def stz_bd_test():
    """The STZ law of the reference's free-boundary driver bd_test.cc:26-38
    (Omega*eps0 with 1.1e9 in place of s_y, bd_test.cc:38)."""
    d = stz_linear_athermal()
    d["Omegaeps0"] = OMEGA * 1e-30 * EPS0 * 1.1e9 / (TZ * KB)
    d["t_grad"] = 0.5 * d["Omegaeps0"] / d["Delta"]
    return d


# bd_test.cc:47-51: Vitreloy 1, E = 101 GPa, nu = 0.35, s_y = 0.85 GPa
E_BD, NU_BD, SY_BD = 101.0, 0.35, 0.85
K_BD = E_BD / (3 * (1 - 2 * NU_BD) * SY_BD)
MU_BD = E_BD / (2 * (1 + NU_BD) * SY_BD)


def bd_params(m, n=None, box=(-1.0, 1.0, -1.0, 1.0), viscosity=0.02,
              filter_stress=0.001, t_scale=1.0, adapt_fac=0.002, chi_len=0.0):
    """bd_sim constructor arguments with the material of bd_test.cc.

    The reference's free-boundary runs all use these moduli (K/mu = 3).  With
    the stiffer simple-shear set (K/mu = 7.2) the free-surface rows of
    qs_multi_bd make Gauss-Seidel V-cycles diverge on plate geometries, the
    the reference's as well as the ones here, so the free-boundary workloads
    keep the reference's material."""
    n = m if n is None else n
    ax, bx, ay, by = box
    return {
        "m": m, "n": n, "ax": ax, "bx": bx, "ay": ay, "by": by,
        "mu": MU_BD, "K": K_BD, "viscosity": viscosity, "chi_len": chi_len,
        "filter_stress": filter_stress, "t_scale": t_scale, "tmult": 0.5,
        "adapt_fac": adapt_fac, "stz": stz_bd_test(),
    }


def plate_hole_phi(m, n=None, box=(-1.0, 1.0, -1.0, 1.0), half_height=0.9,
                   radius=None):
    """Level set of a plate |y| < half_height with a circular hole of radius
    0.2 L (L = box width) at the centre: phi = max(|y| - h, R - r) on the
    cell centres, negative inside the material."""
    n = m if n is None else n
    ax, bx, ay, by = box
    L = bx - ax
    R = 0.2 * L if radius is None else radius
    dx, dy = (bx - ax) / m, (by - ay) / n
    x = ax + (np.arange(m) + 0.5) * dx - 0.5 * (ax + bx)
    y = ay + (np.arange(n) + 0.5) * dy - 0.5 * (ay + by)
    X, Y = np.meshgrid(x, y)
    return np.maximum(np.abs(Y) - half_height, R - np.sqrt(X * X + Y * Y))


# plate with a hole pulled apart by grips on the left and right edges at
# nominal strain rate 1e-4 (grip speed 1e-4 * L/2), chi = 580 K + 30 K
# Gaussian field with correlation length 4 cells
BD_WORKLOADS = {
    "plate_hole4096": dict(N=4096, steps=2000, chi_mean=580.0, chi_std=30.0, corr=4.0,
               strain_rate=1e-4),
}
