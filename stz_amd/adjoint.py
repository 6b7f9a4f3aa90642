"""Adjoint of the FP64 quasi-static step, and a torch.autograd wrapper.

No counterpart in the reference.  ``Adjoint(sim, checkpoint_every)`` runs
``forward(dt, nsteps)`` on a ``ShearSim`` (FP64 form) while keeping
checkpoints, and ``backward(lam_final)`` returns dL/d(initial fields) and
dL/d(c0, Omegaeps0, chi_inf) for L = <lam_final, final fields>, with every
discrete choice of the forward (upwind side, ENO branch, yield branch,
plastic substep count) frozen.  ``QSRun`` exposes the same as a
torch.autograd.Function of the initial chi field and the three parameters.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

PARAMS = ("c0", "Omegaeps0", "chi_inf")


class Adjoint:
    def __init__(self, sim, checkpoint_every=25):
        if sim.precision != "fp64":
            raise ValueError("the adjoint differentiates the FP64 form")
        self.sim = sim
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.stz_adjoint_create(sim._h, int(checkpoint_every), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.stz_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solve_accuracy(self, rel):
        check(self._lib.stz_adjoint_set_solve_accuracy(self._h, float(rel)))

    def forward(self, dt, nsteps):
        check(self._lib.stz_adjoint_forward(self._h, float(dt), int(nsteps)))

    def backward(self, lam_final):
        lam_final = np.ascontiguousarray(lam_final, dtype=np.float64)
        if lam_final.shape != (10,) + self.sim.plane_shape:
            raise ValueError("lam_final must have shape %s" % ((10,) + self.sim.plane_shape,))
        lam_init = np.zeros_like(lam_final)
        g = np.zeros(3)
        check(self._lib.stz_adjoint_backward(self._h, lam_final, lam_init, g))
        return lam_init, dict(zip(PARAMS, g))

    def stats(self):
        out = np.zeros(6)
        check(self._lib.stz_adjoint_stats(self._h, out))
        return dict(zip(("forward_s", "backward_s", "recomputed_steps", "adjoint_solves",
                         "adjoint_cycles", "forward_cycles"), out))


def set_params(sim, c0=None, Omegaeps0=None, chi_inf=None):
    """change the STZ law's c0 / Omegaeps0 and the plastic term's chi_inf"""
    mp = dict(sim.params["stz"])
    if c0 is not None:
        mp["c0"] = float(c0)
    if Omegaeps0 is not None:
        mp["Omegaeps0"] = float(Omegaeps0)
    sim.params["stz"] = mp
    p = _lib.make_params(sim.params)
    check(_lib.load().stz_sim_set_model(sim._h, C.byref(p.stz)))
    if chi_inf is not None:
        sim.set_chi_inf(float(chi_inf))


def reset_predictor(sim):
    """start the adaptive V-cycle count afresh (tgmg_predict's initial
    limit), so that repeated runs from the same state take the same cycles"""
    init = sim.mg_config().predict_init
    sim.set_mg_config(predict_init=init + 1)
    sim.set_mg_config(predict_init=init)


def chi_planes(chi, planes):
    """planes with chi (n x m, cells) in rows 0..n-1 and the ghost rows
    extrapolated linearly (shear_sim.cc:552-581)"""
    out = np.array(planes, dtype=np.float64, copy=True)
    n = chi.shape[0]
    out[6][1:n + 1] = chi
    out[6][0] = 2 * chi[0] - chi[1]
    out[6][n + 1] = 2 * chi[n - 1] - chi[n - 2]
    return out


def chi_planes_adjoint(lam_planes):
    """adjoint of chi_planes with respect to chi"""
    lc = lam_planes[6]
    n = lc.shape[0] - 2
    g = lc[1:n + 1].copy()
    g[0] += 2 * lc[0]
    g[1] -= lc[0]
    g[n - 1] += 2 * lc[n + 1]
    g[n - 2] -= lc[n + 1]
    return g


def hencky_strain(planes, params):
    """shear_sim::compute_strain (shear_sim.cc:778-824) as a differentiable
    torch function of the field planes [10, n+2, m]: the deviatoric modulus
    of the Hencky strain at the n x m cell centres, in the form with the
    y derivative taken over rows j+1 and j-1 in the interior rows and the
    periodic shift added at both seam columns (the reference's own form
    gives 0 on the interior rows; stz_sim_compute_strain(corrected = 1)
    computes this one on the device).

    With C = F^T F (F the inverse of the reference map's gradient),
    eigenvalues l1 >= l2 > 0, the Hencky strain (1/2) log C has eigenvalues
    (1/2) log l1 and (1/2) log l2 and devmod (mat.hh:51-53) of a symmetric
    2 x 2 tensor is |e1 - e2| / sqrt(2), so the value is
    log(l1 / l2) / (2 sqrt 2) = atanh(sqrt((C00 - C11)^2 + 4 C01^2) / tr C) / sqrt 2,
    which needs no eigenvectors."""
    import torch
    m, n = params["m"], params["n"]
    lx = params["bx"] - params["ax"]
    xsp = m / lx
    ysp = n / (params["by"] - params["ay"])
    X = planes[7][1:n + 2]          # rows 0..n
    Y = planes[8][1:n + 2]
    Xc, Yc = X[:n], Y[:n]
    wrap = torch.zeros(m, dtype=planes.dtype)
    wrap[0] = lx
    wrap[m - 1] = lx
    Xx = (torch.roll(Xc, -1, 1) - torch.roll(Xc, 1, 1) + wrap) * (0.5 * xsp)
    Yx = (torch.roll(Yc, -1, 1) - torch.roll(Yc, 1, 1)) * (0.5 * xsp)

    def ddy(Z):
        return torch.cat([ysp * (Z[1:2] - Z[0:1]), (0.5 * ysp) * (Z[2:n] - Z[0:n - 2]),
                          ysp * (Z[n:n + 1] - Z[n - 1:n])])
    Xy, Yy = ddy(X), ddy(Y)
    J = 1 / (Xx * Yy - Xy * Yx)
    F00, F01, F10, F11 = Yy * J, -Xy * J, -Yx * J, Xx * J
    C00 = F00 * F00 + F10 * F10
    C01 = F00 * F01 + F10 * F11
    C11 = F01 * F01 + F11 * F11
    disc = torch.sqrt((C00 - C11) ** 2 + 4 * C01 * C01)
    return torch.atanh(disc / (C00 + C11)) / np.sqrt(2.0)


def make_qs_run():
    """torch.autograd.Function: (chi [n, m], theta [3], ctx data) -> final
    field planes [10, n+2, m]; backward through the hand-written adjoint"""
    import torch

    class QSRun(torch.autograd.Function):
        @staticmethod
        def forward(ctx, chi, theta, sim, base_planes, vel0, t0, dt, nsteps, every, rel=1e-10):
            set_params(sim, *[float(v) for v in theta.detach().cpu().numpy()])
            sim.set_fields(chi_planes(chi.detach().cpu().numpy(), base_planes))
            sim.set_vel(vel0)
            sim.time = t0
            reset_predictor(sim)
            ctx.adj = None
            if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
                ctx.adj = Adjoint(sim, every)
                ctx.adj.set_solve_accuracy(rel)
                ctx.adj.forward(dt, nsteps)
            else:
                sim.step_forward_quasistatic(dt, nsteps)
            return torch.from_numpy(sim.get_fields())

        @staticmethod
        def backward(ctx, grad_out):
            lam_init, g = ctx.adj.backward(grad_out.detach().cpu().numpy())
            gchi = torch.from_numpy(chi_planes_adjoint(lam_init))
            gth = torch.tensor([g[k] for k in PARAMS], dtype=torch.float64)
            ctx.adj.close()
            return gchi, gth, None, None, None, None, None, None, None, None

    return QSRun
