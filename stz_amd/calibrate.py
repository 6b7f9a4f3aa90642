"""Calibration driver: fit the initial chi field and the STZ parameters
(c0, Omega*eps0, chi_inf) to a target strain field, with gradients from the
hand-written adjoint of the FP64 step (stz_amd.adjoint).

    python -m stz_amd.calibrate [--size 256] [--steps 400] [--every 25]
                                [--iters 8] [--fd-dirs 16] [--fd-step 1e-6]

Workload: simple shear, `size`^2 cells, `steps` steps of dt = 2.5 from
t = 0 to t = 1000 (the end time of the 1024^2 workload: elastic loading,
yield near t = 425, plastic flow).  The target is the final strain field of
a forward run from a hidden chi field (620 K + 25 K seeded Gaussian field,
correlation length 3 cells) and the default parameters, plus 1 % Gaussian
noise.  The fit starts from a uniform 620 K field and parameters 10 % off;
the loss is the mean-square misfit, cell by cell, of the deviatoric modulus
of the Hencky strain that shear_sim::compute_strain writes
(shear_sim.cc:778-824; the form with central y differences,
stz_amd.adjoint.hencky_strain).  Prints one JSON line: loss history, the
cost of an adjoint (backward) run relative to a forward run, and the
agreement of the gradient with FP64 central differences along `fd-dirs`
random directions: over a window of a few steps from a plastic state (with
the number of decisions each difference crosses), and over the whole run.
"""
import argparse
import json
import sys
import time

import numpy as np

# decision bits that are jumps of the step map (include/stz_amd.h): all but
# the plastic substep count in bits 24..31
SWITCHES = np.uint32(0x00FFFFFF)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--dt", type=float, default=2.5)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--fd-dirs", type=int, default=16)
    ap.add_argument("--fd-step", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--adjoint-rel", type=float, default=1e-8,
                    help="relative residual of each adjoint solve")
    ap.add_argument("--window", type=int, default=4, help="steps of the windowed gradient check")
    ap.add_argument("--window-step", type=float, default=1e-7)
    args = ap.parse_args()

    import torch
    from . import problems
    from .adjoint import PARAMS, hencky_strain, make_qs_run
    from .initial import random_chi, simple_shear_state
    from .sim import ShearSim

    torch.set_default_dtype(torch.float64)
    p = problems.simple_shear(args.size)
    TZ = p["stz"]["TZ"]
    n, m = p["n"], p["m"]
    base = simple_shear_state(p, 620.0, 0.0, 3.0, seed=0)
    vel0 = np.zeros((n + 1, m, 2))
    theta_true = np.array([p["stz"]["c0"], p["stz"]["Omegaeps0"], p["stz"]["chi_inf"]])
    sim = ShearSim(p, "fp64")
    sim.set_mg_config(tol_mult=args.tol)
    QSRun = make_qs_run()

    def run(chi, theta):
        return QSRun.apply(chi, theta, sim, base, vel0, 0.0, args.dt, args.steps, args.every, args.adjoint_rel)

    # target: hidden chi, default parameters, 1 % noise
    chi_true = torch.from_numpy(random_chi(p, 620.0, 25.0, 3.0, seed=12345))
    with torch.no_grad():
        g_true = hencky_strain(run(chi_true, torch.from_numpy(theta_true)), p)
    rng = np.random.default_rng(99)
    target = g_true + 0.01 * float(g_true.std()) * torch.from_numpy(rng.standard_normal(tuple(g_true.shape)))

    # unknowns: chi in K, log of the parameters relative to the start
    chi_K = torch.full((n, m), 620.0, requires_grad=True)
    theta0 = torch.from_numpy(theta_true * 1.1)
    logs = torch.zeros(3, requires_grad=True)

    def loss_of(chi_K, logs):
        F = run(chi_K / TZ, theta0 * torch.exp(logs))
        return torch.mean((hencky_strain(F, p) - target) ** 2)

    # gradient check 1: a window of `window` steps from the plastic state the
    # fit starts from at t = 500, f = <lam, fields after the window>, central
    # differences at h and 4h along random chi and parameter directions.  A
    # direction counts as away from kinks when the two differences agree to
    # 1e-5; the number of decisions that differ between the +h and -h runs
    # is reported.  This state sits on kinks almost everywhere: chi started
    # uniform and the reference map is nearly linear, so the ENO comparisons
    # are between second differences at rounding level, and v = 0 up to
    # rounding on every node (the upwind choice along y); the adjoint follows
    # the branches the forward took, a central difference averages across
    # them.
    from .adjoint import Adjoint, reset_predictor, set_params
    from .sim import ShearSim as _S
    theta_start = theta_true * 1.1
    k0 = int(round(500.0 / args.dt))
    pre = _S(p, "fp64")
    pre.set_mg_config(tol_mult=args.tol)
    set_params(pre, *theta_start)
    pre.set_fields(simple_shear_state(p, 620.0, 0.0, 3.0, seed=0))
    pre.set_vel(vel0)
    pre.time = 0.0
    pre.step_forward_quasistatic(args.dt, k0)
    w_planes, w_vel, w_time = pre.get_fields(), pre.get_vel(), pre.time
    pre.close()
    lam_w = np.random.default_rng(5).standard_normal(w_planes.shape)

    def window(planes, theta, record=True):
        g = _S(p, "fp64")
        g.set_mg_config(tol_mult=args.tol)
        set_params(g, *theta)
        g.set_fields(planes)
        g.set_vel(w_vel)
        g.time = w_time
        g.record_decisions(record)
        bits = []
        for _ in range(args.window):
            g.step_forward_quasistatic(args.dt, 1)
            bits.append(g.get_decisions().copy())
        F = g.get_fields()
        g.close()
        return float(np.sum(lam_w * F)), bits

    g = _S(p, "fp64")
    g.set_mg_config(tol_mult=args.tol)
    set_params(g, *theta_start)
    g.set_fields(w_planes)
    g.set_vel(w_vel)
    g.time = w_time
    a = Adjoint(g, args.window)
    a.set_solve_accuracy(args.adjoint_rel)
    a.forward(args.dt, args.window)
    lam0, gth = a.backward(lam_w)
    a.close()
    g.close()
    win, flips = [], []
    for k in range(args.fd_dirs):
        r = np.random.default_rng(2000 + k)
        dx = np.zeros_like(w_planes)
        dx[6] = r.standard_normal(w_planes[6].shape) * (10.0 / TZ)
        dth = r.standard_normal(3) * theta_start * 0.1
        ad = float(np.sum(lam0 * dx)) + float(sum(gth[q] * dth[i] for i, q in enumerate(PARAMS)))
        res, nflip = [], 0
        for hh in (args.window_step, 4 * args.window_step):
            fp, bp = window(w_planes + hh * dx, theta_start + hh * dth)
            fm, bm = window(w_planes - hh * dx, theta_start - hh * dth)
            nflip = max(nflip, int(sum(np.count_nonzero((x ^ y) & SWITCHES) for x, y in zip(bp, bm))))
            res.append((fp - fm) / (2 * hh))
        ok = abs(res[0] - res[1]) <= 1e-5 * abs(res[0])
        rel = abs(res[0] - ad) / max(abs(res[0]), abs(ad))
        win.append(rel if ok else None)
        flips.append(nflip)
        print("calibrate: window direction %2d fd(h) %.10e fd(4h) %.10e ad %.10e rel %.2e consistent %s, "
              "switches flipped %d" % (k, res[0], res[1], ad, rel, ok, nflip), file=sys.stderr, flush=True)

    win_ok = [x for x in win if x is not None]

    # gradient check 2: the loss of the whole run, FP64 central differences
    # along random directions (the run crosses the yield point and the shear
    # band, so such a difference also picks up the jumps of every decision
    # it crosses)
    t0 = time.perf_counter()
    L = loss_of(chi_K, logs)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L.backward()
    t_bwd = time.perf_counter() - t0
    g_chi, g_log = chi_K.grad.clone(), logs.grad.clone()
    fd_rel = []
    h = args.fd_step
    for k in range(args.fd_dirs):
        r = np.random.default_rng(1000 + k)
        v_chi = torch.from_numpy(r.standard_normal((n, m))) * 10.0  # K
        v_log = torch.from_numpy(r.standard_normal(3)) * 0.1
        ad = float(torch.sum(g_chi * v_chi) + torch.sum(g_log * v_log))
        with torch.no_grad():
            lp = float(loss_of(chi_K + h * v_chi, logs + h * v_log))
            lm = float(loss_of(chi_K - h * v_chi, logs - h * v_log))
        fd = (lp - lm) / (2 * h)
        fd_rel.append(abs(fd - ad) / max(abs(fd), abs(ad), 1e-300))
        print("calibrate: direction %2d fd %.10e adjoint %.10e rel %.2e" % (k, fd, ad, fd_rel[-1]),
              file=sys.stderr, flush=True)

    # the fit
    opt = torch.optim.Adam([{"params": [chi_K], "lr": 0.5}, {"params": [logs], "lr": 0.01}])
    history = []
    fwd_s, bwd_s = [], []
    for it in range(args.iters):
        opt.zero_grad()
        t0 = time.perf_counter()
        L = loss_of(chi_K, logs)
        t1 = time.perf_counter()
        L.backward()
        t2 = time.perf_counter()
        fwd_s.append(t1 - t0)
        bwd_s.append(t2 - t1)
        opt.step()
        th = (theta0 * torch.exp(logs)).detach().numpy()
        history.append(L.item())
        print("calibrate: iter %d loss %.6e (fwd %.2f s, bwd %.2f s) theta/true %s chi rms err %.2f K" %
              (it, L.item(), t1 - t0, t2 - t1, np.round(th / theta_true, 4),
               float(torch.sqrt(torch.mean((chi_K.detach() / TZ - chi_true) ** 2))) * TZ),
              file=sys.stderr, flush=True)
    print(json.dumps({
        "workload": "adjoint calibration, %dx%d, %d steps of dt %g, checkpoint every %d, FP64" %
                    (m, n, args.steps, args.dt, args.every),
        "loss_history": history,
        "forward_s_per_run": float(np.median(fwd_s)), "backward_s_per_run": float(np.median(bwd_s)),
        "backward_over_forward": float(np.median(bwd_s) / np.median(fwd_s)),
        "first_forward_s": t_fwd, "first_backward_s": t_bwd,
        "fd_check_window": {"steps": args.window, "from_t": w_time, "directions": args.fd_dirs,
                            "step": args.window_step, "consistent_directions": len(win_ok),
                            "median_switches_flipped": float(np.median(flips)) if flips else None,
                            "max_rel": float(max(win_ok)) if win_ok else None},
        "fd_check_whole_run": {"directions": args.fd_dirs, "step": h,
                               "max_rel": float(max(fd_rel)) if fd_rel else None,
                               "median_rel": float(np.median(fd_rel)) if fd_rel else None},
        "theta_final_over_true": [float(v) for v in (theta0 * torch.exp(logs)).detach().numpy() / theta_true],
        "params": list(PARAMS)}))
    sim.close()


if __name__ == "__main__":
    main()
