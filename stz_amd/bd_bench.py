"""Throughput of the free-boundary quasi-static step (bd_sim) on one GPU.

    python -m stz_amd.bd_bench [--size 4096] [--steps 10] [--warmup 2]
                               [--precision mixed|fp64]

Workload (plate_hole4096 in stz_amd.problems): a plate |y| < 0.9 in
the box [-1,1]^2 with a central hole of radius 0.4 (0.2 L), held by grips on
the left and right edges that move apart at nominal strain rate 1e-4
(u = -+1e-4), material of bd_test.cc, chi = 580 K + 30 K seeded Gaussian
field with correlation length 4 cells inside the plate, dt = 0.5; the
projection's V-cycle preconditioner in FP32 inside FP64 BiCGStab (mixed,
the default) or in FP64.  Prints one JSON line: cell-steps/s over the m*n
cells, ms per step, each timed step's BiCGStab iteration count and the ms
per iteration (the count moves with rounding from build to build; the time
per iteration compares builds), and the work counts of the last step (march
rounds, extrapolation waves, BiCGStab iterations).
"""
import argparse
import json
import time

import numpy as np

from . import problems
from .bd import BdSim
from .initial import random_chi


def setup(N, seed=4):
    p = problems.bd_params(N)
    sim = BdSim(p)
    sim.set_level_set(problems.plate_hole_phi(N))
    st, _ = sim.status()
    w = problems.BD_WORKLOADS["plate_hole4096"]
    chi = random_chi(p, w["chi_mean"], w["chi_std"], w["corr"], seed)
    sim.set_cell("chi", np.where(st < 4, chi, 0.0))
    X, Y = sim.node_coords()
    sim.set_node("X", X)
    sim.set_node("Y", Y)
    u_grip = w["strain_rate"] * 0.5 * (p["bx"] - p["ax"])
    sim.set_edge_velocity(-u_grip, u_grip)
    sim.post_process(True)
    return sim


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dt", type=float, default=0.5)
    ap.add_argument("--nu", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=("mixed", "fp64"))
    a = ap.parse_args()
    sim = setup(a.size)
    sim.set_solver(a.nu, a.nu)
    sim.set_precision(a.precision)
    sim.step_forward_quasistatic(a.dt, a.warmup)
    sim.synchronize()
    # one step per call, so that every step's BiCGStab iteration count is
    # read (the count moves with rounding from build to build, and the step
    # time with it)
    iters = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.step_forward_quasistatic(a.dt, 1)
        iters.append(sim.step_stats()["bicgstab_iterations"])
    sim.synchronize()
    el = time.perf_counter() - t0
    out = {"workload": "plate with hole, grips, %dx%d cells" % (a.size, a.size), "nu": a.nu,
           "precision": a.precision,
           "steps": a.steps, "ms_per_step": 1e3 * el / a.steps,
           "cell_steps_per_s": a.size * a.size * a.steps / el,
           "bicgstab_iterations": iters,
           "ms_per_iteration": 1e3 * el / max(1, sum(iters)),
           "last_solve": sim.last_solve(), "last_step": sim.step_stats()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
