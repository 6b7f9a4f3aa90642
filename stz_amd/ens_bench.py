"""Benchmark of the batched ensemble: many independent simple-shear
simulations of different sizes stepped together on one GPU.

    python -m stz_amd.ens_bench [--members 256] [--steps 200] [--warmup 60]
                                [--form mixed] [--seed 0] [--serial-steps 3]
                                [--one-mode] [--direct-nodes 0]

`--warmup 0 --steps 1000 --one-mode` times the whole 1000-step run of
every member from t = 0.

Workload: `members` simulations, the side of each drawn uniformly from
{64, 128, 256, 384, 512} (numpy PCG64, `seed`), chi = 600 K + 20 K seeded
Gaussian field (correlation length 2 cells, seed per member), zero initial
stress and velocity, dt = 0.2 (the 1024^2 workload's step), V(2,2), residual
threshold 1e-4 x the reference's, FP32 multigrid inside FP64 refinement.
In the mixed form the V-cycles go down to each member's workgroup coarse
solve (`--direct-nodes 0`): the exact Fourier-mode level, which a simulation on its
own stops at, reads each member's mode inverses (9 MB at 64 x 65) and stages
its source once per mode and row block, which costs a batch of 256 members
more than the coarse levels it saves.

Prints one JSON line: aggregate cell-steps/s of the batched step replayed
from HIP graphs (the library's default), the same with direct launches, and
the members stepped one at a time through the single-simulation path
(`serial-steps` steps each) for comparison.
"""
import argparse
import json
import sys
import time

import numpy as np

SIDES = (64, 128, 256, 384, 512)


def build(members, seed, form, nu, tol, direct_nodes=None):
    from . import problems
    from .ensemble import Ensemble
    from .initial import simple_shear_state
    rng = np.random.default_rng(seed)
    sides = [int(s) for s in rng.choice(SIDES, size=members)]
    params = [problems.simple_shear(s) for s in sides]
    ens = Ensemble(params, form)
    for k, p in enumerate(params):
        ens[k].set_fields(simple_shear_state(p, 600.0, 20.0, 2.0, seed=1000 + seed * members + k))
        ens[k].set_vel(np.zeros((p["n"] + 1, p["m"], 2)))
        ens[k].time = 0.0
    kw = {} if direct_nodes is None else {"direct_nodes": direct_nodes}
    ens.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol, **kw)
    return ens, sides


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--form", default="mixed", choices=("mixed", "fp64"))
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--nu", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--serial-steps", type=int, default=3)
    ap.add_argument("--block", type=int, default=10, help="steps per block of one launch mode")
    ap.add_argument("--direct-nodes", type=int, default=None,
                    help="largest coarse level solved exactly (stz_mg_config.direct_nodes); 0: none, the "
                         "V-cycles go down to the workgroup coarse solve, which batches better than the "
                         "exact level's mode inverses (default 0 in the mixed form; the FP64 form keeps "
                         "the library's default, its 384-side members need their exact 48x49 level)")
    ap.add_argument("--tune-sweeps", action="store_true",
                    help="pick V(nu, nu), nu in 1..4, with the ensemble's sweep-count tuner after the warm-up")
    ap.add_argument("--one-mode", action="store_true",
                    help="time every step replayed from graphs (no direct-launch comparison)")
    args = ap.parse_args()
    from . import problems
    dt = problems.T_END / 5000

    t0 = time.perf_counter()
    direct = args.direct_nodes if args.direct_nodes is not None else (0 if args.form == "mixed" else None)
    ens, sides = build(args.members, args.seed, args.form, args.nu, args.tol, direct)
    cells = sum(s * s for s in sides)
    print("ens_bench: %d members, %.2fM cells, set-up %.1f s" % (len(sides), cells / 1e6, time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    def timed(nsteps, graphs):
        ens.use_graphs(graphs)
        ens[0].synchronize()
        t = time.perf_counter()
        ens.step_forward_quasistatic(dt, nsteps)
        return time.perf_counter() - t

    # the predictors start at 8 cycles (tgmg_config.hh:26) and settle within
    # a few tens of steps, and the cycle count keeps drifting with the
    # loading: the two launch modes alternate in blocks of `block` steps so
    # that both see the same phases of the run
    if args.warmup:
        timed(args.warmup, True)
    tuned = None
    if args.tune_sweeps:
        t = time.perf_counter()
        tuned = ens.tune_sweeps(dt)
        print("ens_bench: sweep tuner %s in %.1f s" % (tuned, time.perf_counter() - t), file=sys.stderr, flush=True)
        tuned = {"choice": list(tuned["choice"]),
                 "ms_per_step": {"V(%d,%d)" % k: v for k, v in tuned["ms_per_step"].items()}}
    block = args.steps if args.one_mode else max(1, min(args.block, args.steps // 2))
    rounds = 1 if args.one_mode else max(1, args.steps // (2 * block))
    el, el_d, cyc_g = 0.0, 0.0, []
    s0 = ens.stats()
    for rnd in range(rounds):
        for m in ens:
            m.avg_iters()
        el += timed(block, True)
        cyc_g.append(float(np.mean([m.avg_iters() for m in ens])))
        if not args.one_mode:
            el_d += timed(block, False)
    s1 = ens.stats()
    nsteps = rounds * block
    cyc = float(np.mean(cyc_g))
    timed_steps = nsteps if args.one_mode else 2 * nsteps
    slots = (s1["slots"] - s0["slots"]) / timed_steps
    extra = (s1["extra_slots"] - s0["extra_slots"]) / timed_steps
    print("ens_bench: graphs  %.2f ms/step, %.3e cell-steps/s, %.2f cycles/solve, %.2f slots/step "
          "(%.2f beyond plan), tiles %s" % (1e3 * el / nsteps, cells * nsteps / el, cyc, slots, extra,
                                            ens.tile_choice()), file=sys.stderr, flush=True)
    if not args.one_mode:
        print("ens_bench: direct  %.2f ms/step" % (1e3 * el_d / nsteps), file=sys.stderr, flush=True)

    # the same members one at a time through stz_sim_step_quasistatic
    for m in ens:
        m.avg_iters()
    t = time.perf_counter()
    for m in ens:
        m.step_forward_quasistatic(dt, args.serial_steps)
    el_s = time.perf_counter() - t
    cyc_s = float(np.mean([m.avg_iters() for m in ens]))
    serial = cells * args.serial_steps / el_s
    print("ens_bench: one at a time %.2f ms per ensemble step, %.3e cell-steps/s, %.2f cycles/solve" %
          (1e3 * el_s / args.serial_steps, serial, cyc_s), file=sys.stderr, flush=True)

    value = cells * nsteps / el
    hist = {str(s): sides.count(s) for s in SIDES}
    print(json.dumps({
        "metric": "ensemble cell-steps/s (batched quasi-static step, all members)",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": nsteps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / nsteps, "higher_is_better": True,
        "dtype": "f32+f64" if args.form == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": "ensemble of %d simple-shear simulations, sides %s, chi 600K+20K (corr 2 cells), "
                               "dt %g, V(%d,%d), threshold %g x reference" %
                               (len(sides), hist, dt, *(tuned["choice"] if tuned else (args.nu, args.nu)),
                                args.tol),
                   "cells": cells, "direct_nodes": ens[0].mg_config().direct_nodes,
                   "cycles_per_solve": cyc, "slots_per_step": slots,
                   "slots_beyond_plan_per_step": extra,
                   "direct_launch_ms_per_step": None if args.one_mode else 1e3 * el_d / nsteps,
                   "tiles": ens.tile_choice(),
                   "sweep_tuner": tuned,
                   "one_at_a_time_cycles_per_solve": cyc_s,
                   "one_at_a_time_cell_steps_per_s": serial,
                   "batched_over_one_at_a_time": value / serial}}))
    ens.close()


if __name__ == "__main__":
    main()
