"""Benchmark: cell-steps/s of the quasi-static STZ step at 1024^2 on one GPU.

Workload: simple shear 1024x1024 cells (x periodic, walls at y = +/-1 moving
at +/-U), chi = 600 K + 25 K seeded Gaussian field (correlation 3 cells),
zero initial stress, dt = 0.2 (5000 steps reach T = 1000: elastic loading,
yield near step 2100, then plastic flow and a shear band), V(2,2) cycles to
the reference's own residual threshold.  By default the timed steps are the
whole run after 10 warm-up steps, so the number covers both phases.  A step is one
shear_sim::step_forward_quasistatic: explicit ENO/STZ update, projection
source, adaptive multigrid solve, stress correction and boundaries.

The headline value is the mixed-precision form (FP32 V-cycles inside FP64
iterative refinement); the plain FP64 form runs the same steps from the same
state, and the two are compared field by field.  The roofline object is for
the dominant kernel (the top-level smoother colour pass of the mixed form),
timed live with HIP events on the simulation's own stream.  The CPU baseline
is the reference's own code (oracle/_ref) on a bounded sample of the same
workload, or the FP64 restatement when the reference build is absent.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s at 1024^2: FP32-MG+FP64 refine vs FP64 @1e-8 rel; % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def group_err(a, b, ix):
    return float(np.max(np.abs(a[ix] - b[ix])) / max(np.max(np.abs(b[ix])), 1e-300))


def run_form(params, planes, form, dt, warmup, steps, nu, tol, stream, keep=()):
    import torch
    from stz_amd import ShearSim
    sim = ShearSim(params, form)
    sim.set_stream(stream.cuda_stream)
    sim.set_fields(planes)
    sim.set_vel(np.zeros((params["n"] + 1, params["m"], 2)))
    sim.time = 0.0
    sim.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol)
    sim.step_forward_quasistatic(dt, warmup)
    sim.avg_iters()
    torch.cuda.synchronize()
    # the timed run is split where states are kept for the checks; every
    # split is a synchronisation the unsplit run would not have, so the
    # figure is, if anything, slightly pessimistic
    kept, done, elapsed = {}, warmup, 0.0
    for stop in sorted(k for k in keep if warmup < k < warmup + steps) + [warmup + steps]:
        t0 = time.perf_counter()
        sim.step_forward_quasistatic(dt, stop - done)
        sim.synchronize()
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        done = stop
        if stop in keep:
            kept[stop] = (sim.get_fields(), sim.get_vel(), sim.time)
    cycles = sim.avg_iters()
    info = sim.last_solve()
    return sim, elapsed, cycles, info, kept


def smoother_roofline(sim, nodes, bytes_per_node_sweep, stream, launches=100):
    """average duration of one launch of the top-level smoother (two fused
    four-colour sweeps per launch), from HIP events recorded on the stream
    the kernels run on; algorithmic bytes = 2 sweeps x (read z, read b,
    write z) per node"""
    import torch
    sim.mg_smooth(0, 2)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sim.mg_smooth(0, 2 * launches)
    e1.record(stream)
    e1.synchronize()
    per_launch_s = e0.elapsed_time(e1) * 1e-3 / launches
    algo_bytes = 2 * bytes_per_node_sweep * nodes
    achieved = algo_bytes / per_launch_s / 1e9
    return per_launch_s, algo_bytes, achieved


def cpu_baseline(params, samples, dt, tol, phase_steps, budget_s=10.0, max_steps=30):
    """The reference's own code (or the FP64 restatement) timed on bounded
    samples of the same run: a few steps from each kept GPU state, one per
    phase (elastic loading, plastic flow).  The full-run figure weights the
    per-step times by how many of the run's steps fall in each phase."""
    from oracle import pyoracle
    kind = "reference" if pyoracle.reference_available() else "port"
    cls = pyoracle.Reference if kind == "reference" else pyoracle.Oracle
    per_step, desc, finals = {}, [], {}
    for phase, (fields, vel, t) in samples.items():
        sim = cls(params)
        sim.set_fields(fields)
        sim.set_vel(vel)
        sim.time = t
        sim.set_tol_mult(tol)
        sim.step(dt, 1)  # multigrid setup and OpenMP thread tuning
        steps, t0 = 0, time.perf_counter()
        while steps < max_steps and time.perf_counter() - t0 < budget_s:
            sim.step(dt, 1)
            steps += 1
        el = time.perf_counter() - t0
        per_step[phase] = el / steps
        desc.append("%s: %d steps from t=%.0f, %.3f s/step" % (phase, steps, t, el / steps))
        finals[phase] = (1 + steps, sim.get_fields())
        sim.close()
    total_t = sum(phase_steps[p] * per_step[p] for p in per_step)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    if kind == "port":
        cores = 1
    cells = params["m"] * params["n"]
    return {"value": cells * sum(phase_steps.values()) / total_t, "unit": "cell-steps/s", "cores": cores,
            "kind": kind,
            "sample": "%dx%d simple shear at the same threshold; %s; weighted by %s steps per phase" %
                      (params["m"], params["n"], "; ".join(desc), phase_steps)}, finals


def gpu_vs_cpu(params, samples, finals, dt, nu, tol):
    """the GPU FP64 form stepped from each CPU sample's start for the same
    number of steps: relative difference per field group from the CPU's
    final state (parity at the full size)"""
    from stz_amd import ShearSim
    out = {}
    for phase, (fields, vel, t) in samples.items():
        nsteps, ref = finals[phase]
        g = ShearSim(params, "fp64")
        g.set_fields(fields)
        g.set_vel(vel)
        g.time = t
        g.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol)
        g.step_forward_quasistatic(dt, nsteps)
        f = g.get_fields()
        g.close()
        out[phase] = {"steps": nsteps, "stress": group_err(f, ref, [2, 3, 4, 5]), "chi": group_err(f, ref, [6]),
                      "velocity": group_err(f, ref, [0, 1]), "refmap": group_err(f, ref, [7, 8])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--nu", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-4,
                    help="residual threshold as a multiple of the reference's")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU only (N=%d requested); the step "
              "would shard into y-slabs across GPUs, which is not built" % args.gpus)
        sys.exit(2)

    import torch
    from stz_amd import problems
    from stz_amd.initial import simple_shear_state

    N = args.size
    params = problems.simple_shear(N)
    dt = problems.T_END / 5000
    planes = simple_shear_state(params, 600.0, 25.0, 3.0, seed=1)
    stream = torch.cuda.Stream()
    cells = N * N

    log("bench: %dx%d, warmup %d, steps %d, V(%d,%d), threshold %g x reference" %
        (N, N, args.warmup, args.steps, args.nu, args.nu, args.tol))
    total = args.warmup + args.steps
    checks = [c for c in (500, 1000, 2000, 3000, 3500, 3750, 4000, 4500) if c < total]
    elastic_at, plastic_at = min(1000, total // 4), min(3000, 3 * total // 5)
    keep = sorted(set(checks + [elastic_at, plastic_at]))
    res = {}
    for form in ("fp64", "mixed"):
        sim, el, cyc, info, kept = run_form(params, planes, form, dt, args.warmup, args.steps, args.nu, args.tol,
                                            stream, keep)
        res[form] = dict(sim=sim, elapsed=el, cycles=cyc, info=info, kept=kept,
                         value=cells * args.steps / el, ms=1e3 * el / args.steps)
        log("bench: %-5s %.3f ms/step, %.3e cell-steps/s, %.2f cycles/solve, last %s" %
            (form, res[form]["ms"], res[form]["value"], cyc, info))

    def errs_of(fa, fb):
        return {"stress": group_err(fa, fb, [2, 3, 4, 5]), "chi": group_err(fa, fb, [6]),
                "velocity": group_err(fa, fb, [0, 1]), "refmap": group_err(fa, fb, [7, 8])}
    lockstep = {k: errs_of(res["mixed"]["kept"][k][0], res["fp64"]["kept"][k][0])
                for k in checks if k in res["mixed"]["kept"] and k in res["fp64"]["kept"]}
    final = errs_of(res["mixed"]["sim"].get_fields(), res["fp64"]["sim"].get_fields())
    for k, e in lockstep.items():
        log("bench: mixed vs fp64 at step %d: %s" % (k, e))
    log("bench: mixed vs fp64 at step %d: %s" % (total, final))
    # one step from the same FP64 state in both forms
    one = {}
    st = res["fp64"]["kept"].get(plastic_at)
    for form in (("fp64", "mixed") if st else ()):
        fields, vel, t = st
        g = res[form]["sim"].__class__(params, form)
        g.set_fields(fields)
        g.set_vel(vel)
        g.time = t
        g.set_mg_config(nu1=args.nu, nu2=args.nu, tol_mult=args.tol)
        g.step_forward_quasistatic(dt, 1)
        one[form] = g.get_fields()
        g.close()
    onestep = errs_of(one["mixed"], one["fp64"]) if one else None
    log("bench: one step from the step-%d state, mixed vs fp64: %s" % (plastic_at, onestep))
    errs = {"one_step_from_plastic_state": onestep,
            "lockstep_before_instability": max((max(e["stress"], e["chi"]) for k, e in lockstep.items()
                                                if k <= 3750), default=None),
            "final_after_shear_band": final}

    nodes = N * (N + 1)
    per_launch, algo_bytes, achieved = smoother_roofline(res["mixed"]["sim"], nodes, 24.0, stream)
    per_launch64, algo_bytes64, achieved64 = smoother_roofline(res["fp64"]["sim"], nodes, 48.0, stream)
    log("bench: tiled smoother launch (2 sweeps) fp32 %.2f us (%.0f GB/s), fp64 %.2f us (%.0f GB/s)" %
        (per_launch * 1e6, achieved, per_launch64 * 1e6, achieved64))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        # HBM bytes per launch from FETCH_SIZE / WRITE_SIZE, scaled by a copy kernel
        if tj.get("kernel") == "smooth_tile_kernel<float,64,32,2,0>" and tj.get("level", "").startswith("%dx" % N):
            traffic = tj.get("traffic_bytes_per_launch")

    cpu, vs_cpu = None, None
    if not args.no_cpu and elastic_at in res["fp64"]["kept"] and plastic_at in res["fp64"]["kept"]:
        # steps before the first yield (t ~ 425 at this loading rate) are elastic
        n_el = min(total, int(425.0 / dt))
        samples = {"elastic": res["fp64"]["kept"][elastic_at], "plastic": res["fp64"]["kept"][plastic_at]}
        cpu, finals = cpu_baseline(params, samples, dt, args.tol, {"elastic": n_el, "plastic": total - n_el})
        log("bench: cpu baseline %s" % cpu)
        vs_cpu = gpu_vs_cpu(params, samples, finals, dt, args.nu, args.tol)
        log("bench: GPU FP64 vs the CPU (%s) from the same states: %s" % (cpu["kind"], vs_cpu))

    mixed, fp64 = res["mixed"], res["fp64"]
    line = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "simple shear %dx%d, chi 600K+25K seeded (corr 3 cells), dt %.3g, %d steps, V(%d,%d), "
                        "residual threshold %g x reference" % (N, N, dt, total, args.nu, args.nu, args.tol),
            "form": "FP32 V-cycles in FP64 iterative refinement",
            "fp64_value": fp64["value"],
            "fp64_ms_per_step": fp64["ms"],
            "mixed_over_fp64": mixed["value"] / fp64["value"],
            "mixed_vs_fp64_rel_err": errs,
            "cycles_per_solve": {"mixed": mixed["cycles"], "fp64": fp64["cycles"]},
            "fp64_vs_cpu_same_start": vs_cpu,
            "smoother_fp64_GBps": achieved64,
        },
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "smooth_tile_kernel<float,64,32,2,*> (top level, 2 sweeps)",
                     "algorithmic_bytes_per_launch": algo_bytes, "avg_launch_us": per_launch * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
