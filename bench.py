"""Benchmark: cell-steps/s of the quasi-static STZ step at 1024^2 on one GPU.

Workload: simple shear 1024x1024 cells (x periodic, walls at y = +/-1 moving
at +/-U), chi = 600 K + 25 K seeded Gaussian field (correlation 3 cells),
zero initial stress, dt = 0.2 (5000 steps reach T = 1000: elastic loading,
first yield near t = 425, then plastic flow and a shear band), V(2,2) cycles
to 1e-4 x the reference's residual threshold.  The timed steps are the
`--steps` after `--warmup` warm-up steps from t = 0.  A step is one
shear_sim::step_forward_quasistatic: explicit ENO/STZ update, projection
source, adaptive multigrid solve, stress correction and boundaries.

The headline value is the mixed-precision form (FP32 V-cycles inside FP64
iterative refinement); the plain FP64 form runs the same steps from the same
state, and the two are compared field by field.

Both forms use the reference's adaptive V-cycle count (tgmg::solve with a
tgmg_predict, tgmg.in.cc:266-301: run the predicted number of cycles, test
the residual, add cycles until it is below the threshold, then one extra
cycle) with settings for this solver: first test after 1 cycle, decay 2,
one more cycle per failed test (stz_mg_config.predict_unit), against the
reference's compile-time 8 and 32 (tgmg_config.hh:26-37) and triangular
steps (1, 2, 3, ... cycles between tests), which suit a CPU solver whose
cycles each gain less and whose tests cost a residual sweep.  Here a test
costs a norm launch and the next cycle is already enqueued while the host
reads it.  Every solve still stops only below the same threshold: over the
whole 5000-step run the cycle counts are the reference predictor's to
0.1 %, and at the start of the run (the benchmark's window) the settings
skip cycles the residual no longer needs.  The same steps with the
reference's settings are timed as well and reported under
`with_reference_predictor`.

`roofline` is for the kernel that takes the largest share of a mixed step
(top-level smoothing launches, advection), each timed live with HIP events
on the simulation's own stream; its algorithmic bytes follow the scope
table's per-node figures (24 B FP32 / 48 B FP64 per node and sweep, 4.5 words
(18 B FP32, 36 B FP64) per fine node for a restriction or a prolongation, 32 B for the FP64 read-
modify-write of vel += e, 136 B per cell for the advection).  `roofline_step`
is the whole step: cells x bytes per cell-step (312 + 176 k + 80 r mixed,
312 + 352 k FP64, for k V-cycles and r refinements per solve) / time.

The CPU baseline is the reference's own code (oracle/_ref) timed on the
host's cores on bounded samples of the same run (or the FP64 restatement
when the reference build is absent); from the same samples both step twice
more at a threshold of 1e-15 x the reference's and the GPU FP64 form is
compared with them (parity at the full size).
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
YIELD_T = 425.0      # first yield of the seeded 1024^2 workload (elastic before)
PLASTIC_T = 600.0    # plastic-flow sample time, when the run reaches it
ELASTIC_T = 100.0    # elastic-loading sample time, when the run reaches it
REFERENCE_PREDICT = (8, 32, 0)  # tgmg_predict_init, _decay (tgmg_config.hh:26,37), triangular steps
TIGHT = 1e-15        # threshold multiple of the full-size parity check: the
                     # reference's own solution moves < 2e-11 (stress) from there to 1e-16 x


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def group_err(a, b, ix):
    return float(np.max(np.abs(a[ix] - b[ix])) / max(np.max(np.abs(b[ix])), 1e-300))


def errs_of(fa, fb):
    return {"stress": group_err(fa, fb, [2, 3, 4, 5]), "chi": group_err(fa, fb, [6]),
            "velocity": group_err(fa, fb, [0, 1]), "refmap": group_err(fa, fb, [7, 8])}


def run_form(params, planes, form, dt, warmup, steps, nu, tol, stream, keep=(), predict=(1, 2, 1)):
    import torch
    from stz_amd import ShearSim
    sim = ShearSim(params, form)
    sim.set_stream(stream.cuda_stream)
    sim.set_fields(planes)
    sim.set_vel(np.zeros((params["n"] + 1, params["m"], 2)))
    sim.time = 0.0
    sim.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol, predict_init=predict[0], predict_decay=predict[1],
                      predict_unit=predict[2])
    kept = {}
    if 0 in keep:
        kept[0] = (sim.get_fields(), sim.get_vel(), sim.time)
    sim.step_forward_quasistatic(dt, warmup)
    if warmup in keep:
        kept[warmup] = (sim.get_fields(), sim.get_vel(), sim.time)
    sim.avg_iters()
    torch.cuda.synchronize()
    # the timed run is split where states are kept for the checks; every
    # split is a synchronisation the unsplit run would not have, so the
    # figure is, if anything, slightly pessimistic
    done, elapsed = warmup, 0.0
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


def timed(stream, fn, launches):
    """average seconds per launch of fn(count) between two HIP events on the
    stream the kernels run on"""
    import torch
    fn(1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn(launches)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e-3 / launches


def kernel_rooflines(sims, params, dt, nu, stream, launches=50):
    """live HIP-event timing of the kernels of a step: the top level's
    pre-smoothing launch (nu sweeps + fused restriction) and post-smoothing
    launch (prolongation + nu sweeps, mixed: + FP64 vel += e), and the
    advection kernel; algorithmic bytes per launch from the scope table's
    per-node figures"""
    N, n1 = params["m"], params["n"] + 1
    nodes, cells = N * n1, N * params["n"]
    # the top level's tile shape, as the library chooses it (mg_tiled.hip:
    # big_tiles, small_tiles)
    if N % 64 == 0 and (N // 64) * ((n1 + 31) // 32) >= 256:
        shape = "64,34"
    elif (N // 32) * ((n1 + 15) // 16) < 100:
        shape = "16,16"
    else:
        shape = "32,16"
    out = {}
    for form, sim in sims.items():
        w = 4 if form == "mixed" else 8        # bytes per word of the V-cycle
        sweep = 6 * w                          # read z, b; write z (vec2)
        transfer = 4.5 * w                     # restriction or prolongation: 4.5 words per fine node
        ks = {
            "top_pre": ("smooth_tile_kernel<%s,%s,%d,%d,true,true> (top level: %d sweeps + fused restriction)"
                        % ("float" if w == 4 else "double", shape, nu, 1 if w == 4 else 0, nu),
                        nodes * (nu * sweep + transfer), lambda c, s=sim: s.mg_vcycle_launch(0, 1, c)),
            "top_post": ("smooth_tile_kernel<%s,%s,%d,2,false,true> (top level: prolongation + %d sweeps%s)"
                         % ("float" if w == 4 else "double", shape, nu, nu, " + FP64 vel += e" if w == 4 else ""),
                         nodes * (nu * sweep + transfer + (32 if w == 4 else 0)),
                         lambda c, s=sim: s.mg_vcycle_launch(0, 2, c)),
        }
        if form == "mixed":
            ks["advect"] = ("advect_kernel (ENO2 advection + STZ plastic update)", cells * 136,
                            lambda c, s=sim: s.advection_launch(dt, c))
        for key, (name, algo, fn) in ks.items():
            sec = timed(stream, fn, launches)
            out["%s_%s" % (form, key)] = {"kernel": name, "avg_launch_us": sec * 1e6,
                                          "algorithmic_bytes_per_launch": algo,
                                          "achieved_GBps": algo / sec / 1e9}
    return out


def cpu_baseline(params, samples, dt, tol, phase_steps, budget_s=10.0, max_steps=30):
    """The reference's own code (or the FP64 restatement) timed on bounded
    samples of the same run: a few steps from each kept GPU state, one per
    phase the timed steps cover.  The run's figure weights the per-step
    times by how many of the timed steps fall in each phase."""
    from oracle import pyoracle
    kind = "reference" if pyoracle.reference_available() else "port"
    cls = pyoracle.Reference if kind == "reference" else pyoracle.Oracle
    per_step, desc = {}, []
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
        desc.append("%s: %d steps from t=%.1f, %.3f s/step" % (phase, steps, t + dt, el / steps))
        sim.close()
    total_t = sum(phase_steps[p] * per_step[p] for p in per_step)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    if kind == "port":
        cores = 1
    cells = params["m"] * params["n"]
    return {"value": cells * sum(phase_steps.values()) / total_t, "unit": "cell-steps/s", "cores": cores,
            "kind": kind,
            "sample": "%dx%d simple shear at the benchmark's threshold; %s; weighted by %s timed steps per phase" %
                      (params["m"], params["n"], "; ".join(desc), phase_steps)}


def gpu_vs_reference(params, samples, dt, steps=2):
    """from each sample state, the reference (or the restatement) and the GPU
    FP64 form step `steps` times at 1e-15 x the reference's threshold:
    relative difference per field group (parity at the full size, solves
    converged far below the smoothers' differences)"""
    from oracle import pyoracle
    from stz_amd import ShearSim
    cls = pyoracle.Reference if pyoracle.reference_available() else pyoracle.Oracle
    out = {}
    for phase, (fields, vel, t) in samples.items():
        r = cls(params)
        r.set_fields(fields)
        r.set_vel(vel)
        r.time = t
        r.set_tol_mult(TIGHT)
        r.step(dt, steps)
        ref = r.get_fields()
        r.close()
        g = ShearSim(params, "fp64")
        g.set_fields(fields)
        g.set_vel(vel)
        g.time = t
        g.set_mg_config(nu1=2, nu2=2, tol_mult=TIGHT)
        g.step_forward_quasistatic(dt, steps)
        out[phase] = dict(steps=steps, t0=t, **errs_of(g.get_fields(), ref))
        g.close()
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
    ap.add_argument("--predict-init", type=int, default=1,
                    help="V-cycles before the first convergence test of the first solve (tgmg_predict_init)")
    ap.add_argument("--predict-decay", type=int, default=2,
                    help="the predictor's decay after a converged solve (tgmg_predict_decay)")
    ap.add_argument("--predict-unit", type=int, default=1, choices=(0, 1),
                    help="1: one more V-cycle per failed test; 0: k more (k = 1, 2, ...), as the reference")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    predict = (args.predict_init, args.predict_decay, args.predict_unit)
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

    log("bench: %dx%d, warmup %d, steps %d, V(%d,%d), threshold %g x reference, predictor init %d decay %d unit %d" %
        (N, N, args.warmup, args.steps, args.nu, args.nu, args.tol, *predict))
    total = args.warmup + args.steps
    checks = [c for c in (500, 1000, 2000, 3000, 3500, 3750, 4000, 4500) if c < total]
    # phase samples by simulation time: t = ELASTIC_T when the timed steps
    # reach it, else the end of the timed steps when it is still elastic
    # (t < YIELD_T), else their start if that is; and
    # t = PLASTIC_T when the timed steps reach it.  (Right after the wall ramp,
    # t ~ 1, the stresses are so small that a relative comparison measures the
    # solvers' rounding floor rather than the step.)
    samples_at = {}
    elastic_step = int(round(ELASTIC_T / dt))
    if args.warmup <= elastic_step < total:
        samples_at["elastic"] = elastic_step
    elif total * dt < YIELD_T:
        samples_at["elastic"] = total
    elif args.warmup * dt < YIELD_T:
        samples_at["elastic"] = args.warmup
    plastic_step = int(round(PLASTIC_T / dt))
    if args.warmup <= plastic_step < total:
        samples_at["plastic"] = plastic_step
    elif args.warmup * dt >= YIELD_T:
        samples_at["plastic"] = args.warmup
    keep = sorted(set(checks) | set(samples_at.values()))
    res = {}
    for form in ("fp64", "mixed"):
        sim, el, cyc, info, kept = run_form(params, planes, form, dt, args.warmup, args.steps, args.nu, args.tol,
                                            stream, keep, predict)
        res[form] = dict(sim=sim, elapsed=el, cycles=cyc, info=info, kept=kept,
                         value=cells * args.steps / el, ms=1e3 * el / args.steps)
        log("bench: %-5s %.3f ms/step, %.3e cell-steps/s, %.2f cycles/solve, last %s" %
            (form, res[form]["ms"], res[form]["value"], cyc, info))
    # the same steps with the reference's predictor constants (tgmg_config.hh:26-45)
    ref_pred = {}
    if predict != REFERENCE_PREDICT:
        for form in ("fp64", "mixed"):
            sim, el, cyc, info, _ = run_form(params, planes, form, dt, args.warmup, args.steps, args.nu, args.tol,
                                             stream, (), REFERENCE_PREDICT)
            sim.close()
            ref_pred[form] = {"value": cells * args.steps / el, "ms_per_step": 1e3 * el / args.steps,
                              "cycles_per_solve": cyc}
            log("bench: %-5s with the reference's predictor: %.3f ms/step, %.3e cell-steps/s, %.2f cycles/solve" %
                (form, ref_pred[form]["ms_per_step"], ref_pred[form]["value"], cyc))

    lockstep = {k: errs_of(res["mixed"]["kept"][k][0], res["fp64"]["kept"][k][0])
                for k in checks if k in res["mixed"]["kept"] and k in res["fp64"]["kept"]}
    final = errs_of(res["mixed"]["sim"].get_fields(), res["fp64"]["sim"].get_fields())
    for k, e in lockstep.items():
        log("bench: mixed vs fp64 at step %d: %s" % (k, e))
    log("bench: mixed vs fp64 at step %d: %s" % (total, final))
    # one step from the same FP64 state in both forms (the latest sample)
    one, one_phase = {}, None
    if samples_at:
        one_phase = "plastic" if "plastic" in samples_at else "elastic"
        fields, vel, t = res["fp64"]["kept"][samples_at[one_phase]]
        for form in ("fp64", "mixed"):
            g = res[form]["sim"].__class__(params, form)
            g.set_fields(fields)
            g.set_vel(vel)
            g.time = t
            g.set_mg_config(nu1=args.nu, nu2=args.nu, tol_mult=args.tol, predict_init=predict[0],
                            predict_decay=predict[1], predict_unit=predict[2])
            g.step_forward_quasistatic(dt, 1)
            one[form] = g.get_fields()
            g.close()
    onestep = errs_of(one["mixed"], one["fp64"]) if one else None
    log("bench: one step from the %s state, mixed vs fp64: %s" % (one_phase, onestep))
    errs = {"one_step_from_%s_state" % one_phase: onestep,
            "lockstep_before_instability": max((max(e["stress"], e["chi"]) for k, e in lockstep.items()
                                                if k <= 3750), default=None),
            "final": final}

    # step-level roofline from the scope table's bytes per cell-step
    extra = 1
    step_roof = {}
    for form in ("fp64", "mixed"):
        k = res[form]["cycles"] + extra
        bps = 312 + 176 * k + 80 * k if form == "mixed" else 312 + 352 * k
        ach = cells * bps * args.steps / res[form]["elapsed"] / 1e9
        step_roof[form] = {"bytes_per_cell_step": bps, "cycles_per_solve_run": k, "achieved_GBps": ach,
                           "frac": ach / HBM_PEAK_GBS}
    log("bench: step-level roofline %s" % step_roof)

    kern = kernel_rooflines({"mixed": res["mixed"]["sim"], "fp64": res["fp64"]["sim"]}, params, dt, args.nu,
                            stream)
    for k, v in kern.items():
        log("bench: %-16s %8.2f us  %6.0f GB/s  %s" % (k, v["avg_launch_us"], v["achieved_GBps"], v["kernel"]))
    # the dominant kernel of the mixed step: largest time per step
    per_step = {"mixed_top_pre": res["mixed"]["cycles"] + extra, "mixed_top_post": res["mixed"]["cycles"] + extra,
                "mixed_advect": 1}
    dom = max(per_step, key=lambda k: kern[k]["avg_launch_us"] * per_step[k])
    dk = kern[dom]

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        # HBM bytes per launch from FETCH_SIZE / WRITE_SIZE, scaled by a copy kernel
        entry = tj.get("kernels", {}).get(dom)
        if entry and entry.get("size") == N:
            traffic = entry.get("traffic_bytes_per_launch")

    cpu, vs_ref = None, None
    if not args.no_cpu and samples_at:
        samples = {ph: res["fp64"]["kept"][st] for ph, st in samples_at.items()}
        first = args.warmup
        n_el = max(0, min(total, int(YIELD_T / dt)) - first)
        phase_steps = {"elastic": n_el, "plastic": args.steps - n_el}
        phase_steps = {p: c for p, c in phase_steps.items() if c > 0 and p in samples}
        cpu = cpu_baseline(params, samples, dt, args.tol, phase_steps)
        log("bench: cpu baseline %s" % cpu)
        vs_ref = gpu_vs_reference(params, samples, dt)
        log("bench: GPU FP64 vs the CPU (%s) at %g x threshold: %s" % (cpu["kind"], TIGHT, vs_ref))

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
            "workload": "simple shear %dx%d, chi 600K+25K seeded (corr 3 cells), dt %.3g, steps %d..%d of 5000 "
                        "(t %.1f..%.1f), V(%d,%d), residual threshold %g x reference" %
                        (N, N, dt, args.warmup + 1, total, args.warmup * dt, total * dt, args.nu, args.nu,
                         args.tol),
            "form": "FP32 V-cycles in FP64 iterative refinement",
            "fp64_value": fp64["value"],
            "fp64_ms_per_step": fp64["ms"],
            "mixed_over_fp64": mixed["value"] / fp64["value"],
            "mixed_vs_fp64_rel_err": errs,
            "cycles_per_solve": {"mixed": mixed["cycles"], "fp64": fp64["cycles"]},
            "predictor": {"init": predict[0], "decay": predict[1], "unit": predict[2], "mult": 16},
            "with_reference_predictor": ref_pred or None,
            "fp64_vs_reference_at_1e-15": vs_ref,
            "kernels": kern,
        },
        "roofline": {"bound": "hbm", "achieved": dk["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dk["achieved_GBps"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dk["kernel"], "algorithmic_bytes_per_launch": dk["algorithmic_bytes_per_launch"],
                     "avg_launch_us": dk["avg_launch_us"]},
        "roofline_step": step_roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
