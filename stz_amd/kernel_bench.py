"""Kernel-level timing driver for profiling the step's dominant kernels.

    python -m stz_amd.kernel_bench [--size 1024] [--launches 20] [--copy-mib 256]

Runs, on one HIP stream, for the 1024^2 simple-shear workload of the
benchmark:
  1. `launches` launches each of the mixed form's top-level smoothing
     launches as a V-cycle issues them: the pre-smoothing (two four-colour
     sweeps from zero on the correction equation + the fused restriction of
     the residual) and the post-smoothing (prolongation of the coarse
     correction + two sweeps + the FP64 correction vel += e);
  2. `launches` advection steps (advect_kernel) from a state in plastic
     flow (smooth stresses whose deviatoric magnitude spans the yield
     threshold, the benchmark's seeded chi);
  3. `launches` device-to-device copies of `copy-mib` MiB (torch copy_), the
     kernel of known traffic that scales the hardware byte counters.
Each group is timed with HIP events and printed as one JSON line.  Under
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` the same program gives each
kernel's HBM bytes per dispatch.

Algorithmic bytes per launch (the scope table's per-node figures): pre
2 x 24 B + 18 B (restriction) per node, post 2 x 24 B + 18 B (prolongation)
+ 32 B (vel += e in FP64) per node, advection 136 B per cell.
"""
import argparse
import json
import sys

import numpy as np


def _timed(stream, fn, launches):
    import torch
    fn(1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    fn(launches)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e3 / launches


def plastic_state(p):
    """a smooth state whose deviatoric stress spans the yield threshold
    (sbar from 0.2 to 1.4, about a fifth of the cells above yield), with
    velocities of both signs and the seeded chi of the benchmark"""
    from .initial import simple_shear_state
    m, n = p["m"], p["n"]
    planes = simple_shear_state(p, 600.0, 25.0, 3.0, seed=1)
    x = (np.arange(m) + 0.5) / m
    y = (np.arange(-1, n + 1) + 0.5) / n
    X, Y = np.meshgrid(x, y)
    planes[0] = 1e-4 * np.sin(2 * np.pi * X) * np.cos(np.pi * Y)
    planes[1] = 1e-4 * np.cos(2 * np.pi * 3 * X) * np.sin(2 * np.pi * Y)
    planes[2] = 0.5 * np.sin(2 * np.pi * 2 * X + 1.0)
    planes[3] = 0.3 * np.cos(2 * np.pi * X) * np.sin(np.pi * 3 * Y)
    planes[4] = 1.1 * np.sin(2 * np.pi * 4 * X) * np.cos(np.pi * 2 * Y)
    planes[5] = 0.9 * np.cos(2 * np.pi * 2 * X + 0.5) * np.cos(np.pi * Y)
    vel = np.zeros((n + 1, m, 2))
    vel[..., 0] = planes[0][1:n + 2]
    vel[..., 1] = planes[1][1:n + 2]
    return planes, vel


def kernels(size, launches):
    import torch
    from . import problems
    from .sim import ShearSim
    p = problems.simple_shear(size)
    dt = problems.T_END / 5000
    stream = torch.cuda.Stream()
    sim = ShearSim(p, "mixed")
    sim.set_stream(stream.cuda_stream)
    sim.set_mg_config(nu1=2, nu2=2)
    rng = np.random.default_rng(7)
    sim.mg_setup(p["viscosity"] / dt)
    sim.mg_set_src(rng.standard_normal((size + 1, size, 2)))
    nodes, cells = size * (size + 1), size * size
    out = {}
    for key, mode, algo in (("mixed_top_pre", 1, nodes * (2 * 24 + 18)),
                            ("mixed_top_post", 2, nodes * (2 * 24 + 18 + 32))):
        us = _timed(stream, lambda c, m=mode: sim.mg_vcycle_launch(0, m, c), launches)
        out[key] = {"avg_launch_us": us, "algorithmic_bytes_per_launch": algo, "achieved_GBps": algo / us * 1e-3}
    planes, vel = plastic_state(p)
    sim.set_fields(planes)
    sim.set_vel(vel)
    us = _timed(stream, lambda c: sim.advection_launch(dt, c), launches)
    out["mixed_advect"] = {"avg_launch_us": us, "algorithmic_bytes_per_launch": cells * 136,
                           "achieved_GBps": cells * 136 / us * 1e-3}
    sim.close()
    return out


def copy(mib, launches):
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(launches):
            y.copy_(x)
        e1.record(stream)
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / launches
    return {"kernel": "torch copy_ (float32, %d MiB)" % mib, "launches": launches, "avg_launch_us": us,
            "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
            "achieved_GBps": 8 * n / us * 1e-3}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--launches", type=int, default=20)
    ap.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args(argv)
    out = kernels(a.size, a.launches)
    if a.copy_mib > 0:
        out["copy"] = copy(a.copy_mib, a.launches)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
