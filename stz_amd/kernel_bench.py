"""Kernel-level timing driver for profiling the multigrid smoother.

    python -m stz_amd.kernel_bench [--size 1024] [--precision mixed|fp64]
                                   [--launches 50] [--copy-mib 256]

Runs, on the simulation's own HIP stream:
  1. `launches` launches of the top-level LDS-tiled smoother (two fused
     four-colour Gauss-Seidel sweeps per launch) of the 1024^2 simple-shear
     projection operator, timed with HIP events;
  2. `launches` device-to-device copies of `copy-mib` MiB (torch copy_), the
     reference kernel of known traffic that scales hardware byte counters.
Under `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` the two give the
smoother's HBM bytes per launch and the counter's scale factor.

Algorithmic bytes of the smoother: per node and sweep it reads z and b and
writes z once, 3 vec2 words: 24 B in FP32, 48 B in FP64.
"""
import argparse
import json
import sys

import numpy as np


def smoother(size, precision, launches):
    import torch
    from . import problems
    from .sim import ShearSim
    p = problems.simple_shear(size)
    dt = problems.T_END / 5000
    sim = ShearSim(p, precision)
    stream = torch.cuda.Stream()
    sim.set_stream(stream.cuda_stream)
    rng = np.random.default_rng(7)
    sim.mg_setup(p["viscosity"] / dt)
    sim.mg_set_src(rng.standard_normal((size + 1, size, 2)))
    sim.mg_smooth(0, 2)
    sim.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sim.mg_smooth(0, 2 * launches)
    e1.record(stream)
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / launches
    nodes = size * (size + 1)
    algo = 2 * nodes * (24 if precision == "mixed" else 48)
    sim.close()
    return {"kernel": "smooth_tile_kernel<%s,64,32,2,0>" % ("float" if precision == "mixed" else "double"),
            "launches": launches, "avg_launch_us": us, "algorithmic_bytes_per_launch": algo,
            "achieved_GBps": algo / us * 1e-3}


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
    ap.add_argument("--precision", default="mixed", choices=("mixed", "fp64"))
    ap.add_argument("--launches", type=int, default=50)
    ap.add_argument("--copy-mib", type=int, default=256)
    a = ap.parse_args(argv)
    out = {"smoother": smoother(a.size, a.precision, a.launches)}
    if a.copy_mib > 0:
        out["copy"] = copy(a.copy_mib, a.launches)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
