"""Build the gfx950 shared library ``stz_amd/libstz_amd.so`` in-tree.

hipcc compiles every ``csrc/*.hip`` for gfx950 only, with packed-FP32
instructions disabled and FMA contraction left to each file (the explicit
step turns it off with a pragma to keep the reference's rounding).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libstz_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")
SOURCES = ("explicit.hip", "diag.hip", "mg.hip", "mg_tiled.hip", "mg_direct.hip", "sim.hip",
           "bd_levelset.hip", "bd_step.hip", "bd_mg.hip", "bd_sim.hip", "ensemble.hip",
           "adjoint.hip", "adjoint_run.hip", "law.hip", "noalloc.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = [os.path.join(BUILD_DIR, src.replace(".hip", ".o")) for src in SOURCES]
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)), 8))
    with ThreadPoolExecutor(jobs) as pool:
        outs = list(pool.map(lambda so: _run([HIPCC] + FLAGS + ["-c", os.path.join(CSRC, so[0]),
                                                                "-o", so[1]]),
                             zip(SOURCES, objs)))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
