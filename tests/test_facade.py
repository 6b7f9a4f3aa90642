"""The C++ drop-in headers (stz_amd/host/*.hh): a driver written like the
reference's sb drivers (tests/cpp/facade_demo.cc) compiles against them,
fails loudly without a device, and on the GPU produces the FP64 oracle's
state for the same init_fields -> host edit -> upload -> steps sequence."""
import os
import subprocess

import numpy as np
import pytest

from conftest import group_errors
from oracle.pyoracle import Oracle
from stz_amd.problems import simple_shear

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "stz_amd")


@pytest.fixture(scope="module")
def demo(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libstz_amd.so")):
        pytest.fail("stz_amd/libstz_amd.so is not built")
    exe = str(tmp_path_factory.mktemp("facade") / "facade_demo")
    subprocess.check_call(["g++", "-O2", "-std=c++11", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(ROOT, "stz_amd", "host"),
                           os.path.join(ROOT, "tests", "cpp", "facade_demo.cc"),
                           "-o", exe, "-L", LIBDIR, "-lstz_amd", "-Wl,-rpath," + LIBDIR])
    return exe


def demo_args(p, out, steps, dt, chi_case):
    s = p["stz"]
    vals = [p["u_bdry"], p["mu"], p["K"], p["viscosity"], p["adapt_fac"], p["tmult"],
            s["TZ"], s["c0"], s["tau0"], s["kappa"], s["Delta"], s["Omegaeps0"], s["chi_inf"], s["theta"]]
    return [out, str(p["m"]), str(p["n"]), str(steps), repr(dt), str(chi_case), repr(p["ay"]), repr(p["by"])] + \
        [repr(float(v)) for v in vals]


def test_facade_fails_loudly_without_device(demo, tmp_path):
    import stz_amd
    if stz_amd.device_count() > 0:
        pytest.skip("a device is visible")
    p = simple_shear(32)
    r = subprocess.run([demo] + demo_args(p, str(tmp_path / "o.bin"), 1, 0.5, 0),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "no HIP device visible" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("m,n,chi_case", [(32, 32, 0), (40, 24, 2)])
def test_facade_driver_matches_oracle(demo, tmp_path, m, n, chi_case):
    print("case facade %dx%d chi_case %d" % (m, n, chi_case), flush=True)
    p = simple_shear(m, n, u_bdry=2e-3)
    # equal spacing in x and y: the reference's smoother operator
    # (qs_multi.cc regh/regv) matches its residual operator only then
    p["ay"], p["by"] = -float(n) / m, float(n) / m
    steps, dt = 3, 0.5
    out = str(tmp_path / "state.bin")
    r = subprocess.run([demo] + demo_args(p, out, steps, dt, chi_case),
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    npl = m * (n + 2)
    planes = raw[:10 * npl].reshape(10, n + 2, m)
    vel = raw[10 * npl:10 * npl + 2 * m * (n + 1)].reshape(n + 1, m, 2)
    t = raw[-1]

    o = Oracle(p)
    o.set_tol_mult(1e-12)
    o.init_fields(chi_case, 600.0, 200.0)
    f = o.get_fields()
    jj, ii = np.meshgrid(np.arange(n), np.arange(m), indexing="ij")
    f[6][1:n + 1] += 1e-4 * np.sqrt(1.0 + (ii * 37 + jj * 101) % 97)
    o.set_fields(f)
    o.time = 0.25
    for _ in range(steps):
        assert o.step(dt) == 0
    assert t == pytest.approx(0.25 + steps * dt, rel=1e-15)
    errs = group_errors(planes, o.get_fields())
    print("facade vs oracle", {k: "%.1e" % v for k, v in errs.items()})
    assert max(errs.values()) < 1e-10, errs
    ov = o.get_vel()
    assert np.max(np.abs(vel - ov)) <= 1e-10 * np.max(np.abs(ov))


@pytest.fixture(scope="module")
def frames_demo(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_frames")
    subprocess.check_call(["g++", "-O2", "-std=c++11", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(ROOT, "stz_amd", "host"),
                           os.path.join(ROOT, "tests", "cpp", "facade_frames.cc"),
                           "-o", exe, "-L", LIBDIR, "-lstz_amd", "-Wl,-rpath," + LIBDIR])
    return exe


@pytest.mark.gpu
def test_facade_diagnostics_and_frames(frames_demo, tmp_path):
    """write_files/output, tractions, qs_measure and compute_strain of the C++
    drop-in against the reference's own outputs (tests/golden/diag40x24)"""
    from conftest import load_fixture
    fx = load_fixture("diag40x24")
    print("case facade frames diag40x24", flush=True)
    p = fx["params"]
    m, n = p["m"], p["n"]
    state = tmp_path / "state.bin"
    np.concatenate([fx["in_planes"].ravel(), fx["in_vel"].ravel()]).tofile(state)
    cufile = tmp_path / "cu.bin"
    s = p["stz"]
    args = [str(state), str(tmp_path), str(cufile), str(m), str(n)] + [repr(float(v)) for v in (
        p["ax"], p["bx"], p["ay"], p["by"], p["u_bdry"], p["mu"], p["K"], p["viscosity"], p["adapt_fac"],
        p["tmult"], s["TZ"], s["c0"], s["tau0"], s["kappa"], s["Delta"], s["Omegaeps0"], s["chi_inf"],
        s["theta"])]
    r = subprocess.run([frames_demo] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    vals = {l.split()[0]: np.array([float(x) for x in l.split()[1:]]) for l in r.stdout.splitlines()}
    assert np.allclose(vals["tractions"], fx["tractions"], rtol=1e-12, atol=0)
    assert np.allclose(vals["qs"], np.concatenate([fx["qs_0"], fx["qs_1"]]), rtol=1e-12, atol=0)
    cu = np.fromfile(cufile, dtype=np.float64).reshape(n + 2, m)
    assert np.allclose(cu[1:-1], fx["strain_cu"][1:-1], rtol=1e-12, atol=1e-300)
    for mode, prefix in ((0, "u"), (2, "p"), (7, "tem"), (8, "dev"), (9, "X"), (10, "Y")):
        with open(tmp_path / ("%s.3" % prefix), "rb") as f:
            assert f.read() == fx["gnuplot_%d" % mode].tobytes(), prefix
