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


@pytest.fixture(scope="module")
def tgmg_demo(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("facade") / "facade_tgmg")
    subprocess.check_call(["g++", "-O2", "-std=c++11", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(ROOT, "stz_amd", "host"),
                           os.path.join(ROOT, "tests", "cpp", "facade_tgmg.cc"),
                           "-o", exe, "-L", LIBDIR, "-lstz_amd", "-Wl,-rpath," + LIBDIR])
    return exe


@pytest.mark.gpu
def test_facade_tgmg_members(tgmg_demo, tmp_path):
    """tgmg<qs_multi,vec,mat> of the drop-in (tgmg.hh:286-495): the problem
    class's checks give the reference's verdicts (qs_multi is not symmetric:
    the wall rows hold only their diagonal while their neighbours couple to
    them; mul_a is consistent with the entries when dx = dy;
    tests/test_drivers.py compares the verdicts with the reference's own),
    fmg_pre followed by one V-cycle leaves a smaller residual than one
    V-cycle from zero, and the adaptive and fixed-batch solves both leave a
    solution whose residual, recomputed in numpy with the operator written
    from its definition, is below the problem's threshold q.acc"""
    from dense import top_dense
    m, viscdt = 32, 0.1
    print("case facade tgmg 32", flush=True)
    out = str(tmp_path / "vel.bin")
    r = subprocess.run([tgmg_demo, out, str(m), str(m), repr(viscdt)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "The matrix is not symmetric" in r.stdout
    assert "The mul_a routine is consistent with the matrix terms" in r.stdout
    line = [l for l in r.stdout.splitlines() if l.startswith("residual")][0].split()
    e0, e2, e3, acc, ev = float(line[2]), float(line[6]), float(line[9]), float(line[11]), float(line[13])
    assert e2 < ev < e0 and e3 < acc
    p = simple_shear(m, u_bdry=2e-3)
    A = top_dense(p, viscdt)
    n1 = m + 1
    i, j = np.meshgrid(np.arange(m), np.arange(n1))
    k = i + m * j
    b = np.zeros((n1, m, 2))
    b[..., 0] = np.sin(2 * np.pi * i / m) * np.cos(0.1 * j) + 0.01 * ((k * 37) % 11)
    b[..., 1] = np.cos(2 * np.pi * 2 * i / m) * np.sin(0.05 * j) - 0.01 * ((k * 13) % 7)
    z = np.fromfile(out, dtype=np.float64).reshape(2, n1 * m * 2)
    for zz in z:
        res = np.mean(np.sum((b.ravel() - A @ zz).reshape(-1, 2) ** 2, axis=1))
        assert res <= acc * (1 + 1e-6), (res, acc)
    # dx != dy: the smoother blocks regh, regv of qs_multi::init no longer
    # match mul_a, and the check says so, as the reference's does
    r = subprocess.run([tgmg_demo, out, "32", "24", repr(viscdt)], capture_output=True, text=True, timeout=120)
    assert "The mul_a routine is not consistent with the matrix terms" in r.stdout, r.stdout[-2000:]
