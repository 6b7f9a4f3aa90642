"""The reference's own sb drivers, unchanged, against the C++ drop-in headers
(stz_amd/host/).  CPU only: the drivers are compiled where they lie in the
reference tree, so these tests need that tree and skip without it.

* shear_small.cc, shear_test.cc, shear_compare.cc, shear_runs.cc and
  shear_random.cc (with the reference's common.cc for its file helper)
  compile and link against the headers and libstz_amd.so.  g++'s -I-
  makes every quoted include resolve to stz_amd/host/ first (shear_sim.hh,
  stz_model.hh, qs_multi.hh, tgmg.hh, vec.hh, mat.hh, fields.hh); only what
  the headers do not provide comes from the reference directory.
* dplas_test.cc (the reference's table of stz_dynamics::Dplastic, through
  the reference's own input-file reader fileinfo.cc) built against the
  headers prints byte-identical tables to the same driver built against
  the reference's stz_model.hh, for every law, every truncation model and
  both truncation modifications.
* Without a device the drop-in fails loudly (exit status 1 and the
  library's message), as the reference exits on a failed solve.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "stz_amd", "host")
LIBDIR = os.path.join(ROOT, "stz_amd")
SB = "/root/reference/esim/sb/main"

pytestmark = pytest.mark.skipif(not os.path.isdir(SB), reason="reference tree not present")

DRIVERS = {"shear_small": [], "shear_test": [], "shear_compare": [], "shear_runs": [],
           "shear_random": ["common.cc"]}

FIN_BASE = """chi0 600
T 400
TZ 21000
sim_size 1
gridpoints 65
output p
adapt_fac 0.002
displacement 0.01
viscosity 0.02
notch_radius 0.001
youngs_modulus 101e9
poisson_ratio 0.35
lambda 0.1
gamma 0.1
typical_ki 1e6
rho0 6125
yield_stress 0.85e9
c0 0.4
tau0 1e-13
kappa 4.2
Delta 8000
Omega 300
eps0 0.3
chi_inf 900
tmult 0.5
max_ki 1e7
frho 3.0 2.0 300 600 200
"""

MODELS = ["model athermal linear", "model athermal nonlinear", "model thermal linear",
          "model thermal nonlinear"] + \
    ["model athermal linear\ntruncate_model %d" % t for t in range(5)] + \
    ["model athermal linear\ntruncate_model 3\ntruncate_modify %s 0.5" % k for k in ("scale", "mult")]


def build(out, sources, facade=True):
    inc = ["-I", HOST, "-I-", "-I", SB] if facade else ["-I", SB]
    cmd = ["g++", "-O2", "-std=c++11", "-w"] + inc + [os.path.join(SB, s) for s in sources] + ["-o", out]
    if facade:
        cmd += ["-L", LIBDIR, "-lstz_amd", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


@pytest.mark.parametrize("driver", sorted(DRIVERS))
def test_reference_driver_compiles_against_the_headers(driver, tmp_path):
    exe = build(str(tmp_path / driver), [driver + ".cc"] + DRIVERS[driver])
    assert os.path.exists(exe)


def test_shear_small_fails_loudly_without_a_device(tmp_path):
    import stz_amd
    if stz_amd.device_count() > 0:
        pytest.skip("a device is visible")
    exe = build(str(tmp_path / "shear_small"), ["shear_small.cc"])
    r = subprocess.run([exe, "qs"], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 1
    assert "no HIP device visible" in r.stderr


@pytest.fixture(scope="module")
def dplas(tmp_path_factory):
    d = tmp_path_factory.mktemp("dplas")
    return (build(str(d / "dplas_facade"), ["dplas_test.cc", "fileinfo.cc"]),
            build(str(d / "dplas_reference"), ["dplas_test.cc", "fileinfo.cc"], facade=False), d)


@pytest.mark.parametrize("model", MODELS)
def test_dplas_test_tables_identical(dplas, model):
    facade, reference, d = dplas
    name = str(d / ("case%d" % MODELS.index(model)))
    with open(name + ".fin", "w") as f:
        f.write(FIN_BASE + model + "\n")
    out = []
    for exe in (facade, reference):
        r = subprocess.run([exe, name], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        out.append(r.stdout)
    assert out[0].count("\n") == 1201
    assert out[0] == out[1]


CHECKS = r"""
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <omp.h>
#include <sstream>
#define private public
#include "shear_sim.hh"
#undef private
#include "tgmg.cc"
#include "tgmg_debug.cc"
template void tgmg<qs_multi, vec, mat>::check_consistent();
int main(int argc, char **argv) {
    const int m = std::atoi(argv[1]), n = std::atoi(argv[2]);
    const double TZ = 21000;
    stz_dynamics_linear_athermal stz(TZ, 0.4, 1e-13, 4.2, 8000 / TZ, 0.2638, 900 / TZ, 400 / TZ, 0);
    shear_sim ss(m, n, -1, 1, -1, 1, 20 / 0.85, 143.2 / 0.85, 0.02, 0, 0, 1, 0.5, 0.002, 2e-3, &stz, ".");
    ss.qsm.mg.verbose = 0;
    ss.qsm.init(0.1);
    ss.qsm.mg.check_consistent();
    return 0;
}
"""


def test_reference_tgmg_check_verdicts(tmp_path):
    """The reference's own tgmg::check_consistent (tgmg_debug.cc:69-103) on
    qs_multi, for the verdicts the drop-in's must give (its GPU test,
    test_facade.py::test_facade_tgmg_members): consistent when dx = dy, not
    consistent when dx != dy.  (The reference's check_symmetric does not
    compile for qs_multi: sym_check calls a free mod_sq on the 2x2 block
    type, which shared/mat.hh provides only as a member; the drop-in adds
    the free function.)"""
    ref = os.path.join(ROOT, "oracle", "_ref")
    if not os.path.exists(os.path.join(ref, "libstzref.so")):
        pytest.skip("compiled reference not present")
    src = tmp_path / "checks.cc"
    src.write_text(CHECKS)
    exe = str(tmp_path / "checks")
    esim = os.path.dirname(os.path.dirname(SB))
    cmd = ["g++", "-O2", "-fopenmp", "-w", "-I", SB, "-I", os.path.join(esim, "tgmg"), "-I",
           os.path.join(esim, "shared"), "-I", os.path.join(ref, "gen"), str(src), "-o", exe, "-L", ref,
           "-lstzref", "-Wl,-rpath," + ref]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = {mn: subprocess.run([exe] + mn.split(), capture_output=True, text=True, timeout=120).stdout
           for mn in ("32 32", "32 24")}
    assert "The mul_a routine is consistent with the matrix terms" in out["32 32"]
    assert "The mul_a routine is not consistent with the matrix terms" in out["32 24"]
