"""The FP64 restatement (oracle/stz_oracle.c) against the golden fixtures
written by the reference's own compiled code (tests/golden/make_fixtures.py)."""
import numpy as np
import pytest

from conftest import field_errors, group_errors, load_fixture
from oracle.pyoracle import Oracle

STEP_CASES = ["rand32", "rand64", "rand128", "odd24x31", "odd25x20", "diff32"]


def run_oracle(fx, k):
    o = Oracle(fx["params"])
    o.set_fields(fx["in_planes"])
    o.set_vel(fx["in_vel"])
    o.time = float(fx["t0"])
    o.set_tol_mult(float(fx["tol_mult"]))
    assert o.step(float(fx["dt"]), k) == 0
    return o


@pytest.mark.parametrize("case", STEP_CASES)
def test_oracle_matches_reference_steps(case):
    fx = load_fixture(case)
    for k in [int(s) for s in fx["steps"]]:
        o = run_oracle(fx, k)
        err = field_errors(o.get_fields(), fx["planes_%d" % k])
        # converged solves: agreement to rounding, amplified a little by the
        # plastic update over several steps
        tol = 1e-12 if k <= 3 else 1e-10
        assert max(err) < tol, (case, k, err)
        ev = np.max(np.abs(o.get_vel() - fx["vel_%d" % k])) / np.max(np.abs(fx["vel_%d" % k]))
        assert ev < tol, (case, k, ev)


def test_oracle_advection_bit_identical_to_reference():
    """The explicit part has no solve: same expression order, no FMA, same
    libm -> the restatement is bit-identical to the reference's first
    advection step (rows -1..n of every plane)."""
    from oracle.pyoracle import Reference, reference_available
    if not reference_available():
        pytest.skip("compiled reference not present")
    fx = load_fixture("rand64")
    outs = []
    for cls in (Oracle, Reference):
        s = cls(fx["params"])
        s.set_fields(fx["in_planes"])
        s.time = 1.0
        s.advection_step(0.5)
        outs.append(s.get_fields())
    assert np.array_equal(outs[0], outs[1])


def test_oracle_trajectory_default_tolerance():
    """40 steps at the reference's own tolerance: the restatement takes the
    same number of V-cycles as the reference and stays within rounding
    amplified by the unconverged-at-1e-12 solves."""
    fx = load_fixture("traj64")
    o = Oracle(fx["params"])
    o.set_fields(fx["in_planes"])
    o.set_vel(fx["in_vel"])
    o.time = 0.0
    assert o.step(float(fx["dt"]), int(fx["steps"])) == 0
    assert o.avg_iters() == pytest.approx(float(fx["avg_iters"]), abs=1e-6)
    err = group_errors(o.get_fields(), fx["out_planes"])
    print(err)
    assert max(err.values()) < 1e-8, err


def test_oracle_initialize_random_matches_reference():
    fx = load_fixture("rand_init48")
    o = Oracle(fx["params"])
    o.init_fields(0, 600.0, 200.0)
    o.initialize_random(float(fx["l"]), float(fx["avg"]), float(fx["sd"]), int(fx["seed"]))
    chi = o.get_field(6)
    # rows 0..n-1 are set by initialize_random; ghost rows keep init_fields
    assert np.array_equal(chi, fx["chi"])


def test_oracle_direct_step_bit_identical_to_reference():
    """shear_sim::step_forward: no solve, same expression order and libm ->
    the restatement reproduces the reference's direct steps bit for bit"""
    fx = load_fixture("direct32x24")
    o = Oracle(fx["params"])
    o.set_fields(fx["in_planes"])
    o.time = float(fx["t0"])
    done = 0
    for k in [int(s) for s in fx["steps"]]:
        for _ in range(k - done):
            o.step_forward(float(fx["dt"]))
        done = k
        assert np.array_equal(o.get_fields(), fx["planes_%d" % k]), k
    assert o.time == pytest.approx(float(fx["t0"]) + done * float(fx["dt"]), rel=1e-15)


@pytest.fixture(scope="module")
def diag():
    return load_fixture("diag40x24")


def diag_oracle(fx, planes=None):
    o = Oracle(fx["params"])
    o.set_fields(fx["in_planes"] if planes is None else planes)
    o.set_vel(fx["in_vel"])
    return o


def test_oracle_tractions_and_qs_measure(diag):
    o = diag_oracle(diag)
    # the reference sums tractions serially in the same order: identical
    assert np.array_equal(np.array(o.tractions()), diag["tractions"])
    for st in (0, 1):
        q = np.array(o.qs_measure(bool(st)))
        # qs_measure sums rows in OpenMP order: equal to rounding
        assert np.allclose(q, diag["qs_%d" % st], rtol=1e-13, atol=0), st
        cu = o.get_cu()
        assert np.array_equal(cu[2:-2], diag["qs_cu_%d" % st][2:-2]), st


def test_oracle_compute_strain(diag):
    o = diag_oracle(diag)
    o.compute_strain(False)
    cu = o.get_cu()
    assert np.array_equal(cu[1:-1], diag["strain_cu"][1:-1])
    # the reference's interior rows are 0 (its y difference is row j+1 minus itself)
    assert not np.any(cu[2:-2])
    assert np.all(cu[1] > 0) and np.all(cu[-2] > 0)


def test_oracle_corrected_strain_closed_form():
    """a uniform simple shear X = x - g y, Y = y: F = [[1, g], [0, 1]], whose
    Hencky strain has eigenvalues +-asinh(g/2), so devmod (mat.hh:51-53) is
    sqrt(2) asinh(g/2) on every cell, seam column included"""
    from stz_amd.problems import simple_shear
    p = simple_shear(16, 12)
    o = Oracle(p)
    g = 0.3
    planes = np.zeros((10, 14, 16))
    dx, dy = 2.0 / 16, 2.0 / 12
    x = -1 + dx * np.arange(16)
    y = -1 + dy * (np.arange(14) - 1)
    planes[7] = x[None, :] - g * y[:, None]
    planes[8] = np.broadcast_to(y[:, None], (14, 16))
    o.set_fields(planes)
    o.compute_strain(True)
    cu = o.get_cu()[1:-1]
    assert np.allclose(cu, np.sqrt(2) * np.arcsinh(g / 2), rtol=1e-12)
    # the reference's form: 0 on the interior rows, wrong in the seam column
    o.compute_strain(False)
    cu = o.get_cu()[1:-1]
    assert not np.any(cu[1:-1])
    assert np.allclose(cu[[0, -1], :-1], np.sqrt(2) * np.arcsinh(g / 2), rtol=1e-12)
    assert not np.isclose(cu[0, -1], np.sqrt(2) * np.arcsinh(g / 2))


def test_oracle_l2_comparison(diag):
    a = diag_oracle(diag)
    b = diag_oracle(diag, diag["other_planes"])
    assert np.allclose(a.l2_comparison(b), diag["l2"], rtol=1e-13, atol=0)
    assert np.array_equal(a.l2_comparison(a), np.zeros(4))


def test_oracle_initialize_chi_bicubic(diag):
    o = diag_oracle(diag)
    o.initialize_chi_bicubic(diag["bicubic_arr"], 0.001, 2.0)
    assert np.array_equal(o.get_fields(), diag["bicubic_planes"])


def test_frame_writer_matches_reference_bytes(diag):
    """stz_amd.frames (host code of the product) writes the reference's
    Gnuplot binary matrix byte for byte"""
    from stz_amd import frames
    p = diag["params"]
    for mode in (0, 2, 7, 8, 9, 10):
        b = frames.gnuplot_matrix(p, diag["in_planes"], mode, p["stz"]["TZ"])
        assert b == diag["gnuplot_%d" % mode].tobytes(), mode
