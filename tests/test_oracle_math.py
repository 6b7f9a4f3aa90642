"""Pins of the FP64 restatement that do not go through the reference:
closed forms, invariants, and brute-force linear algebra on tiny grids."""
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from stz_amd.problems import simple_shear
from dense import P1, stencil_from_dense, top_dense


# ---- eno2 (shear_sim.cc:653-655) --------------------------------------------
def test_eno2_quadratic_tie_takes_centred_branch():
    # samples of a quadratic at x = 1, 0, -1, -2: equal second differences,
    # strict '>' keeps the centred difference f(1)-f(-1) = 2*f'(0)
    for a, b, c in [(1.0, 2.0, 3.0), (-0.5, 0.25, 7.0), (3.0, -4.0, 0.0)]:
        f = lambda x: a * x * x + b * x + c
        assert Oracle.eno2(f(1), f(0), f(-1), f(-2)) == 2 * b


def test_eno2_linear_data_both_branches_exact():
    # linear data: both candidate formulas equal 2*slope
    for s in (0.5, -3.0, 1e-3):
        f = lambda x: s * x + 1.0
        assert Oracle.eno2(f(1), f(0), f(-1), f(-2)) == pytest.approx(2 * s, rel=1e-14)


def test_eno2_spike_upstream_takes_one_sided_branch():
    # a spike at p0 makes the left second difference larger: the one-sided
    # three-point formula 3p1-4p2+p3 (the derivative of the parabola through
    # p1,p2,p3) is used and ignores p0 entirely
    for spike in (10.0, -7.0):
        p1, p2, p3 = 0.0, -1.0, -2.0  # linear, slope 1
        assert Oracle.eno2(spike, p1, p2, p3) == 2.0
        # swapping which side is rough flips the branch
        assert Oracle.eno2(1.0, 0.0, -1.0, spike) == 2.0
    # with a quadratic p1..p3 the one-sided result is 2*f'(0) of that parabola
    g = lambda x: 0.5 * x * x + 3 * x
    assert Oracle.eno2(100.0, g(0), g(-1), g(-2)) == pytest.approx(6.0, rel=1e-14)


# ---- STZ law (stz_model.hh:103-151) -----------------------------------------
@pytest.fixture(scope="module")
def orc():
    return Oracle(simple_shear(8))


def test_dplastic_yield_branch(orc):
    below = math.nextafter(1.0, 0.0)
    d, c1, c2 = orc.dplastic(below, 0.03)
    assert d == 0.0 and c1 == 0.0 and c2 == 0.0
    d, c1, c2 = orc.dplastic(1.0, 0.03)
    assert d != 0.0


def test_dplastic_chi_rate_identity(orc):
    # Gamma = 2 tau0 Dpl sbar / bfac  =>  dchi1 = bfac Gamma/(tau0 c0) = 2 Dpl sbar / c0
    c0 = orc.params["stz"]["c0"]
    for sbar in (1.1, 1.5, 2.5):
        for chi in (0.025, 0.03, 0.04):
            d, c1, c2 = orc.dplastic(sbar, chi)
            assert c2 == 0.0
            assert c1 == pytest.approx(2 * d * sbar / c0, rel=1e-13)


def test_dplastic_monotone_in_stress_and_chi(orc):
    ds = [orc.dplastic(s, 0.03)[0] for s in np.linspace(1.05, 2.5, 30)]
    assert np.all(np.diff(ds) > 0)
    dc = [orc.dplastic(1.5, c)[0] for c in np.linspace(0.02, 0.05, 30)]
    assert np.all(np.diff(dc) > 0)


def test_plastic_term_cases(orc):
    # below the small-deviator cut-off: nothing happens
    r, chi, ddev, trips = orc.plastic_term(1e-13, 0.03, 0.5)
    assert (r, chi, ddev, trips) == (0.0, 0.03, 0.0, 0)
    # elastic: one trip, no change
    r, chi, ddev, trips = orc.plastic_term(0.9, 0.03, 0.5)
    assert (r, chi, ddev, trips) == (0.0, 0.03, 0.0, 1)
    # strongly plastic: every full substep lowers sbar by exactly adapt_fac,
    # chi relaxes towards chi_inf, and the scale factor is ddev/sbar
    af = orc.params["adapt_fac"]
    r, chi, ddev, trips = orc.plastic_term(2.0, 0.035, 0.5)
    assert trips > 2
    assert ddev >= (trips - 1) * af * (1 - 1e-12)
    assert ddev <= trips * af * (1 + 1e-12)
    assert r == pytest.approx(ddev / 2.0, rel=1e-15)
    assert 0.035 < chi < orc.params["stz"]["chi_inf"]


# ---- multigrid operator pins --------------------------------------------------
@pytest.mark.parametrize("m,n", [(16, 16), (12, 15), (10, 8)])
def test_galerkin_coarse_operator_equals_dense_PtAP(m, n):
    dx = 2.0 / m
    p = simple_shear(m, n)
    p.update(ax=-1.0, bx=1.0, ay=-0.5 * n * dx, by=0.5 * n * dx)
    o = Oracle(p)
    viscdt = 0.1
    o.mg_setup(viscdt)
    dims = o.mg_levels()
    A = top_dense(p, viscdt)
    for lev in range(1, len(dims)):
        (um, un), (sm, sn) = dims[lev - 1], dims[lev]
        P = np.kron(np.kron(P1(un, sn, False), P1(um, sm, True)), np.eye(2))
        A = P.T @ A @ P
        S = o.mg_stencil(lev)
        ref = stencil_from_dense(A, sm, sn)
        scale = np.abs(ref).max()
        assert np.abs(S[:, :, :9, :] - ref).max() <= 1e-13 * scale, lev
        # the tenth block is the inverse of the centre block
        inv = np.linalg.inv(ref[:, :, 4, :].reshape(sn, sm, 2, 2)).reshape(sn, sm, 4)
        assert np.allclose(S[:, :, 9, :], inv, rtol=1e-12, atol=0)


def test_converged_projection_equals_dense_solve():
    p = simple_shear(8)
    o = Oracle(p)
    viscdt = 0.05
    o.mg_setup(viscdt)
    o.set_tol_mult(1e-14)
    rng = np.random.default_rng(5)
    b = rng.normal(size=(p["n"] + 1, p["m"], 2))
    o.set_src(b)
    o.set_vel(np.zeros_like(b))
    st, cyc, res = o.mg_solve()
    assert st == 0
    z = np.linalg.solve(top_dense(p, viscdt), b.ravel())
    assert np.abs(o.get_vel().ravel() - z).max() <= 1e-12 * np.abs(z).max()
