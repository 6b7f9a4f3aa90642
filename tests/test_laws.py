"""The four STZ plasticity laws and the truncation models, and init_fields,
pinned against the reference's own compiled code on the CPU.

* stz_model_dplastic (the library's host evaluation of the law code its
  kernels run) and the oracle's law against stz_dynamics::Dplastic of the
  reference's classes (stz_model.hh:28, :50-207): bit-identical over a grid
  of (sbar, chi) that crosses the yield threshold sbar = 1 (and the ulps
  around it) and, for the truncated linear athermal law (t_model 0..4,
  :120-139), the crossover Omegaeps0 sbar = Delta; with rho > 0 for the
  thermal laws, so the dchi2 term is exercised.
* one advection step of the oracle with every law against the reference's
  (shear_sim.cc:296-361 with Dplastic inside adaptive_plastic_term
  :513-548): every plane, rows -1..n, bit-identical.
* shear_sim::init_fields (shear_sim.cc:60-114), cases 0, 1 and 2: the
  oracle's planes, all rows, bit-identical to the reference's.
"""
import copy
import ctypes as C
import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle, Reference, reference_available, reference_dplastic
from stategen import random_state
from stz_amd import _lib
from stz_amd.problems import simple_shear

needs_ref = pytest.mark.skipif(not reference_available(), reason="compiled reference not present")

# (kind, t_model, rho): the linear athermal law with every truncation model,
# the other three laws with and without the rho term
LAWS = [(0, t, 0.0) for t in (-1, 0, 1, 2, 3, 4)] + [(1, -1, 0.0), (1, -1, 0.1), (2, -1, 0.0), (2, -1, 0.1),
                                                     (3, -1, 0.0)]


def law(kind, t_model, rho):
    m = copy.deepcopy(simple_shear(8)["stz"])
    m["kind"], m["t_model"], m["rho"] = kind, t_model, rho
    return m


def law_grid(m):
    cross = m["Delta"] / m["Omegaeps0"]
    sb = list(np.linspace(0.0, 3.0, 61)) + [1.0, math.nextafter(1.0, 0), math.nextafter(1.0, 2), cross,
                                            math.nextafter(cross, 0), math.nextafter(cross, 4), 1e-13, 2.5e-12]
    chi = [0.02, 600 / 21000.0, 0.035, 900 / 21000.0, 0.06]
    return [(s, c) for s in sb for c in chi]


def same_bits(a, b):
    """equal bit patterns, or both NaN, element by element"""
    ua = np.array(a, dtype=np.float64).view(np.uint64)
    ub = np.array(b, dtype=np.float64).view(np.uint64)
    return all(x == y or (math.isnan(p) and math.isnan(q)) for x, y, p, q in zip(ua, ub, a, b))


def lib_dplastic(m, sbar, chi):
    lib = _lib.load()
    p = simple_shear(8)
    p["stz"] = m
    mp = _lib.make_params(p).stz
    d, d1, d2 = C.c_double(), C.c_double(), C.c_double()
    _lib.check(lib.stz_model_dplastic(C.byref(mp), float(sbar), float(chi), C.byref(d), C.byref(d1), C.byref(d2)))
    return d.value, d1.value, d2.value


@needs_ref
@pytest.mark.parametrize("kind,t_model,rho", LAWS)
def test_dplastic_bit_identical_to_reference(kind, t_model, rho):
    m = law(kind, t_model, rho)
    o = Oracle(simple_shear(8))
    o._p.stz.kind, o._p.stz.t_model, o._p.stz.rho = kind, t_model, rho
    nonzero = 0
    for sbar, chi in law_grid(m):
        want = reference_dplastic(m, sbar, chi)
        got = lib_dplastic(m, sbar, chi)
        assert same_bits(got, want), (kind, t_model, rho, sbar, chi, got, want)
        assert same_bits(o.dplastic(sbar, chi), want), (kind, t_model, sbar, chi)
        nonzero += want[0] != 0 and not math.isnan(want[0])
    assert nonzero > 0
    o.close()


def test_dplastic_refuses_unknown_law():
    m = law(7, -1, 0.0)
    with pytest.raises(_lib.StzError):
        lib_dplastic(m, 1.5, 0.03)


@needs_ref
@pytest.mark.parametrize("kind,t_model,rho", LAWS)
def test_oracle_advection_with_every_law_matches_reference(kind, t_model, rho):
    p = simple_shear(40, 40, u_bdry=2e-3)
    p["stz"] = law(kind, t_model, rho)
    planes, vel = random_state(p, 11, stress=1.6)
    dt = 0.3
    out = []
    for cls in (Oracle, Reference):
        s = cls(p)
        s.set_fields(planes)
        s.set_vel(vel)
        s.time = 1.3
        s.advection_step(dt)
        out.append(s.get_fields())
        s.close()
    fo, fr = out
    for k in range(10):
        assert np.array_equal(fo[k], fr[k]), (kind, t_model, k)
    # the truncated branch and plastic flow were reached
    sbar = np.sqrt(3 * planes[3] ** 2 + planes[4] ** 2 + planes[5] ** 2)
    assert np.count_nonzero(sbar * p["stz"]["Omegaeps0"] > p["stz"]["Delta"]) > 0
    assert np.count_nonzero(fo[9]) > 0


@needs_ref
@pytest.mark.parametrize("chi_case", [0, 1, 2])
def test_oracle_init_fields_matches_reference(chi_case):
    p = simple_shear(48, 40)
    p["ax"], p["bx"] = -2.0, 2.0
    out = []
    for cls in (Oracle, Reference):
        s = cls(p)
        s.init_fields(chi_case, 600.0, 200.0)
        out.append(s.get_fields())
        s.close()
    fo, fr = out
    for k in range(10):
        assert np.array_equal(fo[k], fr[k]), (chi_case, k)
    assert np.ptp(fr[6]) > 0
