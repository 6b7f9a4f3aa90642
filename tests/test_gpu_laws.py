"""GPU parity of the STZ-law variants and of init_fields against the
reference's own compiled code.

Every law (linear athermal with truncation models -1..4, linear and
nonlinear with and without the rho term, nonlinear athermal;
stz_model.hh:50-207) through the device step: one advection step
(shear_sim.cc:296-361) with the same decision bits as the oracle (which is
bit-identical to the reference for every law, tests/test_laws.py) and fields
within 1e-13 of the reference's (u, v, p, X, Y, which no transcendental
reaches, bit-identical); two whole quasi-static steps at a solve threshold of
1e-12 x the reference's within 1e-10.  init_fields (shear_sim.cc:60-114),
cases 0..2: every plane and row equal to the reference's, chi to within an
ulp (the device's exp)."""
import copy

import numpy as np
import pytest

from conftest import field_errors
from oracle.pyoracle import Oracle, Reference, reference_available
from stategen import random_state
from stz_amd import ShearSim
from stz_amd.problems import simple_shear

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not reference_available(), reason="compiled reference not present")]

LAWS = [(0, t, 0.0) for t in (-1, 0, 1, 2, 3, 4)] + [(1, -1, 0.0), (1, -1, 0.1), (2, -1, 0.0), (2, -1, 0.1),
                                                     (3, -1, 0.0)]
SUBSTEP_MASK = np.uint32(0xFF << 24)


def law_params(kind, t_model, rho, m=40, n=40):
    p = simple_shear(m, n, u_bdry=2e-3)
    st = copy.deepcopy(p["stz"])
    st["kind"], st["t_model"], st["rho"] = kind, t_model, rho
    p["stz"] = st
    return p


@pytest.mark.parametrize("kind,t_model,rho", LAWS)
def test_advection_with_every_law(kind, t_model, rho):
    print("case law advection", kind, t_model, rho, flush=True)
    p = law_params(kind, t_model, rho)
    planes, vel = random_state(p, 11, stress=1.6)
    dt, t1 = 0.3, 1.6
    r = Reference(p)
    r.set_fields(planes)
    r.set_vel(vel)
    r.time = t1
    r.advection_step(dt)
    fr = r.get_fields()
    r.close()
    o = Oracle(p)
    o.set_fields(planes)
    o.record_decisions(True)
    o.time = t1
    o.advection_step(dt)
    do = o.get_decisions()
    o.close()
    g = ShearSim(p)
    g.set_fields(planes)
    g.set_vel(vel)
    g.record_decisions(True)
    g.time = t1
    g.advection_step(dt)
    dg, fg = g.get_decisions(), g.get_fields()
    g.close()
    assert np.array_equal(do & ~SUBSTEP_MASK, dg & ~SUBSTEP_MASK)
    assert np.count_nonzero((do >> 24) != (dg >> 24)) <= 1
    for k in (0, 1, 2, 7, 8):
        assert np.array_equal(fg[k], fr[k]), k
    err = field_errors(fg, fr)
    print(kind, t_model, rho, "advection errors", ["%.1e" % e for e in err])
    assert max(err) < 1e-13, err


@pytest.mark.parametrize("kind,t_model,rho", LAWS)
def test_steps_with_every_law(kind, t_model, rho):
    print("case law steps", kind, t_model, rho, flush=True)
    p = law_params(kind, t_model, rho)
    planes, vel = random_state(p, 12, stress=1.6)
    dt, t0 = 0.3, 1.3
    r = Reference(p)
    r.set_fields(planes)
    r.set_vel(vel)
    r.time = t0
    r.set_tol_mult(1e-12)
    r.step(dt, 2)
    fr = r.get_fields()
    r.close()
    g = ShearSim(p)
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = t0
    g.set_mg_config(tol_mult=1e-12)
    g.step_forward_quasistatic(dt, 2)
    err = field_errors(g.get_fields(), fr)
    g.close()
    print(kind, t_model, rho, "step errors", ["%.1e" % e for e in err])
    assert max(err) < 1e-10, err


@pytest.mark.parametrize("chi_case", [0, 1, 2])
def test_init_fields_against_reference(chi_case):
    print("case init_fields", chi_case, flush=True)
    p = simple_shear(48, 40)
    p["ax"], p["bx"] = -2.0, 2.0
    r = Reference(p)
    r.init_fields(chi_case, 600.0, 200.0)
    fr = r.get_fields()
    r.close()
    g = ShearSim(p)
    g.init_fields(chi_case, 600.0, 200.0)
    fg = g.get_fields()
    g.close()
    for k in range(10):
        if k == 6:
            assert np.max(np.abs(fg[k] - fr[k])) <= 2.3e-16 * np.max(np.abs(fr[k])), k
        else:
            assert np.array_equal(fg[k], fr[k]), k
