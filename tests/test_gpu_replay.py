"""GPU: the frozen-decision forward (stz_sim_replay_decisions) and the adjoint
checked against its central differences on a reduced calibration run.

* Replaying a run's own decision record from the same state repeats the run
  bit for bit, and the record written during the replay is the replayed one.
* Reduced calibration run: simple shear at 64^2, the calibration's start
  (620 K uniform chi) carried to t = 500 in plastic flow, then 40 steps of
  dt = 2.5 differentiated by the adjoint with a checkpoint every 5 steps.
  f = <lam, fields after the 40 steps> for a random lam.  Along random chi
  directions and random (c0, Omega*eps0, chi_inf) directions, the central
  difference of f through the frozen-decision forward (the base run's 40
  records replayed at +h and -h) agrees with the adjoint within 1e-5
  relative, and the differences at h and 4h agree.  The same differences
  without the replay cross many decisions in this state (v = 0 up to
  rounding, ENO comparisons between rounding-level second differences): the
  number crossed is printed, and no bar is put on those differences.
* Every STZ law and truncation model, and chi diffusion, at 32^2: the
  adjoint against frozen-decision central differences within 1e-5.
"""
import numpy as np
import pytest

from stz_amd import ShearSim
from stz_amd.adjoint import PARAMS, Adjoint, set_params
from stz_amd.initial import simple_shear_state
from stz_amd.problems import simple_shear

pytestmark = pytest.mark.gpu

N = 64
DT = 2.5
T_START = 500.0
STEPS = 40
EVERY = 5
TOL = 1e-12         # forward solve threshold, x the reference's
ADJ_REL = 1e-12     # relative residual of each adjoint solve
SWITCHES = np.uint32(0x00FFFFFF)
# difference step: with every decision frozen the step map is smooth, so the
# step can be large enough that solver noise (a threshold-level difference
# between the +h and -h solves, divided by 2h) stays far below the bar
H = 1e-5


@pytest.fixture(scope="module")
def plastic_start():
    p = simple_shear(N)
    g = ShearSim(p, "fp64")
    g.set_mg_config(tol_mult=TOL)
    g.set_fields(simple_shear_state(p, 620.0, 0.0, 3.0, seed=0))
    g.set_vel(np.zeros((N + 1, N, 2)))
    g.time = 0.0
    g.step_forward_quasistatic(DT, int(round(T_START / DT)))
    out = p, g.get_fields(), g.get_vel(), g.time
    g.close()
    return out


def run(start, planes, theta, steps, records=None, record=True):
    """f's fields after `steps` steps from (planes, theta); the decision
    records of every step; records, when given, are replayed"""
    p, _, vel, t = start
    g = ShearSim(p, "fp64")
    g.set_mg_config(tol_mult=TOL)
    set_params(g, *theta)
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = t
    g.record_decisions(record)
    if records is not None:
        g.replay_decisions(records)
    bits = []
    for _ in range(steps):
        g.step_forward_quasistatic(DT, 1)
        bits.append(g.get_decisions().copy())
    F = g.get_fields()
    g.close()
    return F, bits


def theta0(p):
    return np.array([p["stz"][k] for k in PARAMS])


def test_replay_of_own_record_is_bit_identical(plastic_start):
    print("case replay identity", flush=True)
    p, planes, _, _ = plastic_start
    th = theta0(p)
    F, bits = run(plastic_start, planes, th, 6)
    G, rbits = run(plastic_start, planes, th, 6, records=bits)
    assert np.array_equal(F, G)
    for a, b in zip(bits, rbits):
        assert np.array_equal(a, b)
    yielding = int(np.count_nonzero(bits[-1] & np.uint32(1 << 20)))
    print("replay identity: yielding cells %d, max substeps %d" % (yielding, int((bits[-1] >> 24).max())))
    assert yielding > 0


def test_record_is_required_for_every_step(plastic_start):
    print("case replay exhausted", flush=True)
    p, planes, vel, t = plastic_start
    _, bits = run(plastic_start, planes, theta0(p), 2)
    g = ShearSim(p, "fp64")
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = t
    g.replay_decisions(bits)
    g.step_forward_quasistatic(DT, 2)
    with pytest.raises(Exception, match="recorded steps"):
        g.step_forward_quasistatic(DT, 1)
    g.replay_decisions([])
    g.step_forward_quasistatic(DT, 1)
    g.close()


def test_adjoint_matches_frozen_decision_differences(plastic_start):
    print("case replay adjoint 64^2 40 steps", flush=True)
    p, planes, vel, t = plastic_start
    TZ = p["stz"]["TZ"]
    th = theta0(p)
    F0, base = run(plastic_start, planes, th, STEPS)
    sat = int(sum(np.count_nonzero((b >> 24) == 255) for b in base))
    print("base run: yielding cells (last step) %d, max substeps %d, saturated %d" % (
        int(np.count_nonzero(base[-1] & np.uint32(1 << 20))), int(max((b >> 24).max() for b in base)), sat),
        flush=True)
    assert sat == 0
    lam = np.random.default_rng(5).standard_normal(F0.shape)

    g = ShearSim(p, "fp64")
    g.set_mg_config(tol_mult=TOL)
    set_params(g, *th)
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = t
    a = Adjoint(g, EVERY)
    a.set_solve_accuracy(ADJ_REL)
    a.forward(DT, STEPS)
    assert np.array_equal(g.get_fields(), F0)  # the adjoint's forward is the recorded run
    lam0, gth = a.backward(lam)
    a.close()
    g.close()

    def f(planes_, theta_, frozen):
        F, bits = run(plastic_start, planes_, theta_, STEPS, records=base if frozen else None)
        return float(np.sum(lam * F)), bits

    worst = 0.0
    for k in range(4):
        r = np.random.default_rng(300 + k)
        dx = np.zeros_like(planes)
        dth = np.zeros(3)
        if k < 2:
            dx[6] = r.standard_normal(planes[6].shape) * (10.0 / TZ)
        else:
            dth = r.standard_normal(3) * th * 0.1
        ad = float(np.sum(lam0 * dx)) + float(sum(gth[q] * dth[i] for i, q in enumerate(PARAMS)))
        diffs = []
        for h in (H, 4 * H):
            fp, _ = f(planes + h * dx, th + h * dth, True)
            fm, _ = f(planes - h * dx, th - h * dth, True)
            diffs.append((fp - fm) / (2 * h))
        fp, bp = f(planes + H * dx, th + H * dth, False)
        fm, bm = f(planes - H * dx, th - H * dth, False)
        crossed = int(sum(np.count_nonzero((x ^ y) & SWITCHES) for x, y in zip(bp, bm)))
        free = (fp - fm) / (2 * H)
        rel = abs(diffs[0] - ad) / max(abs(diffs[0]), abs(ad))
        worst = max(worst, rel)
        print("direction %d (%s): frozen fd(h) %.12e fd(4h) %.12e adjoint %.12e rel %.2e; unfrozen fd %.6e "
              "crossing %d decisions" % (k, "chi" if k < 2 else "params", diffs[0], diffs[1], ad, rel, free,
                                         crossed), flush=True)
        assert abs(diffs[0] - diffs[1]) <= 1e-6 * abs(diffs[0])
        assert rel < 1e-5, (k, diffs, ad)
    print("worst relative disagreement %.2e" % worst)


# every STZ law and truncation model, and chi diffusion (constant and
# ddev-weighted), on a 32^2 random state: 3 steps of dt 0.3, lam random
LAW_CASES = [(0, -1, 0.0, 0.0, 0.0), (0, 1, 0.0, 0.0, 0.0), (0, 3, 0.0, 0.0, 0.0), (0, 4, 0.0, 0.0, 0.0),
             (1, -1, 0.1, 0.0, 0.0), (2, -1, 0.1, 0.0, 0.0), (3, -1, 0.0, 0.0, 0.0),
             (0, -1, 0.0, 2e-4, 0.02)]


@pytest.mark.parametrize("kind,t_model,rho,chi_diff,chi_len", LAW_CASES)
def test_adjoint_every_law_and_chi_diffusion(kind, t_model, rho, chi_diff, chi_len):
    """The adjoint's plastic Jacobian runs the law code of the step in dual
    numbers, and the chi-diffusion adjoint scatters the transposed stencil;
    both checked against central differences of the frozen-decision forward
    (h = 1e-6 and 4e-6) along a chi direction and a parameter direction,
    within 1e-5."""
    import copy
    from stategen import random_state
    print("case adjoint law", kind, t_model, rho, chi_diff, chi_len, flush=True)
    p = simple_shear(32, u_bdry=2e-3, chi_diff=chi_diff, chi_len=chi_len)
    st = copy.deepcopy(p["stz"])
    st["kind"], st["t_model"], st["rho"] = kind, t_model, rho
    p["stz"] = st
    planes, vel = random_state(p, 21, stress=1.4, ties=False)
    dt, steps, t0 = 0.3, 3, 1.6
    th = np.array([st[k] for k in PARAMS])

    def fwd(pl, theta, records=None):
        g = ShearSim(p, "fp64")
        g.set_mg_config(tol_mult=1e-12)
        set_params(g, *theta)
        g.set_fields(pl)
        g.set_vel(vel)
        g.time = t0
        g.record_decisions(True)
        if records is not None:
            g.replay_decisions(records)
        bits = []
        for _ in range(steps):
            g.step_forward_quasistatic(dt, 1)
            bits.append(g.get_decisions().copy())
        F = g.get_fields()
        g.close()
        return F, bits

    F0, base = fwd(planes, th)
    yielding = int(np.count_nonzero(base[0] & np.uint32(1 << 20)))
    lam = np.random.default_rng(3).standard_normal(F0.shape)
    g = ShearSim(p, "fp64")
    g.set_mg_config(tol_mult=1e-12)
    set_params(g, *th)
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = t0
    a = Adjoint(g, 2)
    a.set_solve_accuracy(1e-12)
    a.forward(dt, steps)
    assert np.array_equal(g.get_fields(), F0)
    lam0, gth = a.backward(lam)
    a.close()
    g.close()
    TZ = st["TZ"]
    r = np.random.default_rng(17)
    for label, dx, dth in (("chi", np.where(np.arange(10)[:, None, None] == 6,
                                             r.standard_normal(planes.shape) * (10.0 / TZ), 0.0), np.zeros(3)),
                           ("params", np.zeros_like(planes), r.standard_normal(3) * th * 0.1)):
        ad = float(np.sum(lam0 * dx)) + float(sum(gth[q] * dth[i] for i, q in enumerate(PARAMS)))
        diffs = []
        for h in (1e-6, 4e-6):
            fp = float(np.sum(lam * fwd(planes + h * dx, th + h * dth, base)[0]))
            fm = float(np.sum(lam * fwd(planes - h * dx, th - h * dth, base)[0]))
            diffs.append((fp - fm) / (2 * h))
        rel = abs(diffs[0] - ad) / max(abs(diffs[0]), abs(ad))
        print("law %d t_model %d rho %g chi_diff %g chi_len %g (%d yielding): %s fd %.12e %.12e adjoint %.12e "
              "rel %.2e" % (kind, t_model, rho, chi_diff, chi_len, yielding, label, diffs[0], diffs[1], ad, rel),
              flush=True)
        assert abs(diffs[0] - diffs[1]) <= 1e-6 * abs(diffs[0])
        assert rel < 1e-5, (label, diffs, ad)
