"""GPU: the batched ensemble (stz_ensemble_*) against single simulations.

An ensemble member runs the single simulation's kernel bodies on its own
arrays, and its adaptive V-cycle count is the same state machine on the same
residuals, so after any number of steps every member must be BIT-identical to
a ShearSim given the same inputs: fields, velocity, time, solve statistics.
The single path is pinned to the reference by test_gpu_parity.py; one
member here is also checked straight against a reference fixture."""
import numpy as np
import pytest

from conftest import field_errors, load_fixture
from stategen import random_state
from stz_amd import Ensemble, ShearSim, StzError
from stz_amd.problems import simple_shear

pytestmark = pytest.mark.gpu

# sides of the members: every multigrid level above the workgroup solver
# must be tiled (a multiple of 32); in FP64 the 48-wide level of a 96 side is
# below that limit only in the FP32 hierarchy
SIZES = {"mixed": [64, 96, 128, 64, 256], "fp64": [64, 128, 32, 64, 256]}


def setup(form, sizes=None, seed=11, tol_mult=1e-2, nu=2, u_bdry=2e-3):
    params = [simple_shear(m, u_bdry=u_bdry) for m in (sizes or SIZES[form])]
    ens = Ensemble(params, form)
    singles = [ShearSim(p, form) for p in params]
    for k, p in enumerate(params):
        planes, vel = random_state(p, seed + k)
        for g in (ens[k], singles[k]):
            g.set_fields(planes)
            g.set_vel(vel)
            g.time = 1.2 + 0.1 * k
    ens.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol_mult)
    for g in singles:
        g.set_mg_config(nu1=nu, nu2=nu, tol_mult=tol_mult)
    return ens, singles


def assert_same(ens, singles, label):
    for k, g in enumerate(singles):
        e = ens[k]
        a, b = e.get_fields(), g.get_fields()
        if not np.array_equal(a, b):
            raise AssertionError("%s member %d (%dx%d) fields differ: %s" %
                                 (label, k, g.m, g.n, ["%.1e" % x for x in field_errors(a, b)]))
        assert np.array_equal(e.get_vel(), g.get_vel()), (label, k, "vel")
        assert e.time == g.time, (label, k, e.time, g.time)
        assert e.last_solve() == g.last_solve(), (label, k, e.last_solve(), g.last_solve())


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_members_equal_single_simulations(form, graphs):
    print("case ensemble", form, "graphs", graphs, flush=True)
    ens, singles = setup(form)
    ens.use_graphs(graphs)
    dt = 0.5
    ens.step_forward_quasistatic(dt, 6)
    for g in singles:
        g.step_forward_quasistatic(dt, 6)
    assert_same(ens, singles, "6 steps")
    assert [e.avg_iters() for e in ens] == [g.avg_iters() for g in singles]
    st = ens.stats()
    print(form, graphs, st, [g.last_solve()["cycles"] for g in singles])
    assert st["steps"] == 6
    # graphs: one per step parity and planned slot count, and the slot and
    # tail graphs of members that need more cycles than planned
    assert (st["captures"] > 0) if graphs else st["captures"] == 0, st
    ens.close()


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_unit_predictor_members_equal_single_simulations(form):
    """with predict_unit = 1 (one more cycle per failed test) and a fast
    decay, from a cold predictor: the device state machine and the single
    simulation's loop (which enqueues the cycle after each test before it
    waits) take the same cycles, and more than one test per solve"""
    print("case ensemble unit predictor", form, flush=True)
    ens, singles = setup(form, tol_mult=1e-6)
    for g in [ens] + singles:
        g.set_mg_config(predict_unit=1, predict_init=1, predict_decay=2)
    dt = 0.5
    for g in [ens] + singles:
        g.step_forward_quasistatic(dt, 4)
    assert_same(ens, singles, "4 steps, unit predictor")
    assert max(g.last_solve()["cycles"] for g in singles) > 1
    ens.close()


def test_sweep_tuner_restores_the_state():
    """the sweep-count tuner times its candidates from the members' state,
    keeps one of them, and leaves every member where it found it (predictor
    included); the ensemble then still equals single simulations with the
    chosen sweeps"""
    print("case ensemble sweep tuner", flush=True)
    ens, singles = setup("mixed", sizes=[64, 128, 96])
    for g in [ens] + singles:
        g.step_forward_quasistatic(0.5, 2)
    before = [(m.get_fields(), m.get_vel(), m.time, m.predictor_limit) for m in ens]
    res = ens.tune_sweeps(0.5, probe_steps=2, candidates=((1, 1), (2, 2), (3, 3)))
    print("sweep tuner", res)
    assert res["choice"] in ((1, 1), (2, 2), (3, 3))
    assert set(res["ms_per_step"]) == {(1, 1), (2, 2), (3, 3)}
    for m, (f, v, t, lim) in zip(ens, before):
        assert np.array_equal(m.get_fields(), f) and np.array_equal(m.get_vel(), v) and m.time == t
        assert m.predictor_limit == lim
    nu1, nu2 = res["choice"]
    for g in singles:
        g.set_mg_config(nu1=nu1, nu2=nu2)
    for g in [ens] + singles:
        g.step_forward_quasistatic(0.5, 3)
    assert_same(ens, singles, "after the sweep tuner")
    ens.close()


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_large_members_equal_single_simulations(form):
    """members of the workload's largest sides, 384 and 512, beside a 64 one:
    their V-cycles stop at the direct levels 48x49 and 64x65 (the 48-wide
    level of a 384 side is not smoothed, so the FP64 form takes it too)"""
    print("case ensemble large", form, flush=True)
    ens, singles = setup(form, sizes=[384, 512, 64])
    dt = 0.5
    ens.step_forward_quasistatic(dt, 3)
    for g in singles:
        g.step_forward_quasistatic(dt, 3)
    assert_same(ens, singles, "3 steps, sides 384 and 512")
    ens.close()


def test_extra_slots_and_table_rebuilds():
    """A member whose predictor underestimates (velocity reset to zero after
    its count has decayed) needs slots beyond the planned ones; a new dt, a
    member stepped on its own and changed settings rebuild the tables."""
    print("case ensemble overflow/rebuild", flush=True)
    ens, singles = setup("mixed", sizes=[64, 128, 96])
    dt = 0.5
    for g in [ens] + singles:
        g.step_forward_quasistatic(dt, 4)
    zero = np.zeros_like(singles[1].get_vel())
    ens[1].set_vel(zero)
    singles[1].set_vel(zero)
    for g in [ens] + singles:
        g.step_forward_quasistatic(dt, 2)
    assert_same(ens, singles, "after velocity reset")
    assert ens.stats()["extra_slots"] > 0, ens.stats()
    # a member stepped alone (its plane sets swap), then a different dt
    ens[2].step_forward_quasistatic(dt, 1)
    singles[2].step_forward_quasistatic(dt, 1)
    for g in [ens] + singles:
        g.step_forward_quasistatic(0.25, 3)
    assert_same(ens, singles, "after member step and new dt")
    # new settings for everyone
    ens.set_mg_config(nu1=1, nu2=1)
    for g in singles:
        g.set_mg_config(nu1=1, nu2=1)
    for g in [ens] + singles:
        g.step_forward_quasistatic(0.25, 2)
    assert_same(ens, singles, "after V(1,1)")
    ens.close()


def test_failed_solve_leaves_member_as_single_does():
    print("case ensemble failure", flush=True)
    ens, singles = setup("mixed", sizes=[64, 64])
    for g in singles:
        g.set_mg_config(tol_mult=1e-40, predict_max_iters=1)
    ens.set_mg_config(tol_mult=1e-40, predict_max_iters=1)
    codes = []
    for g in [ens] + singles:
        with pytest.raises(StzError) as e:
            g.step_forward_quasistatic(0.5, 2)
        codes.append(e.value.code)
    assert codes == [3, 3, 3], codes
    assert_same(ens, singles, "failed solve")
    ens.close()


def test_member_matches_reference_fixture():
    """member 0 from the reference's rand64 fixture, beside random members"""
    fx = load_fixture("rand64")
    print("case ensemble rand64", flush=True)
    k = int(fx["steps"][-1])
    params = [fx["params"], simple_shear(128, u_bdry=2e-3), simple_shear(64, u_bdry=2e-3)]
    ens = Ensemble(params, "fp64")
    for j, p in enumerate(params):
        if j == 0:
            planes, vel, t0 = fx["in_planes"], fx["in_vel"], float(fx["t0"])
        else:
            (planes, vel), t0 = random_state(p, 40 + j), 0.7
        ens[j].set_fields(planes)
        ens[j].set_vel(vel)
        ens[j].time = t0
    ens.set_mg_config(tol_mult=float(fx["tol_mult"]))
    ens.step_forward_quasistatic(float(fx["dt"]), k)
    err = field_errors(ens[0].get_fields(), fx["planes_%d" % k])
    print("rand64 in ensemble", ["%.1e" % x for x in err])
    assert max(err) < 1e-10, err
    ens.close()


def test_unsupported_member_is_rejected():
    print("case ensemble reject", flush=True)
    ens = Ensemble([simple_shear(64), simple_shear(24, 31)], "mixed")
    with pytest.raises(StzError) as e:
        ens.step_forward_quasistatic(0.5, 1)
    assert e.value.code == 1 and "member 1" in str(e.value), str(e.value)
    ens.close()
