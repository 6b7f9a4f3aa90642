"""GPU: the hand-written adjoint of the FP64 step (stz_adjoint_*) against
central finite differences of the same forward step on the GPU.

f(x, theta) = <lam, final fields> for a random lam.  The adjoint gives
df/dx and df/d(c0, Omegaeps0, chi_inf); a central difference along a random
direction v (fields) or e (parameters) must agree within 1e-5 relative,
away from kinks: every step's decision bits (upwind sides, ENO branches,
wall rules, yield branch, plastic substep counts) must be the same in the
+h and -h runs at two step sizes, and the two differences must agree, else
the direction is redrawn.  The adjoint is also independent of the
checkpoint interval (recomputation repeats the forward exactly)."""
import numpy as np
import pytest

from stategen import random_state
from stz_amd import ShearSim
from stz_amd.adjoint import PARAMS, Adjoint, set_params
from stz_amd.problems import simple_shear

pytestmark = pytest.mark.gpu

TOL = 1e-6      # forward solve threshold, x the reference's
DT, STEPS = 0.5, 3


def case(m=32, seed=5):
    p = simple_shear(m, u_bdry=2e-3)
    planes, vel = random_state(p, seed, ties=False)
    return p, planes, vel


def forward(p, planes, vel, theta=None, decisions=True):
    g = ShearSim(p, "fp64")
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = 1.2
    g.set_mg_config(tol_mult=TOL)
    if theta is not None:
        set_params(g, **theta)
    g.record_decisions(decisions)
    bits = []
    for _ in range(STEPS):
        g.step_forward_quasistatic(DT, 1)
        bits.append(g.get_decisions().copy())
    F = g.get_fields()
    g.close()
    return F, bits


def gradient(p, planes, vel, lam, every=2):
    g = ShearSim(p, "fp64")
    g.set_fields(planes)
    g.set_vel(vel)
    g.time = 1.2
    g.set_mg_config(tol_mult=TOL)
    a = Adjoint(g, every)
    a.forward(DT, STEPS)
    lam_init, gth = a.backward(lam)
    st = a.stats()
    a.close()
    g.close()
    return lam_init, gth, st


def smooth(rng, shape, modes=3):
    n2, m = shape
    x = np.arange(m) / m
    y = np.arange(n2) / n2
    f = np.zeros(shape)
    for _ in range(modes):
        a, b = rng.integers(1, 4, 2)
        f += rng.normal() * np.cos(2 * np.pi * a * x[None, :] + rng.uniform(0, 6.3)) * \
            np.cos(np.pi * b * y[:, None] + rng.uniform(0, 6.3))
    return f


def oracle_forward(p, planes, vel, theta):
    """the same steps through the FP64 restatement of the reference's step
    (oracle/, on the CPU), with its own decision bits"""
    from oracle.pyoracle import Oracle
    q = dict(p)
    q["stz"] = dict(p["stz"], c0=theta["c0"], Omegaeps0=theta["Omegaeps0"])
    o = Oracle(q)
    o.set_fields(planes)
    o.set_vel(vel)
    o.time = 1.2
    o.set_tol_mult(TOL)
    o.set_chi_inf(theta["chi_inf"])
    o.record_decisions(True)
    bits = []
    for _ in range(STEPS):
        assert o.step(DT, 1) == 0
        bits.append(o.get_decisions().copy())
    F = o.get_fields()
    o.close()
    return F, bits


def fd(p, planes, vel, lam, dx, dth, h, fwd=forward):
    """central difference of <lam, final fields> along (dx, dth); None when
    a decision of any step differs between the +h and -h runs"""
    base = {k: p["stz"][k] for k in PARAMS}
    th_p = {k: base[k] + h * dth.get(k, 0.0) for k in base}
    th_m = {k: base[k] - h * dth.get(k, 0.0) for k in base}
    Fp, bp = fwd(p, planes + h * dx, vel, th_p)
    Fm, bm = fwd(p, planes - h * dx, vel, th_m)
    if not all(np.array_equal(a, b) for a, b in zip(bp, bm)):
        return None
    return (np.sum(lam * Fp) - np.sum(lam * Fm)) / (2 * h)


def check_direction(p, planes, vel, lam, dirs, label, h=1e-7, fwd=forward):
    """away from kinks: the decisions agree at h and 4h and the two central
    differences agree to 1e-6 (no hidden switch inside the plastic
    substeps; a switch moves a difference by far more)"""
    for attempt, (dx, dth, ad) in enumerate(dirs):
        f1 = fd(p, planes, vel, lam, dx, dth, h, fwd)
        f4 = fd(p, planes, vel, lam, dx, dth, 4 * h, fwd)
        smooth_here = f1 is not None and f4 is not None and abs(f1 - f4) <= 1e-6 * max(abs(f1), 1e-300)
        print("%s attempt %d: fd(h) %s fd(4h) %s ad %.10e smooth %s" % (label, attempt, f1, f4, ad, smooth_here),
              flush=True)
        if not smooth_here:
            continue
        rel = abs(f1 - ad) / max(abs(f1), abs(ad), 1e-300)
        print("%s: rel %.2e" % (label, rel), flush=True)
        assert rel < 1e-5, (label, f1, ad, rel)
        return
    pytest.skip("%s: every direction drawn is near a kink" % label)


@pytest.mark.parametrize("which", ["chi", "fields", "params"])
def test_adjoint_matches_central_differences(which):
    print("case adjoint", which, flush=True)
    p, planes, vel = case()
    rng = np.random.default_rng(17)
    lam = rng.standard_normal(planes.shape)
    lam_init, gth, st = gradient(p, planes, vel, lam)
    print("adjoint stats", st, "grad theta", gth, flush=True)
    scales = np.array([np.max(np.abs(planes[k])) for k in range(10)])

    def field_dirs():
        for _ in range(6):
            dx = np.zeros_like(planes)
            for k in ([6] if which == "chi" else range(9)):
                dx[k] = (scales[k] if scales[k] > 0 else 1.0) * smooth(rng, planes[k].shape)
            yield dx, {}, float(np.sum(lam_init * dx))

    if which == "params":
        for k in PARAMS:
            e = {k: p["stz"][k]}
            check_direction(p, planes, vel, lam, iter([(np.zeros_like(planes), e, gth[k] * e[k])]), "param " + k)
    else:
        # every field at once meets more near-ties (the reference map is
        # nearly linear), so a smaller step
        check_direction(p, planes, vel, lam, field_dirs(), which, h=1e-7 if which == "chi" else 2e-8)


def test_adjoint_matches_differences_of_the_reference_step():
    """the GPU adjoint against central differences of the oracle's FP64
    restatement of the reference's step (the oracle, CPU): an independent
    FP64 reference gradient along a chi direction and each parameter, away
    from kinks by the oracle's own decision bits"""
    print("case adjoint vs oracle differences", flush=True)
    p, planes, vel = case()
    rng = np.random.default_rng(29)
    lam = rng.standard_normal(planes.shape)
    lam_init, gth, _ = gradient(p, planes, vel, lam)
    scale = np.max(np.abs(planes[6]))

    def chi_dirs():
        for _ in range(6):
            dx = np.zeros_like(planes)
            dx[6] = scale * smooth(rng, planes[6].shape)
            yield dx, {}, float(np.sum(lam_init * dx))

    check_direction(p, planes, vel, lam, chi_dirs(), "oracle chi", fwd=oracle_forward)
    for k in PARAMS:
        e = {k: p["stz"][k]}
        check_direction(p, planes, vel, lam, iter([(np.zeros_like(planes), e, gth[k] * e[k])]), "oracle " + k,
                        fwd=oracle_forward)


def test_adjoint_checkpoint_interval_does_not_matter():
    """recomputing from checkpoints repeats the forward exactly: the
    gradient is the same for every checkpoint interval"""
    print("case adjoint checkpoints", flush=True)
    p, planes, vel = case()
    lam = np.random.default_rng(2).standard_normal(planes.shape)
    ref = None
    for every in (1, 2, 3):
        li, g, _ = gradient(p, planes, vel, lam, every=every)
        if ref is None:
            ref = (li, g)
            continue
        assert np.array_equal(li, ref[0]) or np.max(np.abs(li - ref[0])) <= 1e-12 * np.max(np.abs(ref[0]))
        for k in PARAMS:
            assert abs(g[k] - ref[1][k]) <= 1e-12 * abs(ref[1][k])


def test_autograd_run_with_strain_loss():
    """QSRun (the torch.autograd.Function over the initial chi field and the
    three parameters) under the calibration's loss, a weighted sum of the
    Hencky strain of the final state: its gradient along a chi direction and
    along each parameter against central differences of the same loss, away
    from kinks as above.  The strain of a state near the identity map is
    computed with a cancellation (C00 - C11 of C = F^T F near I), which
    leaves a rounding noise of about 1e-15 in the loss: the steps tried
    here start 20 times longer than above and shrink when the longer step
    crosses a kink."""
    import torch
    from stz_amd.adjoint import chi_planes, hencky_strain, make_qs_run
    print("case adjoint autograd strain loss", flush=True)
    p, planes, vel = case()
    n, m = p["n"], p["m"]
    rng = np.random.default_rng(23)
    w = torch.from_numpy(rng.standard_normal((n, m)))
    theta = np.array([p["stz"][k] for k in PARAMS])
    chi0 = planes[6][1:n + 1].copy()
    sim = ShearSim(p, "fp64")
    sim.set_mg_config(tol_mult=TOL)
    chi_t = torch.from_numpy(chi0.copy()).requires_grad_(True)
    th_t = torch.from_numpy(theta.copy()).requires_grad_(True)
    F = make_qs_run().apply(chi_t, th_t, sim, planes, vel, 1.2, DT, STEPS, 2, 1e-10)
    loss = torch.sum(w * hencky_strain(F, p))
    loss.backward()
    sim.close()
    g_chi, g_th = chi_t.grad.numpy(), th_t.grad.numpy()

    def f(chi, th):
        Fk, bits = forward(p, chi_planes(chi, planes), vel, dict(zip(PARAMS, th)))
        return float(torch.sum(w * hencky_strain(torch.from_numpy(Fk), p))), bits

    def central(dchi, dth, h):
        fp, bp = f(chi0 + h * dchi, theta + h * dth)
        fm, bm = f(chi0 - h * dchi, theta - h * dth)
        if not all(np.array_equal(a, b) for a, b in zip(bp, bm)):
            return None
        return (fp - fm) / (2 * h)

    def check(k, dchi, dth):
        ad = float(np.sum(g_chi * dchi) + np.dot(g_th, dth))
        for h in (2e-6, 5e-7, 1e-7):
            f1, f4 = central(dchi, dth, h), central(dchi, dth, 4 * h)
            ok = f1 is not None and f4 is not None and abs(f1 - f4) <= 1e-6 * max(abs(f1), 1e-300)
            print("direction %s h %g: fd(h) %s fd(4h) %s ad %.10e smooth %s" % (k, h, f1, f4, ad, ok), flush=True)
            if ok:
                rel = abs(f1 - ad) / max(abs(f1), abs(ad), 1e-300)
                assert rel < 1e-5, (k, f1, ad, rel)
                return True
        return False

    chi_ok = any(check("chi %d" % k, np.max(np.abs(chi0)) * smooth(rng, (n, m)), np.zeros(3))
                 for k in range(4))
    par_ok = [check(PARAMS[k], np.zeros((n, m)), theta * np.eye(3)[k]) for k in range(3)]
    assert chi_ok and sum(par_ok) >= 2, "too few directions away from kinks"
