"""The exact solve of the direct coarse level (block-circulant inverse of an
x-invariant level, stz_mg_config.direct_nodes).

* With the V-cycle stopping at the direct level, a projection solve to a
  tight threshold gives the solution of the dense FP64 system (numpy LU of
  the operator written from its definition, tests/dense.py) on a 32^2 grid,
  in both precision forms.
* The same solve with the direct level and without it (the V-cycle down to
  the coarsest level, its workgroup solver) converges to the same solution.
  The 1024^2 parity tests (test_gpu_c2.py) run with the default, whose
  direct level there is the 64x65 one."""
import numpy as np
import pytest

from dense import top_dense
from stz_amd import ShearSim
from stz_amd.problems import simple_shear

pytestmark = pytest.mark.gpu


def solve(p, form, viscdt, b, tol, direct_nodes, nu=2):
    g = ShearSim(p, form)
    g.set_mg_config(tol_mult=tol, direct_nodes=direct_nodes, nu1=nu, nu2=nu)
    g.mg_setup(viscdt)
    g.mg_set_src(b)
    g.set_vel(np.zeros_like(b))
    info = g.mg_solve()
    z = g.get_vel()
    g.close()
    return z, info


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_direct_level_solve_equals_dense_solve(form):
    print("case direct dense32", form, flush=True)
    p = simple_shear(32)
    viscdt = 0.05
    rng = np.random.default_rng(3)
    b = rng.normal(size=(33, 32, 2))
    z, info = solve(p, form, viscdt, b, 1e-14, 4200)
    print(info)
    assert info["status"] == 0
    zd = np.linalg.solve(top_dense(p, viscdt), b.ravel())
    err = np.abs(z.ravel() - zd).max() / np.abs(zd).max()
    print(form, "vs dense %.2e" % err)
    assert err < 1e-11


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("size", [64, 256])
def test_direct_and_workgroup_coarse_solves_agree(form, size):
    print("case direct vs workgroup", form, size, flush=True)
    p = simple_shear(size)
    viscdt = 0.1
    rng = np.random.default_rng(4)
    b = rng.normal(size=(size + 1, size, 2))
    za, ia = solve(p, form, viscdt, b, 1e-14, 4200)
    zb, ib = solve(p, form, viscdt, b, 1e-14, 0)
    err = np.abs(za - zb).max() / np.abs(zb).max()
    print(form, size, "cycles direct %d workgroup %d, difference %.2e" % (ia["cycles"], ib["cycles"], err))
    assert ia["status"] == 0 and ib["status"] == 0
    assert err < 1e-10

