"""The free-boundary workload: a plate |y| < 0.9 with a central hole of
radius 0.2 L, pulled by grips on its left and right edges at nominal strain
rate 1e-4, chi = 580 K + 30 K seeded Gaussian field (correlation 4 cells),
dt 0.5, the mixed form (FP32 V-cycle preconditioner inside FP64 BiCGStab).

The reference cannot run this loading (bd_sim has no moving edges,
bd_sim.cc:732-764), so the checks are against mathematics:

* at 48^2, after an advection step and the regular-field post-processing,
  the projection's velocity equals scipy's sparse direct solution of the
  system the library assembles (stz_bd_operator: qs_multi_bd::init,
  qs_multi_bd.cc:26-156), in both precision forms;
* at 4096^2 (the workload's size), one step from rest: the solve reaches
  its threshold; the solution has the loading's mirror symmetries (u odd in
  x and even in y, v even in x and odd in y, statuses symmetric); and the
  horizontal force carried through vertical sections of the plate, the
  column sums of sigma_xx = -p - q + s over the body's cells (the stress
  combination of net_force, shear_sim.cc:447-453), is the same at the grip,
  half-way to the hole and through the hole's centre (equilibrium with free
  hole and plate surfaces), up to the viscous term and the
  cell-by-cell description of the curved surface."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from stz_amd.bd import BdSim
from stz_amd.initial import random_chi
from stz_amd.problems import BD_WORKLOADS, bd_params, plate_hole_phi

pytestmark = pytest.mark.gpu

W = BD_WORKLOADS["plate_hole4096"]
DT = 0.5


def workload(N, precision):
    p = bd_params(N)
    g = BdSim(p)
    g.set_level_set(plate_hole_phi(N))
    st, _ = g.status()
    chi = random_chi(p, W["chi_mean"], W["chi_std"], W["corr"], 4)
    g.set_cell("chi", np.where(st < 4, chi, 0.0))
    X, Y = g.node_coords()
    g.set_node("X", X)
    g.set_node("Y", Y)
    u = W["strain_rate"] * 0.5 * (p["bx"] - p["ax"])
    g.set_edge_velocity(-u, u)
    g.set_precision(precision)
    g.post_process(True)
    return p, g


def sparse_operator(d, blocks):
    """the 2x2-block 9-point operator of every node as a sparse matrix"""
    ne, me = d.shape
    idx = np.arange(ne * me)
    I, J = idx % me, idx // me
    B = blocks.reshape(ne * me, 9, 2, 2)
    rows, cols, vals = [], [], []
    for k in range(9):
        ii, jj = I + k % 3 - 1, J + k // 3 - 1
        ok = (ii >= 0) & (ii < me) & (jj >= 0) & (jj < ne)
        for a in range(2):
            for c in range(2):
                v = B[:, k, a, c]
                s = ok & (v != 0)
                rows.append(2 * idx[s] + a)
                cols.append(2 * (ii[s] + me * jj[s]) + c)
                vals.append(v[s])
    n2 = 2 * ne * me
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n2, n2))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_projection_equals_sparse_solve_48(precision):
    print("case c3 48", precision, flush=True)
    p, g = workload(48, precision)
    g.set_solver(1, 1, 20, 500, 1e-26)
    g.step_forward_quasistatic(DT, 2)
    g.advection_step(DT)
    g.post_process(False)
    d, blocks, b = g.operator(DT)
    x = spla.spsolve(sparse_operator(d, blocks).tocsc(), b.ravel())
    g.projection_step(DT)
    info = g.last_solve()
    v = g.get_vel().ravel()
    err = np.abs(v - x).max() / np.abs(x).max()
    print(precision, "vs sparse direct %.2e" % err, info)
    assert info["status"] == 0
    assert np.abs(x).max() > 1e-6
    assert err < 1e-9, err
    g.close()


def test_one_step_at_4096():
    N = 4096
    print("case c3 4096 setup", flush=True)
    p, g = workload(N, "mixed")
    print("case c3 4096 step", flush=True)
    g.step_forward_quasistatic(DT, 1)
    info = g.last_solve()
    print(info, g.step_stats(), flush=True)
    assert info["status"] == 0 and info["residual"] <= 1e-20
    st, nst = g.status()
    assert np.array_equal(st, st[:, ::-1]) and np.array_equal(st, st[::-1, :])
    u, v = g.get_node("u"), g.get_node("v")
    su = np.abs(u).max()
    assert su > 1e-5
    for a, sx, sy in ((u, -1, 1), (v, 1, -1)):
        assert np.abs(a - sx * a[:, ::-1]).max() < 1e-6 * su
        assert np.abs(a - sy * a[::-1, :]).max() < 1e-6 * su
    sxx = -g.get_cell("p") - g.get_cell("q") + g.get_cell("s")
    body = st < 4
    dy = (p["by"] - p["ay"]) / N
    F = {i: float(np.sum(np.where(body[:, i], sxx[:, i], 0.0)) * dy) for i in (0, N // 4, N // 2)}
    print("section forces", F, flush=True)
    assert F[0] > 0
    for i in (N // 4, N // 2):
        assert abs(F[i] - F[0]) < 2e-2 * abs(F[0]), F
    g.close()
