"""Checks that pin the free-boundary fixtures (tests/golden/bd_*.npz, written
from the reference's own bd_sim by tests/golden/make_bd_fixtures.py) to
what the mathematics fixes, so that the GPU tests compare against outputs
whose meaning is established: status codes follow the sign pattern of phi
(levelset.cc:217-273) and the node rule (bd_sim.cc:125-172); the interface
cells hold the distance to the zero contour; every extrapolated cell value is
the weighted plane fit through the band cells of smaller phi
(traverse.hh:176-240); the free-surface rows give zero traction on a
uniform velocity; the projection's velocity solves the assembled system."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import load_fixture
from oracle.pyoracle import bd_remarch


@pytest.fixture(scope="module")
def plate():
    return load_fixture("bd_plate96")


def node_status(phi, st):
    n, m = st.shape
    out = np.zeros((n + 1, m + 1), np.int32)
    for je in range(n + 1):
        for ie in range(m + 1):
            k = l = 0
            cells = []
            for ci, cj in ((ie - 1, je - 1), (ie - 1, je), (ie, je - 1), (ie, je)):
                if 0 <= ci < m and 0 <= cj < n:
                    k += 1
                    cells.append((ci, cj))
                    if st[cj, ci] > 3:
                        l += 1
            if l == 0:
                out[je, ie] = 0
            elif l == k:
                out[je, ie] = 7
            else:
                out[je, ie] = 7 if sum(phi[cj, ci] for ci, cj in cells) > 0 else 0
    return out


def test_build_status_follows_sign_pattern(plate):
    phi0, phi, st = plate["phi0"], plate["build_phi"], plate["build_status"]
    dx = 2.0 / phi.shape[1]
    pos = phi0 > 0
    opp = np.zeros_like(pos)
    opp[:, 1:] |= pos[:, 1:] != pos[:, :-1]
    opp[:, :-1] |= pos[:, :-1] != pos[:, 1:]
    opp[1:, :] |= pos[1:, :] != pos[:-1, :]
    opp[:-1, :] |= pos[:-1, :] != pos[1:, :]
    assert np.array_equal((st == 3) | (st == 4), opp)
    assert np.array_equal(st >= 4, pos)
    band = (st >= 2) & (st <= 5)
    assert np.array_equal(band, np.abs(phi) < 5 * dx)
    assert np.all(np.abs(phi[~band]) == 5 * dx)
    assert np.array_equal(plate["build_nstatus"], node_status(phi, st))


def test_interface_cells_hold_contour_distance(plate):
    """plate_hole_phi is a signed distance away from its corners, so the
    Newton closest point on the bicubic interpolant must land within a small
    fraction of a cell of it"""
    phi0, phi, st = plate["phi0"], plate["build_phi"], plate["build_status"]
    dx = 2.0 / phi.shape[1]
    iface = (st == 3) | (st == 4)
    n, m = phi.shape
    x = -1 + (np.arange(m) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    far_from_corner = (np.abs(np.abs(X) - 1) > 3 * dx)
    sel = iface & far_from_corner
    assert np.all(np.sign(phi[sel]) == np.sign(phi0[sel]))
    assert np.max(np.abs(phi[sel] - phi0[sel])) < 0.05 * dx


def test_extrapolated_cells_are_plane_fits(plate):
    """restatement of traverse::extrapolate: every outside band cell holds
    the plane fitted, with weights phi(c)-phi(k), through the band cells k of
    smaller phi within a 7x7 window (growing while the fit is singular)"""
    phi, st = plate["build_phi"], plate["build_status"]
    n, m = phi.shape
    band = (st >= 2) & (st <= 5)
    names = ("p", "q", "s", "tau", "chi")
    post = {k: plate["post_" + k] for k in names}
    checked = 0
    for j, i in np.argwhere(band & (st >= 4)):
        for r in range(3, 8):
            ilo, ihi = max(-i, -r), min(m - 1 - i, r)
            jlo, jhi = max(-j, -r), min(n - 1 - j, r)
            A = np.zeros((3, 3))
            rhs = {k: np.zeros(3) for k in names}
            for dj in range(jlo, jhi + 1):
                for di in range(ilo, ihi + 1):
                    jj, ii = j + dj, i + di
                    if not band[jj, ii] or not phi[jj, ii] < phi[j, i]:
                        continue
                    w = phi[j, i] - phi[jj, ii]
                    v = np.array([1.0, di, dj])
                    A += w * np.outer(v, v)
                    for k in names:
                        rhs[k] += w * v * post[k][jj, ii]
            if abs(np.linalg.det(A)) >= 1e-10 * (2.0 / m) ** 0:
                break
        for k in names:
            plane = np.linalg.solve(A, rhs[k])
            assert abs(plane[0] - post[k][j, i]) <= 1e-9 * max(1e-300, np.abs(post[k]).max()), (k, i, j)
        checked += 1
    assert checked > 500


def assemble(plate):
    """the projection operator: interior rows from the constants of
    qs_multi_bd::init (qs_multi_bd.cc:29-51), box edges and outside nodes
    diagonal, free-surface rows from the fixture's blocks"""
    P = plate["params"]
    dt = float(plate["dt"])
    d = plate["op_rowtype"]
    me = d.shape[1]
    ne = d.shape[0]
    dx = (P["bx"] - P["ax"]) / P["m"]
    xsp = ysp = 1 / dx
    Kp = P["K"] + P["mu"] / 3.0
    Kmu, mu = Kp + P["mu"], P["mu"]
    viscdt = P["viscosity"] / dt
    xsp2 = ysp2 = xsp * xsp
    Kmx, Kmy = (Kmu + viscdt) * xsp2, (Kmu + viscdt) * ysp2
    mx, my = (mu + viscdt) * xsp2, (mu + viscdt) * ysp2
    qK = .25 * Kp * xsp * ysp
    cxx = -2 * (Kmu * xsp2 + mu * ysp2 + viscdt * (xsp2 + ysp2))
    cyy = -2 * (mu * xsp2 + Kmu * ysp2 + viscdt * (xsp2 + ysp2))
    interior = [[[0, qK], [qK, 0]], [[mx, 0], [0, Kmy]], [[0, -qK], [-qK, 0]],
                [[Kmx, 0], [0, my]], [[cxx, 0], [0, cyy]], [[Kmx, 0], [0, my]],
                [[0, -qK], [-qK, 0]], [[mx, 0], [0, Kmy]], [[0, qK], [qK, 0]]]
    blocks = np.zeros((ne * me, 9, 2, 2))
    dd = d.ravel()
    blocks[dd == -1] = np.array(interior)
    blocks[dd == -2, 4] = np.diag([cxx, cyy])
    blocks[dd == -3, 4] = np.diag([1e-3 * cxx, 1e-3 * cyy])
    blocks[plate["op_free_index"]] = plate["op_free_blocks"]
    rows, cols, vals = [], [], []
    idx = np.arange(me * ne)
    I, J = idx % me, idx // me
    for k in range(9):
        ii, jj = I + k % 3 - 1, J + k // 3 - 1
        ok = (ii >= 0) & (ii < me) & (jj >= 0) & (jj < ne)
        for a in range(2):
            for c in range(2):
                v = blocks[:, k, a, c]
                s = ok & (v != 0)
                rows.append(2 * idx[s] + a)
                cols.append(2 * (ii[s] + me * jj[s]) + c)
                vals.append(v[s])
    return blocks, sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                                 shape=(2 * me * ne, 2 * me * ne))


def test_free_rows_have_zero_traction_on_uniform_velocity(plate):
    """a uniform velocity has no strain, so each free-surface row (the
    traction sigma.n = 0 of qs_multi_bd.cc:74-149, preconditioned) annihilates
    it: the nine blocks sum to zero"""
    blocks, _ = assemble(plate)
    free = plate["op_free_index"]
    s = blocks[free].sum(axis=1)
    scale = np.abs(blocks[free]).max()
    assert free.size > 200
    assert np.abs(s).max() < 1e-12 * scale


def test_projection_solves_the_assembled_system(plate):
    _, A = assemble(plate)
    x = plate["proj_vel"].ravel()
    b = plate["op_src"].ravel()
    r = b - A @ x
    n = plate["op_rowtype"].size
    assert (r @ r) / n < 1e-19   # qs_multi_bd.cc:16: mean-square residual 1e-20
    assert np.abs(x).max() > 1e-6


@pytest.mark.parametrize("name", ["bd_bar1", "bd_bar2", "bd_bar3"])
def test_init_bar_statuses(name):
    f = load_fixture(name)
    phi, st = f["phi"], f["status"]
    dx = (f["params"]["bx"] - f["params"]["ax"]) / f["params"]["m"]
    band = (st >= 2) & (st <= 5)
    assert np.array_equal(band, np.abs(phi) < 5 * dx)
    assert np.array_equal(f["nstatus"], node_status(phi, st))
    inside = st < 4
    assert np.all(f["chi"][inside] == 620 / 21000)


def _planar(N, y0):
    """statuses and phi of a flat interface y = y0: the distance is exact"""
    dx = 2.0 / N
    x = -1 + (np.arange(N) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    phi = Y - y0
    pos = phi > 0
    opp = np.zeros_like(pos)
    opp[1:, :] |= pos[1:, :] != pos[:-1, :]
    opp[:-1, :] |= pos[:-1, :] != pos[1:, :]
    st = np.where(pos, np.where(opp, 4, 7), np.where(opp, 3, 0)).astype(np.int32)
    return phi, st, dx


@pytest.mark.parametrize("N,y0", [(32, 0.1234), (40, -0.377)])
def test_remarch_planar_is_the_distance(N, y0):
    """on a flat interface the rounds march must reproduce the signed
    distance y - y0 in the whole band (to rounding), with band statuses
    exactly where |y - y0| < 5 dx"""
    phi, st, dx = _planar(N, y0)
    p, s, rounds = bd_remarch(phi, st, dx, dx, 0.3 * dx)
    band = (s >= 2) & (s <= 5)
    assert np.array_equal(band, np.abs(phi) < 5 * dx)
    assert np.abs(p - phi)[band].max() < 1e-13
    assert np.all(np.abs(p[~band]) == 5 * dx)
    assert 5 <= rounds <= 40


def test_remarch_follows_reference_march(plate):
    """against the reference's heap march (the fixture's band, built by
    build_band from the analytic phi): same statuses, most band cells equal
    to rounding, none off by more than a tenth of a cell"""
    phi, st = plate["build_phi"], plate["build_status"]
    dx = 2.0 / phi.shape[1]
    p, s, _ = bd_remarch(phi, st, dx, dx, 0.3 * dx)
    assert np.array_equal(s, st)
    band = (st >= 2) & (st <= 5)
    d = np.abs(p - phi)[band] / dx
    assert (d <= 1e-12).mean() > 0.8
    assert d.max() < 0.1
    iface = (st == 3) | (st == 4)
    assert np.array_equal(p[iface], phi[iface])
