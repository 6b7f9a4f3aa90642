"""Dense restatements of the projection operator and prolongation for tiny
grids, written from their definitions (qs_multi.hh:30-50, tgmg.in.cc:552-621)."""
import numpy as np


def P1(uf, sc, periodic):
    P = np.zeros((uf, sc))
    for f in range(uf):
        if periodic:
            if f % 2 == 0:
                P[f, f // 2] = 1
            else:
                P[f, f // 2] += 0.5
                P[f, (f // 2 + 1) % sc] += 0.5
        else:
            if uf % 2 == 0 and f == uf - 1:
                P[f, sc - 1] = 1
            elif f % 2 == 0:
                P[f, f // 2] = 1
            else:
                P[f, f // 2] = 0.5
                P[f, f // 2 + 1] = 0.5
    return P


def top_dense(p, viscdt):
    """qs_multi's operator written from its definition (qs_multi.hh:30-50):
    interior x row: Kmx (l+r) + my (d+u) + qK (dl+ur-dr-ul).y + cxx z,
    y row symmetric; wall rows bcs*cxx, bcs*cyy on the diagonal only."""
    m, n = p["m"], p["n"] + 1
    dx = (p["bx"] - p["ax"]) / p["m"]
    dy = (p["by"] - p["ay"]) / p["n"]
    xsp, ysp = 1 / dx, 1 / dy
    Kp = p["K"] + p["mu"] / 3.0
    Kmu, mu = Kp + p["mu"], p["mu"]
    qK = 0.25 * Kp * xsp * ysp
    Kmx, Kmy = (Kmu + viscdt) * xsp ** 2, (Kmu + viscdt) * ysp ** 2
    mx, my = (mu + viscdt) * xsp ** 2, (mu + viscdt) * ysp ** 2
    cxx, cyy = -2 * (Kmx + my), -2 * (mx + Kmy)
    N = 2 * m * n
    A = np.zeros((N, N))
    idx = lambda i, j, c: 2 * ((i % m) + m * j) + c
    for j in range(n):
        for i in range(m):
            rx, ry = idx(i, j, 0), idx(i, j, 1)
            if j == 0 or j == n - 1:
                A[rx, rx] = 4 * cxx
                A[ry, ry] = 4 * cyy
                continue
            A[rx, rx] += cxx
            A[ry, ry] += cyy
            for di, dj, wx, wy in ((-1, 0, Kmx, mx), (1, 0, Kmx, mx), (0, -1, my, Kmy), (0, 1, my, Kmy)):
                A[rx, idx(i + di, j + dj, 0)] += wx
                A[ry, idx(i + di, j + dj, 1)] += wy
            for di, dj, sgn in ((-1, -1, 1), (1, 1, 1), (1, -1, -1), (-1, 1, -1)):
                A[rx, idx(i + di, j + dj, 1)] += sgn * qK
                A[ry, idx(i + di, j + dj, 0)] += sgn * qK
    return A


def stencil_from_dense(Ac, m, n):
    S = np.zeros((n, m, 9, 4))
    for j in range(n):
        for i in range(m):
            for k in range(9):
                di, dj = k % 3 - 1, k // 3 - 1
                jj = j + dj
                if jj < 0 or jj >= n:
                    continue
                ii = (i + di) % m
                r, c = 2 * (i + m * j), 2 * (ii + m * jj)
                S[j, i, k] = [Ac[r, c], Ac[r, c + 1], Ac[r + 1, c], Ac[r + 1, c + 1]]
    return S
