"""Seeded synthetic simulation states for parity tests.

The states are built to exercise every branch of the explicit step: smooth
random stresses whose deviatoric magnitude spans the yield threshold (sbar
from ~0 to ~2, so the plastic loop takes many substeps in places),
velocities of both signs (both upwind sides), a reference map perturbed away
from the identity, chi around 620 K, and a band of cells where the fields are
exactly linear in x (equal second differences, so eno2's strict '>' tie rule
is hit).
"""
import numpy as np

NFIELDS = 10


def _smooth(rng, m, n, modes=4):
    x = (np.arange(m) + 0.5) / m
    y = (np.arange(-1, n + 1) + 0.5) / n
    f = np.zeros((n + 2, m))
    for _ in range(modes):
        a, b = rng.integers(1, 4, 2)
        f += rng.normal() * np.cos(2 * np.pi * a * x[None, :] + rng.uniform(0, 2 * np.pi)) * \
            np.cos(np.pi * b * y[:, None] + rng.uniform(0, 2 * np.pi))
    return f


def random_state(params, seed, stress=1.0, vel=3e-3, chi_mean=620.0, chi_amp=40.0, ties=True):
    """Return (planes[10, n+2, m], vel[n+1, m, 2]) for shear_sim params."""
    rng = np.random.default_rng(seed)
    m, n = params["m"], params["n"]
    TZ = params["stz"]["TZ"]
    planes = np.zeros((NFIELDS, n + 2, m))
    planes[0] = vel * _smooth(rng, m, n)
    planes[1] = vel * _smooth(rng, m, n)
    planes[2] = stress * _smooth(rng, m, n)
    planes[3] = 0.4 * stress * _smooth(rng, m, n)
    planes[4] = stress * _smooth(rng, m, n)
    planes[5] = stress * _smooth(rng, m, n)
    planes[6] = (chi_mean + chi_amp * _smooth(rng, m, n)) / TZ
    dx = (params["bx"] - params["ax"]) / m
    dy = (params["by"] - params["ay"]) / n
    xs = params["ax"] + dx * np.arange(m)[None, :]
    ys = params["ay"] + dy * np.arange(-1, n + 1)[:, None]
    planes[7] = xs + 0.1 * dx * _smooth(rng, m, n)
    planes[8] = ys + 0.1 * dy * _smooth(rng, m, n)
    if ties and m >= 16:
        # a band where p, q, s, tau, chi are exactly linear in x, so the two
        # second differences compared by eno2 are equal
        c0, c1 = m // 4, m // 4 + 6
        i = np.arange(c0 - 2, c1 + 2, dtype=np.float64)
        for k, a in ((2, 0.125), (3, 0.0625), (4, -0.25), (5, 0.5)):
            planes[k][:, c0 - 2:c1 + 2] = a * (i - c0)[None, :]
        planes[6][:, c0 - 2:c1 + 2] = (0.03125 + 2.0 ** -12 * (i - c0))[None, :]
    v = np.zeros((n + 1, m, 2))
    v[..., 0] = planes[0][1:n + 2]
    v[..., 1] = planes[1][1:n + 2]
    return planes, v
