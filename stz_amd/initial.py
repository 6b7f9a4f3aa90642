"""Seeded initial effective-temperature fields.

``random_chi`` follows the construction of shear_sim::initialize_random
(shear_sim.cc:932-984): unit white noise on a grid padded by cut = int(5 l)
rows above and below, smoothed by the Gaussian kernel exp(-r^2 / (2 l^2))
truncated at |di|,|dj| <= cut (periodic in x), rescaled to standard deviation
sigma and shifted to the mean.  The noise comes from numpy's PCG64 instead of
glibc rand(), the kernel is applied as two separable 1D passes, and the
rescaling uses the kernel's exact L2 norm.
"""
import numpy as np


def random_chi(params, mean_K, std_K, corr_cells, seed):
    m, n = params["m"], params["n"]
    TZ = params["stz"]["TZ"]
    cut = int(corr_cells * 5.0)
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n + 2 * cut, m))
    d = np.arange(-cut, cut + 1)
    w = np.exp(-0.5 / corr_cells ** 2 * d * d)
    # x: periodic
    gx = np.zeros_like(noise)
    for k, wk in zip(d, w):
        gx += wk * np.roll(noise, -k, axis=1)
    # y: padded
    g = np.zeros((n, m))
    for k, wk in zip(d, w):
        g += wk * gx[cut + k: cut + k + n]
    g /= np.sqrt(np.sum(np.outer(w, w) ** 2))
    return (mean_K + std_K * g) / TZ


def simple_shear_state(params, mean_K, std_K, corr_cells, seed):
    """Planes (10, n+2, m) at t=0: zero stress and velocity, reference map at
    the node coordinates, seeded chi on rows 0..n-1 extrapolated linearly
    into the ghost rows (shear_sim.cc:552-581)."""
    m, n = params["m"], params["n"]
    dx = (params["bx"] - params["ax"]) / m
    dy = (params["by"] - params["ay"]) / n
    planes = np.zeros((10, n + 2, m))
    planes[7] = params["ax"] + dx * np.arange(m)[None, :]
    planes[8] = params["ay"] + dy * np.arange(-1, n + 1)[:, None]
    chi = random_chi(params, mean_K, std_K, corr_cells, seed)
    planes[6][1:n + 1] = chi
    planes[6][0] = 2 * chi[0] - chi[1]
    planes[6][n + 1] = 2 * chi[n - 1] - chi[n - 2]
    return planes
