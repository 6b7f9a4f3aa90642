"""CPU: the host side of the adjoint driver.  The differentiable Hencky
strain used as a calibration target against the oracle's
compute_strain (corrected form, itself pinned by the closed form of a
uniform simple shear and, in its reference form, by the reference's
fixture), and the chi-plane embedding against its adjoint (dot-product
test)."""
import numpy as np
import pytest

from conftest import load_fixture
from oracle.pyoracle import Oracle

torch = pytest.importorskip("torch")

from stz_amd.adjoint import chi_planes, chi_planes_adjoint, hencky_strain  # noqa: E402
from stz_amd.problems import simple_shear  # noqa: E402


def test_hencky_strain_matches_compute_strain():
    fx = load_fixture("diag40x24")
    o = Oracle(fx["params"])
    o.set_fields(fx["in_planes"])
    o.compute_strain(True)
    ref = o.get_cu()[1:-1]
    got = hencky_strain(torch.from_numpy(fx["in_planes"]), fx["params"]).numpy()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_hencky_strain_uniform_shear_closed_form():
    """X = x - g y, Y = y: devmod = sqrt(2) asinh(g/2) on every cell"""
    p = simple_shear(16, 12)
    g = 0.3
    planes = np.zeros((10, 14, 16))
    x = -1 + (2.0 / 16) * np.arange(16)
    y = -1 + (2.0 / 12) * (np.arange(14) - 1)
    planes[7] = x[None, :] - g * y[:, None]
    planes[8] = np.broadcast_to(y[:, None], (14, 16))
    got = hencky_strain(torch.from_numpy(planes), p).numpy()
    assert np.allclose(got, np.sqrt(2) * np.arcsinh(g / 2), rtol=1e-13, atol=0)


def test_hencky_strain_gradient():
    fx = load_fixture("diag40x24")
    planes = torch.from_numpy(fx["in_planes"][:, :8, :10].copy())
    p = dict(fx["params"], m=10, n=6)
    planes.requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: hencky_strain(t, p), (planes,), eps=1e-7, atol=1e-6)


def test_chi_planes_adjoint_is_the_transpose():
    rng = np.random.default_rng(3)
    n, m = 12, 16
    base = rng.standard_normal((10, n + 2, m))
    dchi = rng.standard_normal((n, m))
    lam = rng.standard_normal((10, n + 2, m))
    d = chi_planes(dchi, base) - chi_planes(np.zeros((n, m)), base)
    assert np.isclose(np.sum(lam * d), np.sum(dchi * chi_planes_adjoint(lam)), rtol=1e-13)
