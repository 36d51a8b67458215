"""Synthetic material: the semantics of the reference's Material
(openbte/material.py:105-189) on the workloads' grids."""
import numpy as np
import pytest

from openbte_amd.material import synthetic_material_2d, synthetic_material_3d, gray_2d, kappa_weights


@pytest.mark.parametrize("M,nphi,mfp_max", [(30, 48, 1e4), (60, 96, 1e5), (20, 48, 1e5)])
def test_material_2d(M, nphi, mfp_max):
    mat = synthetic_material_2d(M, nphi, 1.0, mfp_max)
    assert np.all(np.isfinite(mat.thermal_conductivity))
    np.testing.assert_allclose(mat.thermal_conductivity, 150.0 * np.eye(2), atol=1e-9)
    assert abs(mat.t_coeff.sum() - 1) < 1e-14
    # directions: s_l = [sin phi, cos phi], l = 0 along +y
    np.testing.assert_allclose(mat.vmfp[0], [0.0, 1.0], atol=1e-15)
    # sum_l t_ml s_l = 0 per bin (symmetric direction set)
    assert np.abs(np.einsum('ml,li->mi', mat.t_coeff, mat.vmfp)).max() < 1e-16
    # sigma = gamma * Lambda * s * 1e-9 with Lambda in m  <=>  sigma / t ∝ Lambda_nm s
    g = mat.t_coeff
    nz = g > 0
    ratio = mat.sigma[..., 0][nz] / (g[nz] * (mat.mfps[:, None] * mat.vmfp[None, :, 0] * np.ones_like(g))[nz] + 1e-300)
    ok = np.abs(mat.vmfp[None, :, 0] * np.ones_like(g))[nz] > 1e-12
    assert np.ptp(ratio[ok]) < 1e-9 * np.abs(ratio[ok]).max()


def test_kappa_weights_lognormal_median():
    mfps, w = kappa_weights(60, 1.0, 1e5)
    assert abs(w.sum() - 150.0) < 1e-9
    c = np.cumsum(w) / w.sum()
    med = mfps[np.searchsorted(c, 0.5)]
    assert 200 < med < 450


def test_material_3d_isotropic():
    mat = synthetic_material_3d(10, 8, 8)
    k = mat.thermal_conductivity
    assert np.all(np.isfinite(k))
    assert np.abs(k - np.diag(np.diag(k))).max() < 1e-9 * k.max()
    assert abs(mat.t_coeff.sum() - 1) < 1e-14


def test_gray():
    mat = gray_2d(10.0, 16, 2.0)
    k = np.einsum('mli,ml,mlj->ij', mat.sigma, 1.0 / (mat.t_coeff * 0 + 1), mat.sigma)
    assert mat.sigma.shape == (1, 16, 2)
