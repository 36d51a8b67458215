"""GPU interpolation of phonon modes onto (MFP, direction) grids equals the
oracle (reference index rule) on a subset of the Si database."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _data():
    from openbte_amd.objects import MaterialRTA
    d = np.load(os.path.join(HERE, 'golden', 'si_rta_subset.npz'))
    return MaterialRTA(d['heat_capacity'], d['scattering_time'], d['group_velocity'], d['frequency'], d['kappa_db'])


def test_rta2dsym_matches_oracle():
    from oracle import material_oracle as mo
    from openbte_amd.rta import RTA2DSym
    d = _data()
    mat = RTA2DSym(d, n_phi=48, n_mfp=30)
    Ws, Ss, grid, kappa, h, coeff = mo.rta2d_sym(d.heat_capacity, d.scattering_time, d.group_velocity,
                                                 n_phi=48, n_mfp=30)
    assert np.allclose(mat.mfps, grid * 1e9, rtol=1e-14)
    assert np.allclose(mat.t_coeff, Ws / Ws.sum(), rtol=1e-11, atol=1e-16)
    assert np.allclose(mat.sigma, Ss * 1e-9, rtol=1e-11, atol=1e-12 * np.abs(Ss).max() * 1e-9)
    assert np.allclose(mat.thermal_conductivity, kappa, rtol=1e-11)
    assert np.isclose(mat.h, h) and np.isclose(mat.heat_source_coeff, coeff)
    assert mat.grid == (30, 48)


def test_rta3d_matches_oracle():
    from oracle import material_oracle as mo
    from openbte_amd.rta import RTA3D
    d = _data()
    mat = RTA3D(d, n_phi=12, n_mfp=20, n_theta=8)
    Ws, Ss, grid, kappa, h, coeff = mo.rta3d(d.heat_capacity, d.scattering_time, d.group_velocity, 12, 20, 8)
    assert np.allclose(mat.t_coeff, Ws / Ws.sum(), rtol=1e-11, atol=1e-16)
    assert np.allclose(mat.sigma, Ss * 1e-9, rtol=1e-11, atol=1e-12 * np.abs(Ss).max() * 1e-9)
    assert np.allclose(mat.thermal_conductivity, kappa, rtol=1e-11)
    assert mat.grid == (20, 8, 12) and mat.vmfp.shape == (96, 3)
    assert np.allclose(np.linalg.norm(mat.vmfp, axis=1), 1.0)


def test_bte_with_interpolated_si_matches_oracle():
    """End to end with an interpolated ab-initio material (Si subset)."""
    from oracle import bte_oracle as bo, fourier_oracle as fo
    from openbte_amd.bte import BTE_RTA
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    from openbte_amd.rta import RTA2DSym
    mat = RTA2DSym(_data(), n_phi=16, n_mfp=10)
    geo = pore_cell_2d(20.0, 0.2, 600, seed=5)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    f = fo.fourier(geo, mat.thermal_conductivity, bcs, eff)
    ref = bo.bte_rta(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, bte_max_iter=6)
    out = BTE_RTA(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, bte_max_iter=6)
    print('Si kappa', out.kappa_eff, ref.kappa_eff)
    assert abs(out.kappa_eff / ref.kappa_eff - 1) < 1e-6
    dT = np.abs(out.variables['Temperature_BTE']['data'] - ref.variables['Temperature_BTE']['data']).max()
    assert dT < 1e-8
