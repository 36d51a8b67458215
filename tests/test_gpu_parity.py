"""GPU parity of the HIP path against the oracle (FP64 SuperLU restatement).

Gates: kappa_eff within 1e-6 relative, temperature within 1e-8 K max-abs
(same inputs, same Fourier guess, same outer-iteration count)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KAPPA_RTOL = 1e-6
T_ATOL = 1e-8


def _case(L, porosity, n, M, nphi, seed, mfp_max=1e4):
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(L, porosity, n, seed=seed)
    mat = synthetic_material_2d(M, nphi, 1.0, mfp_max)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def test_fourier_matches_oracle(small_case):
    from oracle import fourier_oracle as fo
    from openbte_amd.fourier import Fourier
    geo, mat, bcs, eff = small_case
    ref = fo.fourier(geo, mat.thermal_conductivity, bcs, eff)
    out = Fourier(geo, mat.thermal_conductivity, bcs, eff, verbose=False)
    print('fourier iters', ref.aux['n_iter'], out.aux['n_iter'], 'cg', out.aux['cg'][:3])
    assert out.aux['n_iter'] == ref.aux['n_iter']
    dT = np.abs(out.variables['Temperature_Fourier']['data'] - ref.variables['Temperature_Fourier']['data']).max()
    assert dT < T_ATOL
    assert abs(out.aux['kappa_diffusive'] / ref.aux['kappa_diffusive'] - 1) < 1e-9
    assert out.kappa_eff == 0.0
    dJ = np.abs(out.variables['Flux_Fourier']['data'] - ref.variables['Flux_Fourier']['data']).max()
    assert dJ < 1e-6 * np.abs(ref.variables['Flux_Fourier']['data']).max()


@pytest.mark.parametrize("mc,block", [(8, 64), (16, 256), (32, 128)])
def test_bte_matches_oracle_small(small_case, mc, block):
    from oracle import bte_oracle as bo, fourier_oracle as fo
    from openbte_amd.bte import BTE_RTA, SweepConfig
    geo, mat, bcs, eff = small_case
    f = fo.fourier(geo, mat.thermal_conductivity, bcs, eff)
    ref = bo.bte_rta(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, bte_max_iter=6)
    out = BTE_RTA(geo, mat, bcs, fourier=f, effective_thermal_conductivity=eff, bte_max_iter=6,
                  config=SweepConfig(mc=mc, block=block))
    assert out.aux['n_iter'] == ref.aux['n_iter']
    rel = abs(out.kappa_eff / ref.kappa_eff - 1)
    dT = np.abs(out.variables['Temperature_BTE']['data'] - ref.variables['Temperature_BTE']['data']).max()
    print('kappa', out.kappa_eff, ref.kappa_eff, 'rel', rel, 'dT', dT)
    assert rel < KAPPA_RTOL
    assert dT < T_ATOL
    np.testing.assert_allclose(out.aux['suppression'], ref.aux['suppression'], rtol=1e-8, atol=1e-12)
    Jr = ref.variables['Flux_BTE']['data']
    assert np.abs(out.variables['Flux_BTE']['data'] - Jr).max() < 1e-7 * np.abs(Jr).max()
