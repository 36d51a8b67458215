"""Pins the Fourier oracle (oracle/fourier_oracle.py).

* linear fields: the LSQ gradient (openbte/objects.py:322-405) is exact for
  T = g.x with the periodic jumps, on quads and on jittered triangles;
* pore-free periodic meshes: the two-point flux plus the non-orthogonal
  correction is exact for linear T, so the converged Fourier temperature is
  linear with slope jump/L;
* uniform quads (orthogonal): kappa_F = kappa_xx exactly;
* V F is a symmetric weighted graph Laplacian except the pinned row.
"""
import numpy as np
import pytest

from oracle import fourier_oracle as fo
from openbte_amd.geometry import pore_cell_2d, quad_cell_2d
from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity

EFF = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)


@pytest.mark.parametrize("geo", [quad_cell_2d(10.0, 9), pore_cell_2d(10.0, 0.0, 400, seed=4)])
def test_gradient_exact_for_linear_fields(geo):
    g = np.array([0.37, -0.81])
    L = 10.0
    bcs = BoundaryConditions(periodic={'Periodic_x': L * g[0], 'Periodic_y': L * g[1]})
    T = geo.elem_centroids @ g
    grad = fo.gradient(geo, bcs, T)
    assert np.abs(grad - g).max() < 1e-11


@pytest.mark.parametrize("geo", [quad_cell_2d(10.0, 9), pore_cell_2d(10.0, 0.0, 500, seed=8)])
def test_fourier_linear_solution(geo):
    kappa = np.array([[150.0, 0.0], [0.0, 150.0]])
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    r = fo.fourier(geo, kappa, bcs, EFF, fourier_max_iter=200, fourier_rtol=1e-14)
    T = r.variables['Temperature_Fourier']['data']
    x = geo.elem_centroids[:, 0]
    lin = x / 10.0
    d = T - lin
    assert np.abs(d - d.mean()).max() < 1e-9
    J = r.variables['Flux_Fourier']['data']
    np.testing.assert_allclose(J[:, 0], -150.0 / 10.0 * 1e9, rtol=1e-9)
    assert np.abs(J[:, 1]).max() < 1e-9 * 150.0 / 10.0 * 1e9
    assert r.kappa_eff == 0.0  # the reference's while-else resets it


def test_kappa_orthogonal_quads():
    geo = quad_cell_2d(10.0, 7)
    kappa = np.array([[2.5, 0.0], [0.0, 4.0]])
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0})
    r = fo.fourier(geo, kappa, bcs, EFF)
    assert abs(r.aux['kappa_diffusive'] - 2.5) < 1e-12


def test_symmetric_laplacian_structure():
    geo = pore_cell_2d(10.0, 0.25, 300, seed=6)
    kappa = 150.0 * np.eye(2)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    F, B, _, _ = fo.assemble(geo, kappa, bcs)
    A = (F.multiply(geo.elem_volumes[:, None])).toarray()
    keep = np.ones(geo.n_elems, bool)
    keep[fo.PINNED_CELL] = False
    Ak = A[np.ix_(keep, keep)]
    assert np.abs(Ak - Ak.T).max() < 1e-12 * np.abs(A).max()
    assert np.abs(A[keep].sum(axis=1)).max() < 1e-12 * np.abs(A).max()
    assert A[fo.PINNED_CELL, fo.PINNED_CELL] == geo.elem_volumes[fo.PINNED_CELL]


@pytest.mark.parametrize("h", [100.0, 1.0e9])
def test_mixed_contacts_closed_form(h, isothermal_mesh):
    """Isothermal contacts (mixed BCs) at x = -L/2 (T_L) and +L/2 (T_R) of a
    pore-free quad strip, periodic in y: the two-point flux scheme is exact for
    1D conduction, the contact being the conductance h in series with the
    half cell.  Flux q = (T_L - T_R) / (L / k + 2 / h); cell centres
    T(x) = T_L - q (1 / h + (x + L / 2) / k); kappa_diffusive = q L_x / dT
    with normalization -L_x / (L_y dT).  A wrong contact conductance, sign of
    the wall term or pinned row fails it."""
    L, Nx, Ny, k = 10.0, 10, 4, 150.0
    TL, TR = 0.7, -0.3
    geo = isothermal_mesh(L, Nx, Ny)
    Ly = L * Ny / Nx
    bcs = BoundaryConditions(periodic={'Periodic_y': 0.0},
                             mixed={'Left': {'value': TL, 'boundary_conductance': h},
                                    'Right': {'value': TR, 'boundary_conductance': h}})
    eff = EffectiveThermalConductivity(contact='Left', normalization=-L / (Ly * (TL - TR)))
    r = fo.fourier(geo, k * np.eye(2), bcs, eff)
    q = (TL - TR) / (L / k + 2.0 / h)
    x = geo.elem_centroids[:, 0]
    expect = TL - q * (1.0 / h + (x + L / 2) / k)
    np.testing.assert_allclose(r.variables['Temperature_Fourier']['data'], expect, rtol=0, atol=1e-12)
    assert abs(r.aux['kappa_diffusive'] / (q * L / (TL - TR)) - 1) < 1e-12
