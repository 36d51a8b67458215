import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libopenbte_amd.so")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_case():
    """~600-cell periodic pore cell, 6 MFP bins x 8 directions."""
    from openbte_amd.geometry import pore_cell_2d
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo = pore_cell_2d(10.0, 0.25, 600, seed=3)
    mat = synthetic_material_2d(6, 8, 1.0, 1.0e4)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff


def two_wall_quads(L: float = 10.0, N: int = 10):
    """Uniform N x N periodic quad cell with a pore made of removed cells: a
    diagonal staircase (cells beside two of its steps have walls on two
    adjacent sides) and two columns around a one-cell-wide slot (walls on two
    opposite sides).  Returns the mesh and the number of cells with two walls."""
    from openbte_amd.geometry import _REGIONS, _edge_region
    from openbte_amd.mesh import build_mesh
    xs = np.linspace(-L / 2, L / 2, N + 1)
    X, Y = np.meshgrid(xs, xs, indexing='ij')
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.arange((N + 1) ** 2).reshape(N + 1, N + 1)
    q = np.stack([idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]], axis=-1)  # [i][j][4]
    removed = np.zeros((N, N), bool)
    for k in range(2, 6):
        removed[k, k] = True
    removed[6, 2:6] = True
    removed[8, 2:6] = True
    geo = build_mesh(pts, q[~removed].reshape(-1, 4), _edge_region(L, L), _REGIONS,
                     periodic={'Periodic_x': np.array([L, 0.0]), 'Periodic_y': np.array([0.0, L])})
    walls = np.bincount(geo.side_elem_map[np.asarray(geo.side_physical_regions['Boundary']), 0],
                        minlength=geo.n_elems)
    return geo, int((walls >= 2).sum())


@pytest.fixture(scope="session")
def two_wall_case():
    """two_wall_quads with 3 MFP bins x 8 directions and diffuse pore walls."""
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    geo, n2 = two_wall_quads()
    mat = synthetic_material_2d(3, 8, 1.0, 1.0e3)
    bcs = BoundaryConditions(periodic={'Periodic_x': 1.0, 'Periodic_y': 0.0}, diffuse='Boundary')
    eff = EffectiveThermalConductivity(contact='Periodic_x', normalization=-1.0)
    return geo, mat, bcs, eff, n2


def isothermal_quads(L: float = 10.0, Nx: int = 10, Ny: int = 6, pore: bool = False):
    """Uniform Nx x Ny quad cell, periodic in y, with the x = -L/2 and x = +L/2
    sides in regions 'Left' and 'Right' (isothermal contacts); with ``pore``
    a staircase of removed cells adds diffuse walls ('Boundary')."""
    from openbte_amd.geometry import _edge_region
    from openbte_amd.mesh import build_mesh
    Ly = L * Ny / Nx
    xs = np.linspace(-L / 2, L / 2, Nx + 1)
    ys = np.linspace(-Ly / 2, Ly / 2, Ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing='ij')
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.arange((Nx + 1) * (Ny + 1)).reshape(Nx + 1, Ny + 1)
    q = np.stack([idx[:-1, :-1], idx[1:, :-1], idx[1:, 1:], idx[:-1, 1:]], axis=-1)
    removed = np.zeros((Nx, Ny), bool)
    if pore:
        for k in range(2, min(Nx, Ny) - 1):
            removed[k, k - 1] = True
    return build_mesh(pts, q[~removed].reshape(-1, 4), _edge_region(L, Ly),
                      ['Left', 'Right', 'Periodic_y_a', 'Periodic_y_b', 'Boundary'],
                      periodic={'Periodic_y': np.array([0.0, Ly])})


@pytest.fixture(scope="session")
def isothermal_case():
    """isothermal_quads with a pore: contacts at +0.5 / -0.5, kappa at 'Left'."""
    from openbte_amd.material import synthetic_material_2d
    from openbte_amd.objects import BoundaryConditions, EffectiveThermalConductivity
    L, Nx, Ny = 10.0, 12, 8
    geo = isothermal_quads(L, Nx, Ny, pore=True)
    mat = synthetic_material_2d(4, 8, 1.0, 1.0e3)
    bcs = BoundaryConditions(periodic={'Periodic_y': 0.0}, diffuse='Boundary',
                             mixed={'Left': {'value': 0.5, 'boundary_conductance': 100.0},
                                    'Right': {'value': -0.5, 'boundary_conductance': 100.0}})
    eff = EffectiveThermalConductivity(contact='Left', normalization=-Nx / Ny)
    return geo, mat, bcs, eff


@pytest.fixture(scope="session")
def isothermal_mesh():
    """The isothermal_quads builder, for tests that choose its sizes."""
    return isothermal_quads
