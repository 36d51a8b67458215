"""Data contracts of the steady-state BTE path.

These NamedTuples keep the field names, shapes and units of the reference's
contracts so that code written against the reference can hand its inputs to
this package unchanged:

* ``Mesh``                          -- openbte/objects.py:244-268
* ``Material``                      -- openbte/objects.py:500-509
* ``BoundaryConditions``            -- openbte/objects.py:220-241
* ``EffectiveThermalConductivity``  -- openbte/objects.py:128-135
* ``SolverResults``                 -- openbte/objects.py:493-497

Units follow the reference: lengths in nm, ``Material.mfps`` in nm,
``Material.sigma`` is C*v scaled by 1e-9, temperatures in K.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np


class Mesh(NamedTuple):
    """Finite-volume mesh (field contract of openbte/objects.py:244-268).

    Semantics (openbte/mesh.py:180-320):
      * ``side_elem_map[s] = [e0, e1]``; a boundary side has ``e0 == e1``.
      * ``normal_areas[s]`` points out of ``e0`` and has length ``side_areas[s]``.
      * A periodic pair keeps the side of the ``<name>_b`` region; the kept side
        is internal and ``side_elem_map[kept, 1]`` is the partner element on the
        discarded ``<name>_a`` side.
      * ``internal`` = every side that is neither a remaining boundary side nor
        a discarded periodic side.
    """

    n_elems: int
    n_nodes: int
    dim: int
    size: np.ndarray
    nodes: np.ndarray
    sides: np.ndarray
    elems: np.ndarray
    elem_side_map: np.ndarray
    side_elem_map: np.ndarray
    normal_areas: np.ndarray
    gradient_weights: np.ndarray
    interp_weights: np.ndarray
    elem_centroids: np.ndarray
    side_centroids: np.ndarray
    elem_volumes: np.ndarray
    side_areas: np.ndarray
    dists: np.ndarray
    second_order_correction: np.ndarray
    internal: np.ndarray
    boundary_sides: np.ndarray
    periodic_sides: dict = {}
    elem_physical_regions: dict = {}
    side_physical_regions: dict = {}


class Material(NamedTuple):
    """Mode-resolved RTA material (openbte/objects.py:500-509).

    ``sigma[m, l, :]`` = C v (x1e-9), ``vmfp[l, :]`` = unit direction s_l,
    ``t_coeff[m, l]`` = gamma_ml / sum(gamma), ``mfps[m]`` = Lambda_m in nm,
    ``grid`` = (n_mfp, n_dirs) in 2D or (n_mfp, n_theta, n_phi) in 3D
    (openbte/material.py:100,183-189).
    """

    thermal_conductivity: np.ndarray
    sigma: np.ndarray
    vmfp: np.ndarray
    t_coeff: np.ndarray
    mfps: np.ndarray
    grid: tuple
    h: float
    heat_source_coeff: float


class MaterialRTA(NamedTuple):
    """Phonon modes of a relaxation-time material (openbte/objects.py:114-120):
    per mode heat capacity (W s/m^3/K), scattering time (s), group velocity
    (m/s, 3 components) and frequency; the bulk conductivity tensor (W/m/K)."""

    heat_capacity: np.ndarray
    scattering_time: np.ndarray
    group_velocity: np.ndarray
    frequency: np.ndarray = np.zeros(0)
    thermal_conductivity: np.ndarray = np.zeros(0)


class BoundaryConditions(NamedTuple):
    """Boundary conditions (openbte/objects.py:220-241).

    ``diffuse``: name of the diffuse-wall side region; ``mixed``:
    ``{region: {'value': T_wall, 'boundary_conductance': h}}``; ``periodic``:
    ``{region: jump}``.
    """

    diffuse: str = ''
    mixed: dict = {}
    periodic: dict = {}
    is_nanowire: bool = False

    def get_bc_by_name(self, name: str) -> str:
        """Mirror of openbte/objects.py:232-241, including its quirk that any
        non-mixed name resolves to 'Periodic' once a periodic BC exists."""
        if name in self.mixed.keys():
            return 'Mixed'
        for _ in self.periodic.keys():
            return 'Periodic'
        raise ValueError("{} failed (no region associated to side).".format(name))


class EffectiveThermalConductivity(NamedTuple):
    """Contact used to integrate the heat flux (openbte/objects.py:128-135).

    ``normalization`` is alpha = -Lx/Ly/DeltaT (docs/effective_kappa.rst:10).
    """

    contact: str
    normalization: float = 0.0


class SolverResults(NamedTuple):
    """Solver output (openbte/objects.py:493-497)."""

    kappa_eff: float = 0.0
    aux: dict = {}
    variables: dict = {}
