"""MI355X-native steady-state mode-resolved phonon BTE solver."""
from .objects import (Mesh, Material, BoundaryConditions,
                      EffectiveThermalConductivity, SolverResults)
