"""Boundary-condition enums (same members as boundary/bc_types.py:22-47)."""

import enum


class BoundaryCondition(enum.Enum):
  PERIODIC = 0
  DIRICHLET = 1
  NEUMANN = 2
  REFLECTIVE = 3


class BoundaryFluxType(enum.Enum):
  CONVECTIVE = 'convective'
  DIFFUSIVE = 'diffusive'
  TOTAL = 'total'


class BoundaryDictionaryType(enum.Enum):
  CELL_AVERAGES = 'cell_averages'
  CELL_FACES = 'cell_faces'
  INTERCELL_FLUXES = 'intercell_fluxes'
