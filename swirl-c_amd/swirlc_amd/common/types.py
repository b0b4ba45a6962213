"""Variable names and dtypes of the flow-field dictionaries.

Same names and values as the reference's common/types.py:24-142 (they are the
dictionary keys of the drop-in surface).
"""

import re

import numpy as np
import torch

DTYPE = torch.float32
NP_DTYPE = np.float32
SMALL = float(np.finfo(np.float32).resolution)

DIMS = ('x', 'y', 'z')
RHO = 'rho'
U, V, W = 'u', 'v', 'w'
VELOCITY = (U, V, W)
E = 'e'
BASE_PRIMITIVES = (RHO, U, V, W, E)
CONS_VAR_PREFIX = 'rho_'
RHO_U, RHO_V, RHO_W = 'rho_u', 'rho_v', 'rho_w'
MOMENTUM = (RHO_U, RHO_V, RHO_W)
RHO_E = 'rho_e'
BASE_CONSERVATIVE = (RHO, RHO_U, RHO_V, RHO_W, RHO_E)
P, T, POTENTIAL_T, H = 'p', 't', 'theta', 'h'


def primitive_to_conservative_name(primitive: str) -> str:
  """common/types.py:88-105."""
  if not primitive:
    raise ValueError('Primitive variable name must not be empty.')
  return RHO if primitive == RHO else CONS_VAR_PREFIX + primitive


def is_conservative_name(varname: str) -> bool:
  """common/types.py:133-142."""
  return bool(re.match(f'^({CONS_VAR_PREFIX}).', varname))


def conservative_to_primitive_name(conservative: str) -> str:
  """common/types.py:108-130."""
  if conservative == RHO:
    return RHO
  if is_conservative_name(conservative):
    return conservative[len(CONS_VAR_PREFIX):]
  raise ValueError(
      f'Invalid conserved variable name: {conservative}. Must start with'
      f' "{CONS_VAR_PREFIX}".')
