"""Simulation configuration object with the reference's option keys.

`SwirlCParameters(user_cfg)` accepts the same dictionary as the reference's
core/parameter.py:213-248: defaults from `_DEFAULT_OPTIONS` (:44-131)
overridden by the user's keys, every key set as an attribute.  The TPU
stencil options (`kernel_op_type`, `kernel_size`) are accepted and ignored:
they select numerically equivalent TPU strategies (numerics/kernel_op_types.py).
"""

import copy
import os
from typing import Any, Dict, Optional

import numpy as np
import yaml

from swirlc_amd.boundary import bc_types
from swirlc_amd.common import types


def _get_cell_center_coordinates(start: float, stop: float, n_cells: int) -> np.ndarray:
  """Cell-centre coordinates of n_cells uniform cells on [start, stop]."""
  xx = np.linspace(start, stop, n_cells + 1, dtype=types.NP_DTYPE)
  return xx[:-1] + 0.5 * np.diff(xx)[0]


_XX = _get_cell_center_coordinates(0.0, 1.0, 10)
_DX = np.diff(_XX)[0]

_DEFAULT_OPTIONS = {
    'apply_preprocess': False,
    'apply_postprocess': False,
    'states_to_file': None,
    'states_from_file': None,
    'cx': 1,
    'cy': 1,
    'cz': 1,
    'use_3d_tf_tensor': False,
    'halo_width': 3,
    'nx': 10,
    'ny': 10,
    'nz': 10,
    'core_nx': 10,
    'core_ny': 10,
    'core_nz': 10,
    'lx': 1.0,
    'ly': 1.0,
    'lz': 1.0,
    'x': _XX,
    'dx': _DX,
    'y': _XX,
    'dy': _DX,
    'z': _XX,
    'dz': _DX,
    'g': {'x': 0.0, 'y': 0.0, 'z': 0.0},
    'start_step': 0,
    'num_steps': 50,
    'num_cycles': 5,
    'dt': 1.0e-3,
    'solver': 'compressible_navier_stokes',
    'interpolation_scheme': 'WENO_5',
    'numeric_flux_scheme': 'HLL',
    'include_diffusion': False,
    'time_integration_scheme': 'rk3',
    'source_functions': None,
    'kernel_op_type': 'KERNEL_OP_SLICE',
    'kernel_size': 8,
    'thermodynamics_model': 'generic',
    'p_0': 101325.0,
    'transport_model': 'simple',
    'transport_parameters': {'nu': 1.0e-4, 'pr': 0.7},
    'nu': 1e-1,
    'nu_t': None,
    'conservative_variable_names': list(types.BASE_CONSERVATIVE),
    'primitive_variable_names': list(types.BASE_PRIMITIVES),
    'additional_state_names': [],
    'helper_variable_names': [],
    'debug_variables': [],
    'bc': {
        'cell_averages': {
            var_name: {
                dim: {
                    0: (bc_types.BoundaryCondition.PERIODIC,),
                    1: (bc_types.BoundaryCondition.PERIODIC,),
                }
                for dim in types.DIMS
            }
            for var_name in types.BASE_CONSERVATIVE
        }
    },
    # Extension (no reference counterpart): Courant number for a
    # device-computed time step; None keeps the fixed `dt`.
    'cfl': None,
}


class SwirlCParameters:
  """The simulation context object (core/parameter.py:134-255)."""

  def __init__(self, user_cfg: Optional[Dict[str, Any]] = None):
    self._simulation_cfg = copy.deepcopy(_DEFAULT_OPTIONS)
    if user_cfg:
      for key, val in user_cfg.items():
        self._simulation_cfg[key] = val
    for key, val in self._simulation_cfg.items():
      setattr(self, key, val)

  def save_to_file(self, prefix: str) -> None:
    """Saves the configuration dictionary as YAML (core/parameter.py:250-255)."""
    output_dir, _ = os.path.split(prefix)
    if output_dir:
      os.makedirs(output_dir, exist_ok=True)
    with open(f'{prefix}_swirl_c_cfg.yml', 'w') as f:
      yaml.dump(_yaml_safe(self._simulation_cfg), f, default_flow_style=False)


def _yaml_safe(obj):
  if isinstance(obj, dict):
    return {k: _yaml_safe(v) for k, v in obj.items()}
  if isinstance(obj, (list, tuple)):
    return [_yaml_safe(v) for v in obj]
  if isinstance(obj, np.ndarray):
    return obj.tolist()
  if isinstance(obj, np.generic):
    return obj.item()
  if isinstance(obj, bc_types.BoundaryCondition):
    return obj.name
  if hasattr(obj, 'detach'):
    return obj.detach().cpu().tolist()
  return obj
