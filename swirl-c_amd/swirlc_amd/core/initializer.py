"""Initial-state helpers with the reference's interface (core/initializer.py).

Host utilities (numpy) that build the conservative fields handed to
`core.driver.solver` / `core.solver.Solver`:

  default_initializer(replica_id, coordinates, cfg)   all fields zero
  partial_field_for_core(cfg, coordinate, value_fn)   one field of one core
  constant_initial_state_fn(value)                    value functions
  step_function_initial_state_fn(left, right, loc, dim)

A value function takes the local cell-centre mesh (x, y, z order, `ij`
indexing), the domain lengths and the core coordinate, and returns the
interior values; `partial_field_for_core` pads the halo (mode 'SYMMETRIC' by
default, like tf.pad) and permutes to the [z, x, y] field layout.  Arrays are
float32 (the reference's types.DTYPE).
"""

from typing import Callable, Optional, Sequence

import numpy as np

from swirlc_amd.common import types

ValueFunction = Callable[..., np.ndarray]

_DEFAULT_PERMUTATION = (2, 0, 1)
# tf.pad modes -> numpy.pad modes.
_PAD_MODES = {'CONSTANT': 'constant', 'REFLECT': 'reflect', 'SYMMETRIC': 'symmetric'}


def default_initializer(replica_id, coordinates, cfg):
  """Every variable of cfg.conservative_variable_names, zero on the local
  mesh halos included (core/initializer.py:41-71): the state a restart then
  overwrites from a save file."""
  del replica_id
  return {
      name: partial_field_for_core(cfg, coordinates, constant_initial_state_fn(0.0))
      for name in cfg.conservative_variable_names
  }


def _get_local_1d_mesh(core_n: int, core_id: int, global_1d_mesh) -> np.ndarray:
  """The core's part of a global 1-D mesh (core/initializer.py:226-251): the
  points core_id * core_n .. (core_id + 1) * core_n - 1."""
  idx = np.linspace(core_id * core_n, (core_id + 1) * core_n - 1, core_n).astype(np.int32)
  return np.asarray(global_1d_mesh)[idx]


def partial_field_for_core(cfg, coordinate: Sequence[int], value_fn: ValueFunction,
                           perm: Optional[Sequence[int]] = _DEFAULT_PERMUTATION,
                           pad_mode: Optional[str] = 'SYMMETRIC') -> np.ndarray:
  """One flow field on the local mesh of core `coordinate` (x, y, z core
  indices) (core/initializer.py:74-223).

  The value function sees the core's cell-centre mesh and returns its
  interior values; the halo of width cfg.halo_width is filled by `pad_mode`
  ('CONSTANT' zeros, 'REFLECT', 'SYMMETRIC', or None for no padding) and the
  result is transposed by `perm` ((2, 0, 1): [x, y, z] -> [z, x, y]; None:
  no transpose).  Raises ValueError for an out-of-range core coordinate or a
  local mesh smaller than the halo width.
  """
  core_n = (cfg.core_nx, cfg.core_ny, cfg.core_nz)
  counts = (cfg.cx, cfg.cy, cfg.cz)
  for name, g, c in zip('xyz', coordinate, counts):
    if g < 0:
      raise ValueError(f'Invalid subgrid coordinate specified with negative {name} core index.')
    if g >= c:
      raise ValueError(f'Invalid subgrid coordinate specified with {name} core index. Must be '
                       f'smaller than total number of core partitioning in {name} direction.')
  for name, n in zip('xyz', core_n):
    if n < cfg.halo_width:
      raise ValueError(f'Local mesh size `core_n{name} = {n}` must be at least the halo '
                       f'width `cfg.halo_width = {cfg.halo_width}`.')
  xs = _get_local_1d_mesh(cfg.core_nx, coordinate[0], cfg.x)
  ys = _get_local_1d_mesh(cfg.core_ny, coordinate[1], cfg.y)
  zs = _get_local_1d_mesh(cfg.core_nz, coordinate[2], cfg.z)
  xx, yy, zz = np.meshgrid(xs, ys, zs, indexing='ij')
  dt = types.NP_DTYPE
  val = np.asarray(value_fn(xx, yy, zz, dt(cfg.lx), dt(cfg.ly), dt(cfg.lz), coordinate))
  if pad_mode:
    if pad_mode not in _PAD_MODES:
      raise ValueError(f'pad_mode must be one of {sorted(_PAD_MODES)} or None, got {pad_mode}')
    h = cfg.halo_width
    val = np.pad(val, [(h, h)] * 3, mode=_PAD_MODES[pad_mode])
  if perm:
    val = np.transpose(val, perm)
  return np.ascontiguousarray(val, dtype=types.NP_DTYPE)


def constant_initial_state_fn(value: float) -> ValueFunction:
  """Value function of a constant field (core/initializer.py:255-274)."""

  def init_fn(xx, yy, zz, lx, ly, lz, coord):
    del yy, zz, lx, ly, lz, coord
    return value * np.ones_like(xx)

  return init_fn


def step_function_initial_state_fn(left_value: float, right_value: float, step_loc: float,
                                   dim: str) -> ValueFunction:
  """Value function of a step along `dim` (core/initializer.py:277-322):
  `left_value` where |coordinate / length| <= step_loc, else `right_value`."""

  def init_fn(xx, yy, zz, lx, ly, lz, coord):
    del coord
    if dim == 'x':
      l, v = lx, xx
    elif dim == 'y':
      l, v = ly, yy
    else:
      l, v = lz, zz
    return np.where(np.abs(v / l) <= step_loc, left_value * np.ones_like(v),
                    right_value * np.ones_like(v))

  return init_fn
