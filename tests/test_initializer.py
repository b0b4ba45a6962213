"""Host initializers against the expectations of the reference's
core/initializer_test.py (default zeros, symmetric halo padding and the
[z, x, y] permutation on one and several cores, local 1-D meshes, constant
and step value functions)."""

import numpy as np
import pytest

from swirlc_amd.common import types
from swirlc_amd.core import initializer
from swirlc_amd.core.parameter import SwirlCParameters

DIMS = ('x', 'y', 'z')


def _to_3d(v, dim, size):
  """A 1-D vector broadcast along `dim` to a [nz, nx, ny] array."""
  nx, ny, nz = size
  v = np.asarray(v)
  if dim == 'x':
    return np.broadcast_to(v[np.newaxis, :, np.newaxis], (nz, nx, ny)).copy()
  if dim == 'y':
    return np.broadcast_to(v[np.newaxis, np.newaxis, :], (nz, nx, ny)).copy()
  return np.broadcast_to(v[:, np.newaxis, np.newaxis], (nz, nx, ny)).copy()


def _symmetric_halo(expected):
  """The reference test's halo fill: mirror the 3 interior layers outward."""
  for i in range(3):
    expected[3:-3, 3 - i - 1, 3:-3] = expected[3:-3, 3 + i, 3:-3]
    expected[3:-3, -(3 - i), 3:-3] = expected[3:-3, -(3 + 1 + i), 3:-3]
  for j in range(3):
    expected[3:-3, :, 3 - j - 1] = expected[3:-3, :, 3 + j]
    expected[3:-3, :, -(3 - j)] = expected[3:-3, :, -(3 + 1 + j)]
  for k in range(3):
    expected[3 - k - 1, :, :] = expected[3 + k, :, :]
    expected[-(3 - k), :, :] = expected[-(3 + 1 + k), :, :]
  return expected


def _coord_fn(dim):
  def init_fn(xx, yy, zz, lx, ly, lz, coord):
    del lx, ly, lz, coord
    return {'x': xx, 'y': yy, 'z': zz}[dim]
  return init_fn


def test_default_initializer_returns_zeros_one_replica():
  cfg = SwirlCParameters()
  cfg.conservative_variable_names = list(types.BASE_CONSERVATIVE) + ['rho_y']
  out = initializer.default_initializer(0, [0, 0, 0], cfg)
  assert sorted(out) == sorted(cfg.conservative_variable_names)
  for v in out.values():
    assert v.shape == (16, 16, 16) and v.dtype == np.float32 and not v.any()


def test_default_initializer_returns_zeros_four_replicas():
  cfg = SwirlCParameters()
  cfg.conservative_variable_names = list(types.BASE_CONSERVATIVE) + ['rho_y']
  cfg.cx, cfg.cy, cfg.cz = 2, 2, 1
  cfg.core_nx, cfg.core_ny, cfg.core_nz = 5, 5, 10
  for ic in range(2):
    for jc in range(2):
      out = initializer.default_initializer(0, [ic, jc, 0], cfg)
      for v in out.values():
        assert v.shape == (16, 11, 11) and not v.any()


@pytest.mark.parametrize('dim', DIMS)
def test_partial_field_for_core_one_replica(dim):
  cfg = SwirlCParameters()
  got = initializer.partial_field_for_core(cfg, [0, 0, 0], _coord_fn(dim))
  expected = np.zeros([16, 16, 16])
  expected[3:-3, 3:-3, 3:-3] = _to_3d(getattr(cfg, dim), dim, [10, 10, 10])
  np.testing.assert_allclose(got, _symmetric_halo(expected), rtol=1e-6)


@pytest.mark.parametrize('dim', DIMS)
@pytest.mark.parametrize('comp_shape_dim', DIMS)
def test_partial_field_for_core_several_replicas(dim, comp_shape_dim):
  cfg = SwirlCParameters()
  cfg.cx = cfg.cy = cfg.cz = 2
  cfg.core_nx = cfg.core_ny = cfg.core_nz = 5
  setattr(cfg, 'c' + comp_shape_dim, 1)
  setattr(cfg, 'core_n' + comp_shape_dim, 10)
  interior = _to_3d(getattr(cfg, dim), dim, [10, 10, 10])
  for ic in range(cfg.cx):
    for jc in range(cfg.cy):
      for kc in range(cfg.cz):
        got = initializer.partial_field_for_core(cfg, [ic, jc, kc], _coord_fn(dim))
        shape = [cfg.core_nz + 6, cfg.core_nx + 6, cfg.core_ny + 6]
        expected = np.zeros(shape)
        i0 = 0 if cfg.cx == 1 else 5 * ic
        j0 = 0 if cfg.cy == 1 else 5 * jc
        k0 = 0 if cfg.cz == 1 else 5 * kc
        expected[3:-3, 3:-3, 3:-3] = interior[k0:k0 + cfg.core_nz, i0:i0 + cfg.core_nx,
                                              j0:j0 + cfg.core_ny]
        np.testing.assert_allclose(got, _symmetric_halo(expected), rtol=1e-6)


def test_partial_field_for_core_rejects_bad_coordinates():
  cfg = SwirlCParameters()
  with pytest.raises(ValueError):
    initializer.partial_field_for_core(cfg, [1, 0, 0], _coord_fn('x'))
  with pytest.raises(ValueError):
    initializer.partial_field_for_core(cfg, [0, -1, 0], _coord_fn('x'))
  cfg.core_nz = 2
  with pytest.raises(ValueError):
    initializer.partial_field_for_core(cfg, [0, 0, 0], _coord_fn('x'))


def test_partial_field_for_core_without_padding_or_permutation():
  cfg = SwirlCParameters()
  got = initializer.partial_field_for_core(cfg, [0, 0, 0], _coord_fn('y'), perm=None,
                                           pad_mode=None)
  assert got.shape == (10, 10, 10)
  np.testing.assert_allclose(got[0, :, 0], cfg.y, rtol=1e-6)
  got = initializer.partial_field_for_core(cfg, [0, 0, 0], _coord_fn('y'),
                                           pad_mode='CONSTANT')
  assert got.shape == (16, 16, 16) and got[0].max() == 0.0


def test_get_local_1d_mesh_returns_correct_mesh():
  xx = np.linspace(0.0, 1.0, 64)
  for core_id in range(4):
    np.testing.assert_allclose(initializer._get_local_1d_mesh(16, core_id, xx),
                               xx[16 * core_id:16 * (core_id + 1)])


def test_constant_initial_state_fn_evaluates_correctly():
  s = np.linspace(0.0, 1.0, 16)
  xx, yy, zz = np.meshgrid(s, s, s, indexing='ij')
  got = initializer.constant_initial_state_fn(1.45)(xx, yy, zz, 1.0, 1.0, 1.0, [0, 0, 0])
  np.testing.assert_allclose(got, 1.45 * np.ones([16, 16, 16]))


@pytest.mark.parametrize('dim', DIMS)
def test_step_function_initial_state_fn_evaluates_correctly(dim):
  s = np.linspace(0.25, 0.75, 16)
  xx, yy, zz = np.meshgrid(s, s, s, indexing='ij')
  got = initializer.step_function_initial_state_fn(1.45, 2.5, 0.4, dim)(
      xx, yy, zz, 1.0, 1.0, 1.0, [0, 0, 0])
  expected = 1.45 * np.ones([16, 16, 16])
  if dim == 'x':
    expected[5:, :, :] = 2.5
  elif dim == 'y':
    expected[:, 5:, :] = 2.5
  else:
    expected[:, :, 5:] = 2.5
  np.testing.assert_allclose(got, expected)
