"""Oracle pins for the face and flux Dirichlet values (cell_faces,
intercell_fluxes): the expectations of the reference's own tests
boundary/boundary_test.py:736-912 (face update on one plane) and :1764-1922
(flux update, every flux type), restated with numpy inputs."""

import numpy as np
import pytest

from oracle import swirlc_oracle as O

SIZE = (16, 16, 16)  # [nz, nx, ny]
AX = {'x': 1, 'y': 2, 'z': 0}


def _tile(f, dim):
  """testing_utils.to_3d_tensor: a 1-D profile along `dim`, tiled."""
  shape = [1, 1, 1]
  shape[AX[dim]] = len(f)
  reps = list(SIZE)
  reps[AX[dim]] = 1
  return np.tile(f.reshape(shape), reps).astype(np.float32)


def _profiles():
  rho = np.linspace(1.0, 2.0, 16)
  u = np.linspace(-1.0, 2.0, 16)
  v = np.linspace(0.0, 5.0, 16)
  w = -np.linspace(0.0, 6.0, 16)
  t = np.linspace(250.0, 500.0, 16)
  e = O.CP * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  return rho, u, v, w, e


@pytest.mark.parametrize('dim', ['x', 'y', 'z'])
@pytest.mark.parametrize('side', [0, 1])
def test_face_update_matches_reference_case(dim, side):
  rho, u, v, w, e = _profiles()
  cons = {'rho': _tile(rho, 'x'), 'rho_u': _tile(rho * u, 'x'), 'rho_v': _tile(rho * v, 'x'),
          'rho_w': _tile(rho * w, 'x'), 'rho_e': _tile(rho * e, 'x')}
  prim = {'rho': _tile(rho, 'x'), 'u': _tile(u, 'x'), 'v': _tile(v, 'x'), 'w': _tile(w, 'x'),
          'e': _tile(e, 'x')}
  spec = {'rho': 1.5, 'u': 2.0, 'rho_e': 12.0, 'rho_v': 6.0, 'rho_w': 9.0}
  loc = 3 if side == 0 else 13
  c, p = O._face_boundary_update_1d(dict(cons), dict(prim), spec, dim, loc)
  assert list(c) == list(cons) and list(p) == list(prim)
  sel = [slice(None)] * 3
  sel[AX[dim]] = loc
  sel = tuple(sel)
  expect_c = {'rho': 1.5, 'rho_u': 3.0, 'rho_v': 6.0, 'rho_w': 9.0, 'rho_e': 12.0}
  expect_p = {'rho': 1.5, 'u': 2.0, 'v': 4.0, 'w': 6.0, 'e': 8.0}
  for k, val in expect_c.items():
    ref = cons[k].copy()
    ref[sel] = val
    np.testing.assert_allclose(c[k], ref, rtol=1e-6, err_msg=k)
  for k, val in expect_p.items():
    ref = prim[k].copy()
    ref[sel] = val
    np.testing.assert_allclose(p[k], ref, rtol=2e-6, atol=1e-6, err_msg=k)


def test_face_update_with_density_drops_pressure_and_enthalpy():
  """The density value replaces the whole primitive dict (boundary.py:284-293)."""
  rho, u, v, w, e = _profiles()
  cons = {'rho': _tile(rho, 'x'), 'rho_u': _tile(rho * u, 'x'), 'rho_v': _tile(rho * v, 'x'),
          'rho_w': _tile(rho * w, 'x'), 'rho_e': _tile(rho * e, 'x')}
  prim = O.conservative_to_primitive(cons)
  prim['p'] = np.ones_like(prim['rho'])
  prim['h'] = np.ones_like(prim['rho'])
  _, p = O._face_boundary_update_1d(dict(cons), dict(prim), {'rho': 1.5}, 'x', 3)
  assert 'p' not in p and 'h' not in p
  _, p = O._face_boundary_update_1d(dict(cons), dict(prim), {'u': 1.5, 'p': 2.0}, 'x', 3)
  assert np.all(p['p'][:, 3, :] == 2.0) and np.all(p['p'][:, 4, :] == 1.0)


@pytest.mark.parametrize('flux_type', ['convective', 'diffusive', 'total'])
def test_update_fluxes_matches_reference_case(flux_type):
  rho, u, v, w, e = _profiles()
  fluxes = {d: {'rho': _tile(rho, d), 'rho_u': _tile(rho * u, d), 'rho_v': _tile(rho * v, d),
                'rho_w': _tile(rho * w, d), 'rho_e': _tile(rho * e, d)} for d in 'xyz'}
  D, N = O.DIRICHLET, (None, None)
  spec = {
      'rho': {'x': {0: (D, 1.0), 1: (D, 2.0)}, 'y': {0: N, 1: N},
              'z': {0: (D, 2.0), 1: (D, 1.0)}},
      'rho_u': {d: {0: (D, 0.0), 1: (D, 0.0)} for d in 'xyz'},
      'rho_v': {d: {0: N, 1: N} for d in 'xyz'},
      'rho_w': {'x': {0: (D, 4.0), 1: N}, 'y': {0: (D, 6.0), 1: N},
                'z': {0: (D, 8.0), 1: N}},
      'rho_e': {'x': {0: N, 1: (D, 3.0)}, 'y': {0: N, 1: (D, 5.0)},
                'z': {0: N, 1: (D, 7.0)}},
  }
  cfg = {'halo_width': 3, 'core_nx': 10, 'core_ny': 10, 'core_nz': 10,
         'conservative_variable_names': list(O.BASE_CONSERVATIVE),
         'bc': {'intercell_fluxes': {flux_type: spec}}}
  out = O.update_fluxes({d: dict(f) for d, f in fluxes.items()}, cfg, flux_type)
  for d in 'xyz':
    for k in O.BASE_CONSERVATIVE:
      ref = fluxes[d][k].copy()
      for side, plane in ((0, 3), (1, 13)):
        t, val = spec[k][d][side]
        if t is not None:
          sel = [slice(None)] * 3
          sel[AX[d]] = plane
          ref[tuple(sel)] = val
      np.testing.assert_array_equal(out[d][k], ref, err_msg=f'{d} {k}')
  # Only conserved variables take flux values.
  bad = {'intercell_fluxes': {flux_type: {'u': spec['rho_u']}}}
  with pytest.raises(AssertionError):
    O.update_fluxes(fluxes, dict(cfg, bc=bad), flux_type)
