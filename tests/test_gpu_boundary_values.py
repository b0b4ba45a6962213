"""Face and flux Dirichlet values on the domain boundary (cfg.bc['cell_faces']
and cfg.bc['intercell_fluxes'], boundary/boundary.py:89-461) through the GPU
RHS, against the fp32 oracle (relative L-infinity <= 1e-5)."""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd.boundary import bc_types
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import grid_cfg, interior, oracle_cfg, random_state, rel_linf

pytestmark = pytest.mark.gpu
BC = bc_types.BoundaryCondition
D, NONE = BC.DIRICHLET, (None, None)
NAMES = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')
TOL = 1e-5


def _cell_bcs():
  wall = {'rho': 1.0, 'rho_u': 0.0, 'rho_v': 0.05, 'rho_w': 0.0, 'rho_e': 6.0}
  return {k: {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.DIRICHLET, wall[k])},
              'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
              'z': {0: (BC.DIRICHLET, wall[k]), 1: (BC.NEUMANN, 0.0)}} for k in NAMES}


def _cfg(n, scheme, flux, visc, faces=None, fluxes=None):
  cfg = grid_cfg(n, 3)
  bc = {'cell_averages': _cell_bcs()}
  if faces is not None:
    bc['cell_faces'] = faces
  if fluxes is not None:
    bc['intercell_fluxes'] = fluxes
  cfg.update({'interpolation_scheme': scheme, 'numeric_flux_scheme': flux,
              'include_diffusion': visc, 'transport_parameters': {'nu': 2e-3, 'pr': 0.7},
              'dt': 2e-3, 'bc': bc})
  return cfg


def _flux_spec():
  return {
      'convective': {
          'rho': {'x': {0: (D, 0.0), 1: NONE}, 'y': {0: NONE, 1: NONE},
                  'z': {0: (D, 0.1), 1: (D, -0.2)}},
          'rho_w': {'x': {0: NONE, 1: NONE}, 'y': {0: NONE, 1: NONE},
                    'z': {0: (D, 1.0), 1: NONE}},
      },
      'diffusive': {
          'rho_e': {'x': {0: (D, 0.5), 1: (D, -0.5)}, 'y': {0: NONE, 1: NONE},
                    'z': {0: NONE, 1: (D, 0.25)}},
      },
      'total': {
          'rho_u': {'x': {0: NONE, 1: (D, 0.3)}, 'y': {0: (D, 0.0), 1: NONE},
                    'z': {0: NONE, 1: NONE}},
      },
  }


def _face_spec():
  spec = {}
  spec['rho_u'] = {'x': {0: (D, 0.0), 1: (D, 0.0)}, 'y': {0: NONE, 1: NONE},
                   'z': {0: NONE, 1: NONE}}
  spec['v'] = {'x': {0: NONE, 1: NONE}, 'y': {0: (D, 0.2), 1: NONE},
               'z': {0: (D, -0.1), 1: NONE}}
  spec['p'] = {'x': {0: (D, 1.5), 1: NONE}, 'y': {0: NONE, 1: NONE},
               'z': {0: NONE, 1: (D, 0.8)}}
  spec['rho_e'] = {'x': {0: NONE, 1: (D, 4.0)}, 'y': {0: NONE, 1: NONE},
                   'z': {0: NONE, 1: NONE}}
  return spec


def _check(cfg, seed, tag):
  n, h = (cfg['nx'], cfg['ny'], cfg['nz']), 3
  st = random_state(n, h, seed=seed)
  ref = O.rhs(st, {}, oracle_cfg(cfg))
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  out = fn(0, {k: torch.from_numpy(v.copy()).cuda() for k, v in st.items()}, {})
  torch.cuda.synchronize()
  errs = {k: rel_linf(interior(out[k].cpu().numpy(), h), interior(ref[k], h)) for k in NAMES}
  print(f'[bvalues] {tag}: ' + ' '.join(f'{k}={v:.1e}' for k, v in errs.items()), flush=True)
  assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize('scheme,flux,visc', [('WENO_5', 'HLLC', True), ('MUSCL', 'HLL', True),
                                              ('WENO_3', 'HLLC', False)])
def test_flux_values_match_oracle(scheme, flux, visc):
  cfg = _cfg((12, 10, 9), scheme, flux, visc, fluxes=_flux_spec())
  _check(cfg, 31, f'fluxes {scheme}/{flux}/visc={visc}')


@pytest.mark.parametrize('scheme,flux', [('WENO_5', 'HLLC'), ('WENO_5', 'HLL'), ('MUSCL', 'HLLC')])
def test_face_values_match_oracle(scheme, flux):
  cfg = _cfg((12, 10, 9), scheme, flux, False, faces=_face_spec(), fluxes=_flux_spec())
  _check(cfg, 32, f'faces {scheme}/{flux}')


def _planes(spec, n, h, seed):
  """spec with every DIRICHLET value v replaced by a plane over the face,
  v * (1 + 0.1 U(-1, 1)) + 0.05 U(-1, 1), [A][B] over the full extents of the
  two other axes: x faces [nz+2h][ny+2h], y faces [nz+2h][nx+2h], z faces
  [nx+2h][ny+2h] (the slice shape of a [z][x][y] field)."""
  rng = np.random.default_rng(seed)
  nx, ny, nz = (m + 2 * h for m in n)
  shape = {'x': (nz, ny), 'y': (nz, nx), 'z': (nx, ny)}
  out = {}
  for var, dims in spec.items():
    out[var] = {}
    for dim, sides in dims.items():
      out[var][dim] = {}
      for side, entry in sides.items():
        if entry[0] is None:
          out[var][dim][side] = entry
          continue
        v = float(entry[1])
        plane = (v * (1 + 0.1 * rng.uniform(-1, 1, shape[dim]))
                 + 0.05 * rng.uniform(-1, 1, shape[dim]))
        out[var][dim][side] = (entry[0], plane.astype(np.float32))
  return out


def _plane_fluxes(n, h, seed):
  return {t: _planes(s, n, h, seed + k) for k, (t, s) in enumerate(_flux_spec().items())}


@pytest.mark.parametrize('scheme,flux,visc', [('WENO_5', 'HLLC', True), ('MUSCL', 'HLL', False)])
def test_plane_flux_values_match_oracle(scheme, flux, visc):
  n = (12, 10, 9)
  cfg = _cfg(n, scheme, flux, visc, fluxes=_plane_fluxes(n, 3, 40))
  _check(cfg, 33, f'plane fluxes {scheme}/{flux}/visc={visc}')


@pytest.mark.parametrize('scheme,flux', [('WENO_5', 'HLLC'), ('WENO_3', 'HLL')])
def test_plane_face_values_match_oracle(scheme, flux):
  """Face planes of a conserved (rho_u, rho_e) and primitive (v, p) value,
  mixed with scalar flux values."""
  n = (12, 10, 9)
  faces = _planes(_face_spec(), n, 3, 50)
  cfg = _cfg(n, scheme, flux, False, faces=faces, fluxes=_flux_spec())
  _check(cfg, 34, f'plane faces {scheme}/{flux}')


def test_face_value_errors_follow_the_reference():
  faces = _face_spec()
  with pytest.raises(NotImplementedError):
    cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(
        _cfg((8, 8, 8), 'WENO_5', 'HLLC', True, faces=faces))))
  # A density face value drops p and h from the face primitives
  # (boundary.py:284-293): a 'p' value on the same face is then rejected
  # (ValueError), and without one the flux step misses the pressure (KeyError).
  rho_x0 = {'x': {0: (D, 1.0), 1: NONE}, 'y': {0: NONE, 1: NONE}, 'z': {0: NONE, 1: NONE}}
  for spec, exc in ((dict(faces, rho=rho_x0), ValueError),
                    ({'rho': rho_x0, 'rho_u': faces['rho_u']}, KeyError)):
    cfg = _cfg((8, 8, 8), 'WENO_5', 'HLLC', False, faces=spec)
    with pytest.raises(exc):
      cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
    with pytest.raises(exc):
      O.rhs(random_state((8, 8, 8), 3, seed=1), {}, oracle_cfg(cfg))
