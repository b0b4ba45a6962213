"""GPU parity of the fused RHS / RK kernels against the fp32 oracle.

Bar: per-field relative L-infinity error max|gpu - oracle| / max|oracle| of
at most 1e-5 over interior cells (fp32 arithmetic on both sides), identical
HLLC branch codes, conservation to roundoff on periodic grids.  Every case
goes through the C ABI (libswirlc_amd via swirlc_amd._native).
"""

import copy
import sys

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd.boundary import bc_types
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.equations import compressible_navier_stokes as cns
from swirlc_amd.numerics.time_integration import rk2, rk3

from helpers import grid_cfg, interior, oracle_cfg, random_state, rel_linf

pytestmark = pytest.mark.gpu
BC = bc_types.BoundaryCondition
TOL = 1e-5


def _bc_all(kind, value=0.0):
  entry = (kind,) if kind == BC.PERIODIC else (kind, value)
  return {d: {0: entry, 1: entry} for d in ('x', 'y', 'z')}


def _bcs(name, n, h):
  names = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')
  if name == 'periodic':
    return {'cell_averages': {v: _bc_all(BC.PERIODIC) for v in names}}
  if name == 'neumann':
    return {'cell_averages': {v: _bc_all(BC.NEUMANN, 0.01 * k)
                              for k, v in enumerate(names)}}
  if name == 'mixed':
    nx, ny, nz = n
    plane_z = [np.full((nx + 2 * h, ny + 2 * h), 1.0 + 0.1 * k, np.float32) for k in range(h)]
    return {'cell_averages': {
        'rho': {'x': {0: (BC.DIRICHLET, 1.2), 1: (BC.NEUMANN, 0.0)},
                'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                'z': {0: (BC.DIRICHLET, plane_z), 1: (BC.NEUMANN, -0.01)}},
        'u': {'x': {0: (BC.DIRICHLET, 0.3), 1: (BC.NEUMANN, 0.0)},
              'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
              'z': {0: (BC.DIRICHLET, [0.1, 0.2, 0.3][:h]), 1: (BC.NEUMANN, 0.0)}},
        'rho_v': {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.DIRICHLET, 0.0)},
                  'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                  'z': {0: (BC.NEUMANN, 0.0), 1: (BC.DIRICHLET, -0.1)}},
        'rho_w': {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.NEUMANN, 0.0)},
                  'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                  'z': {0: (BC.DIRICHLET, 0.0), 1: (BC.DIRICHLET, 0.0)}},
        'e': {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.DIRICHLET, 6.0)},
              'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
              'z': {0: (BC.NEUMANN, 0.05), 1: (BC.NEUMANN, 0.0)}},
    }}
  raise ValueError(name)


def _cfg(n, h, scheme, flux, visc, bcs, gravity=False):
  # Under gravity the column is 0.05 high so the potential G z stays below
  # the specific internal energy of the test states (positive pressure).
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05) if gravity else (1.0, 1.0, 1.0))
  cfg.update({'interpolation_scheme': scheme, 'numeric_flux_scheme': flux,
              'include_diffusion': visc, 'bc': _bcs(bcs, n, h),
              'transport_parameters': {'nu': 2e-3, 'pr': 0.7}, 'dt': 2e-3})
  if gravity:
    cfg.update({'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity']})
  return cfg


def _to_gpu(state):
  return {k: torch.from_numpy(v.copy()).cuda() for k, v in state.items()}


def _gpu_rhs(cfg, state):
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  out = fn(0, _to_gpu(state), {})
  torch.cuda.synchronize()
  return {k: v.cpu().numpy() for k, v in out.items()}


def _report(tag, errs):
  print(f'[parity] {tag}: ' + ' '.join(f'{k}={v:.2e}' for k, v in errs.items()), flush=True)


def test_max_wave_speed_matches_oracle():
  n, h = (13, 11, 9), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', False, 'periodic')
  st = random_state(n, h, seed=8)
  ref = O.max_wave_speed(st, oracle_cfg(cfg))
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  ctx = fn.native_context
  q = ctx.fields(_to_gpu(st))
  out = torch.zeros(4, device='cuda')
  dt = torch.zeros(1, device='cuda')
  ctx.max_wave_speed(q, out, 0.5, dt)
  torch.cuda.synchronize()
  got = float(out[0])
  assert abs(got - float(ref)) <= 1e-6 * float(ref)
  hmin = min(cfg['dx'], cfg['dy'], cfg['dz'])
  assert abs(float(dt[0]) - 0.5 * hmin / float(ref)) <= 1e-5 * float(dt[0])


COMBOS = [(s, f, v) for s in ('WENO_5', 'WENO_3', 'MUSCL') for f in ('HLLC', 'HLL')
          for v in (False, True)]


@pytest.mark.parametrize('scheme,flux,visc', COMBOS)
def test_rhs_matches_oracle_periodic(scheme, flux, visc):
  n, h = (13, 11, 9), 3
  cfg = _cfg(n, h, scheme, flux, visc, 'periodic')
  st = random_state(n, h, seed=1)
  ref = O.rhs(st, {}, oracle_cfg(cfg))
  got = _gpu_rhs(cfg, st)
  errs = {k: rel_linf(interior(got[k], h), interior(ref[k], h)) for k in ref}
  _report(f'{scheme}/{flux}/visc={visc}', errs)
  assert max(errs.values()) <= TOL, errs


# Gravity with BCs on conserved fields only runs the gravity-lean build
# (periodic, neumann); with the mixed set's velocity BCs the run-time one.
@pytest.mark.parametrize('bcs,gravity', [('neumann', False), ('mixed', False),
                                         ('mixed', True), ('periodic', True),
                                         ('neumann', True)])
@pytest.mark.parametrize('scheme', ['WENO_5', 'MUSCL'])
def test_rhs_matches_oracle_boundaries(bcs, gravity, scheme):
  n, h = (12, 10, 11), 3
  cfg = _cfg(n, h, scheme, 'HLLC' if scheme == 'WENO_5' else 'HLL', True, bcs, gravity)
  ocfg = oracle_cfg(cfg)
  pe = O.gravitational_potential(ocfg, np.float64) if gravity else None
  st = random_state(n, h, seed=2, pe=pe)
  ref = O.rhs(st, {}, ocfg)
  assert all(np.isfinite(v).all() for v in ref.values())
  got = _gpu_rhs(cfg, st)
  errs = {k: rel_linf(interior(got[k], h), interior(ref[k], h)) for k in ref}
  _report(f'{scheme}/{bcs}/gravity={gravity}', errs)
  assert max(errs.values()) <= TOL, errs


def test_rhs_multi_chunk_grid():
  """Grid larger than one y chunk (58) and one march segment (32)."""
  n, h = (40, 70, 36), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', True, 'periodic')
  st = random_state(n, h, seed=3)
  ref = O.rhs(st, {}, oracle_cfg(cfg))
  got = _gpu_rhs(cfg, st)
  errs = {k: rel_linf(interior(got[k], h), interior(ref[k], h)) for k in ref}
  _report('multi-chunk', errs)
  assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize('scheme,h,n,bcs', [
    ('WENO_3', 2, (11, 7, 6), 'periodic'),
    ('MUSCL', 2, (9, 61, 5), 'neumann'),
    ('WENO_5', 4, (10, 9, 7), 'neumann'),
    ('WENO_5', 3, (6, 1, 5), 'neumann'),
    ('WENO_5', 3, (5, 130, 4), 'periodic'),
])
def test_rhs_halo_widths_and_row_lengths(scheme, h, n, bcs):
  """Halo widths 2 and 4 (the y sweep's first output sits at position h of
  the plane's run of rows) and rows much shorter or longer than one 64-lane
  window (ny = 1: a window spans many rows; ny = 130: a row spans windows)."""
  cfg = _cfg(n, h, scheme, 'HLLC', True, bcs)
  st = random_state(n, h, seed=9)
  ref = O.rhs(st, {}, oracle_cfg(cfg))
  got = _gpu_rhs(cfg, st)
  errs = {k: rel_linf(interior(got[k], h), interior(ref[k], h)) for k in ref}
  _report(f'{scheme}/h={h}/n={n}', errs)
  assert max(errs.values()) <= TOL, errs


def test_halo_fill_matches_oracle_bitwise():
  n, h = (12, 10, 11), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', False, 'mixed')
  st = random_state(n, h, seed=4)
  ocfg = oracle_cfg(cfg)
  ref = O.update_conservative_cell_averages(st, ocfg)
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  ctx = fn.native_context
  g = _to_gpu(st)
  q = ctx.fields(g)
  ctx.fill_halos(q)
  torch.cuda.synchronize()
  hm = np.ones(ref['rho'].shape, bool)
  hm[h:-h, h:-h, h:-h] = False
  for k in ref:
    np.testing.assert_array_equal(g[k].cpu().numpy()[hm], ref[k][hm], err_msg=k)


@pytest.mark.parametrize('n,h,scheme', [((12, 10, 11), 3, 'WENO_5'), ((4, 4, 4), 4, 'WENO_5'),
                                         ((7, 2, 5), 2, 'MUSCL'), ((70, 3, 6), 3, 'WENO_5')])
def test_periodic_halo_fill_matches_oracle_bitwise(n, h, scheme):
  """All-periodic grids take the one-launch fill (every halo cell copied from
  its interior cell wrapped along all dims); it must equal the reference's
  sequential x, y, z fill bit for bit, edges and corners included."""
  cfg = _cfg(n, h, scheme, 'HLLC', False, 'periodic')
  st = random_state(n, h, seed=6)
  ref = O.update_conservative_cell_averages(st, oracle_cfg(cfg))
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  ctx = fn.native_context
  g = _to_gpu(st)
  ctx.fill_halos(ctx.fields(g))
  torch.cuda.synchronize()
  for k in ref:
    np.testing.assert_array_equal(g[k].cpu().numpy(), ref[k], err_msg=k)


@pytest.mark.parametrize('scheme', ['WENO_5', 'WENO_3', 'MUSCL'])
def test_hllc_branch_codes_identical(scheme):
  n, h = (12, 10, 11), 3
  cfg = _cfg(n, h, scheme, 'HLLC', False, 'periodic')
  st = random_state(n, h, seed=5)
  ocfg = oracle_cfg(cfg)
  _, br = O.rhs(st, {}, ocfg, return_branch=True)
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  ctx = fn.native_context
  g = _to_gpu(st)
  q = ctx.fields(g)
  ctx.fill_halos(q)
  for d, dim in enumerate(('x', 'y', 'z')):
    codes = torch.full(st['rho'].shape, 255, dtype=torch.uint8, device='cuda')
    ctx.hllc_branch_codes(q, d, codes)
    torch.cuda.synchronize()
    c = codes.cpu().numpy()
    sl = [slice(h, -h)] * 3
    ax = {'x': 1, 'y': 2, 'z': 0}[dim]
    sl[ax] = slice(h, -h + 1 if -h + 1 < 0 else None)
    sl = tuple(sl)
    counts = np.bincount(br[dim][sl].ravel(), minlength=4)
    print(f'[branch] {scheme} {dim}: oracle branch histogram {counts.tolist()}', flush=True)
    np.testing.assert_array_equal(c[sl], br[dim][sl])


def test_conservation_periodic():
  """Interior sums of the mass/momentum/energy RHS vanish on periodic grids."""
  n, h = (16, 16, 16), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', True, 'periodic')
  st = random_state(n, h, seed=6)
  got = _gpu_rhs(cfg, st)
  for k, v in got.items():
    r = interior(v, h).astype(np.float64)
    scale = np.abs(r).sum()
    assert abs(r.sum()) <= 1e-5 * scale, (k, r.sum(), scale)


@pytest.mark.parametrize('scheme_name', [
    'rk3',
    pytest.param('rk2', marks=pytest.mark.skip(
        reason='the first GPU run of this case (then through the fused SC_RK_AVG stage'
               ' form, now disabled) ended in a device fault whose cause is not'
               ' identified; rk2.integrate now evaluates sc_rhs twice with torch'
               ' updates, tested on the CPU in test_time_integration_host.py, and this'
               ' case stays off the GPU until the cause is shown')),
])
def test_rk_step_matches_oracle(scheme_name):
  n, h = (13, 11, 9), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', True, 'periodic')
  cfg['time_integration_scheme'] = scheme_name
  st = random_state(n, h, seed=7)
  ocfg = oracle_cfg(cfg)
  ref = O.time_integration(st, {}, ocfg)
  c = SwirlCParameters(copy.deepcopy(cfg))
  fn = cns.get_rhs_fn(None, c)
  mod = rk3 if scheme_name == 'rk3' else rk2
  out = mod.integrate(0, None, _to_gpu(st), {}, c, None, fn)
  torch.cuda.synchronize()
  errs = {k: rel_linf(interior(out[k].cpu().numpy(), h), interior(ref[k], h)) for k in ref}
  _report(f'{scheme_name} step', errs)
  assert max(errs.values()) <= TOL, errs


def test_unphysical_ghost_states_follow_reference_nan_semantics():
  """Periodic-y density with Neumann-y energy gives negative-pressure ghost
  cells and NaN wave speeds; tf.minimum/maximum propagate NaN (selecting the
  F_R branch), and the kernels must do the same."""
  n, h = (14, 11, 9), 3
  cfg = _cfg(n, h, 'WENO_5', 'HLLC', True, 'neumann')
  for v in ('rho', 'rho_u'):
    cfg['bc']['cell_averages'][v]['y'] = {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)}
  st = random_state(n, h, seed=21)
  with np.errstate(all='ignore'):
    ref = O.rhs(st, {}, oracle_cfg(cfg))
  got = _gpu_rhs(cfg, st)
  for k in ref:
    a, b = interior(got[k], h), interior(ref[k], h)
    assert np.array_equal(np.isnan(a), np.isnan(b)), k
    fin = ~np.isnan(b)
    assert rel_linf(a[fin], b[fin]) <= TOL, k
