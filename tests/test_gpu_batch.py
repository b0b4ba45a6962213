"""GPU tests of batched ensembles (one launch per phase over all instances).

The batched kernels run the same per-tile bodies as the single-grid kernels,
so a batched RK3 step must equal per-instance single-grid steps bit for bit
(including the per-instance CFL time step); one instance is also checked
against the CPU oracle.
"""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import configs
from swirlc_amd.core.batch import BatchSolver
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.core.solver import Solver

from helpers import grid_cfg, interior, oracle_cfg, random_state, rel_linf

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12, 10), (9, 70, 14), (20, 8, 33), (12, 12, 12), (40, 6, 7)]


def _cfg_for(shape, visc):
  cfg = grid_cfg(shape)
  BC = configs.BC
  cfg.update({'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
              'include_diffusion': visc, 'transport_parameters': {'nu': 2e-3, 'pr': 0.7},
              'dt': 1e-3, 'bc': {'cell_averages': {
                  v: {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.NEUMANN, 0.0)},
                      'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                      'z': ({0: (BC.DIRICHLET, 0.5 if v == 'rho' else 0.0),
                             1: (BC.NEUMANN, 0.0)} if v in ('rho', 'rho_u') else
                            {0: (BC.NEUMANN, 0.0), 1: (BC.NEUMANN, 0.0)})}
                  for v in ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')}}})
  return cfg


@pytest.mark.parametrize('visc', [False, True])
def test_batch_equals_per_instance_bitwise(visc):
  members = []
  for k, shape in enumerate(SHAPES):
    cfg = _cfg_for(shape, visc)
    st = random_state(shape, 3, seed=10 + k)
    members.append(((*shape, cfg['dx'], cfg['dy'], cfg['dz']), st, cfg))
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(members[0][2])),
                      [(g, s) for g, s, _ in members], cfl=0.5, use_graph=True)
  batch.step(2)
  torch.cuda.synchronize()
  dts = batch.dt.cpu().numpy()
  for k, (g, st, cfg) in enumerate(members):
    solo = Solver(SwirlCParameters(copy.deepcopy(cfg)),
                  {n: torch.from_numpy(v).cuda() for n, v in st.items()}, cfl=0.5,
                  use_graph=False)
    solo.step(2)
    torch.cuda.synchronize()
    assert float(solo.dt[0]) == float(dts[k]), (k, float(solo.dt[0]), dts[k])
    got = batch.state(k)
    for n, t in solo.state.items():
      a = interior(got[n].cpu().numpy(), 3)
      b = interior(t.cpu().numpy(), 3)
      assert np.array_equal(a, b), (k, n, float(np.abs(a - b).max()))
  print(f'[batch] {len(SHAPES)} instances bitwise equal to single-grid steps; dt={dts}', flush=True)


def test_batch_instance_matches_oracle():
  shape = (14, 11, 9)
  cfg = _cfg_for(shape, True)
  st = random_state(shape, 3, seed=21)
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(cfg)),
                      [((*shape, cfg['dx'], cfg['dy'], cfg['dz']), st)], cfl=None)
  batch.step(1)
  torch.cuda.synchronize()
  ref = O.rk3(st, {}, oracle_cfg(cfg))
  got = batch.state(0)
  errs = {n: rel_linf(interior(got[n].cpu().numpy(), 3), interior(ref[n], 3)) for n in ref}
  print(f'[batch] one instance vs oracle: {errs}', flush=True)
  assert max(errs.values()) <= 1e-5, errs


def test_ensemble_generator_sizes():
  cfg, members = configs.riemann_ensemble(16, seed=0)
  assert len(members) == 16
  for g, st in members:
    assert g[:3] in configs.ENSEMBLE_SHAPES
    assert st['rho'].shape == (g[2] + 6, g[0] + 6, g[1] + 6)


def _wide_cfg(shape, visc, faces):
  """Gravity (-z) with its source, two transported scalars (the first's BC on
  its primitive), u's BC on the primitive, scalar flux values and (inviscid
  runs) face values."""
  BC = configs.BC
  D, NONE = BC.DIRICHLET, (None, None)
  cfg = grid_cfg(shape, 3, (1.0, 1.0, 0.5))
  names = ['rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e', 'rho_y0', 'rho_y1']
  wall = {'rho': 1.0, 'u': 0.0, 'rho_v': 0.0, 'rho_w': 0.0, 'rho_e': 6.0, 'y0': 0.3,
          'rho_y1': 0.2}
  cells = {k: {'x': {0: (BC.NEUMANN, 0.0), 1: (BC.NEUMANN, 0.0)},
               'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
               'z': {0: (BC.DIRICHLET, v), 1: (BC.NEUMANN, 0.0)}} for k, v in wall.items()}
  fluxes = {'convective': {
      'rho': {'x': {0: NONE, 1: NONE}, 'y': {0: NONE, 1: NONE}, 'z': {0: (D, 0.0), 1: NONE}},
      'rho_y1': {'x': {0: NONE, 1: (D, 0.05)}, 'y': {0: NONE, 1: NONE},
                 'z': {0: NONE, 1: NONE}}}}
  bc = {'cell_averages': cells, 'intercell_fluxes': fluxes}
  if faces:
    bc['cell_faces'] = {
        'v': {'x': {0: (D, 0.1), 1: NONE}, 'y': {0: NONE, 1: NONE}, 'z': {0: NONE, 1: NONE}},
        'rho_y0': {'x': {0: NONE, 1: NONE}, 'y': {0: NONE, 1: NONE},
                   'z': {0: NONE, 1: (D, 0.25)}}}
  cfg.update({'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
              'include_diffusion': visc, 'transport_parameters': {'nu': 2e-3, 'pr': 0.7},
              'dt': 1e-3, 'g': {'x': 0.0, 'y': 0.0, 'z': -1.0},
              'source_functions': ['gravity'], 'conservative_variable_names': names,
              'primitive_variable_names': ['rho', 'u', 'v', 'w', 'e', 'y0', 'y1'], 'bc': bc})
  return cfg


def _wide_state(cfg, shape, seed):
  pe = O.gravitational_potential(oracle_cfg(cfg), np.float64)
  return random_state(shape, 3, seed=seed, nscalar=2, pe=pe)


@pytest.mark.parametrize('visc,faces', [(True, False), (False, True)])
def test_widened_batch_equals_per_instance_bitwise(visc, faces):
  """Gravity, primitive-variable BCs, scalar boundary values and two scalars
  in one ensemble: every member bitwise equal to its single-grid steps.
  Member 0 takes the default coordinates (centres of [0, n d]), the others
  their own."""
  members = []
  for k, shape in enumerate(SHAPES):
    cfg = _wide_cfg(shape, visc, faces)
    grid6 = (*shape, cfg['dx'], cfg['dy'], cfg['dz'])
    if k == 0:
      cfg = configs.member_cfg(cfg, grid6)
      grid = grid6
    else:
      grid = (*grid6, cfg['x'], cfg['y'], cfg['z'])
    members.append((grid, _wide_state(cfg, shape, 30 + k), cfg))
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(members[0][2])),
                      [(g, s) for g, s, _ in members], cfl=0.5, use_graph=True)
  assert batch.nf == 7
  batch.step(2)
  torch.cuda.synchronize()
  dts = batch.dt.cpu().numpy()
  for k, (g, st, cfg) in enumerate(members):
    solo = Solver(SwirlCParameters(copy.deepcopy(cfg)),
                  {n: torch.from_numpy(v).cuda() for n, v in st.items()}, cfl=0.5,
                  use_graph=False)
    solo.step(2)
    torch.cuda.synchronize()
    assert float(solo.dt[0]) == float(dts[k]), (k, float(solo.dt[0]), dts[k])
    got = batch.state(k)
    for n, t in solo.state.items():
      a = interior(got[n].cpu().numpy(), 3)
      b = interior(t.cpu().numpy(), 3)
      assert np.all(np.isfinite(a)), (k, n)
      assert np.array_equal(a, b), (k, n, float(np.abs(a - b).max()))
  print(f'[batch] widened ensemble (visc={visc}, faces={faces}): {len(SHAPES)} members'
        f' bitwise equal; dt={dts}', flush=True)


def test_widened_batch_instance_matches_oracle():
  shape = (14, 11, 9)
  cfg = _wide_cfg(shape, True, False)
  st = _wide_state(cfg, shape, 22)
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(cfg)),
                      [((*shape, cfg['dx'], cfg['dy'], cfg['dz'], cfg['x'], cfg['y'],
                         cfg['z']), st)], cfl=None)
  batch.step(1)
  torch.cuda.synchronize()
  ref = O.rk3(st, {}, oracle_cfg(cfg))
  got = batch.state(0)
  errs = {n: rel_linf(interior(got[n].cpu().numpy(), 3), interior(ref[n], 3)) for n in ref}
  print('[batch] widened instance vs oracle: ' +
        ' '.join(f'{k}={v:.1e}' for k, v in errs.items()), flush=True)
  assert max(errs.values()) <= 1e-5, errs
