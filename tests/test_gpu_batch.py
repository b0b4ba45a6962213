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
