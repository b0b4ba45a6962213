"""Host driver on the GPU: a run split into a saved first part and a
restarted second part (start_step, states_from_file) ends bit for bit where
the uninterrupted run ends."""

import copy

import numpy as np
import pytest
import torch

from swirlc_amd import configs
from swirlc_amd.core import driver
from swirlc_amd.core.parameter import SwirlCParameters

pytestmark = pytest.mark.gpu


def test_restart_reproduces_the_uninterrupted_run(tmp_path):
  cfg_d, st = configs.kh(n=24)
  names = list(st)
  init = lambda replica_id, coordinates, cfg: st
  prefix = str(tmp_path / 'kh24')
  full = SwirlCParameters(dict(copy.deepcopy(cfg_d), num_cycles=2, num_steps=3, start_step=0,
                               states_to_file=names))
  print('[driver] uninterrupted 2 x 3 steps', flush=True)
  final, step = driver.solver(init, full, dump_prefix=prefix, cfl=0.5)
  assert step == 6
  ref = {k: v.cpu().numpy() for k, v in final.items()}
  saved = driver.load_states(prefix, 6, names)
  for k in names:
    np.testing.assert_array_equal(saved[k], ref[k])
  # Restart from the save file of step 3 with the default (zero) initializer.
  part = SwirlCParameters(dict(copy.deepcopy(cfg_d), num_cycles=1, num_steps=3, start_step=3,
                               states_from_file=names))
  print('[driver] restart from step 3', flush=True)
  final2, step2 = driver.solver(None, part, load_prefix=prefix, cfl=0.5)
  torch.cuda.synchronize()
  assert step2 == 6
  sl = (slice(3, -3),) * 3
  for k in names:
    np.testing.assert_array_equal(final2[k].cpu().numpy()[sl], ref[k][sl])


def test_driver_cycles_match_oracle_rk3_steps(tmp_path):
  """driver.solver (core/driver.py:135-160 with swirl_lm's cycle loop):
  2 cycles x 2 steps from start_step 5, the state loaded from the save file
  of step 5, fixed cfg.dt; every saved cycle and the final state against the
  oracle's SSP-RK3 steps (numerics/time_integration/rk3.py:34-122) with the
  same dt."""
  from oracle import swirlc_oracle as O
  from helpers import interior, oracle_cfg, random_state, rel_linf
  import test_gpu_rhs_parity as P
  n, h = (13, 11, 9), 3
  cfg_d = P._cfg(n, h, 'WENO_5', 'HLLC', True, 'mixed')
  st = random_state(n, h, seed=41)
  names = list(st)
  load = str(tmp_path / 'in')
  dump = str(tmp_path / 'out')
  driver.save_states(load, 5, st, names, h)
  cfg = SwirlCParameters(dict(copy.deepcopy(cfg_d), num_cycles=2, num_steps=2, start_step=5,
                              states_from_file=names, states_to_file=names))
  print('[driver] 2 cycles x 2 steps from step 5 vs the oracle', flush=True)
  final, step = driver.solver(None, cfg, load_prefix=load, dump_prefix=dump, cfl=None,
                              use_graph=True)
  torch.cuda.synchronize()
  assert step == 9
  ocfg = oracle_cfg(cfg_d)
  ref = dict(st)
  for k_step in (7, 9):
    for _ in range(2):
      ref = O.rk3(ref, {}, ocfg)
    saved = driver.load_states(dump, k_step, names)
    errs = {k: rel_linf(interior(saved[k], h), interior(ref[k], h)) for k in names}
    print(f'[driver] step {k_step}: ' + ' '.join(f'{k}={v:.1e}' for k, v in errs.items()),
          flush=True)
    assert max(errs.values()) <= 1e-5, (k_step, errs)
  for k in names:
    np.testing.assert_array_equal(interior(final[k].cpu().numpy(), h),
                                  interior(driver.load_states(dump, 9, names)[k], h))
