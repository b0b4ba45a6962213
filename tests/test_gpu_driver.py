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
