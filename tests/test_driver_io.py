"""Save files of the host driver: round trip, naming, safe loading."""

import numpy as np
import pytest
import torch

from swirlc_amd.core import driver
from swirlc_amd.core.parameter import SwirlCParameters


def test_save_load_round_trip(tmp_path):
  rng = np.random.default_rng(0)
  states = {'rho': rng.random((8, 10, 9), dtype=np.float32),
            'rho_u': torch.from_numpy(rng.random((8, 10, 9), dtype=np.float32))}
  prefix = str(tmp_path / 'out' / 'kh')
  path = driver.save_states(prefix, 12, states, ['rho', 'rho_u'], 3)
  assert path == f'{prefix}-0000000012.npz'
  got = driver.load_states(prefix, 12, ['rho_u', 'rho'])
  np.testing.assert_array_equal(got['rho'], states['rho'])
  np.testing.assert_array_equal(got['rho_u'], states['rho_u'].numpy())
  with np.load(path, allow_pickle=False) as f:
    assert int(f['step']) == 12 and int(f['halo_width']) == 3
  with pytest.raises(KeyError):
    driver.load_states(prefix, 12, ['rho_e'])
  with pytest.raises(FileNotFoundError):
    driver.load_states(prefix, 13, ['rho'])


def test_default_initial_state_is_zero():
  cfg = SwirlCParameters({'core_nx': 6, 'core_ny': 5, 'core_nz': 4, 'nx': 6, 'ny': 5, 'nz': 4})
  st = driver.init_state(None, cfg, device='cpu')
  assert set(st) == set(cfg.conservative_variable_names)
  for v in st.values():
    assert tuple(v.shape) == (10, 12, 11) and not v.any()


def test_config_is_saved_as_yaml(tmp_path):
  import yaml
  cfg = SwirlCParameters({'num_steps': 7})
  cfg.save_to_file(str(tmp_path / 'run'))
  with open(tmp_path / 'run_swirl_c_cfg.yml') as f:
    d = yaml.safe_load(f)
  assert d['num_steps'] == 7
