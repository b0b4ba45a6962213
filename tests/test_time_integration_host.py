"""The host Runge-Kutta integrators with an RHS callable that is not native.

rk2.integrate / rk3.integrate / time_integration take any
`rhs(replica_id, state, helper_vars)` (numerics/time_integration/rhs_type.py:
22-29); given one that is not a get_rhs_fn closure they integrate with torch
element-wise ops on the state's device.  These tests run that path on CPU
tensors: the reference's staged known answers (rk2_test.py:85-207,
rk3_test.py:94-212) and agreement with the oracle's rk2 / rk3 on a real RHS
(the oracle's fp32 RHS wrapped as a torch callable).
"""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.numerics.time_integration import rk2, rk3, time_integration

from helpers import grid_cfg, interior, oracle_cfg, random_state

NAMES = ['rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e', 'rho_y']


def _state():
  return {k: torch.full((4, 4, 4), float(i + 1)) for i, k in enumerate(NAMES)}


def _cfg(scheme):
  return SwirlCParameters({'dt': 0.1, 'core_nx': 4, 'core_ny': 4, 'core_nz': 4,
                           'halo_width': 0, 'time_integration_scheme': scheme,
                           'conservative_variable_names': NAMES})


def _zero_rhs(replica_id, state, helper_vars):
  del replica_id, helper_vars
  return {k: torch.zeros_like(v) for k, v in state.items()}


def _staged_rhs2(replica_id, state, helper_vars):
  """rk2_test.py:49-79: x 20 at stage 1 (mean rho < 2), x -10 at stage 2."""
  del replica_id, helper_vars
  f = 20.0 if float(state['rho'].mean()) < 2.0 else -10.0
  return {k: v * f for k, v in state.items()}


def test_rk2_constant_solution():
  out = rk2.integrate(0, None, _state(), {}, _cfg('rk2'), None, _zero_rhs)
  for i, k in enumerate(NAMES):
    assert torch.equal(out[k], torch.full((4, 4, 4), float(i + 1))), k


def test_rk2_staged_known_answer():
  """rk2_test.py:145-207: u1 = 3 u0, u2 = u1 - u1 = 0, u = u0 / 2."""
  out = rk2.integrate(0, None, _state(), {}, _cfg('rk2'), None, _staged_rhs2)
  for i, k in enumerate(NAMES):
    np.testing.assert_allclose(out[k].numpy(), 0.5 * (i + 1), rtol=1e-6, err_msg=k)


def test_time_integration_dispatches_rk2():
  a = time_integration.time_integration(0, None, _state(), {}, _cfg('rk2'), None,
                                        _staged_rhs2)
  for i, k in enumerate(NAMES):
    np.testing.assert_allclose(a[k].numpy(), 0.5 * (i + 1), rtol=1e-6, err_msg=k)
  with pytest.raises(NotImplementedError):
    time_integration.time_integration(0, None, _state(), {}, _cfg('rk4'), None, _zero_rhs)


def _oracle_rhs_fn(cfg_d):
  ocfg = oracle_cfg(cfg_d)

  def fn(replica_id, state, helper_vars):
    del replica_id
    out = O.rhs({k: v.numpy() for k, v in state.items()}, helper_vars, ocfg)
    return {k: torch.from_numpy(v) for k, v in out.items()}
  return fn


@pytest.mark.parametrize('scheme', ['rk2', 'rk3'])
def test_host_integrators_match_the_oracle_step(scheme):
  n, h = (6, 5, 4), 3
  cfg_d = grid_cfg(n, h)
  cfg_d.update({'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
                'include_diffusion': True, 'dt': 2e-3, 'time_integration_scheme': scheme,
                'transport_parameters': {'nu': 2e-3, 'pr': 0.7}})
  st = random_state(n, h, seed=17)
  ref = O.time_integration(st, {}, oracle_cfg(cfg_d))
  mod = rk2 if scheme == 'rk2' else rk3
  out = mod.integrate(0, None, {k: torch.from_numpy(v.copy()) for k, v in st.items()}, {},
                      SwirlCParameters(copy.deepcopy(cfg_d)), None, _oracle_rhs_fn(cfg_d))
  for k in ref:
    np.testing.assert_array_equal(interior(out[k].numpy(), h), interior(ref[k], h), err_msg=k)
