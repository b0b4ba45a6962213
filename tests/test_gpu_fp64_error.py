"""Error of the GPU fp32 path against an fp64 run of the oracle, for each
benchmark configuration: the Sod shock tube at full size and length (fp64
profiles in tests/golden/sod_fp64.npz), and the Kelvin-Helmholtz layer, the
buoyant bubble and a batched Riemann ensemble at reduced sizes (fp64 oracle
run in the test).  Both sides take the same fixed time steps, so the figures
are the fp32 arithmetic error of the whole step sequence.  (The adjoint's
fp64 check is the finite-difference test of test_gpu_adjoint.py.)

Bound: relative L-infinity error per field of at most 1e-5 for the short
3-D runs, 1e-4 for the bubble (its momenta are the small residual of the
pressure gradient against gravity, both large, so fp32 rounding of the
balance shows at 1e-5 of them), 1e-3 for Sod's 410 steps (the shock and
contact amplify rounding differences).  A momentum component is measured against the largest momentum
component (max|gpu - fp64| / max_d max|rho u_d|): components that start at
zero (w in the shear layer, the bubble's horizontal momenta) would otherwise
divide rounding noise of the flow by their own near-zero size.  Their own
normwise figures are printed beside.
"""

import copy
import os

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import configs
from swirlc_amd.core.batch import BatchSolver
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.core.solver import Solver

from helpers import interior, oracle_cfg, rel_linf

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _report(tag, errs):
  print(f'[fp64] {tag}: ' + ' '.join(f'{k}={v:.2e}' for k, v in errs.items()), flush=True)


def _gpu_run(cfg_d, st, steps, dt):
  cfg_d = copy.deepcopy(cfg_d)
  cfg_d['dt'] = float(dt)
  cfg_d['cfl'] = None  # the fixed step dt, as the oracle takes
  sim = Solver(SwirlCParameters(cfg_d), {k: torch.from_numpy(v.copy()).cuda()
                                        for k, v in st.items()}, cfl=None, use_graph=False)
  sim.step(steps)
  torch.cuda.synchronize()
  return {k: v.cpu().numpy() for k, v in sim.state.items()}


def _oracle64(cfg_d, st, steps, dt):
  ocfg = oracle_cfg(cfg_d)
  s = {k: v.astype(np.float64) for k, v in st.items()}
  for _ in range(steps):
    s = O.rk3(s, {}, ocfg, dt=dt)
  return s


def _dt(cfg_d, st, cfl=0.5):
  with np.errstate(divide='ignore', invalid='ignore'):
    speed = float(O.max_wave_speed({k: v.astype(np.float64) for k, v in st.items()},
                                   oracle_cfg(cfg_d)))
  return np.float32(cfl * min(cfg_d['dx'], cfg_d['dy'], cfg_d['dz']) / speed)


MOMENTA = ('rho_u', 'rho_v', 'rho_w')


def _errors(got, ref, h):
  """Per field max|gpu - fp64| over the field's scale (momenta: the largest
  momentum component), and the plain normwise figures."""
  g = {k: interior(got[k], h).astype(np.float64) for k in ref}
  r = {k: interior(ref[k], h) for k in ref}
  mom = max(float(np.max(np.abs(r[k]))) for k in MOMENTA)
  errs = {}
  for k in ref:
    scale = mom if k in MOMENTA else float(np.max(np.abs(r[k])))
    errs[k] = float(np.max(np.abs(g[k] - r[k]))) / (scale if scale > 0 else 1.0)
  return errs, {k: rel_linf(g[k], r[k]) for k in ref}


def test_sod_full_length_vs_fp64():
  ref = np.load(os.path.join(HERE, 'golden', 'sod_fp64.npz'))
  cfg_d, st = configs.sod(int(ref['n']))
  h = cfg_d['halo_width']
  got = _gpu_run(cfg_d, st, int(ref['steps']), float(ref['dt']))
  with np.errstate(divide='ignore', invalid='ignore'):
    prim = O.conservative_to_primitive({k: v.astype(np.float64) for k, v in got.items()})
    p = O.pressure(prim, 0.0)
  sl = (h, slice(h, -h), h)
  errs = {'rho': rel_linf(prim['rho'][sl], ref['rho']), 'u': rel_linf(prim['u'][sl], ref['u']),
          'p': rel_linf(p[sl], ref['p'])}
  l1 = float(np.mean(np.abs(prim['rho'][sl] - ref['rho'])))
  _report(f'sod 512 x {int(ref["steps"])} steps (t=0.2), profiles, rho L1 {l1:.1e}', errs)
  assert max(errs.values()) <= 1e-3, errs


def test_kelvin_helmholtz_vs_fp64():
  cfg_d, st = configs.kh(n=32)
  dt = _dt(cfg_d, st)
  got = _gpu_run(cfg_d, st, 10, dt)
  ref = _oracle64(cfg_d, st, 10, float(dt))
  errs, plain = _errors(got, ref, cfg_d['halo_width'])
  _report('kelvin-helmholtz 32^3 viscous x 10 steps', errs)
  _report('  (each field normwise)', plain)
  assert max(errs.values()) <= 1e-5, errs


def test_buoyant_bubble_vs_fp64():
  cfg_d, st = configs.bubble(32, 32, 16)
  dt = _dt(cfg_d, st)
  got = _gpu_run(cfg_d, st, 5, dt)
  ref = _oracle64(cfg_d, st, 5, float(dt))
  errs, plain = _errors(got, ref, cfg_d['halo_width'])
  _report('buoyant bubble 32x32x16 MUSCL/HLL viscous gravity x 5 steps', errs)
  _report('  (each field normwise)', plain)
  assert max(errs.values()) <= 1e-4, errs


def test_batched_ensemble_vs_fp64():
  shapes = ((16, 16, 16), (24, 16, 12), (12, 20, 16))
  cfg, members = configs.riemann_ensemble(3, seed=0, shapes=shapes)
  cfg = copy.deepcopy(cfg)
  cfg['dt'] = 2e-3
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(cfg)), members, cfl=None, use_graph=False)
  batch.step(5)
  torch.cuda.synchronize()
  worst = {}
  for k, (grid, st) in enumerate(members):
    mcfg = configs.member_cfg(cfg, grid)
    ref = _oracle64(mcfg, st, 5, 2e-3)
    got = {n: t.cpu().numpy() for n, t in batch.state(k).items()}
    errs, _ = _errors(got, ref, cfg['halo_width'])
    _report(f'batched member {k} {grid[:3]} x 5 steps', errs)
    for n, e in errs.items():
      worst[n] = max(worst.get(n, 0.0), e)
  assert max(worst.values()) <= 1e-5, worst
