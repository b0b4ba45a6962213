"""The batched-ensemble and adjoint workloads at their configured sizes.

* Batching: the full ensemble of 256 3-D Riemann problems drawn with
  default_rng(0) (WENO5 + HLLC, inviscid, Neumann walls, per-instance CFL-0.5
  time steps), advanced by the graph-replayed batched kernels; sampled members
  must equal single-grid steps bit for bit, per-instance dt included.
* Differentiation: the viscous Kelvin-Helmholtz setup at 128^3 rolled out
  200 SSP-RK3 steps with checkpoints every 20 steps: the gradient does not
  depend on the checkpoint interval (20, 50, or every stage state kept; bit
  for bit) and repeats exactly.  The
  gradient's agreement with FP64 central finite differences of the oracle is
  checked on a KH-like state with w != 0 (so z faces leave the HLLC
  contact-speed snap of hllc.py:339-341) along 16 directions from
  default_rng(2).
"""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import configs
from swirlc_amd.core import adjoint as A
from swirlc_amd.core.batch import BatchSolver
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.core.solver import Solver
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import interior, oracle_cfg
import test_gpu_adjoint as T

pytestmark = pytest.mark.gpu
NAMES = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')


def test_full_riemann_ensemble_matches_single_grid_steps():
  cfg, members = configs.riemann_ensemble(256, seed=0)
  cells = sum(g[0] * g[1] * g[2] for g, _ in members)
  print(f'[ensemble] 256 instances, {cells} cells', flush=True)
  assert len(members) == 256
  batch = BatchSolver(SwirlCParameters(dict(cfg)), members, cfl=0.5, use_graph=True)
  batch.step(2)
  torch.cuda.synchronize()
  dts = batch.dt.cpu().numpy()
  assert np.all(np.isfinite(dts)) and np.all(dts > 0)
  for f in range(5):
    assert bool(torch.isfinite(batch.u[f]).all())
  sizes = [g[0] * g[1] * g[2] for g, _ in members]
  picks = sorted({0, int(np.argmin(sizes)), int(np.argmax(sizes)), 131, 255})
  for k in picks:
    g, st = members[k]
    solo = Solver(SwirlCParameters(configs.member_cfg(cfg, g)),
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
    del solo
  print(f'[ensemble] members {picks} bitwise equal to single-grid steps', flush=True)


def test_kh128_checkpointed_rollout_is_interval_invariant_and_deterministic():
  cfg_d, st = configs.kh(n=128)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg_d))).native_context
  h = cfg_d['halo_width']
  u0 = [torch.from_numpy(st[k]).cuda() for k in ctx.names]
  speed = torch.zeros(4, device='cuda')
  dt = torch.zeros(1, device='cuda')
  ctx.max_wave_speed(u0, speed, 0.5, dt)
  rho_t = torch.from_numpy(np.random.default_rng(1).uniform(
      0.9, 2.1, tuple(u0[0].shape)).astype(np.float32)).cuda()

  def grad(every):
    x = [t.clone().requires_grad_(True) for t in u0]
    J = A.kinetic_density_loss(A.rollout(ctx, dt, 200, *x, every=every), rho_t, h)
    J.backward()
    torch.cuda.synchronize()
    return float(J.detach()), [t.grad for t in x]

  print('[adjoint] KH 128^3, 200 steps, checkpoints every 20', flush=True)
  J20, g20 = grad(20)
  J50, g50 = grad(50)
  J20b, g20b = grad(20)
  J0, g0 = grad(0)  # every stage state kept, nothing recomputed
  assert J20 == J50 == J20b == J0
  sl = (slice(h, -h),) * 3
  for a, b, c, d in zip(g20, g50, g20b, g0):
    assert torch.equal(a, b) and torch.equal(a, c) and torch.equal(a, d)
    assert bool(torch.isfinite(a).all())
    halo = a.clone()
    halo[sl] = 0.0
    assert not bool(halo.any())
  print(f'[adjoint] J = {J20:.6e}, gradient identical for checkpoint intervals 20, 50 and'
        ' for the full tape', flush=True)


def _kh_3d(n, h=3):
  """A smoothed KH shear layer with v and w perturbations (w != 0)."""
  cfg, st = T._kh_small(n, nu=1e-2)
  x = cfg['x'].astype(np.float64)[None, :, None]
  z = cfg['z'].astype(np.float64)[:, None, None]
  rng = np.random.default_rng(3)
  shape = (n[2], n[0], n[1])
  w = 0.05 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z) + rng.uniform(-1e-2, 1e-2, shape)
  sl = (slice(h, -h),) * 3
  rho = st['rho'][sl].astype(np.float64)
  e = st['rho_e'][sl].astype(np.float64) / rho
  st['rho_w'][sl] = (rho * w).astype(np.float32)
  st['rho_e'][sl] = (rho * (e + 0.5 * w * w)).astype(np.float32)
  return cfg, st


def test_kh_like_rollout_gradient_matches_fd_along_16_directions():
  n, h, steps, every, dt = (12, 12, 10), 3, 3, 2, 2e-3
  cfg, st = _kh_3d(n, h)
  ctx = T._ctx(cfg)
  rho_t_np = np.random.default_rng(1).uniform(0.9, 2.1, st['rho'].shape).astype(np.float32)
  rho_t = torch.from_numpy(rho_t_np).cuda()
  print('[adjoint] KH-like 12x12x10 (w != 0), 3 steps, 16 directions', flush=True)
  J32, grad = T._rollout_grad(ctx, st, rho_t, h, dt, steps, every)
  ocfg = oracle_cfg(cfg)
  sl = (slice(h, -h),) * 3

  def J(s):
    for _ in range(steps):
      s = O.rk3(s, {}, ocfg, dt=dt)
    rho = s['rho'][sl]
    ke = 0.5 * (s['rho_u'][sl] ** 2 + s['rho_v'][sl] ** 2 + s['rho_w'][sl] ** 2) / rho
    return float(ke.sum() + 0.1 * ((rho - rho_t_np[sl].astype(np.float64)) ** 2).sum())

  assert abs(J32 - J({k: st[k].astype(np.float64) for k in NAMES})) <= 1e-5 * abs(J32)
  errs = []
  for v in T._directions(st, h, 16, seed=2):
    fd = T._fd(J, st, v)
    ad = T._dot(grad, v)
    errs.append(abs(ad - fd) / abs(fd))
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert max(errs) <= T.FD_TOL
