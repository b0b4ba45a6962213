"""Chained RK stages (sc_rk_stage_ex): halo cells and the CFL wave speed
produced by the last sweep of a stage.

With SC_STAGE_FILL_OUT a stage leaves its output with the halo cells filled
(on all-periodic grids the last sweep writes every output cell's periodic
images); they must equal what sc_fill_halos writes, bit for bit, corners
included.  With speed_dev the stage max-accumulates the wave speed of the
cells it writes; it must equal sc_max_wave_speed on the output bit for bit
(both take the same per-cell function and a maximum, which is exact in any
order).  A Solver stepping through chained stages must reproduce the plain
stepping (halo fill at every stage, a separate wave-speed pass every step),
state and dt, bit for bit.  The oracle pins the wave speed itself
(test_gpu_rhs_parity.py::test_max_wave_speed_matches_oracle) and the halo fill
(test_gpu_rhs_parity.py, tests/test_oracle_reference_cases.py).
"""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import _native as N
from swirlc_amd import configs
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.core.solver import Solver
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import grid_cfg, oracle_cfg, random_state
from test_gpu_rhs_parity import _bcs
from test_gpu_scalars import _cfg as scalar_cfg

pytestmark = pytest.mark.gpu


def _ctx_cfg(n, h, gravity, bcs, scheme='WENO_5', flux='HLLC', visc=True, nscalar=0):
  if nscalar:
    return scalar_cfg(n, scheme, flux, visc, bcs, nscalar=nscalar)
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05) if gravity else (1.0, 1.0, 1.0))
  cfg.update({'interpolation_scheme': scheme, 'numeric_flux_scheme': flux,
              'include_diffusion': visc, 'bc': _bcs(bcs, n, h),
              'transport_parameters': {'nu': 2e-3, 'pr': 0.7}, 'dt': 2e-3})
  if gravity:
    cfg.update({'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity']})
  return cfg


@pytest.mark.parametrize('gravity,bcs,scheme,flux,nscalar', [
    (False, 'periodic', 'WENO_5', 'HLLC', 0),
    (True, 'mixed', 'MUSCL', 'HLL', 0),
    (False, 'bounded', 'WENO_3', 'HLLC', 1),
])
def test_stage_speed_equals_separate_pass(gravity, bcs, scheme, flux, nscalar):
  n, h = (70, 13, 11), 3  # a y row longer than one wave, a ragged last wave
  cfg = _ctx_cfg(n, h, gravity, bcs, scheme, flux, nscalar=nscalar)
  pe = O.gravitational_potential(oracle_cfg(cfg), np.float64) if gravity else None
  st = random_state(n, h, seed=21, pe=pe, nscalar=nscalar)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  q = ctx.fields({k: torch.from_numpy(v.copy()).cuda() for k, v in st.items()})
  u0 = [t.clone() for t in q]
  dt = torch.full((1,), 1e-3, device='cuda')
  for form, c0, c1 in ((N.SC_RK_EULER, 0.0, 0.0), (N.SC_RK_BLEND, 0.75, 0.25)):
    out = [torch.zeros_like(t) for t in q]
    acc = torch.zeros(4, device='cuda')
    # Two calls into one accumulator: the second's maximum is the same, so
    # the accumulation must leave it unchanged.
    ctx.rk_stage_ex(form, c0, c1, dt, u0, q, out, speed_dev=acc)
    ctx.rk_stage_ex(form, c0, c1, dt, u0, q, out, speed_dev=acc)
    ref_out = [torch.zeros_like(t) for t in q]
    ctx.rk_stage(form, c0, c1, dt, u0, q, ref_out)
    sep = torch.zeros(4, device='cuda')
    ctx.max_wave_speed(out, sep)
    torch.cuda.synchronize()
    for a, b in zip(out, ref_out):
      assert torch.equal(a, b)
    got, want = acc[0].item(), sep[0].item()
    print(f'[cfl] form {form}: fused {got!r} separate {want!r}', flush=True)
    assert got == want and want > 0.0


def test_cfl_dt_resets_the_maximum():
  n, h = (16, 16, 16), 3
  cfg = _ctx_cfg(n, h, False, 'periodic')
  st = random_state(n, h, seed=22)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  q = ctx.fields({k: torch.from_numpy(v.copy()).cuda() for k, v in st.items()})
  speed = torch.zeros(4, device='cuda')
  dt = torch.zeros(1, device='cuda')
  ctx.max_wave_speed(q, speed, 0.5, dt)
  torch.cuda.synchronize()
  want_dt = dt.item()
  dt.zero_()
  ctx.cfl_dt(speed, 0.5, dt)
  torch.cuda.synchronize()
  assert dt.item() == want_dt
  assert speed[0].item() == 0.0


@pytest.mark.parametrize('n,bcs', [
    ((70, 13, 11), 'periodic'),   # images written by the last sweep
    ((9, 12, 6), 'periodic'),     # an extent below 2h: fill after the sweeps
    ((14, 11, 9), 'mixed'),       # not periodic: fill after the sweeps
])
def test_stage_output_halos_equal_the_fill(n, bcs):
  h = 3
  cfg = _ctx_cfg(n, h, False, bcs)
  st = random_state(n, h, seed=23)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  q = ctx.fields({k: torch.from_numpy(v.copy()).cuda() for k, v in st.items()})
  u0 = [t.clone() for t in q]
  dt = torch.full((1,), 1e-3, device='cuda')
  fused = ctx.out_images_fused()
  assert fused == (bcs == 'periodic' and min(n) >= 2 * h)
  for form, c0, c1 in ((N.SC_RK_EULER, 0.0, 0.0), (N.SC_RK_BLEND, 1.0 / 3.0, 2.0 / 3.0)):
    out = [torch.full_like(t, float('nan')) for t in q]
    ctx.rk_stage_ex(form, c0, c1, dt, u0, q, out, N.SC_STAGE_FILL_OUT)
    ref = [t.clone() for t in out]
    ctx.fill_halos(ref)
    plain = [torch.zeros_like(t) for t in q]
    ctx.rk_stage(form, c0, c1, dt, u0, q, plain)
    torch.cuda.synchronize()
    for a, b, c in zip(out, ref, plain):
      assert torch.equal(a, b)  # every cell, halos and corners included
      sl = (slice(h, -h),) * 3
      assert torch.equal(a[sl], c[sl])
    # The next stage with SC_STAGE_HALOS_READY on that output equals a plain
    # stage that fills the halos itself.
    nxt = [torch.zeros_like(t) for t in q]
    ctx.rk_stage_ex(N.SC_RK_EULER, 0.0, 0.0, dt, None, out, nxt, N.SC_STAGE_HALOS_READY)
    nxt_ref = [torch.zeros_like(t) for t in q]
    ctx.rk_stage(N.SC_RK_EULER, 0.0, 0.0, dt, None, ref, nxt_ref)
    torch.cuda.synchronize()
    sl = (slice(h, -h),) * 3
    for a, b in zip(nxt, nxt_ref):
      assert torch.equal(a[sl], b[sl])
  print(f'[halo] {n} {bcs}: fused images {fused}', flush=True)


def test_stage_ex_rejects_unknown_flags():
  n, h = (8, 8, 8), 3
  cfg = _ctx_cfg(n, h, False, 'periodic')
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  q = [torch.zeros(ctx.shape, device='cuda') for _ in range(5)]
  dt = torch.zeros(1, device='cuda')
  with pytest.raises(Exception):
    ctx.rk_stage_ex(N.SC_RK_EULER, 0.0, 0.0, dt, None, q, [t.clone() for t in q], 4)


@pytest.mark.parametrize('use_graph', [True, False])
@pytest.mark.parametrize('cfl', [0.5, None])
def test_chained_solver_matches_plain_stepping(use_graph, cfl):
  cfg_d, st = configs.kh(n=32)
  cfg = SwirlCParameters(copy.deepcopy(cfg_d))
  gpu = {k: torch.from_numpy(v).cuda() for k, v in st.items()}
  chained = Solver(cfg, gpu, cfl=cfl, use_graph=use_graph)
  plain = Solver(cfg, gpu, cfl=cfl, use_graph=use_graph)
  assert chained._chained and chained.ctx.out_images_fused()
  plain._chained = False  # halo fill at every stage, separate wave-speed pass
  for _ in range(3):
    chained.step(2)
    plain.step(2)
    torch.cuda.synchronize()
    assert chained.dt.item() == plain.dt.item()
    h = chained.ctx.hw
    sl = (slice(h, -h),) * 3
    for a, b in zip(chained.u, plain.u):
      assert torch.equal(a[sl], b[sl])
  print(f'[chain] cfl {cfl}: dt after 6 steps {chained.dt.item()!r}', flush=True)
