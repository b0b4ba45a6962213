"""Launch-geometry tuning changes timings only: every tuning gives bit-identical
results (single grid and batched), the adjoint passes agree across their
geometries, and the autotuners return valid ones."""

import copy

import numpy as np
import pytest
import torch

from swirlc_amd import _native as N
from swirlc_amd import configs
from swirlc_amd.core import autotune
from swirlc_amd.core.batch import BatchSolver
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.equations import compressible_navier_stokes as cns

pytestmark = pytest.mark.gpu

TUNINGS = [
    dict(seg_x=16, by_x=2, seg_z=64, by_z=1, by_y=2, xcd_remap=1, seg_y=1),
    dict(seg_x=64, by_x=4, seg_z=16, by_z=4, by_y=1, xcd_remap=0, seg_y=3),
    dict(seg_x=5, by_x=3, seg_z=7, by_z=3, by_y=3, xcd_remap=1, seg_y=8),
    dict(seg_x=32, by_x=4, seg_z=32, by_z=4, by_y=4, xcd_remap=1, seg_y=64),
    dict(seg_x=32, by_x=4, seg_z=32, by_z=4, by_y=4, xcd_remap=1, seg_y=7),
]


def _stage(ctx, q, dt):
  out = [torch.zeros_like(t) for t in q]
  ctx.rk_stage(N.SC_RK_BLEND, 0.75, 0.25, dt, q, q, out)
  torch.cuda.synchronize()
  return out


def test_tuning_is_bitwise_neutral_single_grid():
  cfg_d, st = configs.kh(n=40)
  cfg_d = copy.deepcopy(cfg_d)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(cfg_d)).native_context
  q = [torch.from_numpy(st[k]).cuda() for k in ctx.names]
  dt = torch.full((1,), 1e-3, device='cuda')
  base = ctx.get_tuning()
  ref = _stage(ctx, q, dt)
  for t in TUNINGS:
    print(f'[tuning] single grid {t}', flush=True)
    ctx.set_tuning(t)
    assert {k: ctx.get_tuning()[k] for k in t} == t
    got = _stage(ctx, q, dt)
    for a, b in zip(got, ref):
      assert torch.equal(a, b)
  ctx.set_tuning(base)
  with pytest.raises(N.NativeError):
    ctx.set_tuning(dict(base, by_y=8))
  with pytest.raises(N.NativeError):
    ctx.set_tuning(dict(base, seg_x=2))
  with pytest.raises(N.NativeError):
    ctx.set_tuning(dict(base, seg_y=0))
  with pytest.raises(N.NativeError):
    ctx.set_tuning(dict(base, adj_by=5))
  with pytest.raises(N.NativeError):
    ctx.set_tuning(dict(base, adj_seg=3))


def test_autotuner_picks_a_valid_geometry():
  cfg_d, st = configs.kh(n=48)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(cfg_d)).native_context
  q = [torch.from_numpy(st[k]).cuda() for k in ctx.names]
  out = [torch.zeros_like(t) for t in q]
  dt = torch.full((1,), 1e-3, device='cuda')
  best, timings = autotune.tune_context(ctx, q, out, dt, reps=2)
  print(f'[tuning] autotuned {best}', flush=True)
  assert ctx.get_tuning() == best
  assert len(timings) == 2 * len(autotune.MARCH_SPACE) + len(autotune.Y_SPACE) + 4
  assert all(ms > 0 for _, _, ms in timings)


def test_tuning_is_bitwise_neutral_batched():
  cfg, members = configs.riemann_ensemble(6, seed=3, shapes=((20, 18, 16), (33, 24, 17)))
  batch = BatchSolver(SwirlCParameters(dict(cfg)), members, cfl=0.5, use_graph=False)
  batch.step(1)
  torch.cuda.synchronize()
  start = [t.clone() for t in batch.u]

  def run():
    for t, s in zip(batch.u, start):
      t.copy_(s)
    batch.refresh()  # the state was written: refill its halos, reseed the CFL maxima
    batch.step(2)
    torch.cuda.synchronize()
    return [t.clone() for t in batch.u]

  ref = run()
  for t in TUNINGS:
    print(f'[tuning] batched {t}', flush=True)
    batch.set_tuning(t)
    got = run()
    for a, b in zip(got, ref):
      assert torch.equal(a, b)
  best, _ = autotune.tune_batch(batch, reps=1)
  assert batch.get_tuning() == best
  got = run()
  for a, b in zip(got, ref):
    assert torch.equal(a, b)


ADJ_TUNINGS = [
    dict(adj_seg=4, adj_by=1, adj_by_y=2),
    dict(adj_seg=7, adj_by=3, adj_by_y=3),
    dict(adj_seg=64, adj_by=2, adj_by_y=1),
]


def test_adjoint_geometry_changes_no_result():
  """The RHS VJP under several adjoint launch geometries matches the default
  one to a few ulps (the same per-cell sums in the same order; only the
  compiler's multiply-add contraction may differ between inlined copies)."""
  cfg_d, st = configs.kh(n=40)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg_d))).native_context
  q = [torch.from_numpy(st[k]).cuda() for k in ctx.names]
  g = [torch.randn(t.shape, generator=torch.Generator().manual_seed(i)).cuda()
       for i, t in enumerate(q)]

  def vjp():
    lq = [torch.zeros_like(t) for t in q]
    ctx.rhs_vjp(q, g, lq)
    torch.cuda.synchronize()
    return lq

  base = ctx.get_tuning()
  ref = vjp()
  for t in ADJ_TUNINGS:
    ctx.set_tuning(t)
    got = vjp()
    err = max(float((a - b).abs().max() / b.abs().max()) for a, b in zip(got, ref))
    print(f'[tuning] adjoint {t}: max rel diff {err:.1e}', flush=True)
    assert err <= 1e-6
  ctx.set_tuning(base)
  lq = [torch.zeros_like(t) for t in q]
  best, timings = autotune.tune_adjoint(ctx, q, g, lq, reps=1)
  print(f'[tuning] adjoint autotuned {best}', flush=True)
  assert ctx.get_tuning() == best
  assert len(timings) == len(autotune.ADJ_MARCH_SPACE) + len(autotune.ADJ_Y_SPACE)
