"""sc_rk_stage_noalloc (RHS accumulator in the library's static device array)
against sc_rk_stage (the context's scratch fields), through the C ABI.

Both run the same kernels, so every stage output must be bit-for-bit equal.
Two grids in a row, the second smaller, show that the array's leftover
contents from the larger call do not leak into the next one.
"""

import copy
import ctypes

import numpy as np
import pytest
import torch

from swirlc_amd import _native as N
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import grid_cfg, random_state
import test_gpu_rhs_parity as P

pytestmark = pytest.mark.gpu

STAGES = ((N.SC_RK_EULER, 0.0, 0.0), (N.SC_RK_BLEND, 0.75, 0.25),
          (N.SC_RK_BLEND, 1.0 / 3.0, 2.0 / 3.0))


def _stage(name, ctx, form, c0, c1, dt, u0, uk, out):
  fn = getattr(N.lib(), name)
  return fn(ctx._ctx, form, ctypes.c_float(c0), ctypes.c_float(c1),
            ctypes.c_void_p(dt.data_ptr()), N.ptr_array(u0) if u0 is not None else None,
            N.ptr_array(uk), N.ptr_array(out), N.stream_handle())


def _run_both(n, bcs, visc, seed):
  h = 3
  cfg = P._cfg(n, h, 'WENO_5', 'HLLC', visc, bcs)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  st = random_state(n, h, seed=seed)
  u0 = [torch.from_numpy(st[k].copy()).cuda() for k in ctx.names]
  dt = torch.full((1,), 2e-3, device='cuda')
  print(f'[no-alloc] grid {n} {bcs} visc={visc}', flush=True)
  ref = [t.clone() for t in u0]
  got = [t.clone() for t in u0]
  for form, c0, c1 in STAGES:
    ref_out = [torch.zeros_like(t) for t in u0]
    got_out = [torch.zeros_like(t) for t in u0]
    base = None if form == N.SC_RK_EULER else u0
    N.check(_stage('sc_rk_stage', ctx, form, c0, c1, dt, base, ref, ref_out), 'sc_rk_stage')
    N.check(_stage('sc_rk_stage_noalloc', ctx, form, c0, c1, dt, base, got, got_out),
            'sc_rk_stage_noalloc')
    torch.cuda.synchronize()
    for f, name in enumerate(ctx.names):
      assert torch.equal(ref_out[f], got_out[f]), (n, form, name)
      assert torch.equal(ref[f], got[f]), (n, form, name)  # halo fill of uk
    ref, got = ref_out, got_out
  return ctx


def test_noalloc_stage_matches_rk_stage_on_two_calls_in_a_row():
  _run_both((24, 20, 18), 'periodic', True, seed=31)
  _run_both((9, 8, 7), 'neumann', False, seed=32)


def test_noalloc_stage_rejects_grids_beyond_its_array():
  n, h = (200, 200, 200), 3
  cfg = grid_cfg(n, h)
  cfg.update({'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
              'include_diffusion': False, 'bc': P._bcs('periodic', n, h)})
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  st = N.lib().sc_rk_stage_noalloc(ctx._ctx, N.SC_RK_EULER, ctypes.c_float(0.0),
                                   ctypes.c_float(0.0), None, None, None, None, None)
  assert st == 2
  assert 'too large' in N.lib().sc_last_error().decode()
  # The AVG form stays refused, as in sc_rk_stage.
  small = _run_both((8, 7, 6), 'periodic', False, seed=33)
  dummy = [torch.zeros(small.shape, device='cuda') for _ in range(5)]
  dt = torch.zeros(1, device='cuda')
  assert _stage('sc_rk_stage_noalloc', small, N.SC_RK_AVG, 0.0, 0.0, dt, dummy, dummy[::-1],
                [torch.zeros_like(t) for t in dummy]) == 2
  torch.cuda.synchronize()
