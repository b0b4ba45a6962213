"""Chained RK stages (sc_rk_stage_ex): halo cells and the CFL wave speed
produced by the last sweep of a stage.

With SC_STAGE_FILL_OUT a stage leaves its output with the halo cells filled
(where the fill is a pure copy along every dim - periodic, or NEUMANN with
value 0 - the last sweep writes every output cell's images); they must equal
what sc_fill_halos writes, bit for bit, corners included.  The batched
stages (sc_batch_rk_stage_ex) are held to the same, and to per-instance
single-grid stepping.  With speed_dev the stage max-accumulates the wave speed of the
cells it writes; it must equal sc_max_wave_speed on the output bit for bit
(both take the same per-cell function and a maximum, which is exact in any
order).  A Solver stepping through chained stages must reproduce the plain
stepping (halo fill at every stage, a separate wave-speed pass every step),
state and dt, bit for bit.  The oracle pins the wave speed itself
(test_gpu_rhs_parity.py::test_max_wave_speed_matches_oracle) and the halo fill
(test_gpu_rhs_parity.py, tests/test_oracle_reference_cases.py).
"""

import copy
import ctypes

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import _native as N
from swirlc_amd import configs
from swirlc_amd.core.batch import BatchSolver
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.core.solver import Solver
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import grid_cfg, oracle_cfg, random_state
from test_gpu_rhs_parity import _bcs
from test_gpu_scalars import _cfg as scalar_cfg

pytestmark = pytest.mark.gpu


NAMES = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')


def _copy_bcs(name):
  """BC sets whose halo fill is a pure copy along every dim (or, for
  'neumann', one that is not: nonzero NEUMANN increments)."""
  BC = configs.BC
  per, neu0 = {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)}, {0: (BC.NEUMANN, 0.0), 1: (BC.NEUMANN, 0.0)}
  if name == 'neumann0':
    dims = {'x': neu0, 'y': neu0, 'z': neu0}
  elif name == 'copy-mixed':
    dims = {'x': per, 'y': neu0, 'z': neu0}
  elif name == 'copy-mixed-2':
    dims = {'x': neu0, 'y': per, 'z': per}
  else:
    return None
  return {'cell_averages': {v: dims for v in NAMES}}


def _ctx_cfg(n, h, gravity, bcs, scheme='WENO_5', flux='HLLC', visc=True, nscalar=0):
  if nscalar:
    return scalar_cfg(n, scheme, flux, visc, bcs, nscalar=nscalar)
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05) if gravity else (1.0, 1.0, 1.0))
  bc = _copy_bcs(bcs) or _bcs(bcs, n, h)
  cfg.update({'interpolation_scheme': scheme, 'numeric_flux_scheme': flux,
              'include_diffusion': visc, 'bc': bc,
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


@pytest.mark.parametrize('n,bcs,want_fused', [
    ((70, 13, 11), 'periodic', True),    # images written by the last sweep
    ((9, 12, 5), 'periodic', False),     # an extent below 2h: fill after the sweeps
    ((14, 11, 9), 'mixed', False),       # Dirichlet faces: fill after the sweeps
    ((14, 11, 9), 'neumann', False),     # nonzero NEUMANN increments: fill after
    # NEUMANN 0 (clamped images, taken by the batched stages only): the fill
    # after the sweeps.
    ((14, 11, 9), 'neumann0', False),
    ((70, 13, 11), 'copy-mixed', False),
    ((33, 9, 4), 'periodic-h2', True),    # halo width 2 (MUSCL): extents >= 4
])
def test_stage_output_halos_equal_the_fill(n, bcs, want_fused):
  h = 2 if bcs.endswith('-h2') else 3
  if h == 2:
    cfg = _ctx_cfg(n, h, False, 'periodic', scheme='MUSCL', flux='HLL')
  else:
    cfg = _ctx_cfg(n, h, False, bcs)
  st = random_state(n, h, seed=23)
  ctx = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg))).native_context
  q = ctx.fields({k: torch.from_numpy(v.copy()).cuda() for k, v in st.items()})
  u0 = [t.clone() for t in q]
  dt = torch.full((1,), 1e-3, device='cuda')
  fused = ctx.out_images_fused()
  assert fused == want_fused
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
@pytest.mark.parametrize('case', ['kh', 'sod'])
def test_chained_solver_matches_plain_stepping(use_graph, cfl, case):
  # KH: all-periodic (wrapped images); Sod: NEUMANN 0 (the fill after the
  # sweeps of each stage).
  cfg_d, st = configs.kh(n=32) if case == 'kh' else configs.sod(64)
  cfg = SwirlCParameters(copy.deepcopy(cfg_d))
  gpu = {k: torch.from_numpy(v).cuda() for k, v in st.items()}
  chained = Solver(cfg, gpu, cfl=cfl, use_graph=use_graph)
  plain = Solver(cfg, gpu, cfl=cfl, use_graph=use_graph)
  assert chained._chained and chained.ctx.out_images_fused() == (case == 'kh')
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
  print(f'[chain] {case} cfl {cfl}: dt after 6 steps {chained.dt.item()!r}', flush=True)


# The first grids' rows span one, two and three 64-lane waves; (12, 1, 7) has
# a one-cell extent (both clamped sides from one cell), (9, 2, 6) fewer y
# cells than h.
BATCH_SHAPES = {
    'neumann0': [(16, 12, 10), (9, 70, 14), (20, 150, 33), (12, 1, 7), (9, 2, 6),
                 (40, 6, 7)],
    'copy-mixed': [(16, 12, 10), (9, 70, 14), (20, 150, 33), (12, 12, 12), (40, 6, 7)],
    'copy-mixed-2': [(16, 12, 10), (9, 70, 14), (20, 8, 33), (12, 12, 12), (40, 6, 7)],
}


@pytest.mark.parametrize('bcs,visc,scheme,h', [
    ('neumann0', False, 'WENO_5', 3), ('neumann0', True, 'WENO_5', 3),
    ('copy-mixed', False, 'WENO_5', 3), ('copy-mixed', True, 'WENO_5', 3),
    ('copy-mixed-2', False, 'WENO_5', 3), ('copy-mixed-2', True, 'WENO_5', 3),
    ('neumann0', True, 'MUSCL', 2), ('copy-mixed', False, 'WENO_3', 2),
])
def test_batch_chained_stages(bcs, visc, scheme, h):
  """Batched chained stages (sc_batch_rk_stage_ex, halos written by the z
  sweep as images) against the batched stages that fill the halos at every
  stage, against per-instance single-grid stepping, and the state's halos
  against the batched fill."""
  shapes = BATCH_SHAPES[bcs]
  flux = 'HLL' if scheme == 'MUSCL' else 'HLLC'
  members = []
  for k, n in enumerate(shapes):
    cfg = _ctx_cfg(n, h, False, bcs, scheme=scheme, flux=flux, visc=visc)
    members.append(((*n, cfg['dx'], cfg['dy'], cfg['dz']), random_state(n, h, seed=40 + k), cfg))
  inst = [(g, st) for g, st, _ in members]
  chained = BatchSolver(SwirlCParameters(copy.deepcopy(members[0][2])), inst, cfl=0.5)
  plain = BatchSolver(SwirlCParameters(copy.deepcopy(members[0][2])), inst, cfl=0.5)
  assert chained._chained and chained.out_images_fused()
  plain._chained = False
  sl = (slice(h, -h),) * 3
  for _ in range(2):
    chained.step(2)
    plain.step(2)
    torch.cuda.synchronize()
    assert torch.equal(chained.dt, plain.dt)
    for k in range(len(shapes)):
      a, b = chained.state(k), plain.state(k)
      for name in a:
        assert torch.equal(a[name][sl], b[name][sl]), (k, name)
  # The chained state's halos hold the fill of its interior, corners included.
  ref = [t.clone() for t in chained.u]
  N.check(N.lib().sc_batch_fill_halos(chained._b, N.ptr_array(ref), N.stream_handle()),
          'sc_batch_fill_halos')
  torch.cuda.synchronize()
  for a, b in zip(chained.u, ref):
    assert torch.equal(a, b)
  dts = chained.dt.cpu().numpy()
  for k, (g, st, cfg) in enumerate(members):
    solo = Solver(SwirlCParameters(copy.deepcopy(cfg)),
                  {n: torch.from_numpy(v).cuda() for n, v in st.items()}, cfl=0.5,
                  use_graph=False)
    solo.step(4)
    torch.cuda.synchronize()
    assert float(solo.dt[0]) == float(dts[k]), k
    got = chained.state(k)
    for name, t in solo.state.items():
      assert torch.equal(got[name][sl], t[sl]), (k, name)
  print(f'[batch chain] {bcs} visc {visc} {scheme} h={h}: {len(shapes)} instances, dt {dts}',
        flush=True)


def test_batch_stage_ex_rejects_unknown_flags():
  n, h = (8, 8, 8), 3
  cfg = _ctx_cfg(n, h, False, 'neumann0')
  batch = BatchSolver(SwirlCParameters(copy.deepcopy(cfg)),
                      [((*n, cfg['dx'], cfg['dy'], cfg['dz']), random_state(n, h, seed=3))],
                      cfl=None)
  rc = N.lib().sc_batch_rk_stage_ex(batch._b, N.SC_RK_EULER, 0.0, 0.0,
                                    ctypes.c_void_p(batch.dt.data_ptr()), None,
                                    N.ptr_array(batch.u), N.ptr_array(batch.s1), 4, None,
                                    N.stream_handle())
  assert rc != 0
