"""Adjoint kernels (sc_rhs_vjp, sc_rk_stage_vjp) and the checkpointed rollout.

The reference has no adjoint of its own (it would differentiate its graph with
tf.GradientTape, which is not available here), so the gradient is pinned to
the mathematics: directional derivatives <grad J, v> from the GPU adjoint are
compared with FP64 central finite differences (J(u + eps v) - J(u - eps v)) /
(2 eps) of the fp64 oracle along seeded random directions v, with bound
|adj - fd| / |fd| <= 1e-4.  eps is the largest of 1e-6, 1e-7, 1e-8 (times the
field scale) whose difference quotient agrees with the next smaller one, so a
step across a non-smooth point of the scheme is never used.  At sizes the oracle cannot reach, the tests check
properties that hold at any size: checkpoint-interval invariance (bitwise),
linearity in the cotangent, determinism, zero halo gradients.
"""

import copy

import numpy as np
import pytest
import torch

from oracle import swirlc_oracle as O
from swirlc_amd import _native as N
from swirlc_amd.boundary import bc_types
from swirlc_amd.core import adjoint as A
from swirlc_amd.core.parameter import SwirlCParameters
from swirlc_amd.equations import compressible_navier_stokes as cns

from helpers import grid_cfg, oracle_cfg, random_state

pytestmark = pytest.mark.gpu
BC = bc_types.BoundaryCondition
NAMES = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')
FD_TOL = 1e-4


def _bcs(kind):
  if kind == 'periodic':
    spec = {d: {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)} for d in 'xyz'}
    return {'cell_averages': {k: spec for k in NAMES}}
  if kind == 'neumann':
    return {'cell_averages': {
        k: {d: {0: (BC.NEUMANN, 0.01 * i), 1: (BC.NEUMANN, -0.01 * i)} for d in 'xyz'}
        for i, k in enumerate(NAMES)}}
  if kind == 'mixed':
    wall = {'rho': 1.1, 'rho_u': 0.2, 'rho_v': 0.0, 'rho_w': 0.0, 'rho_e': 6.0}
    return {'cell_averages': {k: {
        'x': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
        'y': {0: (BC.DIRICHLET, [wall[k], wall[k] * 1.01, wall[k] * 1.02]),
              1: (BC.NEUMANN, 0.0)},
        'z': {0: (BC.NEUMANN, 0.0), 1: (BC.DIRICHLET, wall[k])}} for k in NAMES}}
  if kind == 'prim_mixed':
    # rho_u and rho_e with their BCs on the primitives u and e
    # (boundary.py:516-531), the others on the conserved fields.
    wall = {'rho': 1.1, 'u': 0.18, 'rho_v': 0.0, 'rho_w': 0.0, 'e': 5.5}
    return {'cell_averages': {k: {
        'x': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
        'y': {0: (BC.DIRICHLET, [wall[k], wall[k] * 1.01, wall[k] * 1.02]),
              1: (BC.NEUMANN, 0.01)},
        'z': {0: (BC.NEUMANN, -0.02), 1: (BC.DIRICHLET, wall[k])}} for k in wall}}
  if kind == 'prim_neumann':
    return {'cell_averages': {
        k: {d: {0: (BC.NEUMANN, 0.01 * i), 1: (BC.NEUMANN, -0.01 * i)} for d in 'xyz'}
        for i, k in enumerate(('rho', 'u', 'v', 'rho_w', 'e'))}}
  raise ValueError(kind)


def _cfg(n, scheme='WENO_5', flux='HLLC', visc=True, bcs='periodic', gravity=False,
         nu=1e-2):
  cfg = grid_cfg(n, 3, (1.0, 1.0, 0.05) if gravity else (1.0, 1.0, 1.0))
  cfg.update({'interpolation_scheme': scheme, 'numeric_flux_scheme': flux,
              'include_diffusion': visc, 'bc': _bcs(bcs),
              'transport_parameters': {'nu': nu, 'pr': 0.7}, 'dt': 2e-3})
  if gravity:
    cfg.update({'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity']})
  return cfg


def _ctx(cfg):
  fn = cns.get_rhs_fn(None, SwirlCParameters(copy.deepcopy(cfg)))
  return fn.native_context


def _gpu(st):
  return [torch.from_numpy(np.ascontiguousarray(st[k])).cuda() for k in NAMES]


def _directions(st, h, count, seed):
  """Random directions on the interior, each field scaled by its mean size."""
  rng = np.random.default_rng(seed)
  sl = (slice(h, -h),) * 3
  out = []
  for _ in range(count):
    v = {k: np.zeros(st[k].shape) for k in NAMES}
    for k in NAMES:
      v[k][sl] = rng.standard_normal(v[k][sl].shape) * np.abs(st[k][sl]).mean()
    out.append(v)
  return out


def _fd_at(J, st, v, eps):
  sp = {k: st[k].astype(np.float64) + eps * v[k] for k in NAMES}
  sm = {k: st[k].astype(np.float64) - eps * v[k] for k in NAMES}
  return (J(sp) - J(sm)) / (2.0 * eps)


def _fd(J, st, v, steps=(1e-6, 1e-7, 1e-8), agree=1e-5):
  """Central difference at the largest step that agrees with the next smaller
  one.  The RHS is only piecewise smooth (limiter, min/max, HLLC branch and
  contact-speed snap); a step that crosses one of those points gives a
  difference quotient that no derivative matches, and the disagreement
  between consecutive steps detects it."""
  vals = [_fd_at(J, st, v, e) for e in steps]
  for a, b in zip(vals, vals[1:]):
    if abs(a - b) <= agree * abs(b):
      return a
  raise AssertionError(f'finite differences did not converge: {vals}')


def _dot(grad, v):
  return sum(float(np.sum(grad[i].astype(np.float64) * v[k])) for i, k in enumerate(NAMES))


RHS_CASES = [
    ('WENO_5', 'HLLC', True, 'periodic', False),
    ('WENO_5', 'HLLC', False, 'periodic', False),
    ('WENO_5', 'HLL', True, 'neumann', False),
    ('WENO_3', 'HLLC', True, 'mixed', False),
    ('MUSCL', 'HLL', True, 'mixed', True),
    ('MUSCL', 'HLLC', False, 'neumann', False),
    ('WENO_5', 'HLLC', True, 'prim_mixed', False),
    ('MUSCL', 'HLL', False, 'prim_neumann', True),
    ('WENO_3', 'HLL', True, 'prim_neumann', False),
]


@pytest.mark.parametrize('scheme,flux,visc,bcs,gravity', RHS_CASES)
def test_rhs_vjp_matches_fp64_finite_differences(scheme, flux, visc, bcs, gravity):
  n, h = (12, 10, 9), 3
  cfg = _cfg(n, scheme, flux, visc, bcs, gravity)
  pe = O.gravitational_potential(oracle_cfg(cfg), np.float64) if gravity else None
  st = random_state(n, h, seed=3, jump=False, pe=pe)
  sl = (slice(h, -h),) * 3
  rng = np.random.default_rng(11)
  g = {k: np.zeros_like(st[k]) for k in NAMES}
  for k in NAMES:
    g[k][sl] = rng.standard_normal(g[k][sl].shape).astype(np.float32)
  ctx = _ctx(cfg)
  q = _gpu(st)
  lq = [torch.zeros_like(t) for t in q]
  print(f'[adjoint] rhs vjp {scheme} {flux} visc={visc} {bcs} gravity={gravity}', flush=True)
  ctx.rhs_vjp(q, _gpu(g), lq)
  torch.cuda.synchronize()
  grad = [t.cpu().numpy() for t in lq]
  for t in grad:
    halo = t.copy()
    halo[sl] = 0.0
    assert not np.any(halo), 'halo cells of the adjoint must be zero'
  ocfg = oracle_cfg(cfg)

  def J(s):
    r = O.rhs(s, {}, ocfg)
    return sum(float(np.sum(g[k][sl].astype(np.float64) * r[k][sl])) for k in NAMES)

  errs = []
  for v in _directions(st, h, 4, seed=21):
    fd = _fd(J, st, v)
    ad = _dot(grad, v)
    errs.append(abs(ad - fd) / abs(fd))
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert max(errs) <= FD_TOL


BV_CASES = [
    ('WENO_5', 'HLLC', True, False, False),
    ('WENO_3', 'HLLC', False, True, False),
    ('MUSCL', 'HLL', False, True, True),
    ('WENO_5', 'HLL', True, False, True),
]


@pytest.mark.parametrize('scheme,flux,visc,faces,planes', BV_CASES)
def test_rhs_vjp_with_boundary_values_matches_fp64_finite_differences(scheme, flux, visc,
                                                                      faces, planes):
  """Face and flux Dirichlet values (boundary.py:89-461): a fixed flux
  component passes no adjoint, a fixed primitive face value neither, and a
  fixed conserved face value q_f / rho_f reaches the face density; values
  as scalars and as planes."""
  import test_gpu_boundary_values as BV
  n, h = (12, 10, 9), 3
  cfg = _cfg(n, scheme, flux, visc, 'mixed')
  fluxes = BV._plane_fluxes(n, h, 80) if planes else BV._flux_spec()
  cfg['bc']['intercell_fluxes'] = fluxes
  if faces:
    cfg['bc']['cell_faces'] = BV._planes(BV._face_spec(), n, h, 81) if planes else \
        BV._face_spec()
  st = random_state(n, h, seed=4, jump=False)
  sl = (slice(h, -h),) * 3
  rng = np.random.default_rng(12)
  g = {k: np.zeros_like(st[k]) for k in NAMES}
  for k in NAMES:
    g[k][sl] = rng.standard_normal(g[k][sl].shape).astype(np.float32)
  ctx = _ctx(cfg)
  q = _gpu(st)
  lq = [torch.zeros_like(t) for t in q]
  print(f'[adjoint] rhs vjp with boundary values {scheme} {flux} visc={visc} faces={faces}'
        f' planes={planes}', flush=True)
  ctx.rhs_vjp(q, _gpu(g), lq)
  torch.cuda.synchronize()
  grad = [t.cpu().numpy() for t in lq]
  ocfg = oracle_cfg(cfg)

  def J(s):
    r = O.rhs(s, {}, ocfg)
    return sum(float(np.sum(g[k][sl].astype(np.float64) * r[k][sl])) for k in NAMES)

  errs = []
  for v in _directions(st, h, 4, seed=22):
    fd = _fd(J, st, v)
    ad = _dot(grad, v)
    errs.append(abs(ad - fd) / abs(fd))
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert max(errs) <= FD_TOL


@pytest.mark.parametrize('scheme,flux,visc,frac', [('WENO_5', 'HLLC', True, 0.5),
                                                   ('MUSCL', 'HLL', True, 0.9),
                                                   ('WENO_3', 'HLLC', False, 0.7)])
def test_rhs_vjp_with_reference_state_matches_fp64_finite_differences(scheme, flux, visc,
                                                                      frac):
  """Hydrostatic reference state (helper_vars p_ref, rho_ref;
  compressible_navier_stokes.py:273-335): p_dyn = p - p_ref reconstructed and
  used in the momentum flux and, for HLLC, in the wave speeds and star
  states (hllc.py:109-120); rho - rho_ref in the gravity source.  p_ref and
  rho_ref are constants of the derivative."""
  import test_gpu_reference_state as RS
  n, h = (12, 10, 9), 3
  cfg, ocfg, st, helper = RS._setup(n, scheme, flux, visc, 44, frac)
  cfg['transport_parameters'] = {'nu': 1e-2, 'pr': 0.7}
  ocfg = oracle_cfg(cfg)
  sl = (slice(h, -h),) * 3
  rng = np.random.default_rng(13)
  g = {k: np.zeros_like(st[k]) for k in NAMES}
  for k in NAMES:
    g[k][sl] = rng.standard_normal(g[k][sl].shape).astype(np.float32)
  ctx = _ctx(cfg)
  ctx.set_reference_state(*(torch.from_numpy(helper[k]).cuda() for k in ('p_ref', 'rho_ref')))
  q = _gpu(st)
  lq = [torch.zeros_like(t) for t in q]
  print(f'[adjoint] rhs vjp with reference state {scheme} {flux} visc={visc} frac={frac}',
        flush=True)
  ctx.rhs_vjp(q, _gpu(g), lq)
  torch.cuda.synchronize()
  grad = [t.cpu().numpy() for t in lq]
  hv = {k: v.astype(np.float64) for k, v in helper.items()}

  def J(s):
    r = O.rhs(s, hv, ocfg)
    return sum(float(np.sum(g[k][sl].astype(np.float64) * r[k][sl])) for k in NAMES)

  errs = []
  for v in _directions(st, h, 4, seed=23):
    fd = _fd(J, st, v)
    ad = _dot(grad, v)
    errs.append(abs(ad - fd) / abs(fd))
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert max(errs) <= FD_TOL


@pytest.mark.parametrize('nscalar,scheme,flux,visc,bcs', [
    (1, 'WENO_5', 'HLLC', True, 'periodic'),
    (2, 'WENO_5', 'HLL', False, 'bounded'),
    (2, 'WENO_3', 'HLLC', True, 'bounded'),
    (1, 'MUSCL', 'HLL', True, 'periodic'),
    (3, 'WENO_5', 'HLLC', True, 'periodic'),
    (4, 'WENO_3', 'HLL', False, 'bounded'),
])
def test_rhs_vjp_with_scalars_matches_fp64_finite_differences(nscalar, scheme, flux, visc,
                                                               bcs):
  """Transported scalars (compressible_navier_stokes.py:419-427, hllc.py:182-192,
  hll.py): advected as m * y with their own reconstruction; 'bounded' gives
  the first scalar its BC on the primitive.  Three and four scalars go
  through the chunked context (parts of two scalars each)."""
  import test_gpu_scalars as SC
  n, h = (12, 10, 9), 3
  names = SC._names(nscalar)
  cfg = SC._cfg(n, scheme, flux, visc, bcs, nscalar)
  cfg['transport_parameters'] = {'nu': 1e-2, 'pr': 0.7}
  st = random_state(n, h, seed=5, jump=False, nscalar=nscalar)
  sl = (slice(h, -h),) * 3
  rng = np.random.default_rng(14)
  g = {k: np.zeros_like(st[k]) for k in names}
  for k in names:
    g[k][sl] = rng.standard_normal(g[k][sl].shape).astype(np.float32)
  ctx = _ctx(cfg)
  gpu = lambda d: [torch.from_numpy(np.ascontiguousarray(d[k])).cuda() for k in names]
  q = gpu(st)
  lq = [torch.zeros_like(t) for t in q]
  print(f'[adjoint] rhs vjp with {nscalar} scalars {scheme} {flux} visc={visc} {bcs}',
        flush=True)
  ctx.rhs_vjp(q, gpu(g), lq)
  torch.cuda.synchronize()
  grad = [t.cpu().numpy() for t in lq]
  ocfg = oracle_cfg(cfg)

  def J(s):
    r = O.rhs(s, {}, ocfg)
    return sum(float(np.sum(g[k][sl].astype(np.float64) * r[k][sl])) for k in names)

  def fd_at(v, eps):
    sp = {k: st[k].astype(np.float64) + eps * v[k] for k in names}
    sm = {k: st[k].astype(np.float64) - eps * v[k] for k in names}
    return (J(sp) - J(sm)) / (2.0 * eps)

  # Directions along which the difference quotients settle (a direction
  # whose quotients drift crosses a kink of the scheme within 1e-6: limiter,
  # min / max, HLLC branch or contact snap, where no derivative exists);
  # four of at most eight must settle.
  drng = np.random.default_rng(24)
  errs, unsettled = [], 0
  for _ in range(8):
    v = {k: np.zeros(st[k].shape) for k in names}
    for k in names:
      v[k][sl] = drng.standard_normal(v[k][sl].shape) * np.abs(st[k][sl]).mean()
    vals = [fd_at(v, e) for e in (1e-6, 1e-7, 1e-8)]
    fd = next((a for a, b in zip(vals, vals[1:]) if abs(a - b) <= 1e-5 * abs(b)), None)
    if fd is None:
      unsettled += 1
      continue
    ad = sum(float(np.sum(grad[i].astype(np.float64) * v[k])) for i, k in enumerate(names))
    errs.append(abs(ad - fd) / abs(fd))
    if len(errs) == 4:
      break
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}, {unsettled} unsettled',
        flush=True)
  assert len(errs) == 4
  assert max(errs) <= FD_TOL


def test_chunked_rk_stage_vjp_matches_fp64_finite_differences():
  """rk_stage_vjp of a context with three scalars (two parts), BLEND form with
  lam_uk aliasing lam_out as the rollout calls it: lam_uk = c1 (lam +
  dt J^T lam), lam_u0 += c0 lam, against FD of the oracle's stage."""
  import test_gpu_scalars as SC
  n, h, nscalar = (12, 10, 9), 3, 3
  names = SC._names(nscalar)
  cfg = SC._cfg(n, 'WENO_5', 'HLLC', True, 'periodic', nscalar)
  cfg['transport_parameters'] = {'nu': 1e-2, 'pr': 0.7}
  st = random_state(n, h, seed=6, jump=False, nscalar=nscalar)
  sl = (slice(h, -h),) * 3
  rng = np.random.default_rng(15)
  lam = {k: np.zeros_like(st[k]) for k in names}
  for k in names:
    lam[k][sl] = rng.standard_normal(lam[k][sl].shape).astype(np.float32)
  ctx = _ctx(cfg)
  assert len(ctx.parts) == 2
  gpu = lambda d: [torch.from_numpy(np.ascontiguousarray(d[k])).cuda() for k in names]
  dt, c0, c1 = 2e-3, 0.75, 0.25
  dt_dev = torch.full((1,), dt, dtype=torch.float32, device='cuda')
  uk = gpu(st)
  l = gpu(lam)
  l0 = [torch.zeros_like(t) for t in l]
  ctx.rk_stage_vjp(N.SC_RK_BLEND, c0, c1, dt_dev, uk, l, l, l0)
  torch.cuda.synchronize()
  grad = [t.cpu().numpy() for t in l]
  for t, k in zip(l0, names):
    np.testing.assert_allclose(t.cpu().numpy()[sl], c0 * lam[k][sl], rtol=1e-6)
  ocfg = oracle_cfg(cfg)

  def J(s):
    r = O.rhs(s, {}, ocfg)
    return sum(float(np.sum(lam[k][sl].astype(np.float64) *
                            (c1 * (s[k][sl] + dt * r[k][sl])))) for k in names)

  def fd_at(v, eps):
    sp = {k: st[k].astype(np.float64) + eps * v[k] for k in names}
    sm = {k: st[k].astype(np.float64) - eps * v[k] for k in names}
    return (J(sp) - J(sm)) / (2.0 * eps)

  drng = np.random.default_rng(25)
  errs = []
  for _ in range(8):
    v = {k: np.zeros(st[k].shape) for k in names}
    for k in names:
      v[k][sl] = drng.standard_normal(v[k][sl].shape) * np.abs(st[k][sl]).mean()
    vals = [fd_at(v, e) for e in (1e-6, 1e-7, 1e-8)]
    fd = next((a for a, b in zip(vals, vals[1:]) if abs(a - b) <= 1e-5 * abs(b)), None)
    if fd is None:
      continue
    ad = sum(float(np.sum(grad[i].astype(np.float64) * v[k])) for i, k in enumerate(names))
    errs.append(abs(ad - fd) / abs(fd))
    if len(errs) == 4:
      break
  print(f'[adjoint]   chunked rk stage rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert len(errs) == 4
  assert max(errs) <= FD_TOL


def _kh_small(n, nu=1e-2):
  cfg = _cfg(n, 'WENO_5', 'HLLC', True, 'periodic', nu=nu)
  h = 3
  x = cfg['x'].astype(np.float64)[None, :, None]
  y = cfg['y'].astype(np.float64)[None, None, :]
  shape = (n[2], n[0], n[1])
  # Smoothed shear layer (tanh profile) so the rollout stays differentiable.
  prof = 0.5 * (np.tanh((0.25 - np.abs(y - 0.5)) / 0.05) + 1.0)
  rho = np.broadcast_to(1.0 + prof, shape)
  u = np.broadcast_to(-0.5 + prof, shape)
  rng = np.random.default_rng(0)
  v = np.broadcast_to(0.01 * np.sin(4 * np.pi * x) + rng.uniform(-1e-3, 1e-3, shape), shape)
  w = np.zeros(shape)
  p = np.full(shape, 2.5)
  e = p / 0.4 / rho + 0.5 * (u * u + v * v)
  st = {}
  for k, f in zip(NAMES, (rho, rho * u, rho * v, rho * w, rho * e)):
    full = np.zeros((n[2] + 2 * h, n[0] + 2 * h, n[1] + 2 * h), np.float32)
    full[h:-h, h:-h, h:-h] = f
    st[k] = full
  return cfg, st


def _rollout_grad(ctx, st, rho_t, h, dt, steps, every):
  u0 = [t.requires_grad_(True) for t in _gpu(st)]
  out = A.rollout(ctx, dt, steps, *u0, every=every)
  J = A.kinetic_density_loss(out, rho_t, h)
  J.backward()
  torch.cuda.synchronize()
  return float(J.detach()), [t.grad.cpu().numpy() for t in u0]


def test_rollout_gradient_matches_fp64_finite_differences():
  # A generic smooth state: the shear-layer setup has w = 0, which puts every
  # z face's HLLC contact speed at the |s_m| < 1e-6 snap (hllc.py:339-341),
  # a discontinuity that finite differences cannot step over.
  n, h, steps, dt = (12, 12, 10), 3, 6, 2e-3
  cfg = _cfg(n, 'WENO_5', 'HLLC', True, 'periodic')
  st = random_state(n, h, seed=5, jump=False)
  ctx = _ctx(cfg)
  rng = np.random.default_rng(1)
  rho_t_np = rng.uniform(0.9, 2.1, st['rho'].shape).astype(np.float32)
  rho_t = torch.from_numpy(rho_t_np).cuda()
  print('[adjoint] rollout gradient 12x12x10, 6 steps, checkpoints every 4', flush=True)
  J32, grad = _rollout_grad(ctx, st, rho_t, h, dt, steps, 4)
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
  for v in _directions(st, h, 4, seed=2):
    fd = _fd(J, st, v)
    ad = _dot(grad, v)
    errs.append(abs(ad - fd) / abs(fd))
  print(f'[adjoint]   rel errors {["%.1e" % e for e in errs]}', flush=True)
  assert max(errs) <= FD_TOL


def test_rollout_gradient_is_checkpoint_invariant_and_deterministic():
  n, h, steps, dt = (32, 24, 16), 3, 7, 1e-3
  cfg, st = _kh_small(n, nu=1e-4)
  ctx = _ctx(cfg)
  rho_t = torch.from_numpy(
      np.random.default_rng(1).uniform(0.9, 2.1, st['rho'].shape).astype(np.float32)).cuda()
  print('[adjoint] checkpoint invariance 32x24x16, 7 steps', flush=True)
  base = _rollout_grad(ctx, st, rho_t, h, dt, steps, 1)
  for every in (3, 7, 100):
    other = _rollout_grad(ctx, st, rho_t, h, dt, steps, every)
    assert other[0] == base[0]
    for a, b in zip(other[1], base[1]):
      assert np.array_equal(a, b)
  sl = (slice(h, -h),) * 3
  for t in base[1]:
    assert np.all(np.isfinite(t))
    halo = t.copy()
    halo[sl] = 0.0
    assert not np.any(halo)


def test_rhs_vjp_is_linear_in_the_cotangent_at_size():
  n, h = (64, 64, 64), 3
  cfg, st = _kh_small(n, nu=1e-4)
  ctx = _ctx(cfg)
  q = _gpu(st)
  gen = torch.Generator(device='cuda').manual_seed(5)
  sl = (slice(h, -h),) * 3

  def cot():
    out = []
    for t in q:
      g = torch.zeros_like(t)
      g[sl] = torch.randn(g[sl].shape, device='cuda', generator=gen)
      out.append(g)
    return out

  g1, g2 = cot(), cot()
  a, b = 0.75, -1.25
  g12 = [a * x + b * y for x, y in zip(g1, g2)]
  l1 = [torch.zeros_like(t) for t in q]
  l2 = [torch.zeros_like(t) for t in q]
  l12 = [torch.zeros_like(t) for t in q]
  print('[adjoint] linearity 64^3', flush=True)
  ctx.rhs_vjp(q, g1, l1)
  ctx.rhs_vjp(q, g2, l2)
  ctx.rhs_vjp(q, g12, l12)
  again = [torch.zeros_like(t) for t in q]
  ctx.rhs_vjp(q, g1, again)
  torch.cuda.synchronize()
  for x, y in zip(l1, again):
    assert torch.equal(x, y)
  for x, y, z in zip(l1, l2, l12):
    ref = a * x.double() + b * y.double()
    err = float((z.double() - ref).abs().max() / ref.abs().max())
    assert err <= 1e-5


def test_rhs_autograd_function_uses_the_adjoint_kernel():
  n, h = (16, 12, 10), 3
  cfg = _cfg(n, 'WENO_5', 'HLLC', True, 'periodic')
  st = random_state(n, h, seed=4, jump=False)
  ctx = _ctx(cfg)
  q = [t.requires_grad_(True) for t in _gpu(st)]
  sl = (slice(h, -h),) * 3
  w = [torch.randn(t.shape, device='cuda', generator=torch.Generator(device='cuda').manual_seed(i))
       for i, t in enumerate(q)]
  r = A.rhs(ctx, *q)
  loss = sum((ri[sl] * wi[sl]).sum() for ri, wi in zip(r, w))
  loss.backward()
  g = [torch.zeros_like(t) for t in q]
  for gi, wi in zip(g, w):
    gi[sl] = wi[sl]
  lq = [torch.zeros_like(t) for t in q]
  ctx.rhs_vjp([t.detach() for t in q], g, lq)
  torch.cuda.synchronize()
  for t, ref in zip(q, lq):
    assert torch.equal(t.grad, ref)
