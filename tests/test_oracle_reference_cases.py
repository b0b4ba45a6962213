"""Pins the CPU oracle to the reference's own golden vectors and closed forms.

Each test restates a reference test case (cited) with the oracle as the
function under test.  These run on the CPU (no GPU needed).
"""

import numpy as np
import pytest

from oracle import swirlc_oracle as O

DIMS = ('x', 'y', 'z')


def to_3d(f, dim, size):
  """common/testing_utils.py:23-62 (tile a 1-D profile into [nz, nx, ny])."""
  f = np.asarray(f)
  if dim == 'x':
    return np.tile(f[np.newaxis, :, np.newaxis], (size[0], 1, size[2]))
  if dim == 'y':
    return np.tile(f[np.newaxis, np.newaxis, :], (size[0], size[1], 1))
  return np.tile(f[:, np.newaxis, np.newaxis], (1, size[1], size[2]))


def profile(u, dim):
  nz, nx, ny = u.shape
  if dim == 'x':
    return u[nz // 2, :, ny // 2]
  if dim == 'y':
    return u[nz // 2, nx // 2, :]
  return u[:, nx // 2, ny // 2]


# --- numerics/gradient_test.py:53-127, interpolation_test.py:30-52 ----------

def _sine(dim, n=16, n_normal=4):
  x = np.linspace(0, 2.0 * np.pi, n)
  y = np.sin(x)
  if dim == 'x':
    y = np.tile(y[np.newaxis, :, np.newaxis], [n_normal, 1, n_normal])
  elif dim == 'y':
    y = np.tile(y[np.newaxis, np.newaxis, :], [n_normal, n_normal, 1])
  else:
    y = np.tile(y[:, np.newaxis, np.newaxis], [1, n_normal, n_normal])
  return y.astype(np.float32), x[1] - x[0]


@pytest.mark.parametrize('dim', DIMS)
def test_forward_1_golden(dim):
  u, h = _sine(dim)
  expected = [0.97101221, 0.80311538, 0.4963526, 0.10376595, -0.30676278,
              -0.66424943, -0.90688133, -0.9927052, -0.90688133, -0.66424943,
              -0.30676278, 0.10376595, 0.4963526, 0.80311538, 0.97101221]
  np.testing.assert_allclose(profile(O.forward_1(u, h, dim), dim)[:-1], expected,
                             rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize('dim', DIMS)
def test_central_2_golden(dim):
  u, h = _sine(dim)
  expected = [0.88706379, 0.64973399, 0.30005927, -0.10149841, -0.4855061,
              -0.78556538, -0.94979326, -0.94979326, -0.78556538, -0.4855061,
              -0.10149841, 0.30005927, 0.64973399, 0.88706379]
  np.testing.assert_allclose(profile(O.central_2(u, h, dim), dim)[1:-1], expected,
                             rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize('dim', DIMS)
def test_backward_1_closed_form(dim):
  nx = 16
  x = np.linspace(0.0, 1.0, nx)
  f = np.sin(x)
  res = O.backward_1(to_3d(f, dim, [nx] * 3).astype(np.float32), np.diff(x)[0], dim)
  exp = np.zeros(nx)
  exp[1:] = np.diff(f) / np.diff(x)
  np.testing.assert_allclose(res[1:, 1:, 1:], to_3d(exp, dim, [nx] * 3)[1:, 1:, 1:],
                             rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize('dim', DIMS)
def test_linear_interpolation_closed_form(dim):
  nx = 16
  f = np.logspace(0.0, 2.0, nx)
  u = to_3d(f, dim, [nx] * 3).astype(np.float32)
  for face_dim in DIMS:
    exp = np.zeros(nx)
    if face_dim == dim:
      exp[1:] = f[:-1] + 0.5 * np.diff(f)
    else:
      exp = f
    np.testing.assert_allclose(O.linear_interpolation(u, face_dim)[1:, 1:, 1:],
                               to_3d(exp, dim, [nx] * 3)[1:, 1:, 1:], rtol=1e-6)


# --- boundary/boundary_test.py:61-341 (single-replica expectations) ---------

def test_exchange_halos_golden():
  n = 16
  u_bc_z0 = [np.ones((16, 16)), 2.0 * np.ones((16, 16)), 3.0 * np.ones((16, 16))]
  p_bc_x1 = [[60.0 * np.ones((1, n))] * n, [80.0 * np.ones((1, n))] * n,
             [120.0 * np.ones((1, n))] * n]
  p_bc_x1 = [np.stack(layer) for layer in p_bc_x1]  # [nz, 1, ny] planes
  bc_u = {'x': {0: (O.DIRICHLET, 8.0), 1: (O.NEUMANN, 0.0)},
          'y': {0: (O.PERIODIC,), 1: (O.PERIODIC,)},
          'z': {0: (O.DIRICHLET, u_bc_z0), 1: (O.NEUMANN, 1.0)}}
  bc_p = {'x': {0: (O.NEUMANN, 8.0), 1: (O.DIRICHLET, p_bc_x1)},
          'y': {0: (O.PERIODIC,), 1: (O.PERIODIC,)},
          'z': {0: (O.DIRICHLET, 100.0), 1: (O.NEUMANN, 0.0)}}
  u0 = np.zeros((n, n, n), np.float32)
  u0[3:-3, 3:-3, 3:-3] = 6.0
  p0 = np.zeros((n, n, n), np.float32)
  p0[3:-3, 3:-3, 3:-3] = 200.0
  u = O.exchange_halos(u0, bc_u, 3)
  p = O.exchange_halos(p0, {'x': {0: bc_p['x'][0], 1: (O.DIRICHLET, [l[:, 0, :] for l in p_bc_x1])},
                             'y': bc_p['y'], 'z': bc_p['z']}, 3)
  c = slice(3, -3)
  assert (u[c, c, c] == 6.0).all()
  assert (u[c, :3, c] == 8.0).all()
  assert (u[c, -3:, c] == 6.0).all()
  assert (u[c, c, :3] == 6.0).all() and (u[c, c, -3:] == 6.0).all()
  np.testing.assert_array_equal(u[:3, c, c][:, 0, 0], [1.0, 2.0, 3.0])
  np.testing.assert_array_equal(u[-3:, c, c][:, 0, 0], [7.0, 8.0, 9.0])
  assert (p[c, c, c] == 200.0).all()
  np.testing.assert_array_equal(p[c, :3, c][0, :, 0], [176.0, 184.0, 192.0])
  np.testing.assert_array_equal(p[c, -3:, c][0, :, 0], [60.0, 80.0, 120.0])
  assert (p[c, c, :3] == 200.0).all() and (p[c, c, -3:] == 200.0).all()
  assert (p[:3, c, c] == 100.0).all()
  assert (p[-3:, c, c] == 200.0).all()


def test_cell_average_bc_on_primitive_rounds_interior():
  """boundary.py:516-531: q -> rho * divide_no_nan(q, rho) everywhere."""
  cfg = O.default_cfg()
  rng = np.random.default_rng(0)
  st = {k: rng.uniform(0.5, 2.0, (16, 16, 16)).astype(np.float32) for k in O.BASE_CONSERVATIVE}
  bc = dict(cfg['bc']['cell_averages'])
  bc['u'] = bc.pop('rho_u')
  cfg['bc'] = {'cell_averages': bc}
  out = O.update_conservative_cell_averages(st, cfg)
  c = slice(3, -3)
  exp = st['rho'][c, c, c] * (st['rho_u'][c, c, c] / st['rho'][c, c, c])
  np.testing.assert_array_equal(out['rho_u'][c, c, c], exp)
  with pytest.raises(ValueError):
    bc.pop('u')
    O.update_conservative_cell_averages(st, cfg)


# --- equations/compressible_navier_stokes_test.py:286-779 -------------------

def normal_shock(m):
  p1 = (2.0 * O.GAMMA * m ** 2 - (O.GAMMA - 1.0)) / (O.GAMMA + 1.0)
  r1 = (O.GAMMA + 1.0) * m ** 2 / ((O.GAMMA - 1.0) * m ** 2 + 2.0)
  return p1, r1, p1 / r1


def _states(rho, u, v, w, p, t, y=None, dim='x', size=(16, 16, 16)):
  e = O.CV * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  h = O.CP * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  prim = {'rho': rho, 'u': u, 'v': v, 'w': w, 'e': e, 'p': p, 'h': h}
  cons = {'rho': rho, 'rho_u': rho * u, 'rho_v': rho * v, 'rho_w': rho * w,
          'rho_e': rho * e}
  if y is not None:
    prim['y'] = y
    cons['rho_y'] = rho * y
  T = lambda f: to_3d(f, dim, list(size)).astype(np.float64)
  return {k: T(v) for k, v in cons.items()}, {k: T(v) for k, v in prim.items()}


@pytest.mark.parametrize('dim', DIMS)
def test_physical_flux_closed_form(dim):
  n = 16
  u = np.linspace(-3.0, 12.0, n)
  v = np.linspace(6.0, -24.0, n)
  w = np.linspace(-9.0, 36.0, n)
  p = np.logspace(3.0, 6.0, n)
  t = np.linspace(250.0, 3000.0, n)
  y = np.linspace(0.0, 1.0, n)
  rho = p / (O.R * t)
  cons, prim = _states(rho, u, v, w, p, t, y, dim)
  e = O.CV * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  un = {'x': u, 'y': v, 'z': w}[dim]
  exp = {'rho': rho * un, 'rho_u': rho * u * un + (p if dim == 'x' else 0),
         'rho_v': rho * v * un + (p if dim == 'y' else 0),
         'rho_w': rho * w * un + (p if dim == 'z' else 0),
         'rho_e': rho * (e + p / rho) * un, 'rho_y': rho * y * un}
  for var, ev in exp.items():
    np.testing.assert_allclose(O.physical_flux_1d(cons, prim, {}, dim, var),
                               to_3d(ev, dim, [n] * 3), rtol=1e-6)


NAMES = ['rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e', 'rho_y']


@pytest.mark.parametrize('flux', ['HLL', 'HLLC'])
@pytest.mark.parametrize('dim', DIMS)
def test_intercell_flux_quiescent_gas(flux, dim):
  n = 16
  z = np.zeros(n)
  p = 101325.0 * np.ones(n)
  t = 500.0 * np.ones(n)
  cons, prim = _states(p / (O.R * t), z, z, z, p, t, 0.5 * np.ones(n), dim)
  pe = np.zeros((1, 1, 1))
  for fd in DIMS:
    res = O.numeric_flux_fn(flux)(cons, cons, prim, prim, {}, {}, fd, NAMES, pe)
    for var in NAMES:
      exp = p if var == 'rho_' + {'x': 'u', 'y': 'v', 'z': 'w'}[fd] else np.zeros(n)
      np.testing.assert_allclose(res[var][3:-2, 3:-2, 3:-2],
                                 to_3d(exp, dim, [n] * 3)[3:-2, 3:-2, 3:-2],
                                 rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize('flux', ['HLL', 'HLLC'])
@pytest.mark.parametrize('dim', DIMS)
def test_intercell_flux_plug_flow(flux, dim):
  n = 16
  z = np.zeros(n)
  vc = {'x': 10.0, 'y': -25.0, 'z': 15.0}[dim]
  u, v, w = z.copy(), z.copy(), z.copy()
  {'x': u, 'y': v, 'z': w}[dim][:] = vc
  p = 101325.0 * np.ones(n)
  t = np.linspace(250.0, 3000.0, n)
  y = np.linspace(0.0, 1.0, n)
  rho = p / (O.R * t)
  cons, prim = _states(rho, u, v, w, p, t, y, dim)
  h = O.CP * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  pe = np.zeros((1, 1, 1))
  for fd in DIMS:
    res = O.numeric_flux_fn(flux)(cons, cons, prim, prim, {}, {}, fd, NAMES, pe)
    if fd == dim:
      exp = {'rho': rho * vc, 'rho_u': rho * u * vc, 'rho_v': rho * v * vc,
             'rho_w': rho * w * vc, 'rho_e': rho * h * vc, 'rho_y': rho * y * vc}
    else:
      exp = {k: np.zeros(n) for k in NAMES}
    exp = dict(exp)
    key = 'rho_' + {'x': 'u', 'y': 'v', 'z': 'w'}[fd]
    exp[key] = exp[key] + p
    for var in NAMES:
      np.testing.assert_allclose(res[var][3:-2, 3:-2, 3:-2],
                                 to_3d(exp[var], dim, [n] * 3)[3:-2, 3:-2, 3:-2],
                                 rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize('flux', ['HLL', 'HLLC'])
@pytest.mark.parametrize('dim', DIMS)
def test_intercell_flux_standing_shock(flux, dim):
  n = 16
  m, t0, p0 = 6.0, 300.0, 101325.0
  rho0 = p0 / (O.R * t0)
  u0 = m * np.sqrt(O.GAMMA * O.R * t0)
  h0 = O.CP * t0 + 0.5 * u0 ** 2
  _, r1, t1 = normal_shock(m)
  hv = lambda a, b: np.where(np.arange(n) < n // 2, a, b)
  rho = hv(rho0, r1 * rho0)
  t = hv(t0, t1 * t0)
  p = rho * O.R * t
  z = np.zeros(n)
  u, v, w = z.copy(), z.copy(), z.copy()
  {'x': u, 'y': v, 'z': w}[dim][:] = hv(u0, u0 / r1)
  cons, prim = _states(rho, u, v, w, p, t, 0.5 * np.ones(n), dim)
  pe = np.zeros((1, 1, 1))
  for fd in DIMS:
    res = O.numeric_flux_fn(flux)(cons, cons, prim, prim, {}, {}, fd, NAMES, pe)
    key = 'rho_' + {'x': 'u', 'y': 'v', 'z': 'w'}[fd]
    if fd == dim:
      exp = {'rho': rho0 * u0 * np.ones(n), 'rho_u': z, 'rho_v': z, 'rho_w': z,
             'rho_e': rho0 * h0 * u0 * np.ones(n), 'rho_y': rho0 * 0.5 * u0 * np.ones(n)}
      exp = dict(exp)
      exp[key] = exp[key] + (rho0 * u0 ** 2 + p0)
    else:
      exp = {k: z for k in NAMES}
      exp = dict(exp)
      exp[key] = exp[key] + p
    for var in NAMES:
      np.testing.assert_allclose(res[var][3:-2, 3:-2, 3:-2],
                                 to_3d(exp[var], dim, [n] * 3)[3:-2, 3:-2, 3:-2],
                                 rtol=1e-6, atol=1e-3)


def _cfg16(**kw):
  x = O.cell_centres(0, 1, 10)
  cfg = O.default_cfg(halo_width=3, core_nx=10, core_ny=10, core_nz=10,
                      conservative_variable_names=list(O.BASE_CONSERVATIVE) + ['rho_y'],
                      primitive_variable_names=list(O.BASE_PRIMITIVES) + ['y'])
  cfg['bc'] = {'cell_averages': {v: {d: {0: (O.PERIODIC,), 1: (O.PERIODIC,)} for d in DIMS}
                                 for v in cfg['conservative_variable_names']}}
  cfg.update(kw)
  return cfg


@pytest.mark.parametrize('scheme', ['WENO_3', 'WENO_5', 'MUSCL'])
@pytest.mark.parametrize('flux', ['HLL', 'HLLC'])
def test_rhs_quiescent_gas_is_zero(scheme, flux):
  """compressible_navier_stokes_test.py:927-974 (real WENO instead of a mock)."""
  n = 16
  z = np.zeros(n)
  p = 101325.0 * np.ones(n)
  t = 500.0 * np.ones(n)
  cons, _ = _states(p / (O.R * t), z, z, z, p, t, 0.5 * np.ones(n), 'x')
  cfg = _cfg16(interpolation_scheme=scheme, numeric_flux_scheme=flux)
  res = O.rhs({k: v.astype(np.float64) for k, v in cons.items()}, {}, cfg)
  for var, r in res.items():
    scale = {'rho': 1.0, 'rho_e': 1e5}.get(var, 1.0)
    assert np.abs(r[3:-3, 3:-3, 3:-3]).max() <= 1e-9 * max(scale, 1.0) * 1e3, var


@pytest.mark.parametrize('dim', DIMS)
def test_rhs_plug_flow_is_zero(dim):
  """compressible_navier_stokes_test.py:976-1061: uniform advection, dRHS=0."""
  n = 16
  z = np.zeros(n)
  u, v, w = z.copy(), z.copy(), z.copy()
  {'x': u, 'y': v, 'z': w}[dim][:] = 10.0
  p = 101325.0 * np.ones(n)
  t = 500.0 * np.ones(n)
  cons, _ = _states(p / (O.R * t), u, v, w, p, t, 0.5 * np.ones(n), dim)
  cfg = _cfg16(interpolation_scheme='WENO_5', numeric_flux_scheme='HLLC')
  res = O.rhs(cons, {}, cfg)
  for var, r in res.items():
    assert np.abs(r[3:-3, 3:-3, 3:-3]).max() <= 1e-6 * 1e5, var


def test_rhs_divergence_of_specified_flux():
  """compressible_navier_stokes_test.py:1063-1147: rhs = src - sum_d forward_1(F_d)."""
  n = 16
  rng = np.random.default_rng(3)
  flux = {d: {v: rng.standard_normal((n, n, n)) for v in NAMES} for d in DIMS}
  hs = {'x': 0.1, 'y': 0.2, 'z': 0.3}
  for v in NAMES:
    got = np.zeros((n, n, n))
    for d in DIMS:
      got = got - O.forward_1(flux[d][v], hs[d], d)
    ax = {'x': 1, 'y': 2, 'z': 0}
    exp = -sum((np.roll(flux[d][v], -1, ax[d]) - flux[d][v]) / hs[d] for d in DIMS)
    np.testing.assert_allclose(got[:-1, :-1, :-1], exp[:-1, :-1, :-1], rtol=1e-12, atol=1e-12)


def test_diffusive_flux_isothermal_uniform_is_zero():
  """compressible_navier_stokes_test.py:1979-2027."""
  n = 16
  z = np.zeros(n)
  p = 101325.0 * np.ones(n)
  t = 300.0 * np.ones(n)
  cons, prim = _states(p / (O.R * t), z + 3.0, z - 2.0, z + 1.0, p, t, None, 'x')
  cfg = _cfg16(include_diffusion=True, conservative_variable_names=list(O.BASE_CONSERVATIVE))
  fl = O.diffusive_flux(prim, cfg)
  for d in DIMS:
    for v, f in fl[d].items():
      assert np.abs(f[3:-3, 3:-3, 3:-3]).max() <= 1e-9, (d, v)


@pytest.mark.parametrize('dim', DIMS)
def test_diffusive_flux_heat_conduction(dim):
  """compressible_navier_stokes_test.py:2029-2102: q = -kappa dT/dx only."""
  n = 16
  z = np.zeros(n)
  t = np.linspace(300.0, 600.0, n)
  p = 101325.0 * np.ones(n)
  rho = p / (O.R * t)
  cons, prim = _states(rho, z, z, z, p, t, None, dim)
  cfg = _cfg16(include_diffusion=True, conservative_variable_names=list(O.BASE_CONSERVATIVE))
  h = cfg['d' + dim]
  fl = O.diffusive_flux(prim, cfg)
  nu, pr = cfg['transport_parameters']['nu'], cfg['transport_parameters']['pr']
  kappa = O.CP * nu * rho / pr
  kf = np.zeros(n)
  kf[1:] = 0.5 * (kappa[1:] + kappa[:-1])
  dt = np.zeros(n)
  dt[1:] = np.diff(t) / h
  exp = to_3d(-kf * dt, dim, [n] * 3)
  np.testing.assert_allclose(fl[dim]['rho_e'][3:-3, 3:-3, 3:-3], exp[3:-3, 3:-3, 3:-3],
                             rtol=1e-9)
  for v in ('rho', 'rho_u', 'rho_v', 'rho_w'):
    assert np.abs(fl[dim][v][3:-3, 3:-3, 3:-3]).max() <= 1e-9


def test_strain_rate_taylor_green():
  """physics/fluid_test.py:56-144: S of the Taylor-Green vortex (atol 5e-2)."""
  n = 64
  L = 2 * np.pi
  x = (np.arange(n) + 0.5) * L / n
  Z, X, Y = np.meshgrid(x, x, x, indexing='ij')
  u = np.sin(X) * np.cos(Y) * np.cos(Z)
  v = -np.cos(X) * np.sin(Y) * np.cos(Z)
  w = np.zeros_like(u)
  cfg = {'dx': L / n, 'dy': L / n, 'dz': L / n}
  s = O.strain_rate({'u': u, 'v': v, 'w': w}, cfg)
  # S_xy = 0.5 (du/dy + dv/dx) on y faces; du/dy = -sin x sin y cos z,
  # dv/dx = sin x sin y cos z -> S_xy = 0.
  c = slice(2, -2)
  assert np.abs(s[0][1][c, c, c]).max() < 5e-2
  # S_xx on x faces = du/dx - div/3 = cos x cos y cos z at x faces.
  xf = X - 0.5 * L / n
  exp = np.cos(xf) * np.cos(Y) * np.cos(Z)
  assert np.abs(s[0][0][c, c, c] - exp[c, c, c]).max() < 5e-2


def test_gravity_source_closed_form():
  """physics/additional_models/gravity_test.py:36-101."""
  n = 8
  rho = np.linspace(0.5, 1.5, n ** 3).reshape(n, n, n)
  cons = {k: rho.copy() for k in O.BASE_CONSERVATIVE}
  for g, mask in (({'x': 0.0, 'y': 0.0, 'z': -1.0}, ('rho_w',)),
                  ({'x': 1.0, 'y': 0.0, 'z': 0.0}, ('rho_u',)),
                  ({'x': 0.6, 'y': 0.8, 'z': 0.0}, ('rho_u', 'rho_v', 'rho_w'))):
    src = O.source_function(cons, {}, {'source_functions': ['gravity'], 'g': g})
    for k in O.BASE_CONSERVATIVE:
      if k in mask:
        d = {'rho_u': 'x', 'rho_v': 'y', 'rho_w': 'z'}[k]
        np.testing.assert_allclose(src[k], rho * O.G * g[d], rtol=1e-12)
      else:
        assert (src[k] == 0).all()
  with pytest.raises(ValueError):
    O.source_function(cons, {}, {'source_functions': ['gravity'],
                                 'g': {'x': 0.0, 'y': 0.0, 'z': 0.0}})


# --- numerics/time_integration/rk3_test.py:94-212, rk2_test.py:85-207 ------

def _staged_rhs3(state):
  m = state['rho'].mean()
  f = 20.0 if m < 2.0 else (40.0 if m > 2.75 else -10.0)
  return {k: np.ones_like(v) * f for k, v in state.items()}


def _staged_rhs2(state):
  m = state['rho'].mean()
  f = 20.0 if m < 2.0 else -10.0
  return {k: v * f for k, v in state.items()}


def test_rk3_staged_known_answer():
  st = {k: np.full((4, 4, 4), float(i + 1), np.float32) for i, k in enumerate(NAMES)}
  out = O.rk3(st, {}, {'dt': 0.1}, rhs_fn=_staged_rhs3)
  for i, k in enumerate(NAMES):
    np.testing.assert_allclose(out[k], i + 1 + 1.0 / 3.0, rtol=1e-6)
  zero = O.rk3(st, {}, {'dt': 0.1}, rhs_fn=lambda s: {k: np.zeros_like(v) for k, v in s.items()})
  for k in NAMES:
    np.testing.assert_array_equal(zero[k], st[k])


def test_rk2_staged_known_answer():
  st = {k: np.full((4, 4, 4), float(i + 1), np.float32) for i, k in enumerate(NAMES)}
  out = O.rk2(st, {}, {'dt': 0.1}, rhs_fn=_staged_rhs2)
  for i, k in enumerate(NAMES):
    np.testing.assert_allclose(out[k], 0.5 * (i + 1), rtol=1e-6)


# --- physics/thermodynamics/generic_test.py closed forms -------------------

def test_eos_closed_forms():
  t = np.linspace(250.0, 3000.0, 11)
  p = np.logspace(3.0, 6.0, 11)
  u, v, w = np.linspace(-3, 12, 11), np.linspace(6, -24, 11), np.linspace(-9, 36, 11)
  rho = p / (O.R * t)
  e = O.CV * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  s = {'rho': rho, 'u': u, 'v': v, 'w': w, 'e': e}
  np.testing.assert_allclose(O.pressure(s, 0.0), p, rtol=1e-10)
  np.testing.assert_allclose(O.total_enthalpy(s, 0.0), O.CP * t + 0.5 * (u**2 + v**2 + w**2),
                             rtol=1e-10)
  s['p'] = p
  np.testing.assert_allclose(O.sound_speed_p_rho(s), np.sqrt(O.GAMMA * O.R * t), rtol=1e-10)
  s['h'] = O.CP * t + 0.5 * (u ** 2 + v ** 2 + w ** 2)
  np.testing.assert_allclose(O.sound_speed_h(s, 0.0), np.sqrt(O.GAMMA * O.R * t), rtol=1e-10)
  np.testing.assert_allclose(O.temperature_eos(s), t, rtol=1e-10)


def test_conversions_roundtrip():
  """common/utils_test.py:180-297."""
  rng = np.random.default_rng(5)
  cons = {k: rng.uniform(0.5, 2.0, (4, 5, 6)) for k in NAMES}
  prim = O.conservative_to_primitive(cons)
  assert set(prim) == {'rho', 'u', 'v', 'w', 'e', 'y'}
  back = O.primitive_to_conservative(prim)
  for k in NAMES:
    np.testing.assert_allclose(back[k], cons[k], rtol=1e-14)
