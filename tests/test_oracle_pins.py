"""Pins for the oracle functions the reference's own tests leave open.

* the hydrostatic helper path (equations/compressible_navier_stokes.py:273-335,
  numerics/riemann_solver/hllc.py:109-120, physics/additional_models/
  gravity.py:64-102): closed forms a wrong substitution fails;
* the CFL wave speed (no reference counterpart; the KH demo's static formula
  max(|u|) + c, demo/kelvin_helmholtz_instability.py:66-73): closed forms;
* edge and corner halo cells (swirl_lm's sequential exchange, stated in the
  prototype notebook demo/swirl_c_prototype.ipynb cell 2): golden vectors
  written by tests/golden/make_notebook_halo_vectors.py.

Every pin is also run against a deliberately wrong oracle variant (a
mistake planted on the CPU with monkeypatch) to show it would catch it.
"""

import os

import numpy as np
import pytest

from oracle import swirlc_oracle as O

from helpers import grid_cfg, oracle_cfg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


# ---------------------------------------------------------------------------
# Edge and corner halo cells: the notebook's sequential x, y, z fill.
# ---------------------------------------------------------------------------

def _fixture_bc(data, name, h):
  codes, values = data[f'{name}_codes'], data[f'{name}_values']
  bc = {}
  for d, dim in enumerate(('x', 'y', 'z')):
    bc[dim] = {}
    for face in (0, 1):
      code = int(codes[d, face])
      if code == 0:
        bc[dim][face] = (O.PERIODIC,)
      elif code == 1:
        bc[dim][face] = (O.DIRICHLET, float(values[d, face, 0]))
      elif code == 2:
        bc[dim][face] = (O.NEUMANN, float(values[d, face, 0]))
      elif code == 3:
        bc[dim][face] = (O.DIRICHLET, [float(v) for v in values[d, face, :h]])
      else:
        bc[dim][face] = (O.DIRICHLET, list(data[f'{name}_plane_{dim}{face}']))
  return bc


def _halo_cases():
  data = np.load(os.path.join(GOLDEN, 'notebook_halo.npz'))
  names = sorted({k[:-len('_codes')] for k in data.files if k.endswith('_codes')})
  return data, names


def _edge_mask(shape, h):
  """Cells in the halo of at least two dimensions (edges and corners)."""
  nz, nx, ny = shape
  inz = np.zeros(nz, bool)
  inx = np.zeros(nx, bool)
  iny = np.zeros(ny, bool)
  inz[h:-h] = True
  inx[h:-h] = True
  iny[h:-h] = True
  count = ((~inz)[:, None, None].astype(int) + (~inx)[None, :, None] + (~iny)[None, None, :])
  return count >= 2


def test_exchange_halos_matches_notebook_everywhere():
  data, names = _halo_cases()
  assert len(names) == 4
  for name in names:
    h = int(data[f'{name}_halo'])
    u = data[f'{name}_in']
    got = O.exchange_halos(u, _fixture_bc(data, name, h), h)
    np.testing.assert_array_equal(got, data[f'{name}_out'], err_msg=name)
    # The fixture holds edge and corner cells whose value depends on the
    # order of the dimensions (garbage went into every halo cell).
    assert _edge_mask(u.shape, h).sum() > 0


def test_halo_fixture_rejects_another_dimension_order(monkeypatch):
  """Planted mistake: filling z, y, x instead of x, y, z changes edge and
  corner cells of every non-periodic case, so the fixture catches it."""
  data, names = _halo_cases()
  monkeypatch.setattr(O, 'DIMS', ('z', 'y', 'x'))
  caught = 0
  for name in names:
    h = int(data[f'{name}_halo'])
    got = O.exchange_halos(data[f'{name}_in'], _fixture_bc(data, name, h), h)
    ref = data[f'{name}_out']
    m = _edge_mask(ref.shape, h)
    assert np.array_equal(got[~m], ref[~m]), name  # face-interior halos agree
    caught += int(not np.array_equal(got[m], ref[m]))
  assert caught == len(names)


# ---------------------------------------------------------------------------
# Hydrostatic helper path: p_dyn = p - p_ref, d_rho = rho - rho_ref.
# ---------------------------------------------------------------------------

def _quiescent(n, h, seed, cfg):
  """A gas at rest with arbitrary positive rho and p (fp64), gravity in -z:
  rho_e = p / (gamma - 1) + rho * pe, so the EOS returns p."""
  nx, ny, nz = n
  rng = np.random.default_rng(seed)
  shape = (nz + 2 * h, nx + 2 * h, ny + 2 * h)
  rho = rng.uniform(0.5, 2.0, shape)
  p = rng.uniform(0.5, 3.0, shape)
  pe = np.broadcast_to(O.gravitational_potential(cfg, np.float64), shape)
  zero = np.zeros(shape)
  cons = {'rho': rho, 'rho_u': zero.copy(), 'rho_v': zero.copy(), 'rho_w': zero.copy(),
          'rho_e': p / (O.GAMMA - 1.0) + rho * pe}
  return cons, p, rho


def _hydro_cfg(n, h, flux):
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05))
  cfg.update({'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': flux,
              'include_diffusion': False, 'g': {'x': 0.0, 'y': 0.0, 'z': -1.0},
              'source_functions': ['gravity'], 'dt': 1e-3,
              'bc': {'cell_averages': {
                  v: {d: {0: ('PERIODIC',), 1: ('PERIODIC',)} for d in 'xyz'}
                  for v in O.BASE_CONSERVATIVE}}})
  return oracle_cfg(cfg)


def _hydro_residual(flux, n=(6, 5, 7), h=3, seed=3):
  """max |rhs| / scale for the quiescent state with p_ref = p, rho_ref = rho.

  With the reference state equal to the state, p_dyn and d_rho vanish in
  every cell (and in the Dirichlet-0 halos), the face velocity is zero, so
  the aligned momentum flux m u + p_dyn and the buoyancy d_rho G g vanish:
  every momentum RHS is 0.  For HLLC the contact speed is 0 and every
  flux vanishes (F*_L with s_m = v = 0), so all five RHS fields are 0.
  """
  cfg = _hydro_cfg(n, h, flux)
  cons, p, rho = _quiescent(n, h, seed, cfg)
  rhs = O.rhs(cons, {'p_ref': p, 'rho_ref': rho}, cfg)
  core = (slice(h, -h),) * 3
  scale_m = max(np.abs(p).max() / min(cfg['dx'], cfg['dy'], cfg['dz']),
                np.abs(rho).max() * O.G)
  scale_e = scale_m * 10.0
  res = {k: np.abs(v[core]).max() / (scale_e if k in ('rho', 'rho_e') else scale_m)
         for k, v in rhs.items()}
  return res


def test_hydrostatic_helper_hllc_rhs_vanishes():
  res = _hydro_residual('HLLC')
  assert max(res.values()) <= 1e-12, res


def test_hydrostatic_helper_hll_momentum_vanishes():
  res = _hydro_residual('HLL')
  assert max(res[k] for k in O.MOMENTUM) <= 1e-12, res
  # HLL's jump term acts on the density and energy of the two face states.
  assert res['rho'] > 1e-8 and res['rho_e'] > 1e-8


def test_hydrostatic_pin_catches_p_kept_in_the_flux(monkeypatch):
  """Planted mistake: p_dyn taken as p (no p_ref subtraction) leaves the
  face pressure in the aligned momentum flux: RHS_w = -dp/dz."""
  orig = O.update_helper_variables

  def wrong(prim, helper, cfg):
    out, names = orig(prim, helper, cfg)
    out['p_dyn'] = prim['p']
    return out, names

  monkeypatch.setattr(O, 'update_helper_variables', wrong)
  for flux in ('HLLC', 'HLL'):
    res = _hydro_residual(flux)
    assert max(res[k] for k in O.MOMENTUM) > 1e-3, (flux, res)


def test_hydrostatic_pin_catches_rho_kept_in_the_source(monkeypatch):
  """Planted mistake: the buoyancy uses rho instead of d_rho."""
  orig = O.update_helper_variables

  def wrong(prim, helper, cfg):
    out, names = orig(prim, helper, cfg)
    out.pop('d_rho')
    return out, names

  monkeypatch.setattr(O, 'update_helper_variables', wrong)
  res = _hydro_residual('HLLC')
  assert res['rho_w'] > 1e-3, res
  assert res['rho_u'] <= 1e-12 and res['rho_v'] <= 1e-12


# A face between two states moving at the same normal velocity u0 with the
# same dynamic pressure but different density and thermodynamic pressure: an
# isolated contact for the p_dyn system.  HLLC resolves it exactly (Toro,
# "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
# sec. 10.4): s_m = u0, and the flux is the upwind side's physical flux,
# whose aligned momentum carries p_dyn and whose energy flux carries the
# face's own p (compressible_navier_stokes.py:390-414).
RHO_L, RHO_R = 1.3, 0.4
P_L, P_R = 2.0, 0.7
PDYN = 0.6
U0, V0, W0 = 0.35, 0.1, -0.2


def _contact_face(u0):
  def side(rho, p):
    e = p / ((O.GAMMA - 1.0) * rho) + 0.5 * (u0 * u0 + V0 * V0 + W0 * W0)
    prim = {'rho': np.array([rho]), 'u': np.array([u0]), 'v': np.array([V0]),
            'w': np.array([W0]), 'e': np.array([e]), 'p': np.array([p]),
            'h': np.array([e + p / rho])}
    cons = {'rho': prim['rho'], 'rho_u': rho * prim['u'], 'rho_v': rho * prim['v'],
            'rho_w': rho * prim['w'], 'rho_e': rho * prim['e']}
    return cons, prim, {'p_dyn': np.array([PDYN])}
  return side(RHO_L, P_L), side(RHO_R, P_R)


@pytest.mark.parametrize('u0', [U0, -U0])
def test_hllc_isolated_contact_with_p_dyn(u0):
  (cl, pl, hl), (cr, pr, hr) = _contact_face(u0)
  s_l, s_m, s_r = O.hllc_wave_speeds(dict(pl, p=hl['p_dyn']), dict(pr, p=hr['p_dyn']),
                                     'x', 0.0)
  np.testing.assert_allclose(s_m, [u0], rtol=1e-13)
  out, br = O.hllc_flux(cl, cr, pl, pr, hl, hr, 'x', list(O.BASE_CONSERVATIVE), 0.0,
                        return_branch=True)
  rho, p, e = (RHO_L, P_L, pl['e'][0]) if u0 > 0 else (RHO_R, P_R, pr['e'][0])
  assert int(br[0]) == (1 if u0 > 0 else 2)
  expect = {'rho': rho * u0, 'rho_u': rho * u0 * u0 + PDYN, 'rho_v': rho * u0 * V0,
            'rho_w': rho * u0 * W0, 'rho_e': rho * u0 * (e + p / rho)}
  for k, v in expect.items():
    np.testing.assert_allclose(out[k], [v], rtol=1e-12, err_msg=k)


def test_contact_pin_catches_p_in_the_wave_speeds(monkeypatch):
  """Planted mistake: the wave speeds see the thermodynamic p (the p_dyn
  swap of hllc.py:109-120 skipped there): the contact speed and the mass
  flux move off their exact values."""
  orig = O.hllc_wave_speeds
  monkeypatch.setattr(O, 'hllc_wave_speeds', lambda pn, pp, dim, pe: orig(
      dict(pn, p=np.array([P_L])), dict(pp, p=np.array([P_R])), dim, pe))
  (cl, pl, hl), (cr, pr, hr) = _contact_face(U0)
  out = O.hllc_flux(cl, cr, pl, pr, hl, hr, 'x', list(O.BASE_CONSERVATIVE), 0.0)
  assert abs(out['rho'][0] - RHO_L * U0) > 1e-3 * RHO_L * U0


def test_contact_pin_catches_p_dyn_in_the_energy_flux(monkeypatch):
  """Planted mistake: the energy flux taking p_dyn instead of the face p."""
  orig = O.physical_flux_1d

  def wrong(cons, prim, helper, dim, var):
    if var == 'rho_e' and 'p_dyn' in helper:
      prim = dict(prim, p=helper['p_dyn'])
    return orig(cons, prim, helper, dim, var)

  monkeypatch.setattr(O, 'physical_flux_1d', wrong)
  (cl, pl, hl), (cr, pr, hr) = _contact_face(U0)
  out = O.hllc_flux(cl, cr, pl, pr, hl, hr, 'x', list(O.BASE_CONSERVATIVE), 0.0)
  exact = RHO_L * U0 * (pl['e'][0] + P_L / RHO_L)
  assert abs(out['rho_e'][0] - exact) > 1e-3 * abs(exact)


# ---------------------------------------------------------------------------
# CFL wave speed: max over interior cells of max_d |u_d| + sqrt(gamma p / rho).
# ---------------------------------------------------------------------------

def _speed_state(n, h, cfg, rho0=1.4, p0=1.0, vel=(0.3, -0.5, 0.2)):
  nx, ny, nz = n
  shape = (nz + 2 * h, nx + 2 * h, ny + 2 * h)
  pe = np.broadcast_to(O.gravitational_potential(cfg, np.float64), shape)
  rho = np.full(shape, rho0)
  u = [np.full(shape, c) for c in vel]
  p = np.full(shape, p0)
  return rho, u, p, pe


def _cons(rho, u, p, pe):
  ke = 0.5 * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
  return {'rho': rho, 'rho_u': rho * u[0], 'rho_v': rho * u[1], 'rho_w': rho * u[2],
          'rho_e': p / (O.GAMMA - 1.0) + rho * (ke + pe)}


@pytest.mark.parametrize('gravity', [False, True])
def test_max_wave_speed_closed_form(gravity):
  n, h = (5, 6, 4), 3
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05))
  if gravity:
    cfg.update({'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity']})
  cfg = oracle_cfg(cfg)
  rho, u, p, pe = _speed_state(n, h, cfg)
  # Uniform: c = sqrt(1.4 * 1 / 1.4) = 1, max |u_d| = 0.5.
  np.testing.assert_allclose(O.max_wave_speed(_cons(rho, u, p, pe), cfg), 1.5, rtol=1e-13)
  # One interior cell with |w| = 2 and c = sqrt(1.4 * 4 / 1.4) = 2: max 4.
  k, i, j = h + 1, h + 2, h + 3
  u[2][k, i, j] = -2.0
  p[k, i, j] = 4.0
  np.testing.assert_allclose(O.max_wave_speed(_cons(rho, u, p, pe), cfg), 4.0, rtol=1e-13)
  # Faster halo cells (edge, face and corner) do not count.
  for c in ((0, h, h), (h, 1, h), (h, h, n[1] + 2 * h - 1), (0, 0, 0)):
    u[0][c] = 50.0
  np.testing.assert_allclose(O.max_wave_speed(_cons(rho, u, p, pe), cfg), 4.0, rtol=1e-13)


def test_max_wave_speed_pin_catches_a_missing_potential(monkeypatch):
  """Planted mistake: the pressure without the gravitational potential."""
  n, h = (5, 6, 4), 3
  cfg = grid_cfg(n, h, (1.0, 1.0, 0.05))
  cfg.update({'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity']})
  cfg = oracle_cfg(cfg)
  rho, u, p, pe = _speed_state(n, h, cfg)
  cons = _cons(rho, u, p, pe)
  monkeypatch.setattr(O, 'gravitational_potential',
                      lambda c, dtype: np.zeros((1, 1, 1), dtype=dtype))
  assert abs(O.max_wave_speed(cons, cfg) - 1.5) > 1e-2
