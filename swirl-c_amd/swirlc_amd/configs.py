"""Workload definitions: configuration dicts and initial conditions.

Each builder returns (cfg_dict, state) where state maps the conservative
names to float32 numpy arrays shaped [nz+2h, nx+2h, ny+2h] (halo cells are
filled by the first RHS evaluation).  The setups follow the reference demos'
vocabulary (demo/sods_shock_tube, demo/kelvin_helmholtz_instability,
demo/buoyant_bubble) with the sizes of the benchmark workloads:
  * sod:      Sod shock tube, 512 x 1 x 1, WENO_5 + HLLC, Neumann(0), inviscid.
  * kh:       3-D Kelvin-Helmholtz, periodic n^3 on [0,1]^3, WENO_5 + HLLC,
              nu = 1e-4, Pr = 0.7.
  * bubble:   buoyant warm bubble, 512 x 512 x 256 (scalable), MUSCL + HLL,
              gravity -z, nu = 1.8e-5, periodic x/y, walls in z.
  * riemann:  one 3-D Riemann problem (batched-ensemble member).
"""

import numpy as np

from swirlc_amd.boundary import bc_types
from swirlc_amd.common import types
from swirlc_amd.core.parameter import _get_cell_center_coordinates
from swirlc_amd.physics import constant

BC = bc_types.BoundaryCondition


def _bc_all(kind, value=0.0):
  entry = (kind,) if kind == BC.PERIODIC else (kind, value)
  return {d: {0: entry, 1: entry} for d in types.DIMS}


def _state_from_prim(rho, u, v, w, p, h, pe=0.0):
  """Conservative fields (halo-padded) from interior primitives [nz,nx,ny]."""
  e = p / (constant.GAMMA - 1.0) / rho + 0.5 * (u * u + v * v + w * w) + pe
  fields = {'rho': rho, 'rho_u': rho * u, 'rho_v': rho * v, 'rho_w': rho * w,
            'rho_e': rho * e}
  out = {}
  for k, f in fields.items():
    full = np.zeros(tuple(s + 2 * h for s in f.shape), dtype=np.float32)
    full[h:-h, h:-h, h:-h] = f
    out[k] = full
  return out


def _base_cfg(nx, ny, nz, lx, ly, lz, h):
  x = _get_cell_center_coordinates(0.0, lx, nx)
  y = _get_cell_center_coordinates(0.0, ly, ny)
  z = _get_cell_center_coordinates(0.0, lz, nz)
  return {
      'halo_width': h, 'nx': nx, 'ny': ny, 'nz': nz,
      'core_nx': nx, 'core_ny': ny, 'core_nz': nz,
      'lx': lx, 'ly': ly, 'lz': lz,
      'x': x, 'y': y, 'z': z,
      'dx': lx / nx, 'dy': ly / ny, 'dz': lz / nz,
  }


def sod(n=512, h=3):
  cfg = _base_cfg(n, 1, 1, 1.0, 1.0 / n, 1.0 / n, h)
  cfg.update({
      'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
      'include_diffusion': False, 'time_integration_scheme': 'rk3',
      'bc': {'cell_averages': {v: _bc_all(BC.NEUMANN, 0.0)
                               for v in types.BASE_CONSERVATIVE}},
      'dt': 1e-4, 'cfl': 0.5,
  })
  x = cfg['x'].astype(np.float64)
  left = (x < 0.5)[np.newaxis, :, np.newaxis]
  rho = np.where(left, 1.0, 0.125) * np.ones((1, n, 1))
  p = np.where(left, 1.0, 0.1) * np.ones((1, n, 1))
  z = np.zeros_like(rho)
  return cfg, _state_from_prim(rho, z, z, z, p, h)


def kh(n=256, h=3, seed=0, nu=1e-4, pr=0.7, interpolation='WENO_5', flux='HLLC'):
  """Kelvin-Helmholtz shear layer on a periodic n^3 cube."""
  cfg = _base_cfg(n, n, n, 1.0, 1.0, 1.0, h)
  cfg.update({
      'interpolation_scheme': interpolation, 'numeric_flux_scheme': flux,
      'include_diffusion': True, 'time_integration_scheme': 'rk3',
      'transport_parameters': {'nu': nu, 'pr': pr},
      'bc': {'cell_averages': {v: _bc_all(BC.PERIODIC)
                               for v in types.BASE_CONSERVATIVE}},
      'dt': 1e-4, 'cfl': 0.5,
  })
  x = cfg['x'].astype(np.float64)[np.newaxis, :, np.newaxis]
  y = cfg['y'].astype(np.float64)[np.newaxis, np.newaxis, :]
  shape = (n, n, n)
  inner = np.abs(y - 0.5) < 0.25
  rho = np.broadcast_to(np.where(inner, 2.0, 1.0), shape).astype(np.float64)
  u = np.broadcast_to(np.where(inner, 0.5, -0.5), shape).astype(np.float64)
  rng = np.random.default_rng(seed)
  v = 0.01 * np.sin(4.0 * np.pi * x) + rng.uniform(-1e-3, 1e-3, size=shape)
  v = np.broadcast_to(v, shape)
  w = np.zeros(shape)
  p = np.full(shape, 2.5)
  return cfg, _state_from_prim(rho, u, v, w, p, h)


def bubble(nx=512, ny=512, nz=256, h=3, nu=1.8e-5):
  """Buoyant warm bubble in a hydrostatic, isentropic (theta = 300 K) column."""
  lx, ly, lz = 2.0e4, 2.0e4, 1.0e4
  cfg = _base_cfg(nx, ny, nz, lx, ly, lz, h)
  theta0, dtheta, p0 = 300.0, 2.0, constant.P_0
  zc = cfg['z'].astype(np.float64)
  # Hydrostatic pressure of an isentropic atmosphere (demo buoyant_bubble:84-101).
  def p_hyd(z):
    return p0 * (1.0 - constant.G * constant.KAPPA * z / (constant.R * theta0)) ** (
        1.0 / constant.KAPPA)

  def rho_e_of(z, theta):
    p = p_hyd(z)
    t = theta * (p / p0) ** constant.KAPPA
    rho = p / (constant.R * t)
    return rho, rho * (constant.CV * t + constant.G * z), p

  # Dirichlet per-layer density and energy below and above the column
  # (hydrostatic values at the halo cell centres of the reference mesh).
  zf = np.concatenate([(zc[1] - zc[0]) * np.arange(-h, 0), zc,
                       zc[-1] + (zc[-1] - zc[-2]) * np.arange(1, h + 1)])
  rho_lo, re_lo, _ = rho_e_of(zf[:h], theta0)
  rho_hi, re_hi, _ = rho_e_of(zf[-h:], theta0)
  wall = lambda lo, hi: {'x': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                         'y': {0: (BC.PERIODIC,), 1: (BC.PERIODIC,)},
                         'z': {0: (BC.DIRICHLET, lo), 1: (BC.DIRICHLET, hi)}}
  cfg.update({
      'interpolation_scheme': 'MUSCL', 'numeric_flux_scheme': 'HLL',
      'include_diffusion': True, 'time_integration_scheme': 'rk3',
      'transport_parameters': {'nu': nu, 'pr': 0.7},
      'g': {'x': 0.0, 'y': 0.0, 'z': -1.0}, 'source_functions': ['gravity'],
      'bc': {'cell_averages': {
          'rho': wall([float(v) for v in rho_lo], [float(v) for v in rho_hi]),
          'rho_u': wall(0.0, 0.0), 'rho_v': wall(0.0, 0.0), 'rho_w': wall(0.0, 0.0),
          'rho_e': wall([float(v) for v in re_lo], [float(v) for v in re_hi]),
      }},
      'dt': 0.05, 'cfl': 0.5,
  })
  x = cfg['x'].astype(np.float64)[np.newaxis, :, np.newaxis]
  y = cfg['y'].astype(np.float64)[np.newaxis, np.newaxis, :]
  z = zc[:, np.newaxis, np.newaxis]
  rad = 0.2 * lz
  r = np.sqrt((x - lx / 2) ** 2 + (y - ly / 2) ** 2 + (z - 0.25 * lz) ** 2) / rad
  theta = theta0 + dtheta * np.where(r < 1.0, np.cos(0.5 * np.pi * r) ** 2, 0.0)
  shape = (nz, nx, ny)
  theta = np.broadcast_to(theta, shape)
  p = np.broadcast_to(p_hyd(z), shape)
  t = theta * (p / p0) ** constant.KAPPA
  rho = p / (constant.R * t)
  zero = np.zeros(shape)
  # e = cv T + 0.5|u|^2 + G z (the reference's gravitational potential).
  pe = np.broadcast_to(constant.G * z, shape)
  return cfg, _state_from_prim(rho, zero, zero, zero, p, h, pe)


def riemann(shape, rho, p, vel, axis, offset, h=3):
  """One 3-D Riemann problem: two states split by a plane normal to `axis`
  (0 x, 1 y, 2 z) at fraction `offset`; Neumann(0) (transmissive) walls."""
  nx, ny, nz = shape
  cfg = _base_cfg(nx, ny, nz, 1.0, ny / nx, nz / nx, h)
  cfg.update({
      'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLLC',
      'include_diffusion': False, 'time_integration_scheme': 'rk3',
      'bc': {'cell_averages': {v: _bc_all(BC.NEUMANN, 0.0)
                               for v in types.BASE_CONSERVATIVE}},
      'dt': 1e-4, 'cfl': 0.5,
  })
  n_axis = (nx, ny, nz)[axis]
  coord = (np.arange(n_axis) + 0.5) / n_axis
  left = coord < offset
  bshape = [1, 1, 1]
  bshape[[1, 2, 0][axis]] = n_axis
  left = left.reshape(bshape)
  full = (nz, nx, ny)
  pick = lambda a, b: np.broadcast_to(np.where(left, a, b), full).astype(np.float64)
  return cfg, _state_from_prim(pick(rho[0], rho[1]), pick(vel[0][0], vel[1][0]),
                               pick(vel[0][1], vel[1][1]), pick(vel[0][2], vel[1][2]),
                               pick(p[0], p[1]), h)


ENSEMBLE_SHAPES = ((32, 32, 32), (64, 64, 64), (96, 96, 96), (128, 128, 64), (160, 64, 64))


def riemann_ensemble(n=256, seed=0, shapes=ENSEMBLE_SHAPES, h=3):
  """Batched workload: n independent 3-D Riemann problems.

  Sizes (nx, ny, nz) drawn with default_rng(seed) from `shapes`; for each of
  the two states rho ~ logU[0.1, 2], p ~ logU[0.1, 10], velocity ~ U[-1, 1]^3;
  the interface is normal to a random axis at a random offset in [0.3, 0.7].
  WENO_5 + HLLC, inviscid, Neumann(0) walls.  Returns (cfg, [(grid, state)]).
  """
  rng = np.random.default_rng(seed)
  members = []
  cfg = None
  for _ in range(n):
    shape = shapes[int(rng.integers(len(shapes)))]
    rho = np.exp(rng.uniform(np.log(0.1), np.log(2.0), 2))
    p = np.exp(rng.uniform(np.log(0.1), np.log(10.0), 2))
    vel = rng.uniform(-1.0, 1.0, (2, 3))
    axis = int(rng.integers(3))
    offset = float(rng.uniform(0.3, 0.7))
    c, st = riemann(shape, rho, p, vel, axis, offset, h)
    cfg = c if cfg is None else cfg
    members.append(((shape[0], shape[1], shape[2], c['dx'], c['dy'], c['dz']), st))
  return cfg, members


def member_cfg(cfg, grid):
  """The single-grid configuration of one ensemble member (grid keys replaced)."""
  nx, ny, nz, dx, dy, dz = grid
  out = dict(cfg)
  out.update(_base_cfg(nx, ny, nz, dx * nx, dy * ny, dz * nz, cfg['halo_width']))
  return out
