"""Shared helpers of the test suite (states, config conversion, error norms)."""

import copy

import numpy as np

from oracle import swirlc_oracle as O
from swirlc_amd.core.parameter import SwirlCParameters


def oracle_cfg(cfg_dict):
  """The oracle's config dict from a SwirlCParameters-style user dict."""
  c = SwirlCParameters(copy.deepcopy(cfg_dict))
  keys = ('halo_width', 'nx', 'ny', 'nz', 'core_nx', 'core_ny', 'core_nz', 'x', 'y',
          'z', 'dx', 'dy', 'dz', 'g', 'dt', 'interpolation_scheme',
          'numeric_flux_scheme', 'include_diffusion', 'time_integration_scheme',
          'source_functions', 'transport_parameters', 'conservative_variable_names',
          'primitive_variable_names', 'bc')
  return {k: getattr(c, k) for k in keys}


def random_state(n, h=3, seed=0, jump=True, nscalar=0, pe=None):
  """A random, physically valid conservative state [nz+2h, nx+2h, ny+2h] fp32.

  Smooth large-scale fields plus noise, and (jump=True) a density/pressure
  jump across a tilted plane, so limiters and all HLLC branches are hit.
  `pe` (gravitational potential) is added to the specific energy so the
  pressure stays positive under gravity.
  """
  nx, ny, nz = n
  rng = np.random.default_rng(seed)
  z, x, y = np.meshgrid(np.linspace(0, 1, nz + 2 * h), np.linspace(0, 1, nx + 2 * h),
                        np.linspace(0, 1, ny + 2 * h), indexing='ij')
  shape = z.shape
  rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.05 * rng.standard_normal(shape)
  p = 1.0 + 0.2 * np.cos(2 * np.pi * z) + 0.05 * rng.standard_normal(shape)
  u = 0.6 * np.sin(2 * np.pi * y) + 0.1 * rng.standard_normal(shape)
  v = 0.5 * np.cos(2 * np.pi * z) + 0.1 * rng.standard_normal(shape)
  w = -0.4 * np.sin(2 * np.pi * x) + 0.1 * rng.standard_normal(shape)
  if jump:
    side = (x + 0.5 * y + 0.25 * z) > 0.9
    rho = np.where(side, 0.3 * rho, rho)
    p = np.where(side, 0.2 * p, p)
    u = np.where(side, u + 1.5, u)
  rho = np.abs(rho) + 0.05
  p = np.abs(p) + 0.05
  e = p / (O.GAMMA - 1.0) / rho + 0.5 * (u * u + v * v + w * w)
  if pe is not None:  # specific potential energy, broadcast to [nz, nx, ny]
    e = e + pe
  st = {'rho': rho, 'rho_u': rho * u, 'rho_v': rho * v, 'rho_w': rho * w, 'rho_e': rho * e}
  for s in range(nscalar):
    st[f'rho_y{s}'] = rho * rng.uniform(0, 1, shape)
  return {k: v.astype(np.float32) for k, v in st.items()}


def interior(a, h):
  return a[h:-h, h:-h, h:-h]


def rel_linf(a, b):
  """max|a - b| / max|b| (normwise relative L-infinity error)."""
  a = np.asarray(a, dtype=np.float64)
  b = np.asarray(b, dtype=np.float64)
  den = np.max(np.abs(b))
  return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def grid_cfg(n, h=3, lengths=(1.0, 1.0, 1.0)):
  nx, ny, nz = n
  x = (np.arange(nx) + 0.5) / nx * lengths[0]
  y = (np.arange(ny) + 0.5) / ny * lengths[1]
  z = (np.arange(nz) + 0.5) / nz * lengths[2]
  return {'halo_width': h, 'nx': nx, 'ny': ny, 'nz': nz, 'core_nx': nx,
          'core_ny': ny, 'core_nz': nz, 'x': x.astype(np.float32),
          'y': y.astype(np.float32), 'z': z.astype(np.float32),
          'dx': lengths[0] / nx, 'dy': lengths[1] / ny, 'dz': lengths[2] / nz}
