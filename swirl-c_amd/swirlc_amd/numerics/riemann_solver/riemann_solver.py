"""Riemann-solver selection (numerics/riemann_solver/riemann_solver.py:30-62)."""

from swirlc_amd.numerics.riemann_solver import hll
from swirlc_amd.numerics.riemann_solver import hllc


def select_numeric_flux_fn(cfg):
  """hll_convective_flux or hllc_convective_flux per cfg.numeric_flux_scheme."""
  if cfg.numeric_flux_scheme == 'HLL':
    return hll.hll_convective_flux
  if cfg.numeric_flux_scheme == 'HLLC':
    return hllc.hllc_convective_flux
  raise ValueError(
      f'"{cfg.numeric_flux_scheme}" is not implemented as a'
      ' "numeric_flux_scheme". Valid options are: '
      + str(['HLL', 'HLLC']))
