"""Writes tests/golden/sod_fp64.npz: the fp64 oracle's Sod shock tube at t = 0.2.

The Sod configuration (512 x 1 x 1 cells on [0, 1], WENO5-JS, HLLC, SSP-RK3,
Neumann BCs, gamma = 1.4) is advanced with a fixed step dt = 0.2 / 410 (CFL
below 0.5 for the solution's wave speeds, which stay under 2) in float64
arithmetic by the oracle.  The fixture holds the interior density, velocity
and pressure profiles and the step count; the GPU test runs the same steps in
fp32 and states its error against them.

  python tests/golden/make_sod_fp64.py      (about a minute on one core)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'swirl-c_amd'), os.path.dirname(HERE)]

from oracle import swirlc_oracle as O  # pylint: disable=wrong-import-position
from swirlc_amd import configs  # pylint: disable=wrong-import-position
from helpers import oracle_cfg  # pylint: disable=wrong-import-position

N, STEPS, T_END = 512, 410, 0.2
OUT = os.path.join(HERE, 'sod_fp64.npz')


def profiles(state, h):
  """Interior rho, u, p along x of a [1+2h, n+2h, 1+2h] state."""
  with np.errstate(divide='ignore', invalid='ignore'):
    prim = O.conservative_to_primitive(state)
    p = O.pressure(prim, 0.0)
  sl = (h, slice(h, -h), h)
  return prim['rho'][sl], prim['u'][sl], p[sl]


def main():
  cfg, st = configs.sod(N)
  h = cfg['halo_width']
  ocfg = oracle_cfg(cfg)
  dt = T_END / STEPS
  s = {k: v.astype(np.float64) for k, v in st.items()}
  for _ in range(STEPS):
    s = O.rk3(s, {}, ocfg, dt=dt)
  rho, u, p = profiles(s, h)
  np.savez(OUT, rho=rho, u=u, p=p, steps=np.int64(STEPS), dt=np.float64(dt), n=np.int64(N))
  print(OUT, float(rho.min()), float(rho.max()))


if __name__ == '__main__':
  main()
