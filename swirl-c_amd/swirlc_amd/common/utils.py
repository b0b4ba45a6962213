"""Host-side helpers: mesh coordinates incl. halos and gravity direction."""

from typing import Dict

import numpy as np

from swirlc_amd.common import types


def mesh_full(coord, halo_width: int) -> np.ndarray:
  """Cell-centre coordinates of one replica including halo cells (fp32).

  Follows common/utils.py:41-89 for a single replica: the low halo is
  (x[1]-x[0]) * [-h..-1] and the high halo x[-1] + (x[-1]-x[-2]) * [1..h]
  (the low halo is anchored at 0, not at x[0], exactly as the reference
  builds it).  The gravitational potential is evaluated on these points.
  """
  c = np.asarray(coord, dtype=np.float32).reshape(-1)
  if c.size < 2:
    raise ValueError('a mesh coordinate needs at least 2 points to extrapolate halos')
  lo = np.linspace(-halo_width, -1, halo_width).astype(np.float32)
  hi = np.linspace(1, halo_width, halo_width).astype(np.float32)
  return np.concatenate([(c[1] - c[0]) * lo, c, c[-1] + (c[-1] - c[-2]) * hi]).astype(np.float32)


def gravity_direction(g: Dict[str, float]) -> int:
  """common/utils.py:114-142: -1 none, 0/1/2 grid-aligned axis, 3 oblique."""
  g_vec = np.fromiter((g[dim] for dim in types.DIMS), dtype=float)
  if np.abs(np.linalg.norm(g_vec) - 1.0) >= types.SMALL:
    return -1
  for i, dim in enumerate(types.DIMS):
    if np.abs(np.abs(g[dim]) - 1.0) < types.SMALL:
      return i
  return 3
