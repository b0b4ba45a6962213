"""Cell-average boundary update on the GPU (boundary/boundary.py).

  update_conservative_cell_averages(replica_id, replicas, conservative, cfg)
      boundary/boundary.py:464-539

Fills the halo layers of every conservative field from cfg.bc['cell_averages']
(PERIODIC, DIRICHLET with scalar / per-layer values or planes, NEUMANN, and
BCs given on the primitive variable, boundary.py:516-531), in the x, y, z
order of swirl_lm's halo exchange, with libswirlc_amd's halo kernels
(sc_fill_halos).  As in the reference the input dict is left unchanged and a
new dict is returned.  Fields are float32 GPU tensors [nz+2h, nx+2h, ny+2h];
one replica.  Each call builds a native context for cfg (device allocations
included); code that fills halos every step keeps one NativeContext instead.
"""

import torch

from swirlc_amd.equations.native_context import NativeContext


def update_conservative_cell_averages(replica_id, replicas, conservative, cfg):
  """New dict of the conservative fields with their halo cells filled.

  A field whose BC is given on its primitive comes back as
  rho * divide_no_nan(q, rho) in every cell, interior included, as the
  reference computes it (boundary.py:516-531); the in-place native fill
  leaves the interior alone (the sweeps apply that rounding as they load).
  """
  del replica_id, replicas
  ctx = NativeContext(cfg)
  out = {k: v.clone() for k, v in conservative.items()}
  ctx.fill_halos(ctx.fields(out))
  h = int(cfg.halo_width)
  core = (slice(h, -h),) * 3
  rho = out[ctx.names[0]][core]
  for f, name in enumerate(ctx.names):
    if f > 0 and (ctx.prim_bc_mask >> f) & 1:
      q = out[name][core]
      q.copy_(rho * torch.where(rho == 0, torch.zeros_like(q), q / rho))
  return out
