"""RHS of the compressible Navier-Stokes equations on the GPU.

Drop-in for the reference's equations/compressible_navier_stokes.py:46-270:
`get_rhs_fn(replicas, cfg, physics_models)` returns
`rhs(replica_id, conservative, helper_vars) -> FlowFieldMap` with the same
dictionary keys and the same [nz+2h, nx+2h, ny+2h] field layout (torch
float32 tensors on the GPU instead of TF tensors).

The nine reference steps (cell-average BC, primitives, face interpolation,
face BC, convective flux, diffusive flux, total flux, source, divergence)
run as one halo-fill pass and three fused directional sweeps in HIP
(libswirlc_amd, include/swirlc.h).  Differences of contract:
  * the input fields' halo cells are filled in place (the reference fills a
    copy); interior values are never modified;
  * halo cells of the returned RHS are zero (the reference leaves values
    there that are never used).
"""

import enum

import numpy as np
import torch

from swirlc_amd.common import types
from swirlc_amd.equations.native_context import make_context


class InterpolationSchemes(enum.Enum):
  """equations/compressible_navier_stokes.py:39-43."""
  WENO_3 = 'WENO_3'
  WENO_5 = 'WENO_5'
  MUSCL = 'MUSCL'


def _check_helpers(ctx, helper_vars):
  """Hands the hydrostatic reference state to the context when helper_vars
  holds both 'p_ref' and 'rho_ref' (compressible_navier_stokes.py:309), and
  clears it otherwise."""
  if helper_vars and 'p_ref' in helper_vars and 'rho_ref' in helper_vars:
    ctx.set_reference_state(helper_vars['p_ref'], helper_vars['rho_ref'])
  else:
    ctx.set_reference_state(None, None)


def get_rhs_fn(replicas: np.ndarray, cfg, physics_models=None):
  """Builds the RHS closure once (equations/compressible_navier_stokes.py:46-67).

  Like the reference (which traces a tf.function here), call this once per
  configuration, not inside a loop: it creates the native context.
  """
  del replicas, physics_models  # one replica per GPU; models come from cfg
  ctx = make_context(cfg)

  def rhs(replica_id, conservative, helper_vars):
    """d(conservative)/dt over the interior cells (:69-268)."""
    del replica_id
    _check_helpers(ctx, helper_vars)
    q = ctx.fields(conservative)
    out = [torch.zeros_like(t) for t in q]
    ctx.rhs(q, out)
    return dict(zip(ctx.names, out))

  rhs.native_context = ctx
  return rhs
