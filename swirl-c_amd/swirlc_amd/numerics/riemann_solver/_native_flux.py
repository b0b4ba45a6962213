"""Shared driver of the stand-alone HLL / HLLC functions (sc_riemann_flux)."""

import copy
import ctypes

import torch

from swirlc_amd import _native as N
from swirlc_amd.common import types
from swirlc_amd.equations.native_context import cached_context

_DIMS = {'x': 0, 'y': 1, 'z': 2}
_BASE_PRIM = ('rho', 'u', 'v', 'w', 'e', 'p', 'h')


def _ptrs(tensors, n, ctx, what):
  """Device pointers of face-state tensors, each checked against the
  context's grid: the kernel reads and writes every one of them over the
  context's full [nz+2h, nx+2h, ny+2h] extent."""
  arr = (ctypes.c_void_p * n)()
  for k, t in enumerate(tensors):
    _check_field(t, ctx, what)
    arr[k] = t.data_ptr()
  return arr


def _check_field(t, ctx, what):
  N.require_gpu_tensor(t, what)
  if tuple(t.shape) != ctx.shape:
    raise ValueError(f'{what} has shape {tuple(t.shape)}, expected {ctx.shape} (the grid of cfg)')
  if t.device != ctx.device:
    raise ValueError(f'{what} is on {t.device}, the context on {ctx.device}')


def _p_dyn(helper, dim):
  """p_dyn of the faces normal to dim, from a helper-variable dict keyed by
  dimension (the reference's layout) or a flat one; None when absent."""
  if not helper:
    return None
  h = helper.get(dim, helper) if isinstance(helper, dict) else helper
  return h.get('p_dyn') if isinstance(h, dict) else None


def numerical_flux(scheme, conservative_neg, conservative_pos, primitive_neg, primitive_pos,
                   helper_var_neg, helper_var_pos, dim, cfg):
  """Flux dict across face i-1/2 (stored at i) normal to dim, scheme 'HLL' / 'HLLC'."""
  if dim not in _DIMS:
    raise ValueError(f'"{dim}" is not a valid dimension. Available options are: "x", "y", "z".')
  c = copy.copy(cfg)
  c.numeric_flux_scheme = scheme
  ctx = cached_context(c)
  names = ctx.names
  prim_names = list(_BASE_PRIM) + [types.conservative_to_primitive_name(n) for n in names[5:]]
  for d, what in ((conservative_neg, 'conservative_neg'), (conservative_pos, 'conservative_pos')):
    missing = [n for n in names if n not in d]
    if missing:
      raise KeyError(f'{what} lacks {missing}')
  for d, what in ((primitive_neg, 'primitive_neg'), (primitive_pos, 'primitive_pos')):
    missing = [n for n in prim_names if n not in d]
    if missing:
      raise KeyError(f'{what} lacks {missing}')
  # hllc.py:109-120 and compressible_navier_stokes.py:378 substitute p_dyn per
  # side; a side without it keeps its own face pressure, which is what the
  # kernel computes when that side's p_dyn is its p.
  pdn, pdp = _p_dyn(helper_var_neg, dim), _p_dyn(helper_var_pos, dim)
  if pdn is not None or pdp is not None:
    pdn = primitive_neg['p'] if pdn is None else pdn
    pdp = primitive_pos['p'] if pdp is None else pdp
    for t in (pdn, pdp):
      _check_field(t, ctx, 'p_dyn')
  out = {n: torch.empty(ctx.shape, dtype=torch.float32, device=ctx.device) for n in names}
  # More than two scalars: one call per part of the context (five base
  # fields and up to two scalars), the base fluxes kept from the first.
  for c, part in enumerate(getattr(ctx, 'parts', [ctx])):
    pn = part.names
    pp = list(_BASE_PRIM) + [types.conservative_to_primitive_name(n) for n in pn[5:]]
    dst = dict(out) if c == 0 else {**{n: torch.empty_like(out[n]) for n in pn[:5]},
                                    **{n: out[n] for n in pn[5:]}}
    N.check(N.lib().sc_riemann_flux(
        part._ctx, _DIMS[dim],
        _ptrs([conservative_neg[n] for n in pn], N.SC_MAX_FIELDS, ctx, 'conservative_neg'),
        _ptrs([conservative_pos[n] for n in pn], N.SC_MAX_FIELDS, ctx, 'conservative_pos'),
        _ptrs([primitive_neg[n] for n in pp], N.SC_MAX_FIELDS + 2, ctx, 'primitive_neg'),
        _ptrs([primitive_pos[n] for n in pp], N.SC_MAX_FIELDS + 2, ctx, 'primitive_pos'),
        ctypes.c_void_p(pdn.data_ptr()) if pdn is not None else None,
        ctypes.c_void_p(pdp.data_ptr()) if pdp is not None else None,
        _ptrs([dst[n] for n in pn], N.SC_MAX_FIELDS, ctx, 'flux'), N.stream_handle()),
            'sc_riemann_flux')
  return out
