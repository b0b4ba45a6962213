"""Shared Runge-Kutta stage plumbing for the fused (native) and generic paths."""

import torch

from swirlc_amd.equations.compressible_navier_stokes import _check_helpers


def native_context(rhs_fn):
  """The NativeContext behind an RHS closure from get_rhs_fn, else None."""
  return getattr(rhs_fn, 'native_context', None)


def dt_tensor(dt, device):
  """dt as a 1-element float32 device tensor (device tensors pass through)."""
  if torch.is_tensor(dt):
    return dt.reshape(1).to(device=device, dtype=torch.float32)
  return torch.full((1,), float(dt), dtype=torch.float32, device=device)


def empty_like_state(ctx, fields):
  # Halo cells of a stage output are never read before the next halo fill;
  # zeros keep them finite.
  return [torch.zeros_like(t) for t in fields]
