"""Face interpolation of one field on the GPU (numerics/interpolation.py).

  linear_interpolation(f, dim, kernel_op=None)   numerics/interpolation.py:21-53
  weno(v, dim, k=3)                              swirl_lm interpolation.weno
  flux_limiter(v, dim, limiter_type='MUSCL')     swirl_lm interpolation.flux_limiter

weno / flux_limiter follow the swirl_lm contract the reference's
_face_interpolation relies on (equations/compressible_navier_stokes.py:536-615,
567-575): they return (neg, pos), the left and right states of face i - 1/2
stored at index i.  Fields are contiguous float32 GPU tensors shaped
[nz + 2h, nx + 2h, ny + 2h] (the FlowFieldVar layout); the results are new
tensors of the same shape.  Entries whose stencil leaves the array keep the
cell value (WENO) or see a zero slope (MUSCL); linear_interpolation takes
f[-1] as 0.  The kernels are libswirlc_amd's (sc_face_interp,
sc_linear_interp); there is no CPU path.
"""

import ctypes

import torch

from swirlc_amd import _native as N

_DIMS = {'x': 0, 'y': 1, 'z': 2}


def _dim_index(dim):
  if dim not in _DIMS:
    raise ValueError(f'"{dim}" is not a valid dimension. Available options are: "x", "y", "z".')
  return _DIMS[dim]


def _extents(f):
  if f.dim() != 3:
    raise ValueError(f'expected a 3-D field [nz, nx, ny], got shape {tuple(f.shape)}')
  nz, nx, ny = f.shape
  return (ctypes.c_int * 3)(nx, ny, nz)


def linear_interpolation(f, dim, kernel_op=None):
  """0.5 * (f[i-1] + f[i]) on face i - 1/2 along `dim` (numerics/interpolation.py:21-53).

  `kernel_op` (the reference's TPU kernel library) is accepted and unused.
  """
  del kernel_op
  d = _dim_index(dim)
  N.require_gpu_tensor(f, 'f')
  ext = _extents(f)
  out = torch.empty_like(f)
  N.check(N.lib().sc_linear_interp(ctypes.c_void_p(f.data_ptr()), ext, d,
                                   ctypes.c_void_p(out.data_ptr()), N.stream_handle()),
          'sc_linear_interp')
  return out


def _face(v, dim, scheme):
  d = _dim_index(dim)
  N.require_gpu_tensor(v, 'v')
  ext = _extents(v)
  neg = torch.empty_like(v)
  pos = torch.empty_like(v)
  N.check(N.lib().sc_face_interp(ctypes.c_void_p(v.data_ptr()), ext, d, scheme,
                                 ctypes.c_void_p(neg.data_ptr()),
                                 ctypes.c_void_p(pos.data_ptr()), N.stream_handle()),
          'sc_face_interp')
  return neg, pos


def weno(v, dim, k=3):
  """WENO-(2k-1) face states (neg, pos) of face i - 1/2 stored at i (k = 2 or 3)."""
  if k == 3:
    return _face(v, dim, N.SC_WENO_5)
  if k == 2:
    return _face(v, dim, N.SC_WENO_3)
  raise NotImplementedError(f'WENO with k={k} is not implemented; k must be 2 or 3.')


def flux_limiter(v, dim, limiter_type='MUSCL'):
  """Limited face states (neg, pos) of face i - 1/2 stored at i; MUSCL (minmod)."""
  name = getattr(limiter_type, 'name', limiter_type)
  if name != 'MUSCL':
    raise NotImplementedError(f'flux limiter "{name}" is not implemented; use "MUSCL".')
  return _face(v, dim, N.SC_MUSCL)
