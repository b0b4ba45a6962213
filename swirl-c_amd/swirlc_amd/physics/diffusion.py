"""Diffusive terms on the faces on the GPU (physics/diffusion.py).

  shear_stress(states, states_fx, states_fy, states_fz, cfg, physics_models=None)
      physics/diffusion.py:33-67 (strain rate physics/fluid.py:31-111)
  single_component_heat_flux(primitive, cfg, physics_models=None)
      physics/diffusion.py:76-109

With the reference's TransportSimple model (physics/transport/simple.py:44-108:
constant kinematic viscosity cfg.transport_parameters['nu'], conductivity
(1 / pr) cp nu rho), which is the model the reference builds for these
configurations, so `physics_models` is not needed.  Face values of face
i - 1/2 are stored at index i; neighbours outside the array count as 0, as
the reference's kernel ops give in halo cells.  Fields are contiguous
float32 GPU tensors [nz + 2h, nx + 2h, ny + 2h]; the kernels are
libswirlc_amd's (sc_shear_stress, sc_heat_flux).
"""

import ctypes

import torch

from swirlc_amd import _native as N


def _geom(f, cfg):
  if f.dim() != 3:
    raise ValueError(f'expected a 3-D field [nz, nx, ny], got shape {tuple(f.shape)}')
  nz, nx, ny = f.shape
  ext = (ctypes.c_int * 3)(nx, ny, nz)
  h = (ctypes.c_double * 3)(float(cfg.dx), float(cfg.dy), float(cfg.dz))
  return ext, h


def _ptrs(tensors):
  arr = (ctypes.c_void_p * len(tensors))()
  for k, t in enumerate(tensors):
    N.require_gpu_tensor(t, 'field')
    arr[k] = t.data_ptr()
  return arr


def _same_grid(named):
  """The kernels read every field over the first one's extents."""
  (name0, t0), rest = named[0], named[1:]
  for name, t in named:
    N.require_gpu_tensor(t, name)
  for name, t in rest:
    if t.shape != t0.shape or t.device != t0.device:
      raise ValueError(f'{name} has shape {tuple(t.shape)} on {t.device}, expected '
                       f'{tuple(t0.shape)} on {t0.device} (as {name0})')


def shear_stress(states, states_fx, states_fy, states_fz, cfg, physics_models=None):
  """tau[i][j] = 2 rho_j nu S_ij on the j faces (3 x 3 nested list of tensors)."""
  del physics_models
  u, v, w = states['u'], states['v'], states['w']
  rho_f = [states_fx['rho'], states_fy['rho'], states_fz['rho']]
  _same_grid([('u', u), ('v', v), ('w', w), ('rho on the x faces', rho_f[0]),
              ('rho on the y faces', rho_f[1]), ('rho on the z faces', rho_f[2])])
  ext, h = _geom(u, cfg)
  tau = [torch.empty_like(u) for _ in range(9)]
  N.check(N.lib().sc_shear_stress(
      ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(v.data_ptr()),
      ctypes.c_void_p(w.data_ptr()), _ptrs(rho_f), ext, h,
      float(cfg.transport_parameters['nu']), _ptrs(tau), N.stream_handle()), 'sc_shear_stress')
  return [[tau[3 * i + j] for j in range(3)] for i in range(3)]


def single_component_heat_flux(primitive, cfg, physics_models=None):
  """{'x', 'y', 'z'}: -kappa_f dT/dx on the faces of each direction."""
  del physics_models
  rho, p = primitive['rho'], primitive['p']
  _same_grid([('rho', rho), ('p', p)])
  ext, h = _geom(rho, cfg)
  tp = cfg.transport_parameters
  flux = [torch.empty_like(rho) for _ in range(3)]
  N.check(N.lib().sc_heat_flux(
      ctypes.c_void_p(rho.data_ptr()), ctypes.c_void_p(p.data_ptr()), ext, h,
      float(tp['nu']), float(tp['pr']), _ptrs(flux), N.stream_handle()), 'sc_heat_flux')
  return dict(zip(('x', 'y', 'z'), flux))
