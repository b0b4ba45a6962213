"""Translation of a SwirlCParameters config into a native (HIP) context.

`NativeContext(cfg)` validates the configuration the way the reference does
at trace time (get_rhs_fn, PhysicsModels, select_numeric_flux_fn, the
face-interpolation dispatch) and owns one `sc_ctx` of libswirlc_amd.  Its
methods launch the HIP kernels on the current torch stream.
"""

import collections
import copy
import ctypes
import hashlib
import math
import numbers

import numpy as np
import torch

from swirlc_amd import _native as N
from swirlc_amd.boundary import bc_types
from swirlc_amd.common import types
from swirlc_amd.common import utils

_INTERP = {'WENO_3': N.SC_WENO_3, 'WENO_5': N.SC_WENO_5, 'MUSCL': N.SC_MUSCL}
_FLUX = {'HLL': N.SC_HLL, 'HLLC': N.SC_HLLC}
_BCT = {
    bc_types.BoundaryCondition.PERIODIC: N.SC_BC_PERIODIC,
    bc_types.BoundaryCondition.DIRICHLET: N.SC_BC_DIRICHLET,
    bc_types.BoundaryCondition.NEUMANN: N.SC_BC_NEUMANN,
    bc_types.BoundaryCondition.REFLECTIVE: N.SC_BC_REFLECTIVE,
}


def _get(cfg, key, default=None):
  return getattr(cfg, key, default) if not isinstance(cfg, dict) else cfg.get(key, default)


def validate_models(cfg):
  """PhysicsModels.__init__ checks (physics/physics_models.py:45-92)."""
  if _get(cfg, 'thermodynamics_model', 'generic') != 'generic':
    raise NotImplementedError(
        f'Unsupported thermodynamics model: "{_get(cfg, "thermodynamics_model")}".'
        " Supported models are: ['generic']")
  if _get(cfg, 'transport_model', 'simple') != 'simple':
    raise NotImplementedError(
        f'Unsupported transport model: "{_get(cfg, "transport_model")}".'
        " Supported models are: ['simple']")
  for name in _get(cfg, 'source_functions') or []:
    if name != 'gravity':
      raise NotImplementedError(
          f'Unsupported source term model: "{name}". Supported models are:'
          " ['gravity']")


def copy_fill_dims(params, nf, n, h):
  """The dims along which the halo fill of every field is a pure copy, so
  that the last sweep of an RK stage can write its output's halos as images:
  each dim PERIODIC on both faces with extent >= 2h (wraps), or NEUMANN with
  value +0 on both faces (ghost = neighbour - 0 copies the edge cell:
  clamps), the same kind for every field, and no BC on a primitive.
  n = (nx, ny, nz).  Returns None if some dim is neither, else the set of
  clamping dims (0 x, 1 y, 2 z).  Mirrors the library's rule."""
  if params.prim_bc_mask != 0:
    return None
  clamp = set()
  for d in range(3):
    kind = params.bc[0][d][0].type
    for f in range(nf):
      for s in range(2):
        b = params.bc[f][d][s]
        if b.type != kind:
          return None
        if kind == N.SC_BC_NEUMANN and (b.value != 0.0 or math.copysign(1.0, b.value) < 0
                                        or b.n_layer_values != 0 or b.layer_planes):
          return None
    if kind == N.SC_BC_PERIODIC:
      if n[d] < 2 * h:
        return None
    elif kind == N.SC_BC_NEUMANN:
      clamp.add(d)
    else:
      return None
  return clamp


class NativeContext:
  """One native RHS/RK context for a fixed grid and configuration."""

  # Native contexts made by this process (sc_ctx_create calls that succeeded).
  created = 0

  def __init__(self, cfg, device=None, create=True):
    """Validates cfg and builds the native parameters; `create=False` stops
    before the native context is made (the parameters are reused by a batch)."""
    self._ctx = None
    # (struct, pointer field, host planes) uploaded at creation: per-layer
    # DIRICHLET halo planes and plane-valued face / flux boundary values.
    self._pending_planes = []
    validate_models(cfg)
    scheme = _get(cfg, 'interpolation_scheme')
    if scheme not in _INTERP:
      raise NotImplementedError(
          f'"{scheme}" is not implemented as an "interpolation_scheme". Valid'
          " options are: ['WENO_3', 'WENO_5', 'MUSCL']")
    flux = _get(cfg, 'numeric_flux_scheme')
    if flux not in _FLUX:
      raise ValueError(
          f'"{flux}" is not implemented as a "numeric_flux_scheme". Valid'
          " options are: ['HLL', 'HLLC']")
    if (_get(cfg, 'cx', 1), _get(cfg, 'cy', 1), _get(cfg, 'cz', 1)) != (1, 1, 1):
      raise NotImplementedError(
          'multi-replica domain decomposition is not supported: one GPU runs'
          ' one replica (cx = cy = cz = 1)')
    bc = _get(cfg, 'bc')
    if 'cell_faces' in bc and _get(cfg, 'include_diffusion'):
      raise NotImplementedError(
          'Face boundary conditions are not supported with diffusion enabled'
          ' at this time.')

    self.names = list(_get(cfg, 'conservative_variable_names'))
    if self.names[:5] != list(types.BASE_CONSERVATIVE):
      raise ValueError(
          f'conservative_variable_names must start with {list(types.BASE_CONSERVATIVE)}')
    if len(self.names) > 7:
      raise NotImplementedError(
          'one native context holds at most two transported scalars (make_context'
          ' splits more of them over several contexts)')
    # The reference rebuilds face conserved states from the primitives listed
    # in primitive_variable_names (compressible_navier_stokes.py:131-144), so a
    # scalar whose primitive is not listed has no face state there.
    prims = list(_get(cfg, 'primitive_variable_names', types.BASE_PRIMITIVES))
    for name in self.names[5:]:
      if types.conservative_to_primitive_name(name) not in prims:
        raise ValueError(
            f'transported scalar {name} needs its primitive '
            f'{types.conservative_to_primitive_name(name)} in primitive_variable_names')
    self.nf = len(self.names)
    self.hw = int(_get(cfg, 'halo_width'))
    self.n = (int(_get(cfg, 'core_nx')), int(_get(cfg, 'core_ny')), int(_get(cfg, 'core_nz')))
    self.shape = (self.n[2] + 2 * self.hw, self.n[0] + 2 * self.hw, self.n[1] + 2 * self.hw)
    self.device = torch.device('cuda') if device is None else torch.device(device)
    self.h = (float(_get(cfg, 'dx')), float(_get(cfg, 'dy')), float(_get(cfg, 'dz')))
    g = _get(cfg, 'g')
    self.g = tuple(float(g[d]) for d in types.DIMS)
    self.gravity_dir = utils.gravity_direction(g)
    self.gravity_source = 'gravity' in (_get(cfg, 'source_functions') or [])
    if self.gravity_source and self.gravity_dir == -1:
      raise ValueError('Gravity direction must be specified if the gravity source'
                       ' function is to be used.')
    self._keep = []  # device tensors referenced by the native context

    grid = self._grid = N.ScGrid()
    grid.nx, grid.ny, grid.nz = self.n
    grid.halo = self.hw
    grid.dx, grid.dy, grid.dz = self.h
    if self.gravity_dir != -1:
      for attr, key in (('x_full', 'x'), ('y_full', 'y'), ('z_full', 'z')):
        arr = np.ascontiguousarray(utils.mesh_full(_get(cfg, key), self.hw))
        self._keep.append(arr)
        setattr(grid, attr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))

    params = self._params = N.ScParams()
    params.n_fields = self.nf
    params.interpolation = _INTERP[scheme]
    params.numeric_flux = _FLUX[flux]
    params.include_diffusion = 1 if _get(cfg, 'include_diffusion') else 0
    tp = _get(cfg, 'transport_parameters')
    params.nu = float(tp['nu'])
    params.pr = float(tp['pr'])
    params.gravity_source = 1 if self.gravity_source else 0
    for d in range(3):
      params.g[d] = self.g[d]
    cell_bc = bc['cell_averages']
    mask = 0
    for f, name in enumerate(self.names):
      if name in cell_bc:
        spec = cell_bc[name]
      elif types.conservative_to_primitive_name(name) in cell_bc:
        spec = cell_bc[types.conservative_to_primitive_name(name)]
        mask |= 1 << f
      else:
        raise ValueError(
            'A "cells" boundary condition is required for all conservative'
            ' variables, either directly or through their respective primitive.'
            f' No boundary condition was found for {name} or'
            f' {types.conservative_to_primitive_name(name)}.')
      for d, dim in enumerate(types.DIMS):
        if dim not in spec or spec[dim] is None:
          # boundary.py:58-60: no BC along dim, its halo cells are left alone.
          if (mask >> f) & 1:
            raise NotImplementedError(
                f'a BC given on the primitive of {name} must cover every dimension')
          for s in (0, 1):
            params.bc[f][d][s].type = N.SC_BC_NONE
          continue
        self._fill_dim_bc(params.bc[f][d], spec[dim], d)
    params.prim_bc_mask = mask
    self.prim_bc_mask = mask
    self._fill_boundary_values(params, bc)

    if create:
      self.create()

  def create(self):
    """Makes the native context (sc_ctx_create) from the translated config."""
    self.upload_planes()
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      N.check(N.lib().sc_ctx_create(ctypes.byref(self._grid), ctypes.byref(self._params),
                                    ctypes.byref(handle)), 'sc_ctx_create')
    self._ctx = handle
    NativeContext.created += 1

  def upload_planes(self):
    """Copies the per-layer DIRICHLET planes and the plane-valued boundary
    values to the device and points the C tables at them (kept alive by the
    context)."""
    for obj, field, planes in self._pending_planes:
      dev = torch.from_numpy(planes).to(self.device).contiguous()
      self._keep.append(dev)
      setattr(obj, field, dev.data_ptr())
    self._pending_planes = []

  def cache_key(self):
    """Bytes that identify the native context this translation makes: the
    grid and parameter structs (pointer fields excluded), the coordinates
    and the per-layer planes, and the device."""
    digest = hashlib.sha256()
    grid = N.ScGrid.from_buffer_copy(self._grid)
    grid.x_full = grid.y_full = grid.z_full = None
    digest.update(bytes(grid))
    params = N.ScParams.from_buffer_copy(self._params)
    for f in range(N.SC_MAX_FIELDS):
      for d in range(3):
        for s in range(2):
          params.bc[f][d][s].layer_planes = None
    for k in range(N.SC_MAX_BOUNDARY_VALUES):
      params.boundary_values[k].plane = None
    digest.update(bytes(params))
    for arr in self._keep:
      if isinstance(arr, np.ndarray):
        digest.update(arr.tobytes())
    for _, field, planes in self._pending_planes:
      digest.update(field.encode())
      digest.update(planes.tobytes())
    digest.update(str(self.device).encode())
    return digest.digest()

  # -- BC translation ------------------------------------------------------

  @staticmethod
  def _dirichlet_along(spec, dim, side):
    """boundary/boundary.py:186-227: (var -> value) of the DIRICHLET entries
    on one boundary; `None` types are skipped.  A value is a float, or the
    value given (a plane, see _layer_plane)."""
    out = {}
    for var, d in spec.items():
      t, val = d[dim][side]
      if t is None:
        continue
      if t != bc_types.BoundaryCondition.DIRICHLET:
        raise AssertionError(
            f'Unsupported interface (face or flux) BC type {t}. Only "DIRICHLET"'
            ' or `None` is supported.')
      out[var] = float(val) if isinstance(val, numbers.Number) else val
    return out

  def _fill_boundary_values(self, params, bc):
    """cfg.bc['intercell_fluxes'] and cfg.bc['cell_faces'] -> the C table."""
    entries = []
    kinds = {'convective': N.SC_BV_FLUX_CONVECTIVE, 'diffusive': N.SC_BV_FLUX_DIFFUSIVE,
             'total': N.SC_BV_FLUX_TOTAL}
    for ftype, spec in (bc.get('intercell_fluxes') or {}).items():
      if ftype not in kinds:
        raise ValueError(f'unknown flux boundary type {ftype}')
      bad = set(spec) - set(self.names)
      if bad:
        raise AssertionError(
            f'{bad} cannot be specified by a flux boundary condition. Only conservative'
            f' variables {set(self.names)} can be specified with a flux boundary'
            ' condition.')
      for d, dim in enumerate(types.DIMS):
        for side in (0, 1):
          for var, val in self._dirichlet_along(spec, dim, side).items():
            entries.append((kinds[ftype], d, side, self.names.index(var), val))
    prim_index = {'u': 1, 'v': 2, 'w': 3, 'e': 4, 'p': 5, 'h': 6}
    for s, name in enumerate(self.names[5:]):
      prim_index[types.conservative_to_primitive_name(name)] = 7 + s
    faces = bc.get('cell_faces') or {}
    for d, dim in enumerate(types.DIMS):
      for side in (0, 1):
        spec = self._dirichlet_along(faces, dim, side)
        if 'rho' in spec:
          # boundary.py:284-293 replaces the face primitives by
          # conservative_to_primitive_variables (rho, u, v, w, e, scalars: no
          # p, h): a later primitive value of p or h is then rejected there,
          # and otherwise the flux step fails on the missing pressure.
          left = set(self.names) | {'rho', 'u', 'v', 'w', 'e'} | {
              types.conservative_to_primitive_name(n) for n in self.names[5:]}
          for var in spec:
            if var not in left:
              raise ValueError(
                  f'"{var}" is not a valid primitive or conservative variable name for'
                  ' the face boundary condition update.')
          raise KeyError('p')
        for var, val in spec.items():
          if var in self.names:
            entries.append((N.SC_BV_FACE_CONSERVED, d, side, self.names.index(var), val))
          elif var in prim_index:
            entries.append((N.SC_BV_FACE_PRIMITIVE, d, side, prim_index[var], val))
          else:
            raise ValueError(
                f'"{var}" is not a valid primitive or conservative variable name for the'
                ' face boundary condition update.')
    if len(entries) > N.SC_MAX_BOUNDARY_VALUES:
      raise NotImplementedError(
          f'at most {N.SC_MAX_BOUNDARY_VALUES} face / flux boundary values are supported')
    params.n_boundary_values = len(entries)
    for k, (kind, d, side, var, val) in enumerate(entries):
      e = params.boundary_values[k]
      e.kind, e.dim, e.side, e.var = kind, d, side, var
      if isinstance(val, float):
        e.value = val
      else:
        # A plane of values over the face (tensor_scatter_1d_update_global
        # with a tensor update): uploaded with the context.
        e.value = 0.0
        self._pending_planes.append((e, 'plane', self._layer_plane(val, d)))

  def _plane_shape(self, d):
    nz, nx, ny = self.shape
    return [(nz, ny), (nz, nx), (nx, ny)][d]

  def _layer_plane(self, val, d):
    a, b = self._plane_shape(d)
    if isinstance(val, (list, tuple)):
      val = np.stack([np.asarray(v.detach().cpu() if torch.is_tensor(v) else v,
                                 dtype=np.float32) for v in val])
    arr = np.asarray(val.detach().cpu() if torch.is_tensor(val) else val, dtype=np.float32)
    if arr.ndim == 3:
      # A 3-D layer slice with a unit extent along the BC dimension.
      full = [self.shape[0], self.shape[1], self.shape[2]]
      full[[1, 2, 0][d]] = 1
      arr = np.broadcast_to(arr, full).reshape(a, b)
    return np.array(np.broadcast_to(arr, (a, b)), dtype=np.float32)  # a writable copy

  def _fill_dim_bc(self, faces, dim_spec, d):
    types_ = [dim_spec[s][0] for s in (0, 1)]
    periodic = bc_types.BoundaryCondition.PERIODIC in types_
    for s in (0, 1):
      out = faces[s]
      entry = dim_spec[s]
      t = entry[0]
      if t not in _BCT:
        raise NotImplementedError(f'Unsupported BC type {t}.')
      if periodic:
        out.type = N.SC_BC_PERIODIC
        continue
      if t == bc_types.BoundaryCondition.REFLECTIVE:
        raise NotImplementedError(f'Unsupported BC type {t}.')
      out.type = _BCT[t]
      val = entry[1]
      if t == bc_types.BoundaryCondition.NEUMANN:
        if not isinstance(val, numbers.Number):
          raise NotImplementedError('NEUMANN values must be scalars')
        out.value = float(val)
      elif isinstance(val, numbers.Number):
        out.value = float(val)
      elif isinstance(val, (list, tuple)):
        if len(val) != self.hw:
          raise ValueError('a per-layer DIRICHLET value needs halo_width layers')
        if all(isinstance(v, numbers.Number) for v in val):
          out.n_layer_values = self.hw
          for i, v in enumerate(val):
            out.layer_values[i] = float(v)
        else:
          planes = np.ascontiguousarray(np.stack([self._layer_plane(v, d) for v in val]))
          self._pending_planes.append((out, 'layer_planes', planes))
      else:
        raise NotImplementedError(f'unsupported DIRICHLET value of type {type(val)}')

  # -- calls ---------------------------------------------------------------

  def close(self):
    if getattr(self, '_ctx', None) is not None and self._ctx.value:
      N.lib().sc_ctx_destroy(self._ctx)
      self._ctx = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # interpreter shutdown
      pass

  def fields(self, state):
    out = []
    for name in self.names:
      if name not in state:
        raise KeyError(f'{name} missing from the conservative state')
      t = state[name]
      N.require_gpu_tensor(t, name)
      if tuple(t.shape) != self.shape:
        raise ValueError(f'{name} has shape {tuple(t.shape)}, expected {self.shape}')
      out.append(t)
    return out

  def fill_halos(self, q, stream=None):
    N.check(N.lib().sc_fill_halos(self._ctx, N.ptr_array(q), N.stream_handle(stream)),
            'sc_fill_halos')

  def out_images_fused(self):
    """Whether rk_stage_ex with SC_STAGE_FILL_OUT fills out's halos inside its
    last sweep (all-periodic grids, every extent >= 2h: copy_fill_dims with no
    clamping dim) rather than with a halo fill after it."""
    return copy_fill_dims(self._params, self.nf, self.n, self.hw) == set()

  def rhs(self, q, r, stream=None):
    N.check(N.lib().sc_rhs(self._ctx, N.ptr_array(q), N.ptr_array(r),
                           N.stream_handle(stream)), 'sc_rhs')

  def rk_stage(self, form, c0, c1, dt_dev, u0, uk, out, stream=None):
    N.check(N.lib().sc_rk_stage(
        self._ctx, form, ctypes.c_float(c0), ctypes.c_float(c1),
        ctypes.c_void_p(dt_dev.data_ptr()),
        N.ptr_array(u0) if u0 is not None else None,
        N.ptr_array(uk), N.ptr_array(out), N.stream_handle(stream)), 'sc_rk_stage')

  def rk_stage_ex(self, form, c0, c1, dt_dev, u0, uk, out, flags=0, speed_dev=None,
                  stream=None):
    """rk_stage with sc_rk_stage_ex's flags (SC_STAGE_HALOS_READY: uk's halos
    are not refilled; SC_STAGE_FILL_OUT: out leaves with its halos filled)
    and, with speed_dev, out's CFL wave speed max-accumulated into it (as
    max_wave_speed computes it; cleared by cfl_dt)."""
    N.check(N.lib().sc_rk_stage_ex(
        self._ctx, form, ctypes.c_float(c0), ctypes.c_float(c1),
        ctypes.c_void_p(dt_dev.data_ptr()),
        N.ptr_array(u0) if u0 is not None else None,
        N.ptr_array(uk), N.ptr_array(out), int(flags),
        ctypes.c_void_p(speed_dev.data_ptr()) if speed_dev is not None else None,
        N.stream_handle(stream)), 'sc_rk_stage_ex')

  def cfl_dt(self, speed_dev, cfl, dt_dev, stream=None):
    """dt_dev = cfl * min spacing / speed_dev, then speed_dev = 0."""
    N.check(N.lib().sc_cfl_dt(
        self._ctx, ctypes.c_void_p(speed_dev.data_ptr()), ctypes.c_float(cfl),
        ctypes.c_void_p(dt_dev.data_ptr()), N.stream_handle(stream)), 'sc_cfl_dt')

  def sweep_only(self, dim, form, c0, c1, dt_dev, u0, uk, out, stream=None):
    """Launches one directional sweep kernel alone (for kernel timing)."""
    N.check(N.lib().sc_sweep_only(
        self._ctx, int(dim), int(form), ctypes.c_float(c0), ctypes.c_float(c1),
        ctypes.c_void_p(dt_dev.data_ptr()) if dt_dev is not None else None,
        N.ptr_array(u0) if u0 is not None else None, N.ptr_array(uk),
        N.ptr_array(out) if out is not None else None,
        N.stream_handle(stream)), 'sc_sweep_only')

  def sweep_only_timed(self, dim, form, c0, c1, dt_dev, u0, uk, out, ev_start, ev_stop,
                       stream=None, flags=0, speed_dev=None):
    """sweep_only with the kernel's own start / stop times recorded into two
    torch.cuda.Event objects (each recorded once before, so it exists);
    flags and speed_dev as in rk_stage_ex."""
    N.check(N.lib().sc_sweep_only_timed(
        self._ctx, int(dim), int(form), ctypes.c_float(c0), ctypes.c_float(c1),
        ctypes.c_void_p(dt_dev.data_ptr()) if dt_dev is not None else None,
        N.ptr_array(u0) if u0 is not None else None, N.ptr_array(uk),
        N.ptr_array(out) if out is not None else None, int(flags),
        ctypes.c_void_p(speed_dev.data_ptr()) if speed_dev is not None else None,
        ctypes.c_void_p(ev_start.cuda_event), ctypes.c_void_p(ev_stop.cuda_event),
        N.stream_handle(stream)), 'sc_sweep_only_timed')

  def max_wave_speed(self, q, out_dev, cfl=0.0, dt_dev=None, stream=None):
    N.check(N.lib().sc_max_wave_speed(
        self._ctx, N.ptr_array(q), ctypes.c_float(cfl),
        ctypes.c_void_p(out_dev.data_ptr()),
        ctypes.c_void_p(dt_dev.data_ptr()) if dt_dev is not None else None,
        N.stream_handle(stream)), 'sc_max_wave_speed')

  def hllc_branch_codes(self, q, dim, codes, stream=None):
    N.check(N.lib().sc_hllc_branch_codes(
        self._ctx, N.ptr_array(q), int(dim), ctypes.c_void_p(codes.data_ptr()),
        N.stream_handle(stream)), 'sc_hllc_branch_codes')

  # ---- Hydrostatic reference state ------------------------------------------

  def set_reference_state(self, p_ref, rho_ref):
    """p_ref / rho_ref fields for the following calls (None, None clears)."""
    if (p_ref is None) != (rho_ref is None):
      raise ValueError('give both p_ref and rho_ref, or neither')
    if p_ref is not None:
      for name, t in (('p_ref', p_ref), ('rho_ref', rho_ref)):
        N.require_gpu_tensor(t, name)
        if tuple(t.shape) != self.shape:
          raise ValueError(f'{name} has shape {tuple(t.shape)}, expected {self.shape}')
      self._ref_state = (p_ref, rho_ref)  # kept alive while in use
      N.check(N.lib().sc_ctx_set_reference_state(
          self._ctx, ctypes.c_void_p(p_ref.data_ptr()), ctypes.c_void_p(rho_ref.data_ptr())),
              'sc_ctx_set_reference_state')
    else:
      self._ref_state = None
      N.check(N.lib().sc_ctx_set_reference_state(self._ctx, None, None),
              'sc_ctx_set_reference_state')

  # ---- Launch geometry -----------------------------------------------------

  def get_tuning(self):
    """The sweep kernels' launch geometry (dict of the sc_tuning fields)."""
    t = N.ScTuning()
    N.check(N.lib().sc_ctx_get_tuning(self._ctx, ctypes.byref(t)), 'sc_ctx_get_tuning')
    return t.as_dict()

  def set_tuning(self, tuning):
    """Sets the given fields of the launch geometry (others keep their values)."""
    t = N.ScTuning.from_dict(dict(self.get_tuning(), **tuning))
    N.check(N.lib().sc_ctx_set_tuning(self._ctx, ctypes.byref(t)), 'sc_ctx_set_tuning')

  # ---- Adjoint -------------------------------------------------------------

  def enable_adjoint(self):
    """Allocates the context's adjoint work arrays (once)."""
    if not getattr(self, '_adjoint', False):
      N.check(N.lib().sc_ctx_enable_adjoint(self._ctx), 'sc_ctx_enable_adjoint')
      self._adjoint = True

  def rhs_vjp(self, q, g, lq, stream=None):
    """lq = (d rhs / d q)^T g (interior adjoint of q; zero halo cells)."""
    self.enable_adjoint()
    N.check(N.lib().sc_rhs_vjp(self._ctx, N.ptr_array(q), N.ptr_array(g), N.ptr_array(lq),
                               N.stream_handle(stream)), 'sc_rhs_vjp')

  def rk_stage_vjp(self, form, c0, c1, dt_dev, uk, lam_out, lam_uk, lam_u0=None,
                   stream=None):
    """Reverse of rk_stage: lam_uk = c1 (lam_out + dt (d rhs/d uk)^T lam_out),
    lam_u0 += c0 lam_out (BLEND form, if given)."""
    self.enable_adjoint()
    N.check(N.lib().sc_rk_stage_vjp(
        self._ctx, int(form), ctypes.c_float(c0), ctypes.c_float(c1),
        ctypes.c_void_p(dt_dev.data_ptr()), N.ptr_array(uk), N.ptr_array(lam_out),
        N.ptr_array(lam_uk), N.ptr_array(lam_u0) if lam_u0 is not None else None,
        N.stream_handle(stream)), 'sc_rk_stage_vjp')


def _part_cfg(cfg, keep, drop):
  """cfg restricted to the conserved fields `keep`: the scalars in `drop`
  (and their primitives) leave the variable lists and the face / flux value
  tables."""
  src = cfg._simulation_cfg if hasattr(cfg, '_simulation_cfg') else cfg
  out = copy.deepcopy(dict(src))
  gone = set(drop) | {types.conservative_to_primitive_name(n) for n in drop}
  out['conservative_variable_names'] = list(keep)
  prims = out.get('primitive_variable_names')
  if prims is not None:
    out['primitive_variable_names'] = [v for v in prims if v not in gone]
  bc = dict(out.get('bc') or {})
  if bc.get('cell_faces'):
    bc['cell_faces'] = {v: s for v, s in bc['cell_faces'].items() if v not in gone}
  if bc.get('intercell_fluxes'):
    bc['intercell_fluxes'] = {t: {v: s for v, s in spec.items() if v not in gone}
                              for t, spec in bc['intercell_fluxes'].items()}
  out['bc'] = bc
  return type(cfg)(out) if hasattr(cfg, '_simulation_cfg') else out


class ChunkedContext:
  """The native context of a state with more than two transported scalars.

  Scalars are passive: their fluxes depend on the base fields (rho, momentum,
  energy) and on the scalar itself only (compressible_navier_stokes.py:419-
  427, hllc.py:182-192, hll.py).  The scalars are split in groups of two;
  part c is a NativeContext over the five base fields and group c.  Every
  call runs each part over its fields; the parts after the first send their
  copies of the base-field outputs to scratch fields of their own, and run
  before the first part, so an output that aliases u0 is written last.  The
  results equal those of one context over all fields (each field's output
  depends on the base fields and itself only).  Cost: one RHS evaluation per
  part.  No reference counterpart (the reference carries any number of
  scalars in one pass).
  """

  def __init__(self, cfg, device=None, create=True):
    names = list(_get(cfg, 'conservative_variable_names'))
    if names[:5] != list(types.BASE_CONSERVATIVE):
      raise ValueError(
          f'conservative_variable_names must start with {list(types.BASE_CONSERVATIVE)}')
    scalars = names[5:]
    self.groups = [scalars[i:i + 2] for i in range(0, len(scalars), 2)]
    self.parts = [
        NativeContext(_part_cfg(cfg, names[:5] + g, [s for s in scalars if s not in g]),
                      device=device, create=create) for g in self.groups]
    p0 = self.parts[0]
    self.names = names
    self.nf = len(names)
    self.hw, self.n, self.shape, self.h = p0.hw, p0.n, p0.shape, p0.h
    self.prim_bc_mask = 0
    for c, part in enumerate(self.parts):
      for k in range(part.nf):
        if (part.prim_bc_mask >> k) & 1:
          self.prim_bc_mask |= 1 << (k if k < 5 else 5 + 2 * c + (k - 5))
    self._sink = None
    if create:
      self._make_sink()

  def _make_sink(self):
    if self._sink is None and len(self.parts) > 1:
      self._sink = [torch.empty(self.shape, dtype=torch.float32, device=self.device)
                    for _ in range(5)]

  @property
  def device(self):
    return self.parts[0].device

  @device.setter
  def device(self, dev):
    for part in self.parts:
      part.device = dev

  @property
  def _pending_planes(self):
    return [p for part in self.parts for p in part._pending_planes]

  fields = NativeContext.fields

  def create(self):
    for part in self.parts:
      part.create()
    self._make_sink()

  def cache_key(self):
    digest = hashlib.sha256(b'chunked')
    for part in self.parts:
      digest.update(part.cache_key())
    return digest.digest()

  def close(self):
    for part in self.parts:
      part.close()

  def _split(self, fields, c, sink=False):
    """Part c's share of a field list (base fields replaced by the scratch
    fields when `sink`)."""
    if fields is None:
      return None
    base = list(fields[:5])
    if sink and c > 0:
      base = self._sink
    return base + list(fields[5 + 2 * c:7 + 2 * c])

  def _order(self):
    return range(len(self.parts) - 1, -1, -1)

  def fill_halos(self, q, stream=None):
    for c in self._order():
      self.parts[c].fill_halos(self._split(q, c), stream)

  def rhs(self, q, r, stream=None):
    for c in self._order():
      self.parts[c].rhs(self._split(q, c), self._split(r, c, sink=True), stream)

  def rk_stage(self, form, c0, c1, dt_dev, u0, uk, out, stream=None):
    for c in self._order():
      self.parts[c].rk_stage(form, c0, c1, dt_dev, self._split(u0, c), self._split(uk, c),
                             self._split(out, c, sink=True), stream)

  def max_wave_speed(self, q, out_dev, cfl=0.0, dt_dev=None, stream=None):
    self.parts[0].max_wave_speed(self._split(q, 0), out_dev, cfl, dt_dev, stream)

  def hllc_branch_codes(self, q, dim, codes, stream=None):
    self.parts[0].hllc_branch_codes(self._split(q, 0), dim, codes, stream)

  def set_reference_state(self, p_ref, rho_ref):
    for part in self.parts:
      part.set_reference_state(p_ref, rho_ref)

  def get_tuning(self):
    return self.parts[0].get_tuning()

  def set_tuning(self, tuning):
    for part in self.parts:
      part.set_tuning(tuning)

  # -- Adjoint: the base fields' RHS does not depend on the scalars, and a
  # scalar's RHS depends on the base fields and itself, so part c's
  # vector-Jacobian product with the base cotangent set to zero (c > 0)
  # gives exactly its scalars' contribution to the base adjoint, which is
  # summed over the parts.

  def _adj_buffers(self):
    if getattr(self, '_adj', None) is None:
      mk = lambda: [torch.zeros(self.shape, dtype=torch.float32, device=self.device)
                    for _ in range(5)]
      self._adj = {'zero': mk(), 'sink': mk(), 'acc': mk()}
    return self._adj

  def enable_adjoint(self):
    for part in self.parts:
      part.enable_adjoint()
    self._adj_buffers()

  def _with_base(self, fields, c, base):
    return list(base) + list(fields[5 + 2 * c:7 + 2 * c])

  def rhs_vjp(self, q, g, lq, stream=None):
    """lq = (d rhs / d q)^T g over all fields (see NativeContext.rhs_vjp)."""
    self.enable_adjoint()
    b = self._adj
    for t in b['acc']:
      t.zero_()
    for c in self._order():
      if c == 0:
        continue
      self.parts[c].rhs_vjp(self._split(q, c), self._with_base(g, c, b['zero']),
                            self._with_base(lq, c, b['sink']), stream)
      for a, s in zip(b['acc'], b['sink']):
        a.add_(s)
    self.parts[0].rhs_vjp(self._split(q, 0), self._split(g, 0), self._split(lq, 0), stream)
    for t, a in zip(lq[:5], b['acc']):
      t.add_(a)

  def rk_stage_vjp(self, form, c0, c1, dt_dev, uk, lam_out, lam_uk, lam_u0=None,
                   stream=None):
    """Reverse of rk_stage over all fields (see NativeContext.rk_stage_vjp);
    lam_uk may alias lam_out.  The parts after the first run first, with a
    zero base cotangent, so lam_out's base fields are read before they are
    overwritten."""
    self.enable_adjoint()
    b = self._adj
    for t in b['acc']:
      t.zero_()
    for c in self._order():
      if c == 0:
        continue
      self.parts[c].rk_stage_vjp(form, c0, c1, dt_dev, self._split(uk, c),
                                 self._with_base(lam_out, c, b['zero']),
                                 self._with_base(lam_uk, c, b['sink']),
                                 self._split(lam_u0, c), stream)
      for a, s in zip(b['acc'], b['sink']):
        a.add_(s)
    self.parts[0].rk_stage_vjp(form, c0, c1, dt_dev, self._split(uk, 0), self._split(lam_out, 0),
                               self._split(lam_uk, 0), self._split(lam_u0, 0), stream)
    for t, a in zip(lam_uk[:5], b['acc']):
      t.add_(a)

  def _unsupported(self, *args, **kwargs):
    raise NotImplementedError('single-kernel timing takes at most two transported scalars')

  sweep_only = sweep_only_timed = _unsupported


def make_context(cfg, device=None, create=True):
  """The native context for cfg: a NativeContext, or a ChunkedContext when
  the state carries more than two transported scalars."""
  if len(_get(cfg, 'conservative_variable_names') or []) > 7:
    return ChunkedContext(cfg, device=device, create=create)
  return NativeContext(cfg, device=device, create=create)


_CACHE_SIZE = 8
_cache = collections.OrderedDict()


def cached_context(cfg, device=None):
  """A native context (make_context) for cfg, reused across calls with an
  equal configuration (grid, schemes, BC tables and values, device): the
  stand-alone seams call this so repeated calls make no device allocations.
  The least recently used of more than 8 contexts is released."""
  ctx = make_context(cfg, device=device, create=False)
  if ctx.device.type == 'cuda' and ctx.device.index is None:
    ctx.device = torch.device('cuda', torch.cuda.current_device())
  key = ctx.cache_key()
  hit = _cache.get(key)
  if hit is not None:
    _cache.move_to_end(key)
    return hit
  ctx.create()
  _cache[key] = ctx
  while len(_cache) > _CACHE_SIZE:
    _cache.popitem(last=False)[1].close()
  return ctx


def clear_context_cache():
  """Releases every cached context."""
  while _cache:
    _cache.popitem(last=False)[1].close()
