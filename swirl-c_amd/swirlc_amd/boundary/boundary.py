"""Halo updates on the GPU (boundary/boundary.py).

  exchange_halos(replica_id, replicas, u, bc_dict, width)
      boundary/boundary.py:34-86
  update_fluxes(replica_id, replicas, fluxes, cfg, flux_type)
      boundary/boundary.py:89-183
  update_faces(replica_id, replicas, conservative_neg, conservative_pos,
               primitive_neg, primitive_pos, cfg)
      boundary/boundary.py:230-461
  update_boundary(replica_id, replicas, states, cfg)
      boundary/boundary.py:606-628
  update_conservative_cell_averages(replica_id, replicas, conservative, cfg)
      boundary/boundary.py:464-539

Fills the halo layers of every conservative field from cfg.bc['cell_averages']
(PERIODIC, DIRICHLET with scalar / per-layer values or planes, NEUMANN, and
BCs given on the primitive variable, boundary.py:516-531), in the x, y, z
order of swirl_lm's halo exchange, with libswirlc_amd's halo kernels
(sc_fill_halos).  As in the reference the input dict is left unchanged and a
new dict is returned.  Fields are float32 GPU tensors [nz+2h, nx+2h, ny+2h];
one replica.  The native context of a configuration is made once and reused
by later calls with an equal configuration.
"""

import ctypes
import numbers

import torch

from swirlc_amd import _native as N
from swirlc_amd.common import types
from swirlc_amd.equations.native_context import NativeContext, cached_context


class _OneField:
  """The BC translation of NativeContext for one field of a given shape."""
  _plane_shape = NativeContext._plane_shape
  _layer_plane = NativeContext._layer_plane
  _fill_dim_bc = NativeContext._fill_dim_bc
  upload_planes = NativeContext.upload_planes

  def __init__(self, shape, hw, device):
    self.shape = tuple(shape)
    self.hw = hw
    self.device = device
    self._keep = []
    self._pending_planes = []


def exchange_halos(replica_id, replicas, u, bc_dict, width):
  """New tensor: u with its `width` halo layers filled from `bc_dict`
  ({'x' | 'y' | 'z': {0 | 1: (BoundaryCondition, value)}}: PERIODIC,
  DIRICHLET with a scalar, per-layer scalars or planes, NEUMANN), x then y
  then z as swirl_lm's halo exchange.  One replica."""
  del replica_id, replicas
  N.require_gpu_tensor(u, 'u')
  if u.dim() != 3:
    raise ValueError(f'expected a 3-D field [nz, nx, ny], got shape {tuple(u.shape)}')
  hw = int(width)
  nz, nx, ny = (int(x) - 2 * hw for x in u.shape)
  if min(nx, ny, nz) < 1:
    raise ValueError('the field has no interior cells for this halo width')
  t = _OneField(u.shape, hw, u.device)
  faces = (N.ScBcFace * 6)()
  for d, dim in enumerate(('x', 'y', 'z')):
    if bc_dict.get(dim) is None:
      # boundary.py:58-60: a dim without a BC is not exchanged.
      faces[2 * d].type = faces[2 * d + 1].type = N.SC_BC_NONE
      continue
    t._fill_dim_bc((faces[2 * d], faces[2 * d + 1]), bc_dict[dim], d)
  t.upload_planes()
  out = u.clone()
  N.check(N.lib().sc_exchange_halos(ctypes.c_void_p(out.data_ptr()),
                                    (ctypes.c_int * 3)(nx, ny, nz), hw, faces,
                                    N.stream_handle()), 'sc_exchange_halos')
  if t._keep:
    torch.cuda.current_stream().synchronize()  # layer planes stay alive until used
  return out


def update_conservative_cell_averages(replica_id, replicas, conservative, cfg):
  """New dict of the conservative fields with their halo cells filled.

  A field whose BC is given on its primitive comes back as
  rho * divide_no_nan(q, rho) in every cell, interior included, as the
  reference computes it (boundary.py:516-531); the in-place native fill
  leaves the interior alone (the sweeps apply that rounding as they load).
  """
  del replica_id, replicas
  ctx = cached_context(cfg)
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


def _flux_spec(cfg, flux_type):
  bc = cfg.bc or {}
  table = bc.get('intercell_fluxes') or {}
  for key in (flux_type, getattr(flux_type, 'name', None), getattr(flux_type, 'value', None)):
    if key is not None:
      for k in (key, str(key).lower()):
        if k in table:
          return table[k]
  return None


def update_fluxes(replica_id, replicas, fluxes, cfg, flux_type):
  """Fluxes {dim: {var: tensor}} with the DIRICHLET values of
  cfg.bc['intercell_fluxes'][flux_type] written on the boundary flux planes
  (index halo_width on the low face, halo_width + core_n on the high face);
  unchanged when no such entry exists.  Only conservative variables may
  carry a flux BC (AssertionError, as in the reference); a None type is
  skipped; values are scalars or planes (see _plane).  Modified fields are
  new tensors."""
  del replica_id, replicas
  spec = _flux_spec(cfg, flux_type)
  if spec is None:
    return fluxes
  names = list(cfg.conservative_variable_names)
  bad = set(spec) - set(names)
  assert not bad, (f'{bad} cannot be specified by a flux boundary condition. Only '
                   f'conservative variables {set(names)} can be specified with a flux '
                   'boundary condition.')
  hw = int(cfg.halo_width)
  core_n = {'x': cfg.core_nx, 'y': cfg.core_ny, 'z': cfg.core_nz}
  out = {dim: dict(v) for dim, v in fluxes.items()}
  for d, dim in enumerate(('x', 'y', 'z')):
    for side in (0, 1):
      for var, faces in spec.items():
        entry = faces[dim][side]
        kind = getattr(entry[0], 'name', entry[0])
        if kind is None:
          continue
        assert kind == 'DIRICHLET', f'Unsupported interface (face or flux) BC type {kind}.'
        index = hw if side == 0 else hw + int(core_n[dim])
        out[dim][var] = _plane(out[dim][var], d, index, entry[1], 'flux')
  return out


def _ew(op, a, b):
  N.require_gpu_tensor(a, 'face state')
  N.require_gpu_tensor(b, 'face state')
  if a.shape != b.shape or a.device != b.device:
    raise ValueError(f'face states of shapes {tuple(a.shape)} and {tuple(b.shape)} '
                     f'(devices {a.device}, {b.device}) do not match')
  out = torch.empty_like(a)
  N.check(N.lib().sc_face_bc_elementwise(op, ctypes.c_void_p(a.data_ptr()),
                                         ctypes.c_void_p(b.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), a.numel(),
                                         N.stream_handle()), 'sc_face_bc_elementwise')
  return out


def _plane(f, d, index, val, what='face state'):
  """A copy of f with plane `index` normal to dim d set to val: a scalar, or
  a plane of values (2-D over the two other axes, a 3-D slice with unit
  extent along d, or a list of such slices), as
  tensor_scatter_1d_update_global takes them."""
  N.require_gpu_tensor(f, what)
  if f.dim() != 3:
    raise ValueError(f'expected a 3-D {what} [nz, nx, ny], got shape {tuple(f.shape)}')
  f = f.clone()
  nz, nx, ny = f.shape
  ext = (ctypes.c_int * 3)(nx, ny, nz)
  if isinstance(val, numbers.Number):
    N.check(N.lib().sc_fill_plane(ctypes.c_void_p(f.data_ptr()), ext, d, index,
                                  ctypes.c_float(float(val)), N.stream_handle()),
            'sc_fill_plane')
    return f
  host = _OneField(f.shape, 0, f.device)._layer_plane(val, d)
  plane = torch.from_numpy(host).to(f.device)
  N.check(N.lib().sc_fill_plane_values(ctypes.c_void_p(f.data_ptr()), ext, d, index,
                                       ctypes.c_void_p(plane.data_ptr()), N.stream_handle()),
          'sc_fill_plane_values')
  return f


def _face_update_1d(cons, prim, spec, d, index):
  """_face_boundary_update_1d (boundary.py:230-337) on one face plane."""
  if not spec:
    return cons, prim
  cons = dict(cons)
  prim = dict(prim)
  if 'rho' in spec:
    cons['rho'] = _plane(cons['rho'], d, index, spec['rho'])
    # The reference rebuilds the whole primitive dict here (p and h are gone).
    prim = {'rho': cons['rho']}
    for k, q in cons.items():
      if k != 'rho' and k.startswith('rho_') and len(k) > 4:
        prim[types.conservative_to_primitive_name(k)] = _ew(2, q, cons['rho'])
  for var, val in spec.items():
    if var == 'rho':
      continue
    if var in cons:
      cons[var] = _plane(cons[var], d, index, val)
      prim[types.conservative_to_primitive_name(var)] = _ew(0, cons[var], cons['rho'])
    else:
      if var not in prim:
        raise ValueError(f'"{var}" is not a valid primitive or conservative'
                         ' variable name for the face boundary condition.')
      prim[var] = _plane(prim[var], d, index, val)
      cons[types.primitive_to_conservative_name(var)] = _ew(1, cons['rho'], prim[var])
  return cons, prim


def update_faces(replica_id, replicas, conservative_neg, conservative_pos, primitive_neg,
                 primitive_pos, cfg):
  """The four face-state dicts ({dim: {var: tensor}}) with the DIRICHLET
  values of cfg.bc['cell_faces'] written on the boundary face planes (index
  halo_width low, halo_width + core_n high), each value followed by the
  reference's update of its partner variable over the whole array
  (primitive = divide_no_nan(conservative, rho), or conservative = rho *
  primitive); a density value rebuilds the primitive dict from the
  conservative one, as the reference does.  Values are scalars or planes
  (see _plane); inputs are not modified."""
  del replica_id, replicas
  table = (cfg.bc or {}).get('cell_faces')
  if not table:
    return conservative_neg, conservative_pos, primitive_neg, primitive_pos
  hw = int(cfg.halo_width)
  core_n = {'x': cfg.core_nx, 'y': cfg.core_ny, 'z': cfg.core_nz}
  cn = dict(conservative_neg)
  cp = dict(conservative_pos)
  pn = dict(primitive_neg)
  pp = dict(primitive_pos)
  for d, dim in enumerate(('x', 'y', 'z')):
    for side in (0, 1):
      spec = {}
      for var, faces in table.items():
        entry = faces[dim][side]
        kind = getattr(entry[0], 'name', entry[0])
        if kind is None:
          continue
        assert kind == 'DIRICHLET', f'Unsupported interface (face or flux) BC type {kind}.'
        spec[var] = entry[1]
      index = hw if side == 0 else hw + int(core_n[dim])
      cn[dim], pn[dim] = _face_update_1d(cn[dim], pn[dim], spec, d, index)
      cp[dim], pp[dim] = _face_update_1d(cp[dim], pp[dim], spec, d, index)
  return cn, cp, pn, pp


def update_boundary(replica_id, replicas, states, cfg):
  """{key: exchange_halos(u, cfg.bc[key], cfg.halo_width)} for every state
  (boundary.py:606-628): each variable's BC dict sits under its own name in
  cfg.bc."""
  return {key: exchange_halos(replica_id, replicas, val, cfg.bc[key], cfg.halo_width)
          for key, val in states.items()}
