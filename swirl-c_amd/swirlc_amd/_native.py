"""ctypes binding of libswirlc_amd.so (the C ABI in include/swirlc.h).

The shared library is built in-tree by swirl-c_amd/build_native.py.  There is
no fallback: if the library is missing, or no GPU is visible when a compute
entry point is called, an error is raised.
"""

import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lib',
                         'libswirlc_amd.so')

SC_OK = 0
SC_MAX_FIELDS = 7
SC_MAX_HALO = 4

SC_WENO_3, SC_WENO_5, SC_MUSCL = 0, 1, 2
SC_HLL, SC_HLLC = 0, 1
SC_BC_PERIODIC, SC_BC_DIRICHLET, SC_BC_NEUMANN, SC_BC_REFLECTIVE = 0, 1, 2, 3
SC_BC_NONE = -1  # no BC along a dim (its halo cells are left alone)
SC_RK_EULER, SC_RK_BLEND, SC_RK_AVG = 1, 2, 3
SC_STAGE_HALOS_READY, SC_STAGE_FILL_OUT = 1, 2

# Every entry point include/swirlc.h declares.
EXPORTED = (
    'sc_ctx_create', 'sc_ctx_destroy', 'sc_fill_halos', 'sc_rhs',
    'sc_rk_stage', 'sc_max_wave_speed', 'sc_hllc_branch_codes', 'sc_sweep_only',
    'sc_batch_create', 'sc_batch_destroy', 'sc_batch_fill_halos', 'sc_batch_rk_stage',
    'sc_batch_max_wave_speed', 'sc_ctx_enable_adjoint', 'sc_rhs_vjp', 'sc_rk_stage_vjp',
    'sc_ctx_get_tuning', 'sc_ctx_set_tuning', 'sc_batch_get_tuning', 'sc_batch_set_tuning',
    'sc_ctx_set_reference_state', 'sc_face_interp', 'sc_linear_interp', 'sc_riemann_flux',
    'sc_shear_stress', 'sc_heat_flux', 'sc_exchange_halos', 'sc_fill_plane',
    'sc_face_bc_elementwise', 'sc_last_error',
    'sc_version', 'sc_rk_stage_noalloc', 'sc_sweep_only_timed', 'sc_fill_plane_values',
    'sc_rk_stage_ex', 'sc_cfl_dt', 'sc_batch_rk_stage_ex', 'sc_batch_cfl_dt',
)


class ScBcFace(ctypes.Structure):
  _fields_ = [
      ('type', ctypes.c_int),
      ('value', ctypes.c_float),
      ('n_layer_values', ctypes.c_int),
      ('layer_values', ctypes.c_float * SC_MAX_HALO),
      ('layer_planes', ctypes.c_void_p),
  ]


class ScGrid(ctypes.Structure):
  _fields_ = [
      ('nx', ctypes.c_int), ('ny', ctypes.c_int), ('nz', ctypes.c_int),
      ('halo', ctypes.c_int),
      ('dx', ctypes.c_double), ('dy', ctypes.c_double), ('dz', ctypes.c_double),
      ('x_full', ctypes.POINTER(ctypes.c_float)),
      ('y_full', ctypes.POINTER(ctypes.c_float)),
      ('z_full', ctypes.POINTER(ctypes.c_float)),
  ]


SC_MAX_BOUNDARY_VALUES = 64
SC_MAX_FACE_VALUES = 8
(SC_BV_FLUX_CONVECTIVE, SC_BV_FLUX_DIFFUSIVE, SC_BV_FLUX_TOTAL, SC_BV_FACE_CONSERVED,
 SC_BV_FACE_PRIMITIVE) = range(5)


class ScBoundaryValue(ctypes.Structure):
  _fields_ = [
      ('kind', ctypes.c_int), ('dim', ctypes.c_int), ('side', ctypes.c_int),
      ('var', ctypes.c_int), ('value', ctypes.c_float), ('plane', ctypes.c_void_p),
  ]


class ScParams(ctypes.Structure):
  _fields_ = [
      ('n_fields', ctypes.c_int),
      ('interpolation', ctypes.c_int),
      ('numeric_flux', ctypes.c_int),
      ('include_diffusion', ctypes.c_int),
      ('nu', ctypes.c_double), ('pr', ctypes.c_double),
      ('gravity_source', ctypes.c_int),
      ('g', ctypes.c_double * 3),
      ('bc', ((ScBcFace * 2) * 3) * SC_MAX_FIELDS),
      ('prim_bc_mask', ctypes.c_int),
      ('n_boundary_values', ctypes.c_int),
      ('boundary_values', ScBoundaryValue * SC_MAX_BOUNDARY_VALUES),
  ]


class ScInstance(ctypes.Structure):
  _fields_ = [
      ('nx', ctypes.c_int), ('ny', ctypes.c_int), ('nz', ctypes.c_int),
      ('dx', ctypes.c_double), ('dy', ctypes.c_double), ('dz', ctypes.c_double),
      ('offset', ctypes.c_longlong),
      ('x_full', ctypes.POINTER(ctypes.c_float)), ('y_full', ctypes.POINTER(ctypes.c_float)),
      ('z_full', ctypes.POINTER(ctypes.c_float)),
  ]


class ScTuning(ctypes.Structure):
  _fields_ = [
      ('seg_x', ctypes.c_int), ('by_x', ctypes.c_int),
      ('seg_z', ctypes.c_int), ('by_z', ctypes.c_int),
      ('by_y', ctypes.c_int), ('xcd_remap', ctypes.c_int),
      ('adj_seg', ctypes.c_int), ('adj_by', ctypes.c_int), ('adj_by_y', ctypes.c_int),
      ('seg_y', ctypes.c_int),
  ]

  def as_dict(self):
    return {name: getattr(self, name) for name, _ in self._fields_}

  @classmethod
  def from_dict(cls, d):
    t = cls()
    for name, _ in cls._fields_:
      setattr(t, name, int(d[name]))
    return t


class NativeError(RuntimeError):
  pass


_lib = None


def lib():
  """Loads the shared library once; raises if it has not been built."""
  global _lib
  if _lib is None:
    if not os.path.exists(_LIB_PATH):
      raise NativeError(
          f'{_LIB_PATH} is missing: build it with '
          '`python swirl-c_amd/build_native.py` (no CPU fallback exists).')
    L = ctypes.CDLL(_LIB_PATH)
    vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    fpp = ctypes.POINTER(ctypes.c_void_p)
    L.sc_ctx_create.argtypes = [ctypes.POINTER(ScGrid), ctypes.POINTER(ScParams),
                                ctypes.POINTER(vp)]
    L.sc_ctx_destroy.argtypes = [vp]
    L.sc_fill_halos.argtypes = [vp, fpp, vp]
    L.sc_rhs.argtypes = [vp, fpp, fpp, vp]
    L.sc_rk_stage.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, vp]
    L.sc_rk_stage_noalloc.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, vp]
    L.sc_rk_stage_ex.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, i, vp, vp]
    L.sc_cfl_dt.argtypes = [vp, vp, f, vp, vp]
    L.sc_max_wave_speed.argtypes = [vp, fpp, f, vp, vp, vp]
    ip3 = ctypes.POINTER(ctypes.c_int)
    L.sc_face_interp.argtypes = [vp, ip3, i, i, vp, vp, vp]
    L.sc_linear_interp.argtypes = [vp, ip3, i, vp, vp]
    L.sc_riemann_flux.argtypes = [vp, i, fpp, fpp, fpp, fpp, vp, vp, fpp, vp]
    dp3 = ctypes.POINTER(ctypes.c_double)
    L.sc_shear_stress.argtypes = [vp, vp, vp, fpp, ip3, dp3, ctypes.c_double, fpp, vp]
    L.sc_heat_flux.argtypes = [vp, vp, ip3, dp3, ctypes.c_double, ctypes.c_double, fpp, vp]
    L.sc_exchange_halos.argtypes = [vp, ip3, i, ctypes.POINTER(ScBcFace), vp]
    L.sc_fill_plane.argtypes = [vp, ip3, i, i, f, vp]
    L.sc_fill_plane_values.argtypes = [vp, ip3, i, i, vp, vp]
    L.sc_face_bc_elementwise.argtypes = [i, vp, vp, vp, ctypes.c_longlong, vp]
    L.sc_hllc_branch_codes.argtypes = [vp, fpp, i, vp, vp]
    L.sc_sweep_only.argtypes = [vp, i, i, f, f, vp, fpp, fpp, fpp, vp]
    L.sc_sweep_only_timed.argtypes = [vp, i, i, f, f, vp, fpp, fpp, fpp, i, vp, vp, vp, vp]
    L.sc_batch_create.argtypes = [ctypes.POINTER(ScParams), i, i, ctypes.POINTER(ScInstance),
                                  ctypes.POINTER(vp)]
    L.sc_batch_destroy.argtypes = [vp]
    L.sc_batch_fill_halos.argtypes = [vp, fpp, vp]
    L.sc_batch_rk_stage.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, vp]
    L.sc_batch_rk_stage_ex.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, i, vp, vp]
    L.sc_batch_cfl_dt.argtypes = [vp, vp, f, vp, vp]
    L.sc_batch_max_wave_speed.argtypes = [vp, fpp, f, vp, vp, vp]
    L.sc_ctx_set_reference_state.argtypes = [vp, vp, vp]
    L.sc_ctx_get_tuning.argtypes = [vp, ctypes.POINTER(ScTuning)]
    L.sc_ctx_set_tuning.argtypes = [vp, ctypes.POINTER(ScTuning)]
    L.sc_batch_get_tuning.argtypes = [vp, ctypes.POINTER(ScTuning)]
    L.sc_batch_set_tuning.argtypes = [vp, ctypes.POINTER(ScTuning)]
    L.sc_ctx_enable_adjoint.argtypes = [vp]
    L.sc_rhs_vjp.argtypes = [vp, fpp, fpp, fpp, vp]
    L.sc_rk_stage_vjp.argtypes = [vp, i, f, f, vp, fpp, fpp, fpp, fpp, vp]
    L.sc_last_error.restype = ctypes.c_char_p
    L.sc_version.restype = ctypes.c_char_p
    for name in ('sc_ctx_create', 'sc_ctx_destroy', 'sc_fill_halos', 'sc_rhs',
                 'sc_rk_stage', 'sc_max_wave_speed', 'sc_hllc_branch_codes',
                 'sc_sweep_only', 'sc_batch_create', 'sc_batch_destroy',
                 'sc_batch_fill_halos', 'sc_batch_rk_stage', 'sc_batch_max_wave_speed',
                 'sc_ctx_enable_adjoint', 'sc_rhs_vjp', 'sc_rk_stage_vjp',
                 'sc_ctx_get_tuning', 'sc_ctx_set_tuning', 'sc_batch_get_tuning',
                 'sc_batch_set_tuning', 'sc_ctx_set_reference_state', 'sc_face_interp',
                 'sc_linear_interp', 'sc_riemann_flux', 'sc_shear_stress', 'sc_heat_flux',
                 'sc_exchange_halos', 'sc_fill_plane', 'sc_face_bc_elementwise',
                 'sc_rk_stage_noalloc', 'sc_sweep_only_timed', 'sc_fill_plane_values',
                 'sc_rk_stage_ex', 'sc_cfl_dt', 'sc_batch_rk_stage_ex', 'sc_batch_cfl_dt'):
      getattr(L, name).restype = ctypes.c_int
    _lib = L
  return _lib


def check(status, what):
  if status != SC_OK:
    msg = lib().sc_last_error().decode()
    raise NativeError(f'{what} failed (status {status}): {msg}')


def ptr_array(tensors):
  """Host array of device pointers for a list of tensors (None -> NULL)."""
  arr = (ctypes.c_void_p * SC_MAX_FIELDS)()
  for k, t in enumerate(tensors):
    arr[k] = None if t is None else t.data_ptr()
  return arr


def stream_handle(stream=None):
  s = stream if stream is not None else torch.cuda.current_stream()
  return ctypes.c_void_p(s.cuda_stream)


def require_gpu_tensor(t, name):
  if not isinstance(t, torch.Tensor):
    raise TypeError(f'{name} must be a torch.Tensor, got {type(t)}')
  if not t.is_cuda:
    raise ValueError(f'{name} must be a GPU tensor (no CPU path exists)')
  if t.dtype != torch.float32:
    raise ValueError(f'{name} must be float32, got {t.dtype}')
  if not t.is_contiguous():
    raise ValueError(f'{name} must be contiguous')
