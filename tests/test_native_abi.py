"""CPU checks of the C ABI: the library loads, exports every entry point that
include/swirlc.h declares, and rejects invalid configurations the way the
reference rejects them at trace time (no kernel is launched)."""

import ctypes
import os
import re

import pytest

from swirlc_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
  txt = open(os.path.join(ROOT, 'include', 'swirlc.h')).read()
  return sorted(set(re.findall(r'\b(sc_[a-z_]+)\s*\(', txt)))


def test_header_and_binding_agree():
  assert sorted(N.EXPORTED) == _declared()


def test_library_exports_every_declared_symbol():
  lib = N.lib()
  for name in _declared():
    assert hasattr(lib, name), name
  assert lib.sc_version().decode().startswith('swirlc_amd')


def _grid(n=8, h=3):
  g = N.ScGrid()
  g.nx = g.ny = g.nz = n
  g.halo = h
  g.dx = g.dy = g.dz = 1.0 / n
  return g


def _params(interp=N.SC_WENO_5, flux=N.SC_HLLC):
  p = N.ScParams()
  p.n_fields = 5
  p.interpolation = interp
  p.numeric_flux = flux
  p.nu, p.pr = 1e-4, 0.7
  return p


def _create(g, p):
  h = ctypes.c_void_p()
  st = N.lib().sc_ctx_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(h))
  return st, N.lib().sc_last_error().decode()


@pytest.mark.parametrize('interp,flux,needle', [
    (7, N.SC_HLLC, 'interpolation_scheme'),
    (N.SC_WENO_5, 9, 'numeric_flux_scheme'),
])
def test_unknown_schemes_rejected(interp, flux, needle):
  st, msg = _create(_grid(), _params(interp, flux))
  assert st != N.SC_OK and needle in msg


def test_halo_too_small_for_weno5_rejected():
  st, msg = _create(_grid(h=2), _params())
  assert st == 1 and 'halo_width' in msg


def test_gravity_without_direction_rejected():
  p = _params()
  p.gravity_source = 1
  st, msg = _create(_grid(), p)
  assert st != N.SC_OK and 'Gravity direction' in msg


def test_reflective_rejected():
  p = _params()
  p.bc[0][0][0].type = N.SC_BC_REFLECTIVE
  st, msg = _create(_grid(), p)
  assert st == 2 and 'REFLECTIVE' in msg


@pytest.mark.parametrize('nf', [4, 8])
def test_field_count_out_of_range_rejected(nf):
  p = _params()
  p.n_fields = nf
  st, msg = _create(_grid(), p)
  assert st == 2 and 'n_fields' in msg


@pytest.mark.parametrize('nx,ny', [(1 << 20, 8), (4096, (1 << 18) - 6)])
def test_grid_beyond_y_sweep_indexing_rejected(nx, ny):
  g = _grid()
  g.nx, g.ny = nx, ny
  st, msg = _create(g, _params())
  assert st == 2 and 'grid too large' in msg


def test_periodic_with_too_few_cells_rejected():
  st, msg = _create(_grid(n=2), _params(N.SC_MUSCL))
  assert st != N.SC_OK and 'PERIODIC' in msg


def test_null_arguments_rejected():
  L = N.lib()
  assert L.sc_rhs(None, None, None, None) == 1
  assert L.sc_fill_halos(None, None, None) == 1
  assert L.sc_rk_stage(None, 1, 0.0, 0.0, None, None, None, None, None) == 1


def test_interpolation_entry_points_reject_bad_arguments():
  """Argument checks of sc_face_interp / sc_linear_interp run before any GPU use."""
  L = N.lib()
  ext = (ctypes.c_int * 3)(4, 4, 4)
  a, b, c = ctypes.c_void_p(16), ctypes.c_void_p(32), ctypes.c_void_p(48)
  assert L.sc_face_interp(None, ext, 0, N.SC_WENO_5, b, c, None) == 1
  assert L.sc_face_interp(a, ext, 3, N.SC_WENO_5, b, c, None) == 1
  assert L.sc_face_interp(a, ext, 0, N.SC_WENO_5, a, c, None) == 1
  assert L.sc_face_interp(a, ext, 0, 7, b, c, None) == 2
  assert 'interpolation_scheme' in L.sc_last_error().decode()
  assert L.sc_face_interp(a, (ctypes.c_int * 3)(4, 0, 4), 0, N.SC_MUSCL, b, c, None) == 1
  assert L.sc_linear_interp(a, ext, 1, a, None) == 1
  assert L.sc_linear_interp(a, ext, -1, b, None) == 1


def test_interpolation_host_checks():
  import torch
  from swirlc_amd.numerics import interpolation
  with pytest.raises(ValueError, match='valid dimension'):
    interpolation.weno(torch.zeros(3, 3, 3), 'w')
  with pytest.raises(ValueError, match='GPU tensor'):
    interpolation.linear_interpolation(torch.zeros(3, 3, 3), 'x')
  with pytest.raises(NotImplementedError):
    interpolation.weno(torch.zeros(3, 3, 3), 'x', k=4)
  with pytest.raises(NotImplementedError, match='MUSCL'):
    interpolation.flux_limiter(torch.zeros(3, 3, 3), 'x', 'VAN_LEER')


def test_boundary_and_riemann_seams_host_checks():
  """The stand-alone seams validate their arguments before any GPU use."""
  import torch
  from swirlc_amd.boundary import bc_types, boundary
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.numerics.riemann_solver import riemann_solver
  BC = bc_types.BoundaryCondition
  with pytest.raises(ValueError, match='GPU tensor'):
    boundary.exchange_halos(0, None, torch.zeros(8, 8, 8), {}, 2)
  spec = {'u': {d: {0: (BC.DIRICHLET, 0.0), 1: (None, None)} for d in 'xyz'}}
  cfg = SwirlCParameters({'bc': {'intercell_fluxes': {'convective': spec}}})
  with pytest.raises(AssertionError, match='flux boundary condition'):
    boundary.update_fluxes(0, None, {}, cfg, 'convective')
  assert boundary.update_fluxes(0, None, {'x': {}}, cfg, 'total') == {'x': {}}
  with pytest.raises(ValueError, match='numeric_flux_scheme'):
    riemann_solver.select_numeric_flux_fn(SwirlCParameters({'numeric_flux_scheme': 'ROE'}))
  L = N.lib()
  ext = (ctypes.c_int * 3)(4, 4, 4)
  assert L.sc_fill_plane(ctypes.c_void_p(16), ext, 1, 4, ctypes.c_float(1.0), None) == 1
  assert L.sc_exchange_halos(None, ext, 2, None, None) == 1
  assert L.sc_riemann_flux(None, 0, None, None, None, None, None, None, None, None) == 1
  hh = (ctypes.c_double * 3)(0.1, 0.0, 0.1)
  assert L.sc_heat_flux(ctypes.c_void_p(16), ctypes.c_void_p(32), ext, hh, 1e-3, 0.7,
                        (ctypes.c_void_p * 3)(48, 64, 80), None) == 1


def test_python_config_validation_errors():
  """The Python host raises the reference's exception types before any GPU use."""
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.equations.native_context import NativeContext
  with pytest.raises(NotImplementedError, match='interpolation_scheme'):
    NativeContext(SwirlCParameters({'interpolation_scheme': 'WENO_7'}))
  with pytest.raises(ValueError, match='numeric_flux_scheme'):
    NativeContext(SwirlCParameters({'numeric_flux_scheme': 'ROE'}))
  with pytest.raises(NotImplementedError, match='source term'):
    NativeContext(SwirlCParameters({'source_functions': ['coriolis']}))
  with pytest.raises(ValueError, match='cells'):
    NativeContext(SwirlCParameters({'bc': {'cell_averages': {}}}))
  base = ['rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e']
  with pytest.raises(NotImplementedError, match='two transported scalars'):
    NativeContext(SwirlCParameters({
        'conservative_variable_names': base + ['rho_a', 'rho_b', 'rho_c'],
        'primitive_variable_names': ['rho', 'u', 'v', 'w', 'e', 'a', 'b', 'c']}))


def test_plane_values_change_the_context():
  """Contexts built from equal configurations but different face planes are
  different (the plane bytes are part of the cached context's key)."""
  import test_gpu_boundary_values as BV
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.equations.native_context import NativeContext
  n = (8, 8, 8)
  f1 = BV._planes(BV._face_spec(), n, 3, 60)
  f2 = BV._planes(BV._face_spec(), n, 3, 61)
  k1 = NativeContext(SwirlCParameters(BV._cfg(n, 'WENO_5', 'HLLC', False, faces=f1)),
                     create=False).cache_key()
  k1b = NativeContext(SwirlCParameters(BV._cfg(n, 'WENO_5', 'HLLC', False, faces=f1)),
                      create=False).cache_key()
  k2 = NativeContext(SwirlCParameters(BV._cfg(n, 'WENO_5', 'HLLC', False, faces=f2)),
                     create=False).cache_key()
  assert k1 == k1b and k1 != k2


def test_batch_rejects_coordinates_of_the_wrong_length():
  """A gravity ensemble's instance coordinates are read by the host over the
  instance's extents: a wrong length is refused before anything is built."""
  import test_gpu_batch as TB
  from swirlc_amd.core.batch import BatchSolver
  from swirlc_amd.core.parameter import SwirlCParameters
  shape = (8, 6, 5)
  cfg = TB._wide_cfg(shape, True, False)
  grid = (*shape, cfg['dx'], cfg['dy'], cfg['dz'], cfg['x'], cfg['y'][:-1], cfg['z'])
  with pytest.raises(ValueError, match='y coordinates'):
    BatchSolver(SwirlCParameters(cfg), [(grid, {})], cfl=None)
