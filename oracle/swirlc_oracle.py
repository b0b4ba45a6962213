"""CPU oracle: a numpy restatement of Swirl-C's compressible Navier-Stokes RHS.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this module, and only as the
checker (or the timed CPU baseline), never as a product code path.

Every function restates one reference function op by op, in the reference's
order of floating-point operations, with TensorFlow's semantics for
`divide_no_nan`, `reciprocal_no_nan`, `tf.where` and Python-scalar promotion
(a Python float meets an fp32 tensor as fp32).  Arrays are numpy arrays shaped
`[nz, nx, ny]` including halos (y fastest), exactly the reference's
`FlowFieldVar` layout (swirl_c/common/types.py:31-33).  The dtype of the input
arrays (float32 or float64) is the arithmetic type throughout, which gives the
fp32 oracle and the fp64 oracle from one source.

Pinning (what checks each function; see tests/test_oracle_*.py):
  * stencils forward_1/backward_1/central_2/linear_interpolation: the
    reference's golden vectors (numerics/gradient_test.py:53-127) and closed
    forms (numerics/interpolation_test.py:30-52).
  * exchange_halos: the reference's halo golden values
    (boundary/boundary_test.py:61-341).
  * EOS, conversions, physical flux, HLL, HLLC, RHS, source, viscous, RK:
    the reference's closed-form tests (see the test files' docstrings).
  * the p_ref / rho_ref helper path (update_helper_variables, the p_dyn swap
    of hllc_flux, d_rho in the gravity source): a quiescent hydrostatic
    column with p_ref = p, rho_ref = rho gives RHS == 0 (HLLC) and zero
    momentum RHS (HLL); an isolated contact keeps its exact HLLC flux with
    p_dyn; each fails on a planted wrong substitution
    (tests/test_oracle_pins.py).
  * max_wave_speed (no reference counterpart): closed forms max |u| + c on
    uniform states with one extreme cell, with and without gravity.
  * edge and corner halo cells: golden vectors of the prototype notebook's
    sequential x, y, z fill (demo/swirl_c_prototype.ipynb cell 2,
    tests/golden/notebook_halo.npz).
  * weno / muscl: swirl_lm is not vendored (pyproject.toml:31-32) and cannot
    be imported here, so these follow the in-repo prototype notebook
    (demo/swirl_c_prototype.ipynb cell 4, `weno5` and `muscl`), shifted from
    its i+1/2 storage to swirl_lm's i-1/2 storage.  They are pinned to golden
    vectors produced from that notebook (tests/golden/) and to mathematical
    invariants (exactness on quadratics, ENO behaviour at a step).  Against
    swirl_lm itself they are "parity unpinned"; WENO_3 has no notebook
    counterpart and is pinned by textbook invariants only.
"""

import numpy as np

# Physical constants (physics/constant.py:17-41).
G = 9.81
P_0 = 1.01325e5
GAMMA = 1.4
W_AIR = 0.029
R_U = 8.314
R = R_U / W_AIR
CV = R / (GAMMA - 1.0)
CP = CV + R
KAPPA = R / CP

# common/types.py:28
SMALL = float(np.finfo(np.float32).resolution)

DIMS = ('x', 'y', 'z')
VELOCITY = ('u', 'v', 'w')
MOMENTUM = ('rho_u', 'rho_v', 'rho_w')
BASE_CONSERVATIVE = ('rho', 'rho_u', 'rho_v', 'rho_w', 'rho_e')
BASE_PRIMITIVES = ('rho', 'u', 'v', 'w', 'e')
# Array axis of each physical dimension: layout [z][x][y].
AXIS = {'x': 1, 'y': 2, 'z': 0}

PERIODIC, DIRICHLET, NEUMANN, REFLECTIVE = 'PERIODIC', 'DIRICHLET', 'NEUMANN', 'REFLECTIVE'

WENO_EPS = 1e-6
HLLC_SM_SNAP = 1e-6


def prim_name(cons_name):
  """common/types.py:108-130 (conservative_to_primitive_name)."""
  if cons_name == 'rho':
    return 'rho'
  if not cons_name.startswith('rho_') or len(cons_name) <= 4:
    raise ValueError(f'Invalid conserved variable name: {cons_name}.')
  return cons_name[4:]


def cons_name(prim):
  """common/types.py:88-105 (primitive_to_conservative_name)."""
  return 'rho' if prim == 'rho' else 'rho_' + prim


def _dnn(x, y):
  """tf.math.divide_no_nan: 0 where y == 0, else x / y."""
  with np.errstate(divide='ignore', invalid='ignore'):
    q = x / y
  return np.where(y == 0, np.zeros_like(q), q)


def _rnn(x):
  """tf.math.reciprocal_no_nan: 0 where x == 0, else 1 / x."""
  with np.errstate(divide='ignore', invalid='ignore'):
    q = 1.0 / x
  return np.where(x == 0, np.zeros_like(q), q)


def _sqrt(x):
  with np.errstate(invalid='ignore'):
    return np.sqrt(x)


def _c(dtype, value):
  """A Python/numpy scalar as TF would convert it next to a `dtype` tensor."""
  return np.dtype(dtype).type(value)


def _sl(axis, a, b):
  idx = [slice(None)] * 3
  idx[axis] = slice(a, b)
  return tuple(idx)


def _shift(f, axis, s):
  """g[i] = f[i + s] along `axis`, zero where i + s is outside the array.

  The zero padding stands for swirl_lm's kernel-op edge values, which only
  ever land in halo cells (their values are irrelevant to the interior RHS).
  """
  g = np.zeros_like(f)
  n = f.shape[axis]
  if s > 0:
    g[_sl(axis, 0, n - s)] = f[_sl(axis, s, n)]
  elif s < 0:
    g[_sl(axis, -s, n)] = f[_sl(axis, 0, n + s)]
  else:
    g[...] = f
  return g


# ---------------------------------------------------------------------------
# Stencils: numerics/gradient.py:21-94, numerics/interpolation.py:21-53, with
# swirl_lm kernel ops kdx+ (f[i+1]-f[i]), kdx (f[i]-f[i-1]),
# kDx (f[i+1]-f[i-1]) and ksx (f[i-1]+f[i]).
# ---------------------------------------------------------------------------


def forward_1(f, h, dim):
  """numerics/gradient.py:21-43."""
  ax = AXIS[dim]
  return (_shift(f, ax, 1) - f) / _c(f.dtype, h)


def backward_1(f, h, dim):
  """numerics/gradient.py:46-68."""
  ax = AXIS[dim]
  return (f - _shift(f, ax, -1)) / _c(f.dtype, h)


def central_2(f, h, dim):
  """numerics/gradient.py:71-94."""
  ax = AXIS[dim]
  return (_shift(f, ax, 1) - _shift(f, ax, -1)) / _c(f.dtype, 2.0 * h)


def linear_interpolation(f, dim):
  """numerics/interpolation.py:21-53: 0.5 * (f[i-1] + f[i]) on face i-1/2."""
  ax = AXIS[dim]
  return 0.5 * (_shift(f, ax, -1) + f)


# ---------------------------------------------------------------------------
# Face reconstruction (swirl_lm interpolation.weno / flux_limiter contract:
# returns (neg, pos) = left / right states of face i-1/2 stored at index i).
# Follows demo/swirl_c_prototype.ipynb cell 4 (`weno5`, `muscl`), shifted to
# i-1/2 storage.  Entries whose stencil leaves the array keep the cell value,
# as the prototype's `clear_interior` does.
# ---------------------------------------------------------------------------

_WENO_C = {
    2: {-1: (1.5, -0.5), 0: (0.5, 0.5), 1: (-0.5, 1.5)},
    3: {
        -1: (11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0),
        0: (1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0),
        1: (-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0),
        2: (1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0),
    },
}
_WENO_D = {2: (2.0 / 3.0, 1.0 / 3.0), 3: (0.3, 0.6, 0.1)}


def weno_smoothness(win, k):
  """Jiang-Shu smoothness indicators beta_r of the cell at offset 0.

  `win(s)` returns the cell values at offset s.  k=3: prototype cell 4
  l.137-166.  k=2: (v[i+1]-v[i])^2 and (v[i]-v[i-1])^2 (Jiang & Shu 1996).
  """
  if k == 3:
    b0 = (13.0 / 12.0 * (win(0) - 2.0 * win(1) + win(2)) ** 2
          + 0.25 * (3.0 * win(0) - 4.0 * win(1) + win(2)) ** 2)
    b1 = (13.0 / 12.0 * (win(-1) - 2.0 * win(0) + win(1)) ** 2
          + 0.25 * (win(-1) - win(1)) ** 2)
    b2 = (13.0 / 12.0 * (win(-2) - 2.0 * win(-1) + win(0)) ** 2
          + 0.25 * (win(-2) - 4.0 * win(-1) + 3.0 * win(0)) ** 2)
    return [b0, b1, b2]
  if k == 2:
    return [(win(1) - win(0)) ** 2, (win(0) - win(-1)) ** 2]
  raise ValueError(f'Unsupported WENO order k={k}.')


def weno(v, dim, k=3):
  """WENO-(2k-1) face states (neg, pos) at i-1/2 stored at i.

  Prototype cell 4 l.70-211: candidate r at cell i uses v[i-r .. i-r+k-1];
  neg (face i+1/2) weights d[r], pos (face i-1/2) weights d[k-1-r],
  alpha = d / (eps + beta)^2, eps = 1e-6.
  """
  ax = AXIS[dim]
  n = v.shape[ax]
  off = k - 1
  l = n - 2 * off
  if l <= 0:
    return v.copy(), v.copy()
  win = lambda s: v[_sl(ax, off + s, off + s + l)]
  c, d = _WENO_C[k], _WENO_D[k]
  q_neg, q_pos = [], []
  for r in range(k):
    bn = np.zeros_like(win(0))
    bp = np.zeros_like(win(0))
    for j in range(k):
      vv = win(j - r)
      bn += c[r][j] * vv
      bp += c[r - 1][j] * vv
    q_neg.append(bn)
    q_pos.append(bp)
  beta = weno_smoothness(win, k)
  a_neg, a_pos = [], []
  s_neg = np.zeros_like(beta[0])
  s_pos = np.zeros_like(beta[0])
  for r in range(k):
    a_neg.append(d[r] / (WENO_EPS + beta[r]) ** 2)
    a_pos.append(d[k - 1 - r] / (WENO_EPS + beta[r]) ** 2)
    s_neg += a_neg[r]
    s_pos += a_pos[r]
  f_neg = np.zeros_like(beta[0])
  f_pos = np.zeros_like(beta[0])
  for r in range(k):
    f_neg += (a_neg[r] / s_neg) * q_neg[r]
    f_pos += (a_pos[r] / s_pos) * q_pos[r]
  neg = v.copy()
  pos = v.copy()
  neg[_sl(ax, off + 1, off + 1 + l)] = f_neg  # face (i+1/2) of cell i -> i+1
  pos[_sl(ax, off, off + l)] = f_pos          # face (i-1/2) of cell i -> i
  return neg, pos


def _nb_central2(f, h, ax):
  """Prototype cell 4 `central2`: zero at both ends."""
  g = np.zeros_like(f)
  n = f.shape[ax]
  g[_sl(ax, 1, n - 1)] = (f[_sl(ax, 2, n)] - f[_sl(ax, 0, n - 2)]) / (2.0 * h)
  return g


def _nb_forward1(f, h, ax):
  g = np.zeros_like(f)
  n = f.shape[ax]
  g[_sl(ax, 0, n - 1)] = (f[_sl(ax, 1, n)] - f[_sl(ax, 0, n - 1)]) / h
  return g


def _nb_backward1(f, h, ax):
  g = np.zeros_like(f)
  n = f.shape[ax]
  g[_sl(ax, 1, n)] = (f[_sl(ax, 1, n)] - f[_sl(ax, 0, n - 1)]) / h
  return g


def minmod3(a0, a1, a2):
  """Prototype cell 4 l.217-224: sign-agreement minmod of three slopes."""
  s = (np.sign(a0) + np.sign(a1) + np.sign(a2)) / 3
  a_min = np.minimum(np.minimum(np.abs(a0), np.abs(a1)), np.abs(a2))
  return np.where(np.abs(s) == 1, s * a_min, np.zeros_like(a_min))


def muscl(v, dim):
  """MUSCL face states (neg, pos) at i-1/2 stored at i.

  Prototype cell 4 l.214-239: m = minmod(central2(v, 2), forward1(v, 1),
  backward1(v, 1)); left state of face i+1/2 = v[i] + m[i]/2, right state
  = v[i+1] - m[i+1]/2.  Stored at i-1/2: neg[i] = v[i-1] + m[i-1]/2,
  pos[i] = v[i] - m[i]/2.
  """
  ax = AXIS[dim]
  n = v.shape[ax]
  m = minmod3(_nb_central2(v, 2.0, ax), _nb_forward1(v, 1.0, ax),
              _nb_backward1(v, 1.0, ax))
  neg = v.copy()
  neg[_sl(ax, 1, n)] = v[_sl(ax, 0, n - 1)] + 0.5 * m[_sl(ax, 0, n - 1)]
  pos = v - 0.5 * m
  return neg, pos


def face_interp_fn(scheme):
  """equations/compressible_navier_stokes.py:581-595."""
  if scheme == 'WENO_3':
    return lambda v, dim: weno(v, dim, 2)
  if scheme == 'WENO_5':
    return lambda v, dim: weno(v, dim, 3)
  if scheme == 'MUSCL':
    return muscl
  raise NotImplementedError(
      f'"{scheme}" is not implemented as an "interpolation_scheme". Valid'
      " options are: ['WENO_3', 'WENO_5', 'MUSCL']")


# ---------------------------------------------------------------------------
# Mesh, conversions and thermodynamics.
# ---------------------------------------------------------------------------


def mesh_full(coord, hw, dtype):
  """common/utils.py:41-89 for one replica: coordinates incl. halos.

  The low halo is (x[1]-x[0]) * [-hw..-1] (not x[0] - k*dx), as the
  reference writes it (common/utils.py:58-75).
  """
  c = np.asarray(coord, dtype=dtype)
  lo = np.linspace(-hw, -1, hw).astype(dtype)
  hi = np.linspace(1, hw, hw).astype(dtype)
  return np.concatenate([(c[1] - c[0]) * lo, c, c[-1] + (c[-1] - c[-2]) * hi])


def gravity_direction(g):
  """common/utils.py:114-142."""
  g_vec = np.array([g[d] for d in DIMS], dtype=float)
  if abs(np.linalg.norm(g_vec) - 1.0) >= SMALL:
    return -1
  for i, d in enumerate(DIMS):
    if abs(abs(g[d]) - 1.0) < SMALL:
      return i
  return 3


def gravitational_potential(cfg, dtype):
  """physics/thermodynamics/generic.py:282-319: pe = -G g . x at cells."""
  if gravity_direction(cfg['g']) == -1:
    # zeros_like(mesh['z']) broadcasts to the field shape.
    return np.zeros((1, 1, 1), dtype=dtype)
  hw = cfg['halo_width']
  mx = mesh_full(cfg['x'], hw, dtype)[np.newaxis, :, np.newaxis]
  my = mesh_full(cfg['y'], hw, dtype)[np.newaxis, np.newaxis, :]
  mz = mesh_full(cfg['z'], hw, dtype)[:, np.newaxis, np.newaxis]
  gv = G * np.array([cfg['g'][d] for d in DIMS], dtype=float)
  return (_c(dtype, -gv[0]) * mx - _c(dtype, gv[1]) * my) - _c(dtype, gv[2]) * mz


def conservative_to_primitive(cons):
  """common/utils.py:145-183: phi = (1/rho) * q."""
  inv = 1.0 / cons['rho']
  prim = {'rho': cons['rho']}
  for k, q in cons.items():
    if k == 'rho':
      continue
    if k.startswith('rho_') and len(k) > 4:
      prim[prim_name(k)] = inv * q
  return prim


def primitive_to_conservative(prim):
  """common/utils.py:186-216: q = rho * phi."""
  out = {}
  for k, f in prim.items():
    out['rho' if k == 'rho' else cons_name(k)] = (
        f if k == 'rho' else prim['rho'] * f)
  return out


def internal_energy(s, pe):
  """generic.py:357-389."""
  return s['e'] - 0.5 * (s['u'] ** 2 + s['v'] ** 2 + s['w'] ** 2) - pe


def pressure(s, pe):
  """generic.py:27-57."""
  return (GAMMA - 1.0) * s['rho'] * internal_energy(s, pe)


def total_enthalpy(s, pe):
  """generic.py:59-88."""
  return s['e'] + pressure(s, pe) / s['rho']


def sound_speed_p_rho(s):
  """generic.py:149-154 (opt='p_rho')."""
  return _sqrt(GAMMA * s['p'] / s['rho'])


def sound_speed_h(s, pe):
  """generic.py:144-148 (opt='h') via sensible_enthalpy (:391-423)."""
  h_sens = s['h'] - 0.5 * (s['u'] ** 2 + s['v'] ** 2 + s['w'] ** 2) - pe
  return _sqrt((GAMMA - 1.0) * h_sens)


def temperature_eos(s):
  """generic.py:223-228."""
  return s['p'] / (s['rho'] * R)


# ---------------------------------------------------------------------------
# Boundary conditions: boundary/boundary.py + swirl_lm halo semantics as in
# the prototype (cell 2 l.24-163): dims applied in order x, y, z, each over
# the full plane (so edges and corners follow from the sequence).
# ---------------------------------------------------------------------------


def _bct(t):
  """BC type as its name (accepts the strings above or any enum member)."""
  return getattr(t, 'name', t)


def _layer(v, ax, idx):
  s = [slice(None)] * 3
  s[ax] = idx
  return tuple(s)


def exchange_halos(u, bc_dict, hw):
  """boundary/boundary.py:34-86 -> swirl_lm inplace_halo_exchange (1 replica).

  bc_dict: {dim: {0: (type, value), 1: (type, value)}} with type one of
  PERIODIC / DIRICHLET / NEUMANN.  A DIRICHLET value is a scalar or a list of
  hw layers (each a scalar or a plane), layer k going to halo index k on face
  0 and to index n-hw+k on face 1.  NEUMANN fills ghost = neighbour -/+ value
  cumulatively outward.
  """
  u = u.copy()
  for dim in DIMS:
    if dim not in bc_dict:
      continue
    ax = AXIS[dim]
    n = u.shape[ax]
    faces = bc_dict[dim]
    types_ = [_bct(faces[f][0]) for f in (0, 1)]
    if PERIODIC in types_:
      u[_sl(ax, 0, hw)] = u[_sl(ax, n - 2 * hw, n - hw)]
      u[_sl(ax, n - hw, n)] = u[_sl(ax, hw, 2 * hw)]
      continue
    for face in (0, 1):
      t, val = _bct(faces[face][0]), faces[face][1]
      if t == DIRICHLET:
        if isinstance(val, (list, tuple)):
          for i, layer in enumerate(val):
            idx = i if face == 0 else n - hw + i
            u[_layer(u, ax, idx)] = layer
        else:
          if face == 0:
            u[_sl(ax, 0, hw)] = val
          else:
            u[_sl(ax, n - hw, n)] = val
      elif t == NEUMANN:
        for i in range(hw):
          if face == 0:
            u[_layer(u, ax, hw - 1 - i)] = u[_layer(u, ax, hw - i)] - val
          else:
            u[_layer(u, ax, n - hw + i)] = u[_layer(u, ax, n - hw - 1 + i)] + val
      else:
        raise NotImplementedError(f'Unsupported BC type {t}.')
  return u


def update_conservative_cell_averages(cons, cfg):
  """boundary/boundary.py:464-539 (variables in dict order, rho first)."""
  cons = dict(cons)
  bc = cfg['bc']['cell_averages']
  hw = cfg['halo_width']
  for k in list(cons.keys()):
    if k in bc:
      cons[k] = exchange_halos(cons[k], bc[k], hw)
    elif prim_name(k) in bc:
      prim = _dnn(cons[k], cons['rho'])
      prim = exchange_halos(prim, bc[prim_name(k)], hw)
      cons[k] = cons['rho'] * prim
    else:
      raise ValueError(
          'A "cells" boundary condition is required for all conservative'
          f' variables. No boundary condition was found for {k} or'
          f' {prim_name(k)}.')
  return cons


def _bc_dict_along_dim(bc_dict, dim, side):
  """boundary/boundary.py:186-227."""
  out = {}
  for var, d in bc_dict.items():
    t, val = d[dim][side]
    t = _bct(t)
    if t is None:
      continue
    assert t == DIRICHLET, (
        f'Unsupported interface (face or flux) BC type {t}.')
    out[var] = val
  return out


def _plane_idx(cfg, dim, face):
  n = cfg['core_' + 'n' + dim]
  return cfg['halo_width'] if face == 0 else cfg['halo_width'] + n


def update_fluxes(fluxes, cfg, flux_type):
  """boundary/boundary.py:89-183: Dirichlet overwrite of boundary flux planes."""
  bc = cfg['bc']
  if 'intercell_fluxes' not in bc or flux_type not in bc['intercell_fluxes']:
    return fluxes
  spec = bc['intercell_fluxes'][flux_type]
  bad = set(spec) - set(cfg['conservative_variable_names'])
  assert not bad, f'{bad} cannot be specified by a flux boundary condition.'
  for dim in DIMS:
    ax = AXIS[dim]
    for face in (0, 1):
      for var, val in _bc_dict_along_dim(spec, dim, face).items():
        f = fluxes[dim][var].copy()
        f[_layer(f, ax, _plane_idx(cfg, dim, face))] = val
        fluxes[dim][var] = f
  return fluxes


def _face_boundary_update_1d(cons, prim, spec, dim, plane):
  """boundary/boundary.py:230-337."""
  if not spec:
    return cons, prim
  ax = AXIS[dim]
  sel = lambda f: _layer(f, ax, plane)
  if 'rho' in spec:
    r = cons['rho'].copy()
    r[sel(r)] = spec['rho']
    cons['rho'] = r
    # boundary.py:284-293 replaces the whole primitive dict: p and h are gone
    # afterwards (a later flux evaluation then fails on the missing 'p').
    prim = conservative_to_primitive(cons)
  for var, val in spec.items():
    if var == 'rho':
      continue
    if var in cons:
      f = cons[var].copy()
      f[sel(f)] = val
      cons[var] = f
      prim[prim_name(var)] = _dnn(cons[var], cons['rho'])
    else:
      if var not in prim:
        raise ValueError(f'"{var}" is not a valid primitive or conservative'
                         ' variable name for the face boundary condition.')
      f = prim[var].copy()
      f[sel(f)] = val
      prim[var] = f
      cons[cons_name(var)] = cons['rho'] * prim[var]
  return cons, prim


def update_faces(cons_neg, cons_pos, prim_neg, prim_pos, cfg):
  """boundary/boundary.py:340-461."""
  if 'cell_faces' not in cfg['bc']:
    return cons_neg, cons_pos, prim_neg, prim_pos
  for dim in DIMS:
    for face in (0, 1):
      spec = _bc_dict_along_dim(cfg['bc']['cell_faces'], dim, face)
      plane = _plane_idx(cfg, dim, face)
      cons_neg[dim], prim_neg[dim] = _face_boundary_update_1d(
          cons_neg[dim], prim_neg[dim], spec, dim, plane)
      cons_pos[dim], prim_pos[dim] = _face_boundary_update_1d(
          cons_pos[dim], prim_pos[dim], spec, dim, plane)
  return cons_neg, cons_pos, prim_neg, prim_pos


# ---------------------------------------------------------------------------
# Fluxes.
# ---------------------------------------------------------------------------


def physical_flux_1d(cons, prim, helper, dim, var):
  """equations/compressible_navier_stokes.py:338-427."""
  p_upd = helper.get('p_dyn', prim['p'])
  pv = prim_name(var)
  aligned = VELOCITY[DIMS.index(dim)]
  m = cons[cons_name(aligned)]
  if pv == aligned:
    return m * prim[aligned] + p_upd
  if pv == 'e':
    return m * (prim['e'] + prim['p'] / prim['rho'])
  if pv == 'rho':
    return m
  return m * prim[pv]


def roe_wave_speeds(prim_neg, prim_pos, dim, pe):
  """numerics/riemann_solver/hll.py:178-283."""
  aligned = VELOCITY[DIMS.index(dim)]
  c_neg = sound_speed_p_rho(prim_neg)
  c_pos = sound_speed_p_rho(prim_pos)
  sn = _sqrt(prim_neg['rho'])
  sp = _sqrt(prim_pos['rho'])
  inv = _rnn(sn + sp)
  roe = {v: (sn * prim_neg[v] + sp * prim_pos[v]) * inv
         for v in ('u', 'v', 'w', 'h')}
  c_roe = sound_speed_h(roe, pe)
  s_l = np.minimum(prim_neg[aligned] - c_neg, roe[aligned] - c_roe)
  s_r = np.maximum(prim_pos[aligned] + c_pos, roe[aligned] + c_roe)
  return s_l, s_r


def hll_flux(cons_neg, cons_pos, prim_neg, prim_pos, hv_neg, hv_pos, dim,
             var_names, pe):
  """numerics/riemann_solver/hll.py:36-175."""
  s_l, s_r = roe_wave_speeds(prim_neg, prim_pos, dim, pe)
  ds = s_r - s_l
  t1 = _dnn(np.minimum(s_r, 0.0) - np.minimum(s_l, 0.0), ds)
  t2 = 1.0 - t1
  t3 = _dnn(s_r * np.abs(s_l) - s_l * np.abs(s_r), 2.0 * ds)
  out = {}
  for var in var_names:
    fn = physical_flux_1d(cons_neg, prim_neg, hv_neg, dim, var)
    fp = physical_flux_1d(cons_pos, prim_pos, hv_pos, dim, var)
    out[var] = t1 * fp + t2 * fn - t3 * (cons_pos[var] - cons_neg[var])
  return out


def hllc_wave_speeds(prim_neg, prim_pos, dim, pe):
  """numerics/riemann_solver/hllc.py:260-347."""
  aligned = VELOCITY[DIMS.index(dim)]
  v_l, v_r = prim_neg[aligned], prim_pos[aligned]
  rho_l, rho_r = prim_neg['rho'], prim_pos['rho']
  p_l, p_r = prim_neg['p'], prim_pos['p']
  s_l, s_r = roe_wave_speeds(prim_neg, prim_pos, dim, pe)
  s_m = _dnn(p_r - p_l + rho_l * v_l * (s_l - v_l) - rho_r * v_r * (s_r - v_r),
             rho_l * (s_l - v_l) - rho_r * (s_r - v_r))
  s_m = np.where(np.abs(s_m) < HLLC_SM_SNAP, np.zeros_like(s_m), s_m)
  return s_l, s_m, s_r


def _hllc_star(prim_k, s_k, s_m, f_k, var, aligned):
  """numerics/riemann_solver/hllc.py:134-192 (Toro eq. 10.38, rearranged)."""
  pv = prim_name(var)
  rho_k, v_k = prim_k['rho'], prim_k[aligned]
  if pv == 'rho':
    return f_k + rho_k * (s_m - v_k) * _dnn(s_k, s_k - s_m)
  if pv == 'e':
    return f_k + (rho_k * _dnn(s_k, s_k - s_m) * (s_m - v_k)
                  * (prim_k['e'] + prim_k['p'] / rho_k + (s_k - v_k) * s_m))
  if pv == aligned:
    return f_k + rho_k * s_k * (_dnn(s_k - v_k, s_k - s_m) * s_m - v_k)
  return f_k + rho_k * (s_m - v_k) * prim_k[pv] * _dnn(s_k, s_k - s_m)


def hllc_branch(s_l, s_m, s_r):
  """Branch code of hllc.py:227-243: 0 F_L, 1 F*_L, 2 F*_R, 3 F_R."""
  return np.where(s_l >= 0, 0,
                  np.where((s_l <= 0) & (s_m >= 0), 1,
                           np.where((s_m <= 0) & (s_r >= 0), 2, 3))).astype(np.uint8)


def hllc_flux(cons_neg, cons_pos, prim_neg, prim_pos, hv_neg, hv_pos, dim,
              var_names, pe, return_branch=False):
  """numerics/riemann_solver/hllc.py:33-257."""
  def upd(prim, hv):
    p = dict(prim)
    if 'p_dyn' in hv:
      p['p'] = hv['p_dyn']
    return p
  pn_u, pp_u = upd(prim_neg, hv_neg), upd(prim_pos, hv_pos)
  s_l, s_m, s_r = hllc_wave_speeds(pn_u, pp_u, dim, pe)
  aligned = VELOCITY[DIMS.index(dim)]
  br = hllc_branch(s_l, s_m, s_r)
  out = {}
  for var in var_names:
    fl = physical_flux_1d(cons_neg, prim_neg, hv_neg, dim, var)
    fr = physical_flux_1d(cons_pos, prim_pos, hv_pos, dim, var)
    fml = _hllc_star(pn_u, s_l, s_m, fl, var, aligned)
    fmr = _hllc_star(pp_u, s_r, s_m, fr, var, aligned)
    out[var] = np.where(br == 0, fl, np.where(br == 1, fml,
                                              np.where(br == 2, fmr, fr)))
  if return_branch:
    return out, br
  return out


def numeric_flux_fn(scheme):
  """numerics/riemann_solver/riemann_solver.py:30-62."""
  if scheme == 'HLL':
    return hll_flux
  if scheme == 'HLLC':
    return hllc_flux
  raise ValueError(f'"{scheme}" is not implemented as a "numeric_flux_scheme".'
                   " Valid options are: ['HLL', 'HLLC']")


def strain_rate(prim, cfg):
  """physics/fluid.py:31-111."""
  hs = {'x': cfg['dx'], 'y': cfg['dy'], 'z': cfg['dz']}

  def grad(f, gd, fd):
    if gd == fd:
      return backward_1(f, hs[gd], gd)
    return linear_interpolation(central_2(f, hs[gd], gd), fd)

  u, v, w = prim['u'], prim['v'], prim['w']
  diag = lambda g, d: g - 1.0 / 3.0 * d
  off = lambda g0, g1: 0.5 * (g0 + g1)
  div = {fd: grad(u, 'x', fd) + grad(v, 'y', fd) + grad(w, 'z', fd)
         for fd in DIMS}
  return [
      [diag(grad(u, 'x', 'x'), div['x']),
       off(grad(u, 'y', 'y'), grad(v, 'x', 'y')),
       off(grad(u, 'z', 'z'), grad(w, 'x', 'z'))],
      [off(grad(v, 'x', 'x'), grad(u, 'y', 'x')),
       diag(grad(v, 'y', 'y'), div['y']),
       off(grad(v, 'z', 'z'), grad(w, 'y', 'z'))],
      [off(grad(w, 'x', 'x'), grad(u, 'z', 'x')),
       off(grad(w, 'y', 'y'), grad(v, 'z', 'y')),
       diag(grad(w, 'z', 'z'), div['z'])],
  ]


def shear_stress(prim, prim_f, cfg):
  """physics/diffusion.py:33-67 with TransportSimple (simple.py:44-79)."""
  s = strain_rate(prim, cfg)
  nu = cfg['transport_parameters']['nu'] * np.ones_like(prim['rho'])
  rho = [prim_f[d]['rho'] for d in DIMS]
  return [[2.0 * rho[j] * nu * s[i][j] for j in range(3)] for i in range(3)]


def heat_flux(prim, cfg):
  """physics/diffusion.py:76-109 with simple.py:81-108."""
  t = temperature_eos(prim)
  tp = cfg['transport_parameters']
  mu = tp['nu'] * prim['rho']
  cp = CP * np.ones_like(prim['rho'])
  kappa = (1.0 / tp['pr']) * cp * mu
  hs = {'x': cfg['dx'], 'y': cfg['dy'], 'z': cfg['dz']}
  return {d: -linear_interpolation(kappa, d) * backward_1(t, hs[d], d)
          for d in DIMS}


def diffusive_flux(prim, cfg):
  """equations/compressible_navier_stokes.py:618-707."""
  prim_f = {d: {v: linear_interpolation(prim[v], d)
                for v in list(VELOCITY) + ['rho']} for d in DIMS}
  tau = shear_stress(prim, prim_f, cfg)
  q = heat_flux(prim, cfg)
  fl = {d: {v: np.zeros_like(prim['rho'])
            for v in cfg['conservative_variable_names']} for d in DIMS}
  for j, d in enumerate(DIMS):
    for i in range(3):
      fl[d][MOMENTUM[i]] = -tau[i][j]
  for j, d in enumerate(DIMS):
    vel = [prim_f[d][v] for v in VELOCITY]
    fl[d]['rho_e'] = -(vel[0] * tau[0][j] + vel[1] * tau[1][j]
                       + vel[2] * tau[2][j])
  for d in DIMS:
    fl[d]['rho_e'] = q[d] + fl[d]['rho_e']
  return fl


def source_function(cons, helper, cfg):
  """physics/physics_models.py:94-143 and additional_models/gravity.py:64-102."""
  src = {k: np.zeros_like(v) for k, v in cons.items()}
  for name in cfg.get('source_functions') or []:
    if name != 'gravity':
      raise NotImplementedError(f'Unsupported source term model: "{name}".')
    gd = gravity_direction(cfg['g'])
    if gd == -1:
      raise ValueError('Gravity direction must be specified if the gravity'
                       ' source function is to be used.')
    d_rho = helper.get('d_rho', cons['rho'])
    mask = MOMENTUM if gd == 3 else (MOMENTUM[gd],)
    for dim in DIMS:
      var = MOMENTUM[DIMS.index(dim)]
      if var in mask:
        src[var] = src[var] + d_rho * G * cfg['g'][dim]
  return src


def update_helper_variables(prim, helper, cfg):
  """equations/compressible_navier_stokes.py:273-335."""
  helper = dict(helper)
  names = []
  if 'p_ref' in helper and 'rho_ref' in helper:
    zero_bc = {d: {0: (DIRICHLET, 0.0), 1: (DIRICHLET, 0.0)} for d in DIMS}
    hw = cfg['halo_width']
    helper['p_dyn'] = exchange_halos(prim['p'] - helper['p_ref'], zero_bc, hw)
    helper['d_rho'] = exchange_halos(prim['rho'] - helper['rho_ref'], zero_bc, hw)
    names.append('p_dyn')
  return helper, names


def face_interpolation(cell_avg, cfg, names=None):
  """equations/compressible_navier_stokes.py:536-615."""
  fn = face_interp_fn(cfg['interpolation_scheme'])
  interp = {v: {d: fn(f, d) for d in DIMS} for v, f in cell_avg.items()
            if names is None or v in names}
  neg = {d: {v: val[d][0] for v, val in interp.items()} for d in DIMS}
  pos = {d: {v: val[d][1] for v, val in interp.items()} for d in DIMS}
  return neg, pos


def rhs(cons, helper, cfg, return_branch=False):
  """equations/compressible_navier_stokes.py:69-268 (the hot path)."""
  dtype = cons['rho'].dtype
  cons = update_conservative_cell_averages(cons, cfg)
  pe = gravitational_potential(cfg, dtype)
  prim = conservative_to_primitive(cons)
  prim['p'] = pressure(prim, pe)
  prim['h'] = total_enthalpy(prim, pe)
  helper, hv_names = update_helper_variables(prim, helper, cfg)
  prim_neg, prim_pos = face_interpolation(prim, cfg)
  hv_neg, hv_pos = face_interpolation(helper, cfg, hv_names)
  pvn = cfg['primitive_variable_names']
  cons_neg = {d: primitive_to_conservative({v: prim_neg[d][v] for v in pvn})
              for d in DIMS}
  cons_pos = {d: primitive_to_conservative({v: prim_pos[d][v] for v in pvn})
              for d in DIMS}
  cons_neg, cons_pos, prim_neg, prim_pos = update_faces(
      cons_neg, cons_pos, prim_neg, prim_pos, cfg)
  flux_fn = numeric_flux_fn(cfg['numeric_flux_scheme'])
  names = cfg['conservative_variable_names']
  flux_conv, branches = {}, {}
  for d in DIMS:
    if flux_fn is hllc_flux:
      flux_conv[d], branches[d] = hllc_flux(
          cons_neg[d], cons_pos[d], prim_neg[d], prim_pos[d], hv_neg[d],
          hv_pos[d], d, names, pe, return_branch=True)
    else:
      flux_conv[d] = flux_fn(cons_neg[d], cons_pos[d], prim_neg[d],
                             prim_pos[d], hv_neg[d], hv_pos[d], d, names, pe)
  flux_conv = update_fluxes(flux_conv, cfg, 'convective')
  if cfg['include_diffusion']:
    if 'cell_faces' in cfg['bc']:
      raise NotImplementedError('Face boundary conditions are not supported'
                                ' with diffusion enabled at this time.')
    flux_diff = update_fluxes(diffusive_flux(prim, cfg), cfg, 'diffusive')
  else:
    flux_diff = {d: {v: np.zeros_like(f) for v, f in flux_conv[d].items()}
                 for d in DIMS}
  flux = {d: {v: f + flux_diff[d][v] for v, f in flux_conv[d].items()}
          for d in DIMS}
  flux = update_fluxes(flux, cfg, 'total')
  src = source_function(cons, helper, cfg)
  delta = {'x': cfg['dx'], 'y': cfg['dy'], 'z': cfg['dz']}
  for v in names:
    for d in DIMS:
      src[v] = src[v] - forward_1(flux[d][v], delta[d], d)
  if return_branch:
    return src, branches
  return src


# ---------------------------------------------------------------------------
# Time integration: numerics/time_integration/{rk3,rk2,time_integration}.py.
# ---------------------------------------------------------------------------

RK3_C11 = 0.75
RK3_C21 = 1.0 / 3.0
RK3_C12 = 1.0 - RK3_C11
RK3_C22 = 1.0 - RK3_C21


def rk3(state_0, helper, cfg, dt=None, rhs_fn=None):
  """numerics/time_integration/rk3.py:34-122."""
  rhs_fn = rhs_fn or (lambda s: rhs(s, helper, cfg))
  dt = cfg['dt'] if dt is None else dt
  d1 = rhs_fn(state_0)
  s1 = {k: v + d1[k] * dt for k, v in state_0.items()}
  d2 = rhs_fn(s1)
  s2 = {k: v * RK3_C11 + (s1[k] + d2[k] * dt) * RK3_C12
        for k, v in state_0.items()}
  d3 = rhs_fn(s2)
  return {k: v * RK3_C21 + (s2[k] + d3[k] * dt) * RK3_C22
          for k, v in state_0.items()}


def rk2(state_0, helper, cfg, dt=None, rhs_fn=None):
  """numerics/time_integration/rk2.py:30-86."""
  rhs_fn = rhs_fn or (lambda s: rhs(s, helper, cfg))
  dt = cfg['dt'] if dt is None else dt
  d1 = rhs_fn(state_0)
  s1 = {k: v + dt * d1[k] for k, v in state_0.items()}
  d2 = rhs_fn(s1)
  s2 = {k: v + dt * d2[k] for k, v in s1.items()}
  return {k: 0.5 * (v + s2[k]) for k, v in state_0.items()}


def time_integration(state_0, helper, cfg, dt=None):
  """numerics/time_integration/time_integration.py:35-90."""
  scheme = cfg['time_integration_scheme']
  if scheme == 'rk3':
    return rk3(state_0, helper, cfg, dt)
  if scheme == 'rk2':
    return rk2(state_0, helper, cfg, dt)
  raise NotImplementedError(f'"{scheme}" is not implemented as an'
                            ' "time_integration_scheme".')


def max_wave_speed(cons, cfg):
  """CFL extension (no reference counterpart): max over interior cells of
  max_d(|u_d| + c), c = sqrt(gamma p / rho), with the reference's EOS."""
  hw = cfg['halo_width']
  dtype = cons['rho'].dtype
  pe = gravitational_potential(cfg, dtype)
  prim = conservative_to_primitive(cons)
  prim['p'] = pressure(prim, pe)
  c = sound_speed_p_rho(prim)
  sl = (slice(hw, -hw),) * 3
  m = np.maximum(np.maximum(np.abs(prim['u']) + c, np.abs(prim['v']) + c),
                 np.abs(prim['w']) + c)
  return m[sl].max()


# ---------------------------------------------------------------------------
# Config helper mirroring core/parameter.py:_DEFAULT_OPTIONS (the subset the
# hot path reads).
# ---------------------------------------------------------------------------


def cell_centres(lo, hi, n):
  """core/parameter.py:27-32."""
  xx = np.linspace(lo, hi, n + 1, dtype=np.float32)
  return xx[:-1] + 0.5 * np.diff(xx)[0]


def default_cfg(**overrides):
  xx = cell_centres(0.0, 1.0, 10)
  dx = float(np.diff(xx)[0])
  cfg = {
      'halo_width': 3, 'nx': 10, 'ny': 10, 'nz': 10,
      'core_nx': 10, 'core_ny': 10, 'core_nz': 10,
      'x': xx, 'dx': dx, 'y': xx, 'dy': dx, 'z': xx, 'dz': dx,
      'g': {'x': 0.0, 'y': 0.0, 'z': 0.0}, 'dt': 1.0e-3,
      'interpolation_scheme': 'WENO_5', 'numeric_flux_scheme': 'HLL',
      'include_diffusion': False, 'time_integration_scheme': 'rk3',
      'source_functions': None,
      'transport_parameters': {'nu': 1.0e-4, 'pr': 0.7},
      'conservative_variable_names': list(BASE_CONSERVATIVE),
      'primitive_variable_names': list(BASE_PRIMITIVES),
      'bc': {'cell_averages': {
          v: {d: {0: (PERIODIC,), 1: (PERIODIC,)} for d in DIMS}
          for v in BASE_CONSERVATIVE}},
  }
  cfg.update(overrides)
  return cfg
