"""Batched ensembles: many independent grids of different sizes, one launch per phase.

`BatchSolver` advances N simulations that share one configuration (schemes,
BC types and values, transport, gravity, transported scalars) but have their
own grid sizes, spacings, initial states and CFL time steps.  Each field of all instances lives in one packed arena;
every phase of an RK stage (x / y / z sweep, halo fill per dim where one is
needed, wave speed) is one launch over a tile list covering all instances,
and the whole RK3 step is captured once in a HIP graph and replayed.  The
stages are chained: each leaves its output's halos filled (written by the z
sweep as images where the fill is a pure copy), so after writing to the
state between steps call refresh().  No reference counterpart:
the reference runs one simulation per TPU job.
"""

import ctypes

import numpy as np
import torch

from swirlc_amd import _native as N
from swirlc_amd.common import utils
from swirlc_amd.core.parameter import SwirlCParameters, _get_cell_center_coordinates
from swirlc_amd.equations.native_context import NativeContext, copy_fill_dims

_C11, _C21 = 0.75, 1.0 / 3.0
_ALIGN = 64  # elements between instances (256 B alignment of every field)


class BatchSolver:
  """Advances an ensemble of conservative states in place on one GPU.

  Args:
    cfg: configuration shared by every instance (a SwirlCParameters or dict);
      its grid keys are ignored.
    instances: list of (grid, state) where grid = (nx, ny, nz, dx, dy, dz) or
      (nx, ny, nz, dx, dy, dz, x, y, z) and state maps the conserved names to
      arrays shaped [nz+2h, nx+2h, ny+2h].  x, y, z are the instance's
      cell-centre coordinates (used by gravity); by default the centres of
      [0, n * d] along each axis, as a single-grid configuration of lx = nx *
      dx would have them.
    cfl: Courant number of the per-instance time step (device computed), or
      None to use cfg.dt for every instance.
  """

  def __init__(self, cfg, instances, cfl=0.5, use_graph=True, device='cuda'):
    if isinstance(cfg, dict):
      cfg = SwirlCParameters(cfg)
    if cfg.time_integration_scheme != 'rk3':
      raise NotImplementedError('batched ensembles integrate with rk3')
    self.cfg = cfg
    nx0, ny0, nz0 = instances[0][0][:3]
    probe_cfg = SwirlCParameters({**cfg._simulation_cfg, 'core_nx': nx0, 'core_ny': ny0,
                                  'core_nz': nz0})
    probe = NativeContext(probe_cfg, device=device, create=False)
    if probe._pending_planes:
      raise NotImplementedError('batched ensembles do not support per-layer DIRICHLET planes'
                                ' or plane-valued face / flux values')
    self.names = probe.names
    self.nf = probe.nf
    h = self.h = probe.hw
    self.n = len(instances)
    descs = (N.ScInstance * self.n)()
    self.shapes, self.offsets = [], []
    self._coords = []  # host coordinate arrays referenced by descs
    off = 0
    for k, (grid, _) in enumerate(instances):
      nx, ny, nz, dx, dy, dz = grid[:6]
      d = descs[k]
      d.nx, d.ny, d.nz = nx, ny, nz
      d.dx, d.dy, d.dz = float(dx), float(dy), float(dz)
      d.offset = off
      if probe.gravity_dir != -1:
        if len(grid) >= 9:
          axes = grid[6:9]
          for c, m, dim in zip(axes, (nx, ny, nz), 'xyz'):
            if np.asarray(c).reshape(-1).size != m:
              raise ValueError(f'instance {k}: {dim} coordinates have '
                               f'{np.asarray(c).size} entries, expected {m}')
        else:
          axes = [_get_cell_center_coordinates(0.0, sp * m, m)
                  for m, sp in ((nx, dx), (ny, dy), (nz, dz))]
        for attr, c in zip(('x_full', 'y_full', 'z_full'), axes):
          arr = np.ascontiguousarray(utils.mesh_full(c, h), dtype=np.float32)
          self._coords.append(arr)
          setattr(d, attr, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
      shape = (nz + 2 * h, nx + 2 * h, ny + 2 * h)
      self.shapes.append(shape)
      self.offsets.append(off)
      cells = shape[0] * shape[1] * shape[2]
      off += (cells + _ALIGN - 1) // _ALIGN * _ALIGN
    self.total = off
    dev = torch.device(device)
    self.u = [torch.zeros(self.total, dtype=torch.float32, device=dev) for _ in range(self.nf)]
    self.s1 = [torch.zeros_like(t) for t in self.u]
    self.s2 = [torch.zeros_like(t) for t in self.u]
    for k, (_, state) in enumerate(instances):
      for f, name in enumerate(self.names):
        v = state[name]
        v = v if torch.is_tensor(v) else torch.from_numpy(v)
        self.view(self.u[f], k).copy_(v.to(dev, torch.float32))
    self.cfl = cfl
    self.dt = torch.full((self.n,), float(cfg.dt), dtype=torch.float32, device=dev)
    self.speed = torch.zeros(self.n, dtype=torch.float32, device=dev)
    handle = ctypes.c_void_p()
    with torch.cuda.device(dev):
      N.check(N.lib().sc_batch_create(ctypes.byref(probe._params), h, self.n, descs,
                                      ctypes.byref(handle)), 'sc_batch_create')
    self._b = handle
    self._params = probe._params
    self.use_graph = use_graph
    self._graph = None
    self.steps_taken = 0
    # Chained stages: each stage leaves its output's halos filled (written by
    # the z sweep as images where the fill is a pure copy), so the next stage,
    # and the next step, start without a fill.  The first step fills once.
    self._chained = True
    self._seeded = False

  def view(self, arena, k):
    """Instance k's [nz+2h, nx+2h, ny+2h] view of one field arena."""
    shape = self.shapes[k]
    n = shape[0] * shape[1] * shape[2]
    return arena[self.offsets[k]:self.offsets[k] + n].view(shape)

  def state(self, k):
    return {name: self.view(self.u[f], k) for f, name in enumerate(self.names)}

  def close(self):
    if getattr(self, '_b', None) is not None and self._b.value:
      N.lib().sc_batch_destroy(self._b)
      self._b = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # interpreter shutdown
      pass

  def get_tuning(self):
    """Launch geometry of the batched sweeps (dict of the sc_tuning fields)."""
    t = N.ScTuning()
    N.check(N.lib().sc_batch_get_tuning(self._b, ctypes.byref(t)), 'sc_batch_get_tuning')
    return t.as_dict()

  def set_tuning(self, tuning):
    """Sets the given fields of the launch geometry (others keep their values;
    rebuilds the tile lists; drops a captured step)."""
    t = N.ScTuning.from_dict(dict(self.get_tuning(), **tuning))
    N.check(N.lib().sc_batch_set_tuning(self._b, ctypes.byref(t)), 'sc_batch_set_tuning')
    self._graph = None

  def out_images_fused(self):
    """Whether a chained stage writes its output's halos in the z sweep (the
    halo fill a pure copy along every dim for every instance's extents)."""
    return all(copy_fill_dims(self._params, self.nf, (nx, ny, nz), self.h) is not None
               for nz, nx, ny in ((s[0] - 2 * self.h, s[1] - 2 * self.h, s[2] - 2 * self.h)
                                  for s in self.shapes))

  def refresh(self):
    """Fills the halos of every instance's state and seeds the per-instance
    CFL maxima from it (the chained stepping keeps both current from step to
    step; call this after writing to the state)."""
    N.check(N.lib().sc_batch_fill_halos(self._b, N.ptr_array(self.u), N.stream_handle()),
            'sc_batch_fill_halos')
    if self.cfl is not None:
      self._max_wave_speed()
    self._seeded = True

  def _max_wave_speed(self):
    N.check(N.lib().sc_batch_max_wave_speed(
        self._b, N.ptr_array(self.u), ctypes.c_float(self.cfl),
        ctypes.c_void_p(self.speed.data_ptr()), ctypes.c_void_p(self.dt.data_ptr()),
        N.stream_handle()), 'sc_batch_max_wave_speed')

  def _stage(self, form, c0, c1, u0, uk, out, speed=False):
    u0p = N.ptr_array(u0) if u0 is not None else None
    dt = ctypes.c_void_p(self.dt.data_ptr())
    if self._chained:
      flags = N.SC_STAGE_HALOS_READY | N.SC_STAGE_FILL_OUT
      sp = ctypes.c_void_p(self.speed.data_ptr()) if speed else None
      N.check(N.lib().sc_batch_rk_stage_ex(
          self._b, form, ctypes.c_float(c0), ctypes.c_float(c1), dt, u0p, N.ptr_array(uk),
          N.ptr_array(out), flags, sp, N.stream_handle()), 'sc_batch_rk_stage_ex')
    else:
      N.check(N.lib().sc_batch_rk_stage(
          self._b, form, ctypes.c_float(c0), ctypes.c_float(c1), dt, u0p, N.ptr_array(uk),
          N.ptr_array(out), N.stream_handle()), 'sc_batch_rk_stage')

  def _step_body(self):
    cfl = self.cfl is not None
    if cfl:
      if self._chained:
        # dt from the maxima the previous step's last stage accumulated (or
        # refresh seeded), which are then cleared for this step's.
        N.check(N.lib().sc_batch_cfl_dt(
            self._b, ctypes.c_void_p(self.speed.data_ptr()), ctypes.c_float(self.cfl),
            ctypes.c_void_p(self.dt.data_ptr()), N.stream_handle()), 'sc_batch_cfl_dt')
      else:
        self._max_wave_speed()
    self._stage(N.SC_RK_EULER, 0.0, 0.0, None, self.u, self.s1)
    self._stage(N.SC_RK_BLEND, _C11, 1.0 - _C11, self.u, self.s1, self.s2)
    self._stage(N.SC_RK_BLEND, _C21, 1.0 - _C21, self.u, self.s2, self.u, speed=cfl)

  def capture(self):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
      with torch.cuda.graph(g, stream=side):
        self._step_body()
    torch.cuda.current_stream().wait_stream(side)
    self._graph = g

  def step(self, n=1):
    if self._chained and not self._seeded:
      self.refresh()
    for _ in range(n):
      if self.use_graph:
        if self._graph is None:
          self.capture()
        self._graph.replay()
      else:
        self._step_body()
      self.steps_taken += 1
