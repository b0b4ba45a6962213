"""A device-resident time-stepping loop over the native RK stages.

`Solver` owns the state and two stage buffers, computes dt on the device
(fixed cfg.dt, or a CFL time step from the max wave speed), and can capture
one whole RK step in a HIP graph that is replayed per step.  It replaces the
swirl_lm driver's step loop (core/driver.py:202-226, core/simulation.py:68-105)
for one GPU.
"""

import torch

from swirlc_amd import _native as N
from swirlc_amd.equations.compressible_navier_stokes import get_rhs_fn

_C11, _C21 = 0.75, 1.0 / 3.0


class Solver:
  """Advances a conservative state dict in place on one GPU."""

  def __init__(self, cfg, state, cfl=None, use_graph=True, rhs_fn=None):
    self.cfg = cfg
    self.rhs_fn = rhs_fn if rhs_fn is not None else get_rhs_fn(None, cfg)
    self.ctx = self.rhs_fn.native_context
    fields = self.ctx.fields(state)
    self.u = [t.clone() for t in fields]
    self.s1 = [torch.zeros_like(t) for t in fields]
    self.s2 = [torch.zeros_like(t) for t in fields]
    dev = self.u[0].device
    self.cfl = cfl if cfl is not None else getattr(cfg, 'cfl', None)
    self.dt = torch.full((1,), float(cfg.dt), dtype=torch.float32, device=dev)
    self.speed = torch.zeros(4, dtype=torch.float32, device=dev)
    self.scheme = cfg.time_integration_scheme
    if self.scheme not in ('rk3', 'rk2'):
      raise NotImplementedError(
          f'"{self.scheme}" is not implemented as an "time_integration_scheme".'
          " Valid options are: ['rk2', 'rk3']")
    self.use_graph = use_graph
    self._graph = None
    self.steps_taken = 0
    # RK3 stages chained through sc_rk_stage_ex: each stage leaves its output
    # with the halo cells filled (on all-periodic grids its last sweep writes
    # them), so the next stage fills none, and with a CFL time step the last
    # stage also accumulates the wave speed of the state it writes, so the
    # next step's dt needs no separate pass over the state.  The first step
    # seeds both from the state; call refresh() after changing the state in
    # place between steps.
    self._chained = self.scheme == 'rk3' and hasattr(self.ctx, 'rk_stage_ex')
    self._seeded = False

  @property
  def state(self):
    return dict(zip(self.ctx.names, self.u))

  def refresh(self):
    """Fills the state's halos and seeds the CFL maximum from the state (the
    chained stepping keeps both current from step to step)."""
    if self._chained:
      self.ctx.fill_halos(self.u)
      if self.cfl is not None:
        self.ctx.max_wave_speed(self.u, self.speed, float(self.cfl), self.dt)
      self._seeded = True

  def _step_body(self):
    c = self.ctx
    if self._chained:
      flags = N.SC_STAGE_HALOS_READY | N.SC_STAGE_FILL_OUT
      speed = None
      if self.cfl is not None:
        # dt from the maximum the previous step's last stage left, which is
        # then cleared for this step's last stage to accumulate into.
        c.cfl_dt(self.speed, float(self.cfl), self.dt)
        speed = self.speed
      c.rk_stage_ex(N.SC_RK_EULER, 0.0, 0.0, self.dt, None, self.u, self.s1, flags)
      c.rk_stage_ex(N.SC_RK_BLEND, _C11, 1.0 - _C11, self.dt, self.u, self.s1, self.s2, flags)
      c.rk_stage_ex(N.SC_RK_BLEND, _C21, 1.0 - _C21, self.dt, self.u, self.s2, self.u, flags,
                    speed)
      return
    if self.cfl is not None:
      c.max_wave_speed(self.u, self.speed, float(self.cfl), self.dt)
    if self.scheme == 'rk3':
      c.rk_stage(N.SC_RK_EULER, 0.0, 0.0, self.dt, None, self.u, self.s1)
      c.rk_stage(N.SC_RK_BLEND, _C11, 1.0 - _C11, self.dt, self.u, self.s1, self.s2)
      c.rk_stage(N.SC_RK_BLEND, _C21, 1.0 - _C21, self.dt, self.u, self.s2, self.u)
    else:
      # u1 = u0 + dt L(u0); u2 = u1 + dt L(u1); u = 0.5 (u0 + u2).
      c.rk_stage(N.SC_RK_EULER, 0.0, 0.0, self.dt, None, self.u, self.s1)
      c.rk_stage(N.SC_RK_EULER, 0.0, 0.0, self.dt, None, self.s1, self.s2)
      for a, b in zip(self.u, self.s2):
        a.add_(b).mul_(0.5)

  def capture(self):
    """Records one RK step into a HIP graph (replayed by step())."""
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
