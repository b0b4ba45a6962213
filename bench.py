"""Benchmark: SSP-RK3 steps of the 3-D Kelvin-Helmholtz workload on one MI355X.

Workload: periodic 256^3 cube, fp32, gamma = 1.4, WENO_5 + HLLC, viscous
(nu = 1e-4, Pr = 0.7), SSP-RK3 with a device-computed CFL-0.5 time step;
synthetic initial state (shear layer rho = 2 / 1, u = +-0.5, p = 2.5,
v = 0.01 sin(4 pi x) + U(-1e-3, 1e-3) noise from seed 0).

Prints one JSON line: Mcell-updates/s per RK step (interior cells x steps /
seconds / 1e6) over graph-replayed steps, with the roofline of the dominant
kernel from its in-step duration (the step's launch sequence issued with HIP
events between the launches, on the stream they run on) and the CPU oracle's
throughput on one RK3 step of the same workload on the host's cores.

  python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu]
  python bench.py --profile-steps K   (only the graph-replayed steps, for a
                                       rocprofv3 kernel trace of them)
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, 'swirl-c_amd')):
  if _p not in sys.path:
    sys.path.insert(0, _p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# Algorithmic bytes per interior cell per launch of each sweep kernel (fp32,
# 5 conserved fields), in launch order: x reads U, writes R (40); y reads U
# and R, writes R (60); z (last) reads U and R, in RK3 stages 2 and 3 also
# U0, and writes the new state (60 at stage 1, 80 at stages 2-3; the stage
# state at the output cell is the U it reads).
# A step: 120 + 180 + 220 = 520 B per cell.
SWEEP_BYTES = {'x': (40, 40, 40), 'y': (60, 60, 60), 'z': (60, 80, 80)}
STEP_BYTES = 520
# The same step with a separate viscous-flux kernel (read U, read and write
# R: 60 B per stage) would move 700 B per cell.
STEP_BYTES_VISCOUS = 700
# Measured HBM bytes per sweep launch at 256^3 (rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes over `bench.py --pmc-probe`, scaled by a copy kernel's
# counts): the file's `traffic_bytes_per_launch <kernel> remap<0|1> <bytes>`
# lines (XCD-aware block order off / on).
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'sweeps_kh256_hbm_traffic.txt')


def measured_traffic(kernel, remap):
  """HBM bytes per launch of `kernel` in block order `remap` from the
  committed counter summary, or None if it holds no figure for it."""
  import re
  try:
    text = open(TRAFFIC_FILE).read()
  except OSError:
    return None
  m = re.search(r'^traffic_bytes_per_launch\s+' + re.escape(kernel) + r'\s+remap' +
                str(int(remap)) + r'\s+(\d+)\s*$', text, re.M)
  return int(m.group(1)) if m else None


def copy_bandwidth(nbytes=1 << 30, reps=10):
  """Device-to-device copy rate (B/s: bytes read + bytes written per second)
  of a plain copy of `nbytes`, HIP events on the current stream."""
  import torch
  src = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda').fill_(1.0)
  dst = torch.empty_like(src)
  dst.copy_(src)
  stream = torch.cuda.current_stream()
  ev0 = torch.cuda.Event(enable_timing=True)
  ev1 = torch.cuda.Event(enable_timing=True)
  ev0.record(stream)
  for _ in range(reps):
    dst.copy_(src)
  ev1.record(stream)
  ev1.synchronize()
  sec = ev0.elapsed_time(ev1) / 1e3 / reps
  del src, dst
  torch.cuda.empty_cache()
  return 2 * nbytes / sec


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--profile-steps', type=int, default=0,
                  help='run only the tuned, graph-replayed KH steps (for rocprofv3)')
  ap.add_argument('--cpu-workers', type=int, default=0,
                  help='processes of the CPU baseline (default: the affinity cores, at most 16)')
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--n', type=int, default=256)
  ap.add_argument('--kernel-reps', type=int, default=10)
  ap.add_argument('--kernel-steps', type=int, default=10,
                  help='steps of the in-step kernel timing (HIP events between launches)')
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-n', type=int, default=256)
  ap.add_argument('--cpu-single-n', type=int, default=128)
  ap.add_argument('--batch-instances', type=int, default=256)
  ap.add_argument('--batch-steps', type=int, default=10)
  ap.add_argument('--batch-loop-steps', type=int, default=3)
  ap.add_argument('--no-batch', action='store_true')
  ap.add_argument('--adjoint-n', type=int, default=128)
  ap.add_argument('--adjoint-steps', type=int, default=200)
  ap.add_argument('--adjoint-every', type=int, default=20)
  ap.add_argument('--no-adjoint', action='store_true')
  ap.add_argument('--adjoint-only', action='store_true',
                  help='run only the adjoint leg (for a rocprofv3 kernel trace of it)')
  ap.add_argument('--no-large', action='store_true',
                  help='skip the 512 x 512 x 256 buoyant-bubble leg')
  ap.add_argument('--large-steps', type=int, default=10)
  ap.add_argument('--no-tune', action='store_true',
                  help='keep the default launch geometry (no autotuning)')
  ap.add_argument('--tuning', type=str, default='',
                  help='launch geometry as JSON (sc_tuning fields); skips autotuning')
  ap.add_argument('--pmc-probe', action='store_true',
                  help='launch a 256 MiB copy kernel, then each sweep kernel twice on the '
                       'state, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes')
  return ap.parse_args()


def time_step_kernels(solver, steps):
  """Kernel durations inside RK3 steps: the launch sequence of the solver's
  chained step (the CFL dt from the maximum wave speed the previous step's
  last stage accumulated, then per stage the x, y, z sweeps, the last one with
  the stage update, the output's periodic halo images and in stage 3 the
  wave-speed accumulation; on grids that are not all-periodic a halo fill of
  the stage output follows) issued through the C ABI on one stream.  Each
  sweep launch records its own start and stop times into HIP events
  (sc_sweep_only_timed), so nothing is inserted between the kernels; the dt
  kernel and any halo fill are bracketed by events recorded on the stream.
  Returns {kernel: [ms per launch at stage 1, 2, 3]} (the dt kernel as a
  one-element list) averaged over `steps`, and the mean step time (ms)."""
  import torch
  from swirlc_amd import _native as N
  ctx = solver.ctx
  stream = torch.cuda.current_stream()
  u, s1, s2 = solver.u, solver.s1, solver.s2
  stages = ((N.SC_RK_EULER, 0.0, 0.0, None, u, s1), (N.SC_RK_BLEND, 0.75, 0.25, u, s1, s2),
            (N.SC_RK_BLEND, 1.0 / 3.0, 2.0 / 3.0, u, s2, u))
  names = ('sweep_x', 'sweep_y', 'sweep_z')
  acc = {n: [0.0, 0.0, 0.0] for n in names}
  acc['cfl_dt'] = [0.0]
  fused_halo = ctx.out_images_fused()
  if not fused_halo:
    acc['halo'] = [0.0, 0.0, 0.0]
  total = 0.0

  def event():
    e = torch.cuda.Event(enable_timing=True)
    e.record(stream)  # creates the HIP event
    return e

  solver.refresh()
  for _ in range(steps):
    t_begin, t_speed = event(), event()
    halo = [(event(), event()) for _ in range(3)]
    sweeps = [[(event(), event()) for _ in range(3)] for _ in range(3)]
    t_end = event()
    torch.cuda.synchronize()
    t_begin.record(stream)
    ctx.cfl_dt(solver.speed, float(solver.cfl), solver.dt, stream)
    t_speed.record(stream)
    for st, (form, c0, c1, base, uk, out) in enumerate(stages):
      for j, d in enumerate((0, 1, 2)):
        e0, e1 = sweeps[st][j]
        ctx.sweep_only_timed(d, form, c0, c1, solver.dt, base, uk, out, e0, e1, stream,
                             flags=N.SC_STAGE_FILL_OUT,
                             speed_dev=solver.speed if (st == 2 and d == 2) else None)
      if not fused_halo:
        halo[st][0].record(stream)
        ctx.fill_halos(out, stream)
        halo[st][1].record(stream)
    t_end.record(stream)
    t_end.synchronize()
    acc['cfl_dt'][0] += t_begin.elapsed_time(t_speed)
    for st in range(3):
      if not fused_halo:
        acc['halo'][st] += halo[st][0].elapsed_time(halo[st][1])
      for j, n in enumerate(names):
        e0, e1 = sweeps[st][j]
        acc[n][st] += e0.elapsed_time(e1)
    total += t_begin.elapsed_time(t_end)
  return {n: [v / steps for v in vals] for n, vals in acc.items()}, total / steps


def time_sweeps(solver, reps):
  """Average duration (s) of each sweep kernel launched alone, repeatedly,
  HIP events on its stream (a side note beside the in-step figures)."""
  import torch
  from swirlc_amd import _native as N
  ctx = solver.ctx
  stream = torch.cuda.current_stream()
  ctx.fill_halos(solver.u)
  out = {}
  # RK3 stage-2 form with U0 (the last step's stage-1 state) distinct from Uk,
  # as in a real stage; the z sweep reads the x and y sweeps' outputs.
  for d in (0, 1):
    ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.s1, solver.u, solver.s2,
                   stream)
  for d, name in enumerate(('x', 'y', 'z')):
    ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.s1, solver.u,
                   solver.s2, stream)  # warm
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
      ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.s1, solver.u,
                     solver.s2, stream)
    ev1.record(stream)
    ev1.synchronize()
    out[name] = ev0.elapsed_time(ev1) / 1e3 / reps
  return out


def pmc_probe(n):
  """Kernel launches for HBM-traffic counters: a copy of 64 Mi floats (256 MiB
  read + 256 MiB written: the counters' calibration), then the x, y and z
  sweeps (RK3 stage-2 form, U0 distinct from Uk) twice each on the n^3
  Kelvin-Helmholtz state in the default launch geometry with the XCD-aware
  block order on, then the same six launches with it off (the last twelve
  dispatches)."""
  import torch
  from swirlc_amd import _native as N
  from swirlc_amd import configs
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.core.solver import Solver
  src = torch.ones(64 << 20, device='cuda')
  dst = torch.empty_like(src)
  torch.mul(src, 1.0, out=dst)
  torch.cuda.synchronize()
  del src, dst
  cfg_d, st = configs.kh(n=n)
  solver = Solver(SwirlCParameters(cfg_d), {k: torch.from_numpy(v).cuda() for k, v in st.items()},
                  cfl=0.5, use_graph=False)
  del st
  solver.step(1)
  ctx = solver.ctx
  ctx.fill_halos(solver.u)
  torch.cuda.synchronize()
  for remap in (1, 0):
    ctx.set_tuning({'xcd_remap': remap})
    print(f'[pmc-probe] tuning {ctx.get_tuning()}', flush=True)
    for d in (0, 1, 2):
      for _ in range(2):
        ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.s1, solver.u, solver.s2)
    torch.cuda.synchronize()
  print(f'[pmc-probe] copy 268435456 B each way; sweeps x, y, z x2 at {n}^3, '
        'XCD-aware block order on, then off', flush=True)


def bench_batch(n_inst, steps, loop_steps):
  """Batched ensemble (256 3-D Riemann problems, WENO5 + HLLC, inviscid, RK3,
  per-instance CFL dt): one launch per phase over all instances, graph
  replayed, against a per-instance loop of the single-grid kernels."""
  import torch
  from swirlc_amd import configs
  from swirlc_amd.core.batch import BatchSolver
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.core.solver import Solver
  from swirlc_amd.core import autotune
  cfg, members = configs.riemann_ensemble(n_inst, seed=0)
  cells = sum(g[0] * g[1] * g[2] for g, _ in members)
  batch = BatchSolver(SwirlCParameters(dict(cfg)), members, cfl=0.5, use_graph=True)
  tuning, _ = autotune.tune_batch(batch, reps=2)
  batch.step(2)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  batch.step(steps)
  torch.cuda.synchronize()
  t_batch = (time.perf_counter() - t0) / steps
  finite = all(bool(torch.isfinite(t).all()) for t in batch.u)
  del batch
  torch.cuda.empty_cache()
  solvers = [Solver(SwirlCParameters(configs.member_cfg(cfg, g)),
                    {k: torch.from_numpy(v).cuda() for k, v in st.items()}, cfl=0.5,
                    use_graph=False) for g, st in members]
  del members
  for s in solvers:
    s.ctx.set_tuning(tuning)  # same geometry as the batched kernels
    s.step(1)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(loop_steps):
    for s in solvers:
      s.step(1)
  torch.cuda.synchronize()
  t_loop = (time.perf_counter() - t0) / loop_steps
  # The same loop with every launch recorded into one HIP graph (GPU-side
  # cost of per-instance launches, without Python overhead).
  side = torch.cuda.Stream()
  side.wait_stream(torch.cuda.current_stream())
  g = torch.cuda.CUDAGraph()
  with torch.cuda.stream(side):
    with torch.cuda.graph(g, stream=side):
      for s in solvers:
        s._step_body()
  torch.cuda.current_stream().wait_stream(side)
  g.replay()
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  for _ in range(loop_steps):
    g.replay()
  torch.cuda.synchronize()
  t_loop_graph = (time.perf_counter() - t0) / loop_steps
  return {'instances': n_inst, 'cells': cells, 'ms_per_step': round(t_batch * 1e3, 3),
          'Mcell_updates_per_s': round(cells / t_batch / 1e6, 1),
          'loop_ms_per_step': round(t_loop * 1e3, 3),
          'loop_graph_ms_per_step': round(t_loop_graph * 1e3, 3),
          'speedup_vs_loop': round(t_loop / t_batch, 2),
          'speedup_vs_graphed_loop': round(t_loop_graph / t_batch, 2),
          'finite': finite, 'tuning': tuning}


def bench_adjoint(n, steps, every):
  """Gradient of J = sum 0.5 rho |u|^2 + 0.1 sum (rho - rho_target)^2 after a
  `steps`-step SSP-RK3 rollout of the viscous Kelvin-Helmholtz setup at n^3
  (fixed dt = the CFL-0.5 step of the initial state), with respect to all
  five initial conserved fields; RK states checkpointed every `every` steps
  and recomputed.  Reports forward and backward (adjoint incl. recompute)
  times and their ratio, and the same for the rollout that keeps every
  stage state (`full_tape`: nothing recomputed)."""
  import numpy as np
  import torch
  from swirlc_amd import configs
  from swirlc_amd.core import adjoint as A
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.equations.compressible_navier_stokes import get_rhs_fn
  cfg_d, st = configs.kh(n=n)
  cfg = SwirlCParameters(cfg_d)
  nctx = get_rhs_fn(None, cfg).native_context
  h = cfg_d['halo_width']
  u0 = [torch.from_numpy(st[k]).cuda() for k in nctx.names]
  del st
  from swirlc_amd.core import autotune
  scratch = [torch.zeros_like(t) for t in u0]
  dt0 = torch.full((1,), 1e-4, device='cuda')
  autotune.tune_context(nctx, [t.clone() for t in u0], scratch, dt0, reps=3)
  g = [torch.randn_like(t) for t in u0]
  lq = [torch.zeros_like(t) for t in u0]
  q = [t.clone() for t in u0]
  adj_tuning, _ = autotune.tune_adjoint(nctx, q, g, lq)
  del scratch, g, lq, q
  speed = torch.zeros(4, device='cuda')
  dt = torch.zeros(1, device='cuda')
  nctx.max_wave_speed(u0, speed, 0.5, dt)
  rho_t = torch.from_numpy(np.random.default_rng(1).uniform(
      0.9, 2.1, tuple(u0[0].shape)).astype(np.float32)).cuda()
  # Warm-up (kernels loaded, adjoint buffers allocated).
  w = [t.clone().requires_grad_(True) for t in u0]
  A.kinetic_density_loss(A.rollout(nctx, dt, 2, *w, every=1), rho_t, h).backward()
  torch.cuda.synchronize()
  x = [t.clone().requires_grad_(True) for t in u0]
  t0 = time.perf_counter()
  out = A.rollout(nctx, dt, steps, *x, every=every)
  torch.cuda.synchronize()
  t_fwd = time.perf_counter() - t0
  J = A.kinetic_density_loss(out, rho_t, h)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  J.backward()
  torch.cuda.synchronize()
  t_bwd = time.perf_counter() - t0
  J_val = float(J.detach())
  finite = all(bool(torch.isfinite(t.grad).all()) for t in x)
  gnorm = float(sum((t.grad.double() ** 2).sum() for t in x) ** 0.5)
  grads = [t.grad for t in x]
  del out, J, x
  # The same rollout keeping every stage state (every=0: no recomputation in
  # the backward pass); run twice, the second timed (the first allocates the
  # tape, which the allocator then keeps).
  tape = None
  try:
    for rep in range(2):
      xt = [t.clone().requires_grad_(True) for t in u0]
      torch.cuda.synchronize()
      t0 = time.perf_counter()
      out = A.rollout(nctx, dt, steps, *xt, every=0)
      torch.cuda.synchronize()
      tf = time.perf_counter() - t0
      Jt = A.kinetic_density_loss(out, rho_t, h)
      torch.cuda.synchronize()
      t0 = time.perf_counter()
      Jt.backward()
      torch.cuda.synchronize()
      tb = time.perf_counter() - t0
      del out
    same = all(torch.equal(a.grad, b) for a, b in zip(xt, grads))
    tape = {'forward_s': round(tf, 4), 'adjoint_s': round(tb, 4),
            'adjoint_over_forward': round(tb / tf, 3), 'gradient_bitwise_equal': same}
    del xt, Jt
  except MemoryError as e:
    tape = {'skipped': str(e)}
  torch.cuda.empty_cache()
  return {'workload': f'kelvin_helmholtz_{n}^3_viscous_ssprk3_{steps}steps_ckpt{every}',
          'dt': float(dt[0]), 'loss': float(J_val), 'grad_norm': gnorm, 'finite': finite,
          'forward_s': round(t_fwd, 4), 'adjoint_s': round(t_bwd, 4),
          'adjoint_over_forward': round(t_bwd / t_fwd, 3),
          'forward_Mcell_updates_per_s': round(n ** 3 * steps / t_fwd / 1e6, 1),
          'full_tape': tape,
          'tuning': {k: adj_tuning[k] for k in ('adj_seg', 'adj_by', 'adj_by_y')}}


def bench_large(steps):
  """The buoyant bubble at its full 512 x 512 x 256 size (MUSCL + HLL,
  viscous, gravity, Dirichlet z walls, CFL 0.5): RK3 steps timed with HIP
  events after two warm-up steps."""
  import torch
  from swirlc_amd import configs
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.core.solver import Solver
  cfg_d, st = configs.bubble(512, 512, 256)
  sim = Solver(SwirlCParameters(cfg_d), {k: torch.from_numpy(v).cuda() for k, v in st.items()},
               cfl=0.5, use_graph=True)
  del st
  from swirlc_amd.core import autotune
  tuning, _ = autotune.tune_context(sim.ctx, sim.u, sim.s2, sim.dt, u0=sim.s1, reps=3)
  sim.step(2)
  torch.cuda.synchronize()
  ev0 = torch.cuda.Event(enable_timing=True)
  ev1 = torch.cuda.Event(enable_timing=True)
  ev0.record()
  sim.step(steps)
  ev1.record()
  ev1.synchronize()
  ms = ev0.elapsed_time(ev1) / steps
  cells = 512 * 512 * 256
  finite = all(bool(torch.isfinite(v).all()) for v in sim.state.values())
  del sim
  torch.cuda.empty_cache()
  return {'workload': 'buoyant_bubble_512x512x256_muscl_hll_viscous_gravity_ssprk3_cfl0.5',
          'steps': steps, 'ms_per_step': round(ms, 3),
          'Mcell_updates_per_s': round(cells / ms / 1e3, 1), 'finite': finite,
          'tuning': tuning}


def _kh_oracle_cfg(n):
  """The KH workload at n^3 in the oracle's config form, and its state."""
  from oracle import swirlc_oracle as O
  from swirlc_amd import configs
  cfg_d, st = configs.kh(n=n)
  ocfg = O.default_cfg(**{k: cfg_d[k] for k in (
      'halo_width', 'core_nx', 'core_ny', 'core_nz', 'x', 'y', 'z', 'dx', 'dy', 'dz',
      'interpolation_scheme', 'numeric_flux_scheme', 'include_diffusion',
      'transport_parameters', 'bc')})
  return ocfg, st


# State of the slab workers of the parallel CPU oracle (set before the pool
# forks, read by _slab_rhs in the workers).
_SLAB = {}


def _slab_rhs(bounds):
  """Oracle RHS of interior z planes [k0, k1) from the halo-filled state in
  shared memory: the slab with its 2h halo planes is an oracle grid whose z
  dimension has no BC (its halo planes are the filled neighbours)."""
  import numpy as np
  from multiprocessing import shared_memory
  from oracle import swirlc_oracle as O
  k0, k1 = bounds
  g = _SLAB
  h = g['h']
  shm_in = shared_memory.SharedMemory(name=g['in'])
  shm_out = shared_memory.SharedMemory(name=g['out'])
  q = np.ndarray(g['shape'], np.float32, buffer=shm_in.buf)
  r = np.ndarray(g['shape'], np.float32, buffer=shm_out.buf)
  sub = {name: q[f, k0:k1 + 2 * h].copy() for f, name in enumerate(g['names'])}
  cfg = dict(g['cfg'], core_nz=k1 - k0, nz=k1 - k0, z=g['cfg']['z'][k0:k1])
  cfg['bc'] = {'cell_averages': {v: {d: b for d, b in spec.items() if d != 'z'}
                                 for v, spec in g['cfg']['bc']['cell_averages'].items()}}
  out = O.rhs(sub, {}, cfg)
  for f, name in enumerate(g['names']):
    r[f, h + k0:h + k1] = out[name][h:-h]
  del q, r
  shm_in.close()
  shm_out.close()
  return k1 - k0


def parallel_rk3_step(pool, state, ocfg, dt, slabs, shm_in, shm_out, shape):
  """One oracle RK3 step (numerics/time_integration/rk3.py:34-122) with each
  RHS evaluated slab-parallel by `pool`: the parent fills the halos
  (boundary.py:464-539) and writes the state to shared memory, the workers
  compute the RHS of their z slabs.  Equal, bit for bit, to the serial
  oracle's step (the RHS of a plane depends on its +-h neighbour planes only)."""
  import numpy as np
  from oracle import swirlc_oracle as O
  names = list(state)
  q = np.ndarray(shape, np.float32, buffer=shm_in.buf)
  r = np.ndarray(shape, np.float32, buffer=shm_out.buf)

  def rhs(st):
    filled = O.update_conservative_cell_averages(st, ocfg)
    for f, name in enumerate(names):
      q[f] = filled[name]
    pool.map(_slab_rhs, slabs)
    return {name: r[f].copy() for f, name in enumerate(names)}

  return O.rk3(state, {}, ocfg, dt=dt, rhs_fn=rhs)


def _cpu_model():
  """The host CPU's model name with its family / model / stepping numbers and
  clock from /proc/cpuinfo (a virtualised host may report a bare name)."""
  info = {}
  try:
    for line in open('/proc/cpuinfo'):
      if ':' in line:
        k, v = line.split(':', 1)
        info.setdefault(k.strip(), v.strip())
      elif info:
        break
  except OSError:
    return ''
  name = info.get('model name', '')
  extra = [f'{k} {info[k]}' for k in ('cpu family', 'model', 'stepping', 'cpu MHz')
           if k in info]
  return f'{name} ({", ".join(extra)})' if extra else name


def cpu_baseline(n, single_n, workers):
  """The fp32 numpy oracle on the host: one RK3 step of the n^3 workload with
  the RHS split over `workers` processes (z slabs), and one serial RK3 step
  at single_n^3 on one core."""
  import multiprocessing as mp
  from multiprocessing import shared_memory
  import numpy as np
  from oracle import swirlc_oracle as O
  dt = np.float32(1e-4)
  # Single thread.
  ocfg, st = _kh_oracle_cfg(single_n)
  t0 = time.perf_counter()
  O.rk3(st, {}, ocfg, dt=dt)
  el1 = time.perf_counter() - t0
  # All workers.
  ocfg, st = _kh_oracle_cfg(n)
  h = ocfg['halo_width']
  names = list(st)
  shape = (len(names),) + st[names[0]].shape
  nbytes = int(np.prod(shape)) * 4
  shm_in = shared_memory.SharedMemory(create=True, size=nbytes)
  shm_out = shared_memory.SharedMemory(create=True, size=nbytes)
  try:
    _SLAB.update({'in': shm_in.name, 'out': shm_out.name, 'shape': shape, 'h': h,
                  'names': names, 'cfg': ocfg})
    nz = ocfg['core_nz']
    edges = np.linspace(0, nz, workers + 1).round().astype(int)
    slabs = [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    with mp.get_context('fork').Pool(workers) as pool:
      t0 = time.perf_counter()
      parallel_rk3_step(pool, st, ocfg, dt, slabs, shm_in, shm_out, shape)
      el = time.perf_counter() - t0
  finally:
    shm_in.close()
    shm_in.unlink()
    shm_out.close()
    shm_out.unlink()
  return {'value': round(n ** 3 / el / 1e6, 3), 'unit': 'Mcell-updates/s', 'cores': workers,
          'kind': 'port',
          'sample': f'one RK3 step of the same workload at {n}^3 (fp32 numpy oracle, RHS '
                    f'split over {workers} processes by z slab), {el:.1f} s',
          'single_thread_value': round(single_n ** 3 / el1 / 1e6, 3),
          'single_thread_sample': f'one RK3 step at {single_n}^3 on one core, {el1:.1f} s',
          'cpu_model': _cpu_model(),
          'affinity_cores': len(os.sched_getaffinity(0))}


def main():
  args = parse()
  if args.gpus != 1:
    print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); the '
          'multi-GPU decomposition is not part of this build.', file=sys.stderr)
    sys.exit(2)
  if args.pmc_probe:
    pmc_probe(args.n)
    return
  if args.adjoint_only:
    print(json.dumps(bench_adjoint(args.adjoint_n, args.adjoint_steps, args.adjoint_every)),
          flush=True)
    return
  import torch
  from swirlc_amd import configs
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.core.solver import Solver

  n = args.n
  cfg_d, st = configs.kh(n=n)
  cfg = SwirlCParameters(cfg_d)
  state = {k: torch.from_numpy(v).cuda() for k, v in st.items()}
  del st
  solver = Solver(cfg, state, cfl=0.5, use_graph=True)
  del state
  from swirlc_amd.core import autotune
  if args.tuning:
    solver.ctx.set_tuning(json.loads(args.tuning))
    tuning = solver.ctx.get_tuning()
  elif args.no_tune:
    tuning = solver.ctx.get_tuning()
  else:
    tuning, _ = autotune.tune_context(solver.ctx, solver.u, solver.s2, solver.dt, u0=solver.s1)
  print(f'[bench] KH {n}^3 state on GPU; tuning {tuning}; warmup {args.warmup} steps',
        file=sys.stderr, flush=True)
  if args.profile_steps:
    solver.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.step(args.profile_steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(f'[bench] profile: {args.profile_steps} graph-replayed steps, '
          f'{el / args.profile_steps * 1e3:.3f} ms/step', flush=True)
    return
  solver.step(args.warmup)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  solver.step(args.steps)
  torch.cuda.synchronize()
  el = time.perf_counter() - t0
  cells = n ** 3
  value = cells * args.steps / el / 1e6
  ms = el / args.steps * 1e3
  finite = all(bool(torch.isfinite(t).all()) for t in solver.u)
  print(f'[bench] {args.steps} steps in {el:.3f} s ({ms:.3f} ms/step), finite={finite}',
        file=sys.stderr, flush=True)

  # In-step kernel durations (the step's launch sequence with HIP events
  # between launches) and the dominant kernel's roofline from them.
  step_k, step_ev_ms = time_step_kernels(solver, args.kernel_steps)
  per_step = {k: sum(v) for k, v in step_k.items()}
  dom = max((k for k in per_step if k.startswith('sweep_')), key=per_step.get)
  d = dom.split('_')[1]
  bytes_launch = sum(SWEEP_BYTES[d]) / 3.0 * cells  # mean over the three stages
  t_launch = per_step[dom] / 3.0 / 1e3  # s
  achieved = bytes_launch / t_launch / 1e9
  traffic = measured_traffic(dom, tuning['xcd_remap']) if n == 256 else None
  ktimes = time_sweeps(solver, args.kernel_reps)
  copy_bw = copy_bandwidth()
  roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK / 1e9,
              'unit': 'GB/s', 'frac': round(achieved * 1e9 / HBM_PEAK, 4),
              'traffic': traffic, 'kernel': dom,
              'algorithmic_bytes': int(round(bytes_launch)),
              'kernel_ms': round(t_launch * 1e3, 4),
              'in_step_ms': {k: [round(x, 4) for x in v] for k, v in step_k.items()},
              'in_step_sum_ms': round(sum(per_step.values()), 4),
              'in_step_event_step_ms': round(step_ev_ms, 4),
              'isolated_kernel_ms': {k: round(v * 1e3, 4) for k, v in ktimes.items()},
              'copy_GBps_measured': round(copy_bw / 1e9, 1)}
  step_s = el / args.steps
  step_frac = {'520B': round(cells * STEP_BYTES / step_s / HBM_PEAK, 4),
               '700B': round(cells * STEP_BYTES_VISCOUS / step_s / HBM_PEAK, 4)}
  print(f'[bench] in-step kernels (ms per launch, stages 1-3, each launch timing itself): '
        f'{roofline["in_step_ms"]}; sum {roofline["in_step_sum_ms"]} ms vs {ms:.3f} ms per '
        'graph-replayed step', file=sys.stderr, flush=True)

  batched = None
  if not args.no_batch:
    print(f'[bench] batched ensemble: {args.batch_instances} instances', file=sys.stderr,
          flush=True)
    del solver
    torch.cuda.empty_cache()
    batched = bench_batch(args.batch_instances, args.batch_steps, args.batch_loop_steps)
    print(f'[bench] batched: {batched}', file=sys.stderr, flush=True)

  large = None
  if not args.no_large:
    print(f'[bench] buoyant bubble 512x512x256, {args.large_steps} steps', file=sys.stderr,
          flush=True)
    torch.cuda.empty_cache()
    large = bench_large(args.large_steps)
    print(f'[bench] large: {large}', file=sys.stderr, flush=True)

  adjoint = None
  if not args.no_adjoint:
    print(f'[bench] adjoint: {args.adjoint_n}^3, {args.adjoint_steps} steps', file=sys.stderr,
          flush=True)
    torch.cuda.empty_cache()
    adjoint = bench_adjoint(args.adjoint_n, args.adjoint_steps, args.adjoint_every)
    print(f'[bench] adjoint: {adjoint}', file=sys.stderr, flush=True)

  cpu = None
  if not args.no_cpu:
    workers = args.cpu_workers or min(16, len(os.sched_getaffinity(0)))
    print(f'[bench] CPU oracle: one RK3 step at {args.cpu_n}^3 on {workers} processes, one at '
          f'{args.cpu_single_n}^3 on one core', file=sys.stderr, flush=True)
    cpu = cpu_baseline(args.cpu_n, args.cpu_single_n, workers)
    print(f'[bench] cpu: {cpu}', file=sys.stderr, flush=True)

  line = {
      'metric': 'Mcell-updates/s/RK-step', 'value': round(value, 2),
      'unit': 'Mcell-updates/s', 'n_gpus': 1, 'steps': args.steps,
      'warmup': args.warmup, 'ms_per_step': round(ms, 4), 'higher_is_better': True,
      'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
      'config': {'workload': f'kelvin_helmholtz_{n}^3_weno5_hllc_viscous_ssprk3_cfl0.5',
                 'cells': cells, 'parallelism': 'single-gpu', 'tuning': tuning},
      'roofline': roofline,
      'step_hbm_frac': step_frac,
      'cpu_baseline': cpu,
      'batched': batched,
      'large': large,
      'adjoint': adjoint,
      'finite': finite,
  }
  print(json.dumps(line), flush=True)


if __name__ == '__main__':
  main()
