"""Benchmark: SSP-RK3 steps of the 3-D Kelvin-Helmholtz workload on one MI355X.

Workload: periodic 256^3 cube, fp32, gamma = 1.4, WENO_5 + HLLC, viscous
(nu = 1e-4, Pr = 0.7), SSP-RK3 with a device-computed CFL-0.5 time step;
synthetic initial state (shear layer rho = 2 / 1, u = +-0.5, p = 2.5,
v = 0.01 sin(4 pi x) + U(-1e-3, 1e-3) noise from seed 0).

Prints one JSON line: Mcell-updates/s per RK step (interior cells x steps /
seconds / 1e6), with the roofline of the dominant kernel (timed with HIP
events on its own stream) and the CPU oracle's throughput on a bounded
sample of the same workload.

  python bench.py [--steps K] [--warmup W] [--n 256] [--no-cpu]
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
# 5 conserved fields), in launch order: x reads U, writes R; z reads U and R,
# writes R; y (last) in an RK3 stage 2/3 reads U, R and U0 and writes the new
# state.
SWEEP_BYTES = {'x': 40, 'y': 80, 'z': 60}
# Per RK3 step (viscous, per-direction-sweep accounting): 700 B per cell.
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
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--n', type=int, default=256)
  ap.add_argument('--kernel-reps', type=int, default=10)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-n', type=int, default=128)
  ap.add_argument('--cpu-steps', type=int, default=3)
  ap.add_argument('--batch-instances', type=int, default=256)
  ap.add_argument('--batch-steps', type=int, default=10)
  ap.add_argument('--batch-loop-steps', type=int, default=3)
  ap.add_argument('--no-batch', action='store_true')
  ap.add_argument('--adjoint-n', type=int, default=128)
  ap.add_argument('--adjoint-steps', type=int, default=200)
  ap.add_argument('--adjoint-every', type=int, default=20)
  ap.add_argument('--no-adjoint', action='store_true')
  ap.add_argument('--no-large', action='store_true',
                  help='skip the 512 x 512 x 256 buoyant-bubble leg')
  ap.add_argument('--large-steps', type=int, default=10)
  ap.add_argument('--no-tune', action='store_true',
                  help='keep the default launch geometry (no autotuning)')
  ap.add_argument('--pmc-probe', action='store_true',
                  help='launch a 256 MiB copy kernel, then each sweep kernel twice on the '
                       'state, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes')
  return ap.parse_args()


def time_sweeps(solver, reps):
  """Average duration (s) of each sweep kernel, HIP events on its stream."""
  import torch
  from swirlc_amd import _native as N
  ctx = solver.ctx
  stream = torch.cuda.current_stream()
  ctx.fill_halos(solver.u)
  out = {}
  # RK3 stage-2 form with U0 (the last step's stage-1 state) distinct from Uk,
  # as in a real stage.
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
  read + 256 MiB written: the counters' calibration), then the x, z and y
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
    for d in (0, 2, 1):
      for _ in range(2):
        ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.s1, solver.u, solver.s2)
    torch.cuda.synchronize()
  print(f'[pmc-probe] copy 268435456 B each way; sweeps x, z, y x2 at {n}^3, '
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
  times and their ratio."""
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
  finite = all(bool(torch.isfinite(t.grad).all()) for t in x)
  gnorm = float(sum((t.grad.double() ** 2).sum() for t in x) ** 0.5)
  return {'workload': f'kelvin_helmholtz_{n}^3_viscous_ssprk3_{steps}steps_ckpt{every}',
          'dt': float(dt[0]), 'loss': float(J.detach()), 'grad_norm': gnorm, 'finite': finite,
          'forward_s': round(t_fwd, 4), 'adjoint_s': round(t_bwd, 4),
          'adjoint_over_forward': round(t_bwd / t_fwd, 3),
          'forward_Mcell_updates_per_s': round(n ** 3 * steps / t_fwd / 1e6, 1),
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
          'Mcell_updates_per_s': round(cells / ms / 1e3, 1), 'finite': finite}


def cpu_baseline(n, steps):
  """The fp32 numpy oracle on an n^3 sample of the same workload (1 thread)."""
  import numpy as np
  from oracle import swirlc_oracle as O
  from swirlc_amd import configs
  cfg_d, st = configs.kh(n=n)
  ocfg = O.default_cfg(**{k: cfg_d[k] for k in (
      'halo_width', 'core_nx', 'core_ny', 'core_nz', 'x', 'y', 'z', 'dx', 'dy', 'dz',
      'interpolation_scheme', 'numeric_flux_scheme', 'include_diffusion',
      'transport_parameters', 'bc')})
  dt = np.float32(1e-4)
  t0 = time.perf_counter()
  for _ in range(steps):
    st = O.rk3(st, {}, ocfg, dt=dt)
  el = time.perf_counter() - t0
  model = ''
  try:
    for line in open('/proc/cpuinfo'):
      if line.startswith('model name'):
        model = line.split(':', 1)[1].strip()
        break
  except OSError:
    pass
  return {'value': n ** 3 * steps / el / 1e6, 'unit': 'Mcell-updates/s', 'cores': 1,
          'kind': 'port',
          'sample': f'{steps} RK3 steps of the same workload at {n}^3 (fp32 numpy oracle, '
                    f'single thread), {el:.1f} s',
          'cpu_model': model}


def main():
  args = parse()
  if args.gpus != 1:
    print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); the '
          'multi-GPU decomposition is not part of this build.', file=sys.stderr)
    sys.exit(2)
  if args.pmc_probe:
    pmc_probe(args.n)
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
  if args.no_tune:
    tuning = solver.ctx.get_tuning()
  else:
    tuning, _ = autotune.tune_context(solver.ctx, solver.u, solver.s2, solver.dt, u0=solver.s1)
  print(f'[bench] KH {n}^3 state on GPU; tuning {tuning}; warmup {args.warmup} steps',
        file=sys.stderr, flush=True)
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

  ktimes = time_sweeps(solver, args.kernel_reps)
  dom = max(ktimes, key=ktimes.get)
  bytes_launch = SWEEP_BYTES[dom] * cells
  achieved = bytes_launch / ktimes[dom] / 1e9
  traffic = measured_traffic(f'sweep_{dom}', tuning['xcd_remap']) if n == 256 else None
  copy_bw = copy_bandwidth()
  roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK / 1e9,
              'unit': 'GB/s', 'frac': round(achieved * 1e9 / HBM_PEAK, 4),
              'traffic': traffic, 'kernel': f'sweep_{dom}',
              'algorithmic_bytes': bytes_launch,
              'kernel_ms': {k: round(v * 1e3, 4) for k, v in ktimes.items()},
              'copy_GBps_measured': round(copy_bw / 1e9, 1)}
  step_frac = cells * STEP_BYTES_VISCOUS / (el / args.steps) / HBM_PEAK

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
    print(f'[bench] CPU oracle sample {args.cpu_n}^3 x {args.cpu_steps} steps',
          file=sys.stderr, flush=True)
    cpu = cpu_baseline(args.cpu_n, args.cpu_steps)

  line = {
      'metric': 'Mcell-updates/s/RK-step', 'value': round(value, 2),
      'unit': 'Mcell-updates/s', 'n_gpus': 1, 'steps': args.steps,
      'warmup': args.warmup, 'ms_per_step': round(ms, 4), 'higher_is_better': True,
      'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
      'config': {'workload': f'kelvin_helmholtz_{n}^3_weno5_hllc_viscous_ssprk3_cfl0.5',
                 'cells': cells, 'parallelism': 'single-gpu', 'tuning': tuning},
      'roofline': roofline,
      'step_hbm_frac_700B': round(step_frac, 4),
      'cpu_baseline': cpu,
      'batched': batched,
      'large': large,
      'adjoint': adjoint,
      'finite': finite,
  }
  print(json.dumps(line), flush=True)


if __name__ == '__main__':
  main()
