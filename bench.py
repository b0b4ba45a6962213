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
# 5 conserved fields): x reads U, writes R; y reads U and R, writes R; z in an
# RK3 stage 2/3 reads U, R and U0 and writes the new state.
SWEEP_BYTES = {'x': 40, 'y': 60, 'z': 80}
# Per RK3 step (viscous, per-direction-sweep accounting): 700 B per cell.
STEP_BYTES_VISCOUS = 700


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=20)
  ap.add_argument('--warmup', type=int, default=3)
  ap.add_argument('--n', type=int, default=256)
  ap.add_argument('--kernel-reps', type=int, default=10)
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-n', type=int, default=64)
  ap.add_argument('--cpu-steps', type=int, default=2)
  ap.add_argument('--batch-instances', type=int, default=256)
  ap.add_argument('--batch-steps', type=int, default=10)
  ap.add_argument('--batch-loop-steps', type=int, default=3)
  ap.add_argument('--no-batch', action='store_true')
  return ap.parse_args()


def time_sweeps(solver, reps):
  """Average duration (s) of each sweep kernel, HIP events on its stream."""
  import torch
  from swirlc_amd import _native as N
  ctx = solver.ctx
  stream = torch.cuda.current_stream()
  ctx.fill_halos(solver.u)
  out = {}
  for d, name in enumerate(('x', 'y', 'z')):
    ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.u, solver.u,
                   solver.s2, stream)  # warm
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
      ctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, solver.dt, solver.u, solver.u,
                     solver.s2, stream)
    ev1.record(stream)
    ev1.synchronize()
    out[name] = ev0.elapsed_time(ev1) / 1e3 / reps
  return out


def bench_batch(n_inst, steps, loop_steps):
  """Batched ensemble (256 3-D Riemann problems, WENO5 + HLLC, inviscid, RK3,
  per-instance CFL dt): one launch per phase over all instances, graph
  replayed, against a per-instance loop of the single-grid kernels."""
  import torch
  from swirlc_amd import configs
  from swirlc_amd.core.batch import BatchSolver
  from swirlc_amd.core.parameter import SwirlCParameters
  from swirlc_amd.core.solver import Solver
  cfg, members = configs.riemann_ensemble(n_inst, seed=0)
  cells = sum(g[0] * g[1] * g[2] for g, _ in members)
  batch = BatchSolver(SwirlCParameters(dict(cfg)), members, cfl=0.5, use_graph=True)
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
          'finite': finite}


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
  return {'value': n ** 3 * steps / el / 1e6, 'unit': 'Mcell-updates/s', 'cores': 1,
          'kind': 'port',
          'sample': f'{steps} RK3 steps of the same workload at {n}^3 (fp32 numpy oracle, '
                    f'single thread), {el:.1f} s'}


def main():
  args = parse()
  if args.gpus != 1:
    print(f'bench.py runs on one GPU only (asked for --gpus {args.gpus}); the '
          'multi-GPU decomposition is not part of this build.', file=sys.stderr)
    sys.exit(2)
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
  print(f'[bench] KH {n}^3 state on GPU; warmup {args.warmup} steps', file=sys.stderr,
        flush=True)
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
  roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK / 1e9,
              'unit': 'GB/s', 'frac': round(achieved * 1e9 / HBM_PEAK, 4),
              'traffic': None, 'kernel': f'sweep_{dom}',
              'kernel_ms': {k: round(v * 1e3, 4) for k, v in ktimes.items()}}
  step_frac = cells * STEP_BYTES_VISCOUS / (el / args.steps) / HBM_PEAK

  batched = None
  if not args.no_batch:
    print(f'[bench] batched ensemble: {args.batch_instances} instances', file=sys.stderr,
          flush=True)
    del solver
    torch.cuda.empty_cache()
    batched = bench_batch(args.batch_instances, args.batch_steps, args.batch_loop_steps)
    print(f'[bench] batched: {batched}', file=sys.stderr, flush=True)

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
                 'cells': cells, 'parallelism': 'single-gpu'},
      'roofline': roofline,
      'step_hbm_frac_700B': round(step_frac, 4),
      'cpu_baseline': cpu,
      'batched': batched,
      'finite': finite,
  }
  print(json.dumps(line), flush=True)


if __name__ == '__main__':
  main()
