"""Launch-geometry autotuner for the sweep kernels (no reference counterpart).

The reference leaves scheduling to XLA; here each sweep kernel's launch
geometry (cells per thread and waves per block of the marching x / z sweeps,
their register budget, waves per block of the y sweep, and the XCD-aware
block order) is chosen per grid by timing the candidates on the actual state
with HIP events, one direction at a time (a direction's time does not depend
on the other directions' geometry), then the block order over all three.
Results are bit-for-bit independent of the choice, so tuning never changes
the numbers, only the time.

  tune_context(nctx, q, out, dt_dev)   single grid (sc_sweep_only timings)
  tune_batch(batch)                    batched ensemble (full-stage timings)
"""

import itertools

import torch

from swirlc_amd import _native as N

MARCH_SPACE = tuple(itertools.product((16, 32, 64), (2, 4), (0, 4)))  # seg, by, waves
Y_SPACE = (1, 2, 4)  # waves per block


def _timed(fn, reps):
  stream = torch.cuda.current_stream()
  fn()  # warm (first launch of a kernel variant loads it)
  ev0 = torch.cuda.Event(enable_timing=True)
  ev1 = torch.cuda.Event(enable_timing=True)
  ev0.record(stream)
  for _ in range(reps):
    fn()
  ev1.record(stream)
  ev1.synchronize()
  return ev0.elapsed_time(ev1) / reps


def _candidates(base, march_space, y_space):
  """(direction, tuning) pairs: each direction's candidates over `base`."""
  for d, keys in (('x', ('seg_x', 'by_x', 'waves_x')), ('z', ('seg_z', 'by_z', 'waves_z'))):
    for vals in march_space:
      t = dict(base)
      t.update(zip(keys, vals))
      yield d, t
  for by in y_space:
    t = dict(base)
    t['by_y'] = by
    yield 'y', t


def _pick(timings, base):
  best = dict(base)
  keys = {'x': ('seg_x', 'by_x', 'waves_x'), 'z': ('seg_z', 'by_z', 'waves_z'), 'y': ('by_y',)}
  for d, ks in keys.items():
    rows = [(ms, t) for dd, t, ms in timings if dd == d]
    if rows:
      _, t = min(rows, key=lambda r: r[0])
      for k in ks:
        best[k] = t[k]
  return best


def tune_context(nctx, q, out, dt_dev, u0=None, reps=5, march_space=MARCH_SPACE,
                 y_space=Y_SPACE):
  """Times each sweep kernel of an RK3 stage-2 update of state q (base state
  u0, default q; out receives the update) for every candidate geometry and
  keeps the fastest per direction.  Returns (tuning, [(direction, tuning, ms)])."""
  u0 = q if u0 is None else u0
  base = nctx.get_tuning()
  nctx.fill_halos(q)
  dims = {'x': 0, 'y': 1, 'z': 2}
  timings = []
  for d, t in _candidates(base, march_space, y_space):
    nctx.set_tuning(t)
    ms = _timed(lambda: nctx.sweep_only(dims[d], N.SC_RK_BLEND, 0.75, 0.25, dt_dev, u0, q,
                                        out), reps)
    timings.append((d, t, ms))
  best = _pick(timings, base)
  # The block order (XCD-aware remap or not) is shared by the three sweeps:
  # time the whole x, z, y sequence both ways.
  def all_three():
    for d in (0, 2, 1):
      nctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, dt_dev, u0, q, out)

  for remap in (0, 1):
    t = dict(best, xcd_remap=remap)
    nctx.set_tuning(t)
    timings.append(('all', t, _timed(all_three, reps)))
  best = min((r for r in timings if r[0] == 'all'), key=lambda r: r[2])[1]
  nctx.set_tuning(best)
  return best, timings


def tune_batch(batch, reps=3, march_space=MARCH_SPACE, y_space=Y_SPACE):
  """Times one RK3 stage-2 update of the whole ensemble per candidate (the
  other directions held at the base geometry) and keeps the fastest per
  direction.  Returns (tuning, [(direction, tuning, ms)])."""
  base = batch.get_tuning()
  timings = []
  for d, t in _candidates(base, march_space, y_space):
    batch.set_tuning(t)
    ms = _timed(lambda: batch._stage(N.SC_RK_BLEND, 0.75, 0.25, batch.u, batch.u, batch.s2),
                reps)
    timings.append((d, t, ms))
  best = _pick(timings, base)
  batch.set_tuning(best)
  return best, timings
