"""Launch-geometry autotuner for the sweep kernels (no reference counterpart).

The reference leaves scheduling to XLA; here each sweep kernel's launch
geometry (cells per thread and waves per block of the marching x / z sweeps,
waves per block and z planes per wave of the y sweep, and the XCD-aware
block order) is chosen per
grid by timing the candidates on the actual state with HIP events, one
direction at a time (a direction's time does not depend on the other
directions' geometry), then the block order over all three.  Forward results
are bit-for-bit independent of the choice, so tuning never changes the
numbers, only the time; the adjoint passes sum each cell's terms in the same
order under every geometry (equal to a few ulps).

  tune_context(nctx, q, out, dt_dev)   single grid (sc_sweep_only timings)
  tune_batch(batch)                    batched ensemble (full-stage timings)
  tune_adjoint(nctx, q, g, lq)         adjoint passes (sc_rhs_vjp timings)
"""

import itertools

import torch

from swirlc_amd import _native as N

MARCH_SPACE = tuple(itertools.product((16, 32, 64), (2, 4)))  # seg, by
Y_SPACE = tuple(itertools.product((1, 2, 4), (4, 8, 16, 32)))  # by_y, seg_y
ADJ_MARCH_SPACE = tuple(itertools.product((8, 16, 32), (1, 2, 4)))  # adj_seg, adj_by
ADJ_Y_SPACE = (1, 2, 4)  # adj_by_y


MARGIN = 0.02  # a candidate must beat the current geometry by 2 % to replace it


def _timed(fn, reps, rounds=3):
  """Median over `rounds` of the mean time (ms) of `reps` back-to-back calls."""
  stream = torch.cuda.current_stream()
  fn()  # warm (first launch of a kernel variant loads it)
  means = []
  for _ in range(rounds):
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
      fn()
    ev1.record(stream)
    ev1.synchronize()
    means.append(ev0.elapsed_time(ev1) / reps)
  return sorted(means)[len(means) // 2]


def _candidates(base, march_space, y_space):
  """(direction, tuning) pairs: each direction's candidates over `base`."""
  for d, keys in (('x', ('seg_x', 'by_x')), ('z', ('seg_z', 'by_z'))):
    for vals in march_space:
      t = dict(base)
      t.update(zip(keys, vals))
      yield d, t
  for by, nw in y_space:
    t = dict(base)
    t['by_y'] = by
    t['seg_y'] = nw
    yield 'y', t


_KEYS = {'x': ('seg_x', 'by_x'), 'z': ('seg_z', 'by_z'), 'y': ('by_y', 'seg_y')}


def _pick(timings, base):
  """Per direction, the fastest candidate if it beats the base geometry's
  own time by MARGIN, else the base values."""
  best = dict(base)
  for d, ks in _KEYS.items():
    rows = [(ms, t) for dd, t, ms in timings if dd == d]
    if not rows:
      continue
    base_ms = [ms for ms, t in rows if all(t[k] == base[k] for k in ks)]
    ms, t = min(rows, key=lambda r: r[0])
    if not base_ms or ms < (1.0 - MARGIN) * base_ms[0]:
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
  # time the whole x, y, z sequence both ways.
  def all_three():
    for d in (0, 1, 2):
      nctx.sweep_only(d, N.SC_RK_BLEND, 0.75, 0.25, dt_dev, u0, q, out)

  for remap in (base['xcd_remap'], 1 - base['xcd_remap']):
    t = dict(best, xcd_remap=remap)
    nctx.set_tuning(t)
    timings.append(('all', t, _timed(all_three, reps)))
  rows = [r for r in timings if r[0] == 'all']
  best = rows[1][1] if rows[1][2] < (1.0 - MARGIN) * rows[0][2] else rows[0][1]
  # Final check on the fused stage itself (halo fill + three sweeps): keep
  # the base geometry unless the tuned one is faster there too.
  def stage():
    nctx.rk_stage(N.SC_RK_BLEND, 0.75, 0.25, dt_dev, u0, q, out)

  final = []
  for t in (base, best):
    nctx.set_tuning(t)
    final.append(_timed(stage, reps))
    timings.append(('stage', t, final[-1]))
  if not final[1] < (1.0 - MARGIN) * final[0]:
    best = base
  nctx.set_tuning(best)
  return best, timings


def tune_batch(batch, reps=3, march_space=MARCH_SPACE, y_space=Y_SPACE):
  """Times one RK3 stage-2 update of the whole ensemble per candidate (the
  other directions held at the base geometry) and keeps the fastest per
  direction.  Returns (tuning, [(direction, tuning, ms)])."""
  base = batch.get_tuning()
  if batch._chained and not batch._seeded:
    batch.refresh()  # the timed stages take the state's halos as filled
  timings = []
  for d, t in _candidates(base, march_space, y_space):
    batch.set_tuning(t)
    ms = _timed(lambda: batch._stage(N.SC_RK_BLEND, 0.75, 0.25, batch.u, batch.u, batch.s2),
                reps)
    timings.append((d, t, ms))
  best = _pick(timings, base)
  batch.set_tuning(best)
  return best, timings


def tune_adjoint(nctx, q, g, lq, reps=3, march_space=ADJ_MARCH_SPACE, y_space=ADJ_Y_SPACE):
  """Times the RHS VJP of state q (cotangent g, result in lq) for every
  adjoint geometry: the marching x / z passes share (adj_seg, adj_by), the
  y pass has adj_by_y.  Keeps the fastest of each if it beats the current
  geometry by MARGIN.  Returns (tuning, [(knob, tuning, ms)])."""
  base = nctx.get_tuning()
  nctx.enable_adjoint()
  timings = []
  for seg, by in march_space:
    t = dict(base, adj_seg=seg, adj_by=by)
    nctx.set_tuning(t)
    timings.append(('march', t, _timed(lambda: nctx.rhs_vjp(q, g, lq), reps)))
  for by in y_space:
    t = dict(base, adj_by_y=by)
    nctx.set_tuning(t)
    timings.append(('y', t, _timed(lambda: nctx.rhs_vjp(q, g, lq), reps)))
  best = dict(base)
  for knob, ks in (('march', ('adj_seg', 'adj_by')), ('y', ('adj_by_y',))):
    rows = [(ms, t) for kk, t, ms in timings if kk == knob]
    base_ms = [ms for ms, t in rows if all(t[k] == base[k] for k in ks)]
    ms, t = min(rows, key=lambda r: r[0])
    if not base_ms or ms < (1.0 - MARGIN) * base_ms[0]:
      for k in ks:
        best[k] = t[k]
  nctx.set_tuning(best)
  return best, timings
