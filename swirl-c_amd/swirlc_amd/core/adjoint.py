"""Reverse-mode differentiation of RHS evaluations and SSP-RK3 rollouts.

The reference obtains gradients by recording its TensorFlow graph on a
tf.GradientTape.  Here the hand-written adjoint kernels of libswirlc_amd
(sc_rhs_vjp, sc_rk_stage_vjp) are wrapped as torch.autograd.Functions, so a
loss of the final state can be differentiated with `loss.backward()`:

  rhs(nctx, *q)                          -> RHS fields           (RhsFunction)
  rollout(nctx, dt, steps, *u, every=k)  -> state after `steps`  (RolloutFunction)

`rollout` keeps the state every k steps (checkpoints) in its forward pass;
its backward pass recomputes each k-step segment from the segment's
checkpoint, storing the three stage states of every step of the segment, and
then runs the stage adjoints in reverse.  Peak memory is about
(steps / k + 3 k) states instead of 3 * steps; the recomputation costs one
forward rollout.  With k = 0 the forward pass keeps the three stage states
of every step in one device allocation (3 * steps states: 29 GB for 200
steps of a 128^3 grid with halos) and the backward pass recomputes
nothing.  dt is a constant of the rollout (no derivative through a
CFL time step), as cfg.dt is in the reference.

Fields are full-size [nz+2h, nx+2h, ny+2h] fp32 GPU tensors; gradients are
returned for their interior cells (halo gradients are zero, since the halo
fill overwrites them).
"""

import torch

from swirlc_amd import _native as N

_C11, _C21 = 0.75, 1.0 / 3.0


def _dt_tensor(dt, device):
  if isinstance(dt, torch.Tensor):
    if dt.numel() != 1 or dt.dtype != torch.float32 or dt.device != device:
      raise ValueError('dt must be a one-element float32 tensor on the state device')
    return dt.reshape(1).contiguous()
  return torch.full((1,), float(dt), dtype=torch.float32, device=device)


def rk3_forward(nctx, dt_dev, u, s1, s2, out):
  """One SSP-RK3 step u -> out (rk3.py:87-119) with the native fused stages.
  out may alias u; s1, s2 receive the stage states."""
  nctx.rk_stage(N.SC_RK_EULER, 0.0, 0.0, dt_dev, None, u, s1)
  nctx.rk_stage(N.SC_RK_BLEND, _C11, 1.0 - _C11, dt_dev, u, s1, s2)
  nctx.rk_stage(N.SC_RK_BLEND, _C21, 1.0 - _C21, dt_dev, u, s2, out)


def rk3_backward(nctx, dt_dev, u, s1, s2, lam, lam0):
  """Reverse of rk3_forward: lam (adjoint of the step's output) is replaced,
  in place, by the adjoint of u.  lam0 is a work buffer."""
  for t in lam0:
    t.zero_()
  nctx.rk_stage_vjp(N.SC_RK_BLEND, _C21, 1.0 - _C21, dt_dev, s2, lam, lam, lam0)
  nctx.rk_stage_vjp(N.SC_RK_BLEND, _C11, 1.0 - _C11, dt_dev, s1, lam, lam, lam0)
  nctx.rk_stage_vjp(N.SC_RK_EULER, 0.0, 0.0, dt_dev, u, lam, lam)
  for a, b in zip(lam, lam0):
    a.add_(b)


def _grads(nctx, grads, like):
  out = []
  for g, t in zip(grads, like):
    out.append(torch.zeros_like(t) if g is None else g.detach().contiguous().clone())
  return out


class RhsFunction(torch.autograd.Function):
  """rhs(q) with the adjoint kernel as its backward (fills q's halos)."""

  @staticmethod
  def forward(ctx, nctx, *q):
    q = [t.detach() for t in q]
    r = [torch.zeros_like(t) for t in q]
    nctx.rhs(q, r)
    ctx.nctx = nctx
    ctx.save_for_backward(*q)
    return tuple(r)

  @staticmethod
  def backward(ctx, *gr):
    q = list(ctx.saved_tensors)
    g = _grads(ctx.nctx, gr, q)
    lq = [torch.zeros_like(t) for t in q]
    ctx.nctx.rhs_vjp(q, g, lq)
    return (None, *lq)


class RolloutFunction(torch.autograd.Function):
  """`steps` SSP-RK3 steps of fixed dt with checkpointed reverse mode."""

  @staticmethod
  def forward(ctx, nctx, dt_dev, steps, every, *u0):
    if steps < 1 or every < 0:
      raise ValueError('steps must be positive and the checkpoint interval non-negative')
    ctx.nctx = nctx
    ctx.dt = dt_dev
    ctx.steps = steps
    ctx.every = every
    if every == 0:
      return _tape_forward(ctx, nctx, dt_dev, steps, u0)
    u = [t.detach().clone() for t in u0]
    s1 = [torch.empty_like(t) for t in u]
    s2 = [torch.empty_like(t) for t in u]
    ckpts = {0: [t.clone() for t in u]}
    for n in range(steps):
      rk3_forward(nctx, dt_dev, u, s1, s2, u)
      if (n + 1) % every == 0 and n + 1 < steps:
        ckpts[n + 1] = [t.clone() for t in u]
    ctx.ckpts = ckpts
    return tuple(u)

  @staticmethod
  def backward(ctx, *grads):
    nctx, dt_dev, steps, every = ctx.nctx, ctx.dt, ctx.steps, ctx.every
    if every == 0:
      return _tape_backward(ctx, grads)
    base = ctx.ckpts[0]
    lam = _grads(nctx, grads, base)
    lam0 = [torch.zeros_like(t) for t in base]
    seg_cap = min(every, steps)
    pool = [[[torch.empty_like(t) for t in base] for _ in range(3)] for _ in range(seg_cap)]
    for start in reversed(range(0, steps, every)):
      seg = min(every, steps - start)
      # Recompute the segment: states u_n and the stage states of each step.
      for m in range(seg):
        un, a, b = pool[m]
        if m == 0:
          for d, s in zip(un, ctx.ckpts[start]):
            d.copy_(s)
        nctx.rk_stage(N.SC_RK_EULER, 0.0, 0.0, dt_dev, None, un, a)
        nctx.rk_stage(N.SC_RK_BLEND, _C11, 1.0 - _C11, dt_dev, un, a, b)
        if m + 1 < seg:
          nctx.rk_stage(N.SC_RK_BLEND, _C21, 1.0 - _C21, dt_dev, un, b, pool[m + 1][0])
      for m in reversed(range(seg)):
        un, a, b = pool[m]
        rk3_backward(nctx, dt_dev, un, a, b, lam, lam0)
    ctx.ckpts = None
    return (None, None, None, None, *lam)


class _Tape:
  """Stage states u_n, s1, s2 of every step of a rollout, in one device
  allocation; each field starts on a 256-byte boundary."""

  def __init__(self, steps, like):
    self.nf = len(like)
    self.shape = like[0].shape
    self.numel = like[0].numel()
    self.stride = (self.numel + 63) // 64 * 64
    dev = like[0].device
    need = 3 * steps * self.nf * self.stride * 4
    free, _ = torch.cuda.mem_get_info(dev)
    if need > free:
      raise MemoryError(f'rollout with every=0 keeps {need / 2**30:.1f} GiB of stage states;'
                        f' {free / 2**30:.1f} GiB are free (use checkpoints, every > 0)')
    self.flat = torch.empty(3 * steps * self.nf * self.stride, dtype=torch.float32, device=dev)

  def state(self, n, m):
    """Fields of state m (0 u_n, 1 s1, 2 s2) of step n."""
    out = []
    for f in range(self.nf):
      o = ((n * 3 + m) * self.nf + f) * self.stride
      out.append(self.flat[o:o + self.numel].view(self.shape))
    return out


def _tape_forward(ctx, nctx, dt_dev, steps, u0):
  """Forward rollout keeping u_n, s1, s2 of every step."""
  tape = _Tape(steps, u0)
  for d, t in zip(tape.state(0, 0), u0):
    d.copy_(t.detach())
  out = [torch.empty_like(t) for t in u0]
  for n in range(steps):
    nxt = tape.state(n + 1, 0) if n + 1 < steps else out
    rk3_forward(nctx, dt_dev, tape.state(n, 0), tape.state(n, 1), tape.state(n, 2), nxt)
  ctx.tape = tape
  return tuple(out)


def _tape_backward(ctx, grads):
  nctx, dt_dev, tape = ctx.nctx, ctx.dt, ctx.tape
  base = tape.state(0, 0)
  lam = _grads(nctx, grads, base)
  lam0 = [torch.zeros_like(t) for t in base]
  for n in reversed(range(ctx.steps)):
    rk3_backward(nctx, dt_dev, tape.state(n, 0), tape.state(n, 1), tape.state(n, 2), lam,
                 lam0)
  ctx.tape = None
  return (None, None, None, None, *lam)


def rhs(nctx, *q):
  """Differentiable RHS of the conserved fields (list order of nctx.names)."""
  return RhsFunction.apply(nctx, *q)


def rollout(nctx, dt, steps, *u0, every=20):
  """Differentiable `steps`-step SSP-RK3 rollout with fixed dt, checkpointing
  the state every `every` steps (0: keep every stage state, no
  recomputation)."""
  dt_dev = _dt_tensor(dt, u0[0].device)
  return RolloutFunction.apply(nctx, dt_dev, int(steps), int(every), *u0)


def kinetic_density_loss(u, rho_target, h, weight=0.1):
  """J = sum 0.5 rho |u|^2 + weight * sum (rho - rho_target)^2 over the
  interior cells of the final state u = (rho, rho_u, rho_v, rho_w, rho_e)."""
  sl = (slice(h, -h),) * 3
  rho = u[0][sl]
  ke = 0.5 * (u[1][sl] ** 2 + u[2][sl] ** 2 + u[3][sl] ** 2) / rho
  mis = rho - rho_target[sl]
  return ke.sum(dtype=torch.float64) + weight * (mis * mis).sum(dtype=torch.float64)
