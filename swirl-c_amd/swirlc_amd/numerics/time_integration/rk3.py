"""SSP-RK3 time integration (numerics/time_integration/rk3.py:34-122).

u1 = u0 + dt*L(u0); u2 = 3/4 u0 + 1/4 (u1 + dt*L(u1));
u  = 1/3 u0 + 2/3 (u2 + dt*L(u2)).
With an RHS from get_rhs_fn each stage is one native call whose last sweep
applies the stage update (the axpy never makes its own HBM pass); any other
RHS callable is integrated with torch element-wise ops on its device.
"""

from swirlc_amd import _native as N
from swirlc_amd.numerics.time_integration import _stages

_RK3_COEFFS = {'c11': 0.75, 'c21': 1.0 / 3.0}
_RK3_COEFFS['c12'] = 1.0 - _RK3_COEFFS['c11']
_RK3_COEFFS['c22'] = 1.0 - _RK3_COEFFS['c21']


def integrate(replica_id, replicas, state_0, helper_vars, cfg, physics_models,
              rhs_fn, dt=None):
  """Advances state_0 by one step of cfg.dt (or `dt`, a float or device scalar)."""
  del replicas, physics_models
  dt = cfg.dt if dt is None else dt
  ctx = _stages.native_context(rhs_fn)
  if ctx is not None:
    _stages._check_helpers(ctx, helper_vars)
    u0 = ctx.fields(state_0)
    dtd = _stages.dt_tensor(dt, u0[0].device)
    s1 = _stages.empty_like_state(ctx, u0)
    ctx.rk_stage(N.SC_RK_EULER, 0.0, 0.0, dtd, None, u0, s1)
    s2 = _stages.empty_like_state(ctx, u0)
    ctx.rk_stage(N.SC_RK_BLEND, _RK3_COEFFS['c11'], _RK3_COEFFS['c12'], dtd, u0, s1, s2)
    s3 = _stages.empty_like_state(ctx, u0)
    ctx.rk_stage(N.SC_RK_BLEND, _RK3_COEFFS['c21'], _RK3_COEFFS['c22'], dtd, u0, s2, s3)
    return dict(zip(ctx.names, s3))

  def rk_fn(c0, f0, c1, f1, df1_dt):
    if f0 is None:
      return f1 + df1_dt * dt
    return f0 * c0 + (f1 + df1_dt * dt) * c1

  df_dt = rhs_fn(replica_id, state_0, helper_vars)
  state_1 = {k: rk_fn(None, None, None, v, df_dt[k]) for k, v in state_0.items()}
  df_dt = rhs_fn(replica_id, state_1, helper_vars)
  state_2 = {k: rk_fn(_RK3_COEFFS['c11'], v, _RK3_COEFFS['c12'], state_1[k], df_dt[k])
             for k, v in state_0.items()}
  df_dt = rhs_fn(replica_id, state_2, helper_vars)
  return {k: rk_fn(_RK3_COEFFS['c21'], v, _RK3_COEFFS['c22'], state_2[k], df_dt[k])
          for k, v in state_0.items()}
