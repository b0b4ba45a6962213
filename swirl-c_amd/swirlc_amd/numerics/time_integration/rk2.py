"""RK2 time integration (numerics/time_integration/rk2.py:30-86).

u1 = u0 + dt*L(u0); u2 = u1 + dt*L(u1); u = 0.5*(u0 + u2).
"""



def integrate(replica_id, replicas, state_0, helper_vars, cfg, physics_models,
              rhs_fn, dt=None):
  del replicas, physics_models
  dt = cfg.dt if dt is None else dt
  # The native RHS (sc_rhs) is evaluated twice; the stage updates are
  # element-wise tensor ops on the same device.
  step_fn = lambda u, f: u + dt * f
  df_dt = rhs_fn(replica_id, state_0, helper_vars)
  state_1 = {k: step_fn(v, df_dt[k]) for k, v in state_0.items()}
  df_dt = rhs_fn(replica_id, state_1, helper_vars)
  state_2 = {k: step_fn(v, df_dt[k]) for k, v in state_1.items()}
  return {k: 0.5 * (v + state_2[k]) for k, v in state_0.items()}
