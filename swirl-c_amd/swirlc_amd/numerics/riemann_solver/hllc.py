"""HLLC numerical flux on the GPU (numerics/riemann_solver/hllc.py:33-347)."""

from swirlc_amd.numerics.riemann_solver import _native_flux


def hllc_convective_flux(replica_id, replicas, conservative_neg, conservative_pos,
                         primitive_neg, primitive_pos, helper_var_neg, helper_var_pos,
                         flux_fn, dim, cfg, physics_models=None):
  """Intercell flux of the Harten-Lax-van Leer-Contact approximation across
  face i - 1/2 (stored at i) from states already interpolated to that face.

  Same arguments as the reference.  p_dyn from the helper variables replaces
  p in the wave speeds, the contact speed and the star states
  (hllc.py:109-130); the contact speed is snapped to 0 below 1e-6; the
  branch select is that of hllc.py:227-243.  The physical flux is that of
  compressible_navier_stokes.py:338-427, evaluated by the kernel (`flux_fn`
  is not called); the gravitational potential comes from cfg.
  """
  del replica_id, replicas, flux_fn, physics_models
  return _native_flux.numerical_flux('HLLC', conservative_neg, conservative_pos, primitive_neg,
                                     primitive_pos, helper_var_neg, helper_var_pos, dim, cfg)
