"""swirlc_amd: MI355X-native compressible Navier-Stokes RHS and Runge-Kutta step.

Drop-in for Swirl-C's hot path (equations/compressible_navier_stokes.py and
numerics/time_integration/ of google-research/swirl-c).  The module tree
mirrors the reference's (`swirlc_amd.equations.compressible_navier_stokes`,
`swirlc_amd.numerics.time_integration.rk3`, ...); the compute runs in
hand-written HIP kernels for gfx950 behind the C ABI of include/swirlc.h.
"""

__version__ = '0.1.0'
