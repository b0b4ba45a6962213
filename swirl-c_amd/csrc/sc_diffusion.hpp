// Stand-alone diffusive-term launchers (sc_diffusion.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sc_device.hpp"

namespace sc {

// tau[3 i + j] = 2 rho_f[j] nu S_ij on face j (cell u, v, w; face densities
// rho_f[j] of the j faces).  Full [z][x][y] arrays of extents N (x, y, z).
struct ShearArgs {
  const float* u;
  const float* v;
  const float* w;
  const float* rho_f[3];
  float* tau[9];
  int N[3];
  float h[3];
  float nu;
};

// flux[d] = -kappa_f dT/dx_d on face d, T = p / (rho R), kappa = (1/Pr) cp nu rho.
struct HeatArgs {
  const float* rho;
  const float* p;
  float* flux[3];
  int N[3];
  float h[3];
  float nu;
  float inv_pr;
};

hipError_t launch_shear_stress(const ShearArgs& a, hipStream_t stream);
hipError_t launch_heat_flux(const HeatArgs& a, hipStream_t stream);

}  // namespace sc
