// Stand-alone Riemann-solver launcher (sc_riemann.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sc_device.hpp"

namespace sc {

// Face states of face i-1/2 stored at i, full [z][x][y] arrays of extents N
// (x, y, z): conservative cl / cr (5 + NS fields), primitives wl / wr in the
// order rho u v w e p h, then the NS scalars; pdl / pdr the dynamic pressure
// (helper variable p_dyn) or null; pe tables as in the sweeps.
struct RiemannArgs {
  const float* cl[kMaxFields];
  const float* cr[kMaxFields];
  const float* wl[kMaxFields + 2];
  const float* wr[kMaxFields + 2];
  const float* pdl;
  const float* pdr;
  float* out[kMaxFields];
  long long total;
  int N[3];
  int has_pe;
  const float* pe_x;
  const float* pe_y;
  const float* pe_z;
};

// flux kHll / kHllc, dim 0 x, 1 y, 2 z, ns transported scalars (0..2).
hipError_t launch_riemann_flux(const RiemannArgs& a, int dim, int flux, int ns,
                               hipStream_t stream);

}  // namespace sc
