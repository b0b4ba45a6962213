// Host-side launchers shared between the translation units of libswirlc_amd.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "sc_sweep_impl.hpp"

namespace sc {

// Launches a sweep kernel; with the tuning's profiling events set, the
// launch records its own start and stop times into them.
template <typename K>
inline void launch_sweep(K kernel, dim3 grid, dim3 block, hipStream_t stream,
                         const SweepTuning& tune, const SweepArgs& a) {
  if (tune.ev_start != nullptr && tune.ev_stop != nullptr) {
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, tune.ev_start, tune.ev_stop, 0, a);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, 0, stream, a);
  }
}

// Launches the x, z and y sweeps of one RHS evaluation (plus the RK update
// folded into the y sweep when a.rk_form != 0) for interpolation scheme SCHEME.
// only_dir = -1 launches all three; 0, 1 or 2 launches that sweep alone.
template <int SCHEME>
hipError_t launch_rhs_sweeps(const SweepArgs& a, const float h[3], int flux,
                             bool visc, int ns, const SweepTuning& tune,
                             hipStream_t stream, int only_dir = -1);

// Launches the HLLC branch-code kernel for faces normal to dim.
template <int SCHEME>
hipError_t launch_branch_codes(const SweepArgs& a, int dim, int ns,
                               uint8_t* codes, hipStream_t stream);

// Generic body, instantiated once per scheme in sc_sweep_<scheme>.hip.
// PD: reference-state (p_ref / rho_ref) build; BB: boundary face / flux
// value build.
template <int SCHEME, int FLUX, bool VISC, int NS, bool PD = false, bool BB = false,
          int LEAN = 0>
hipError_t launch_rhs_sweeps_t(SweepArgs a, const float h[3],
                               const SweepTuning& tune, hipStream_t stream,
                               int only_dir) {
  a.xcd_remap = tune.xcd_remap;
  // x sweep (first): R = src - div_x.
  if (only_dir < 0 || only_dir == 0) {
    SweepArgs ax = a;
    ax.inv_hdir = 1.0f / h[0];
    ax.seg_len = tune.seg_x;
    const dim3 grid((a.ny + 63) / 64, (a.nz + tune.by_x - 1) / tune.by_x,
                    (a.nx + tune.seg_x - 1) / tune.seg_x);
    launch_sweep(sweep_march_kernel<0, SCHEME, FLUX, VISC, NS, 0, PD, BB, LEAN>, grid,
                 dim3(64, tune.by_x), stream, tune, ax);
  }
  // y sweep (middle): R = R - div_y.
  if (only_dir < 0 || only_dir == 1) {
    SweepArgs ay = a;
    ay.inv_hdir = 1.0f / h[1];
    ay.seg_y = tune.seg_y;
    const int nwin = y_windows(a.nx, a.ny, a.h, YGeom<SCHEME>::OUTW);
    const dim3 grid((nwin + tune.by_y - 1) / tune.by_y, 1,
                    (a.nz + tune.seg_y - 1) / tune.seg_y);
    launch_sweep(sweep_y_kernel<SCHEME, FLUX, VISC, NS, PD, BB, LEAN>, grid,
                 dim3(64, tune.by_y), stream, tune, ay);
  }
  // z sweep (last): R - div_z, then store R or the RK update.
  if (only_dir < 0 || only_dir == 2) {
    SweepArgs az = a;
    az.inv_hdir = 1.0f / h[2];
    az.seg_len = tune.seg_z;
    const dim3 grid((a.ny + 63) / 64, (a.nx + tune.by_z - 1) / tune.by_z,
                    (a.nz + tune.seg_z - 1) / tune.seg_z);
    launch_sweep(sweep_march_kernel<2, SCHEME, FLUX, VISC, NS, 2, PD, BB, LEAN>, grid,
                 dim3(64, tune.by_z), stream, tune, az);
  }
  return hipGetLastError();
}

// Picks the build for the configuration: boundary values (BB) and / or a
// reference state (PD); without either and without primitive-variable BCs,
// a LEAN build: kLeanPlain without gravity, kLeanGrav with both the
// potential and the gravity source.
template <int SCHEME, int FLUX, bool VISC, int NS>
hipError_t launch_rhs_ns(const SweepArgs& a, const float h[3], const SweepTuning& tune,
                         hipStream_t stream, int od) {
  const bool bb = a.bbc_dims != 0;
  if (!bb && a.p_ref == nullptr && a.prim_bc_mask == 0) {
    if (!a.has_pe && !a.grav_src)
      return launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, false, false, kLeanPlain>(a, h, tune,
                                                                                  stream, od);
    if (a.has_pe && a.grav_src)
      return launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, false, false, kLeanGrav>(a, h, tune,
                                                                                 stream, od);
  }
  if (a.p_ref != nullptr)
    return bb ? launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, true, true>(a, h, tune, stream, od)
              : launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, true, false>(a, h, tune, stream, od);
  return bb ? launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, false, true>(a, h, tune, stream, od)
            : launch_rhs_sweeps_t<SCHEME, FLUX, VISC, NS, false, false>(a, h, tune, stream, od);
}

template <int SCHEME, int FLUX, bool VISC>
hipError_t launch_rhs_variant(const SweepArgs& a, const float h[3], int ns,
                              const SweepTuning& tune, hipStream_t stream, int od) {
  if (ns == 0) return launch_rhs_ns<SCHEME, FLUX, VISC, 0>(a, h, tune, stream, od);
  if (ns == 1) return launch_rhs_ns<SCHEME, FLUX, VISC, 1>(a, h, tune, stream, od);
  return launch_rhs_ns<SCHEME, FLUX, VISC, 2>(a, h, tune, stream, od);
}

template <int SCHEME, int NS, bool PD = false>
hipError_t launch_branch_codes_t(const SweepArgs& a, int dim, uint8_t* codes,
                                 hipStream_t stream) {
  const long long nf = (long long)(a.nx + (dim == 0)) * (a.ny + (dim == 1)) *
                       (a.nz + (dim == 2));
  const int blk = 256;
  const unsigned grid = (unsigned)((nf + blk - 1) / blk);
  if (dim == 0) {
    hipLaunchKernelGGL((branch_codes_kernel<0, SCHEME, NS, PD>), grid, blk, 0, stream, a, codes);
  } else if (dim == 1) {
    hipLaunchKernelGGL((branch_codes_kernel<1, SCHEME, NS, PD>), grid, blk, 0, stream, a, codes);
  } else {
    hipLaunchKernelGGL((branch_codes_kernel<2, SCHEME, NS, PD>), grid, blk, 0, stream, a, codes);
  }
  return hipGetLastError();
}

#define SC_INSTANTIATE_SCHEME(SCHEME)                                            \
  template <>                                                                    \
  hipError_t launch_rhs_sweeps<SCHEME>(const SweepArgs& a, const float h[3],     \
                                       int flux, bool visc, int ns,              \
                                       const SweepTuning& tune,                  \
                                       hipStream_t stream, int od) {             \
    if (flux == kHllc)                                                           \
      return visc ? launch_rhs_variant<SCHEME, kHllc, true>(a, h, ns, tune, stream, od)    \
                  : launch_rhs_variant<SCHEME, kHllc, false>(a, h, ns, tune, stream, od);  \
    return visc ? launch_rhs_variant<SCHEME, kHll, true>(a, h, ns, tune, stream, od)        \
                : launch_rhs_variant<SCHEME, kHll, false>(a, h, ns, tune, stream, od);      \
  }                                                                              \
  template <>                                                                    \
  hipError_t launch_branch_codes<SCHEME>(const SweepArgs& a, int dim, int ns,    \
                                         uint8_t* codes, hipStream_t stream) {   \
    const bool pd = a.p_ref != nullptr;                                          \
    if (ns == 0)                                                                 \
      return pd ? launch_branch_codes_t<SCHEME, 0, true>(a, dim, codes, stream)  \
                : launch_branch_codes_t<SCHEME, 0>(a, dim, codes, stream);       \
    if (ns == 1)                                                                 \
      return pd ? launch_branch_codes_t<SCHEME, 1, true>(a, dim, codes, stream)  \
                : launch_branch_codes_t<SCHEME, 1>(a, dim, codes, stream);       \
    return pd ? launch_branch_codes_t<SCHEME, 2, true>(a, dim, codes, stream)    \
              : launch_branch_codes_t<SCHEME, 2>(a, dim, codes, stream);         \
  }

}  // namespace sc
