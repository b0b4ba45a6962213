// Stand-alone numerical flux across face i-1/2 (stored at i) from given face
// states: numerics/riemann_solver/hll.py:36-175 (hll_convective_flux) and
// hllc.py:33-257 (hllc_convective_flux) with the physical flux of
// equations/compressible_navier_stokes.py:338-427, for callers that use the
// Riemann solvers outside the fused sweeps.  One thread per array element;
// the wave speeds and HLLC star states are the sweeps' device functions.
#include <hip/hip_runtime.h>

#include "sc_riemann.hpp"

namespace sc {
namespace {

// _compute_physical_flux_1d of one side: m = the given dim-aligned face
// momentum (conservative state), aligned momentum m u_n + p_upd (p_dyn if
// present), energy m (e + p / rho) with the face p, others m * phi.
template <int D, int NS>
__device__ __forceinline__ void given_physical_flux(float m, const float* w, float p_upd,
                                                    float* f) {
  f[0] = m;
#pragma unroll
  for (int d = 0; d < 3; ++d) f[1 + d] = d == D ? m * w[kU + d] + p_upd : m * w[kU + d];
  f[4] = m * (w[kE] + w[kP] * frcp(w[kRho]));
#pragma unroll
  for (int s = 0; s < NS; ++s) f[5 + s] = m * w[7 + s];
}

template <int D, int FLUX, int NS, bool PD>
__global__ void __launch_bounds__(256) riemann_face_kernel(const RiemannArgs a) {
  constexpr int NF = 5 + NS;
  constexpr int NV = NVar<NS>::value;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.total) return;
  const int j = (int)(idx % a.N[1]);
  const int i = (int)((idx / a.N[1]) % a.N[0]);
  const int k = (int)(idx / ((long long)a.N[1] * a.N[0]));
  const float pe = a.has_pe ? (a.pe_x[i] - a.pe_y[j]) - a.pe_z[k] : 0.0f;
  float L[NV], R[NV], cl[NF], cr[NF];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    L[v] = a.wl[v][idx];
    R[v] = a.wr[v][idx];
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    cl[f] = a.cl[f][idx];
    cr[f] = a.cr[f][idx];
  }
  const float pdl = PD ? a.pdl[idx] : L[kP];
  const float pdr = PD ? a.pdr[idx] : R[kP];
  float fl[NF], fr[NF], F[NF];
  given_physical_flux<D, NS>(cl[1 + D], L, pdl, fl);
  given_physical_flux<D, NS>(cr[1 + D], R, pdr, fr);
  if constexpr (FLUX == kHll) {
    // Wave speeds from the primitives as given (hll.py:107-115).
    float sl, sr;
    roe_speeds<D>(L, R, pe, sl, sr);
    const float ds = sr - sl;
    const float rds = ds == 0.0f ? 0.0f : frcp(ds);
    const float t1 = (tf_min(sr, 0.0f) - tf_min(sl, 0.0f)) * rds;
    const float t2 = 1.0f - t1;
    const float t3 = (sr * fabsf(sl) - sl * fabsf(sr)) * (0.5f * rds);
#pragma unroll
    for (int f = 0; f < NF; ++f) F[f] = (t1 * fr[f] + t2 * fl[f]) - t3 * (cr[f] - cl[f]);
  } else {
    // Wave speeds, contact speed and star states see p_dyn (hllc.py:109-130).
    L[kP] = pdl;
    R[kP] = pdr;
    float sl, sr;
    roe_speeds<D>(L, R, pe, sl, sr);
    const float vl = L[kU + D], vr = R[kU + D];
    const float rl = L[kRho], rr = R[kRho];
    float sm = dnn(((R[kP] - L[kP]) + (rl * vl) * (sl - vl)) - (rr * vr) * (sr - vr),
                   rl * (sl - vl) - rr * (sr - vr));
    sm = fabsf(sm) < kSmSnap ? 0.0f : sm;
    if (sl >= 0.0f) {
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] = fl[f];
    } else if (sl <= 0.0f && sm >= 0.0f) {
      hllc_star<D, NS>(L, frcp(L[kRho]), sl, sm, fl, F);
    } else if (sm <= 0.0f && sr >= 0.0f) {
      hllc_star<D, NS>(R, frcp(R[kRho]), sr, sm, fr, F);
    } else {
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] = fr[f];
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) a.out[f][idx] = F[f];
}

template <int D, int FLUX, int NS>
hipError_t launch_t(const RiemannArgs& a, hipStream_t stream) {
  const dim3 grid((unsigned)((a.total + 255) / 256));
  if (a.pdl != nullptr)
    hipLaunchKernelGGL((riemann_face_kernel<D, FLUX, NS, true>), grid, 256, 0, stream, a);
  else
    hipLaunchKernelGGL((riemann_face_kernel<D, FLUX, NS, false>), grid, 256, 0, stream, a);
  return hipGetLastError();
}

template <int FLUX, int NS>
hipError_t launch_dim(const RiemannArgs& a, int dim, hipStream_t stream) {
  if (dim == 0) return launch_t<0, FLUX, NS>(a, stream);
  if (dim == 1) return launch_t<1, FLUX, NS>(a, stream);
  return launch_t<2, FLUX, NS>(a, stream);
}

template <int FLUX>
hipError_t launch_ns(const RiemannArgs& a, int dim, int ns, hipStream_t stream) {
  if (ns == 0) return launch_dim<FLUX, 0>(a, dim, stream);
  if (ns == 1) return launch_dim<FLUX, 1>(a, dim, stream);
  return launch_dim<FLUX, 2>(a, dim, stream);
}

}  // namespace

hipError_t launch_riemann_flux(const RiemannArgs& a, int dim, int flux, int ns,
                               hipStream_t stream) {
  return flux == kHllc ? launch_ns<kHllc>(a, dim, ns, stream)
                       : launch_ns<kHll>(a, dim, ns, stream);
}

}  // namespace sc
