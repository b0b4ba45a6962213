// Stand-alone face interpolation of one field (the reference's
// numerics/interpolation.py:21-53 linear_interpolation and the swirl_lm
// weno(v, dim, k) / flux_limiter(v, dim, 'MUSCL') contract it calls from
// equations/compressible_navier_stokes.py:536-615), for callers that use
// these building blocks outside the fused sweeps.  One thread per cell of
// the full [z][x][y] array; the per-cell arithmetic is the sweeps' recon<>.
#include <hip/hip_runtime.h>

#include "sc_interp.hpp"

namespace sc {
namespace {

struct LineGeom {
  long long total;  // cells of the array
  long long s;      // element stride along dim
  int n;            // extent along dim
};

__device__ __forceinline__ int coord_along(const LineGeom& g, long long idx) {
  return (int)((idx / g.s) % g.n);
}

// (neg, pos) = left / right states of face i-1/2 stored at index i.  WENO
// (RAD = k - 1): cells RAD .. n-1-RAD reconstruct (their face i+1/2 state
// lands at i+1, their face i-1/2 state at i); every other entry keeps the
// cell value, and nothing is reconstructed when n <= 2 RAD.  MUSCL: the
// minmod slope is zero at both ends of the line, so the end cells give the
// cell value for both faces; neg[0] is the cell value.
template <int SCHEME>
__global__ void __launch_bounds__(256)
face_interp_kernel(const float* __restrict__ v, LineGeom g, float* __restrict__ neg,
                   float* __restrict__ pos) {
  constexpr int RAD = Radius<SCHEME>::value;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.total) return;
  const int c = coord_along(g, idx);
  const int n = g.n;
  const float vc = v[idx];
  if constexpr (SCHEME == kMuscl) {
    float fp = vc, fm = vc;
    if (c >= 1 && c < n - 1) {
      float a[3] = {v[idx - g.s], vc, v[idx + g.s]};
      recon<kMuscl>(a, fp, fm);
    }
    pos[idx] = fm;
    if (c + 1 < n) neg[idx + g.s] = fp;
    if (c == 0) neg[idx] = vc;
  } else {
    const bool active = n > 2 * RAD && c >= RAD && c < n - RAD;
    if (active) {
      float a[2 * RAD + 1];
#pragma unroll
      for (int w = 0; w < 2 * RAD + 1; ++w) a[w] = v[idx + (long long)(w - RAD) * g.s];
      float fp, fm;
      recon<SCHEME>(a, fp, fm);
      pos[idx] = fm;
      neg[idx + g.s] = fp;
    } else {
      pos[idx] = vc;
    }
    // neg entries not written by the cell below: c <= RAD, or c > n - RAD.
    if (n <= 2 * RAD || c <= RAD || c > n - RAD) neg[idx] = vc;
  }
}

// numerics/interpolation.py:21-53: 0.5 * (f[i-1] + f[i]) on face i-1/2,
// with f[-1] = 0 (the kernel op's edge value).
__global__ void __launch_bounds__(256)
linear_interp_kernel(const float* __restrict__ v, LineGeom g, float* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.total) return;
  const float prev = coord_along(g, idx) > 0 ? v[idx - g.s] : 0.0f;
  out[idx] = 0.5f * (prev + v[idx]);
}

// f = value on the plane coordinate[dim] == index (one thread per plane cell).
__global__ void __launch_bounds__(256)
fill_plane_kernel(float* f, int nx, int ny, int nz, int dim, int index, float value,
                  const float* values) {
  const int na = dim == 0 ? nz : (dim == 1 ? nz : nx);  // slow plane axis
  const int nb = dim == 1 ? nx : ny;                     // fast plane axis
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)na * nb) return;
  const int ia = (int)(t / nb), ib = (int)(t % nb);
  int x, y, z;
  if (dim == 0) {
    x = index; z = ia; y = ib;
  } else if (dim == 1) {
    y = index; z = ia; x = ib;
  } else {
    z = index; x = ia; y = ib;
  }
  // values: one per plane cell, [ia][ib] (null: the scalar value everywhere).
  f[((long long)z * nx + x) * ny + y] = values != nullptr ? values[t] : value;
}

// Whole-array steps of the face BC update (boundary.py:230-337), IEEE fp32:
// op 0 out = b == 0 ? 0 : a / b (divide_no_nan), 1 out = a * b,
// 2 out = (1 / b) * a (conservative_to_primitive's reciprocal, then product).
__global__ void __launch_bounds__(256)
elementwise_kernel(int op, const float* a, const float* b, float* out, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float x = a[t], y = b[t];
  float r;
  if (op == 0) {
    r = y == 0.0f ? 0.0f : x / y;
  } else if (op == 1) {
    r = x * y;
  } else {
    r = (1.0f / y) * x;
  }
  out[t] = r;
}

}  // namespace

hipError_t launch_elementwise(int op, const float* a, const float* b, float* out, long long n,
                              hipStream_t stream) {
  hipLaunchKernelGGL(elementwise_kernel, dim3((unsigned)((n + 255) / 256)), 256, 0, stream, op,
                     a, b, out, n);
  return hipGetLastError();
}

hipError_t launch_fill_plane(float* f, const int N[3], int dim, int index, float value,
                             const float* values, hipStream_t stream) {
  const int na = dim == 2 ? N[0] : N[2];
  const int nb = dim == 1 ? N[0] : N[1];
  const long long n = (long long)na * nb;
  hipLaunchKernelGGL(fill_plane_kernel, dim3((unsigned)((n + 255) / 256)), 256, 0, stream, f,
                     N[0], N[1], N[2], dim, index, value, values);
  return hipGetLastError();
}

hipError_t launch_face_interp(const float* v, const int N[3], int dim, int scheme, float* neg,
                              float* pos, hipStream_t stream) {
  LineGeom g;
  g.total = (long long)N[0] * N[1] * N[2];
  g.s = dim == 1 ? 1 : (dim == 0 ? (long long)N[1] : (long long)N[0] * N[1]);
  g.n = N[dim];
  const dim3 grid((unsigned)((g.total + 255) / 256));
  if (scheme == kWeno5)
    hipLaunchKernelGGL(face_interp_kernel<kWeno5>, grid, 256, 0, stream, v, g, neg, pos);
  else if (scheme == kWeno3)
    hipLaunchKernelGGL(face_interp_kernel<kWeno3>, grid, 256, 0, stream, v, g, neg, pos);
  else
    hipLaunchKernelGGL(face_interp_kernel<kMuscl>, grid, 256, 0, stream, v, g, neg, pos);
  return hipGetLastError();
}

hipError_t launch_linear_interp(const float* v, const int N[3], int dim, float* out,
                                hipStream_t stream) {
  LineGeom g;
  g.total = (long long)N[0] * N[1] * N[2];
  g.s = dim == 1 ? 1 : (dim == 0 ? (long long)N[1] : (long long)N[0] * N[1]);
  g.n = N[dim];
  hipLaunchKernelGGL(linear_interp_kernel, dim3((unsigned)((g.total + 255) / 256)), 256, 0,
                     stream, v, g, out);
  return hipGetLastError();
}

}  // namespace sc
