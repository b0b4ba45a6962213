// Stand-alone diffusive terms on the faces (physics/diffusion.py:33-109 with
// physics/fluid.py:31-111 strain rate, physics/transport/simple.py:44-108
// transport, numerics/gradient.py and interpolation.py stencils), for callers
// that use them outside the fused sweeps.  One thread per cell of the full
// [z][x][y] array.  Neighbours outside the array read 0 (the kernel ops'
// edge values, which only reach halo cells); operations are rounded one by
// one in the reference's order (contraction off, IEEE division).
#include <hip/hip_runtime.h>

#include "sc_diffusion.hpp"

namespace sc {
namespace {

struct Grid3 {
  int N[3];           // extents along x, y, z
  long long st[3];    // element strides along x, y, z
  long long total;
};

__device__ __forceinline__ void coords(const Grid3& g, long long idx, int c[3]) {
  c[1] = (int)(idx % g.N[1]);
  c[0] = (int)((idx / g.N[1]) % g.N[0]);
  c[2] = (int)(idx / ((long long)g.N[1] * g.N[0]));
}

// f at cell c + o (o per dim), 0 outside the array.
__device__ __forceinline__ float at(const float* f, const Grid3& g, const int c[3], int ox,
                                    int oy, int oz) {
  const int x = c[0] + ox, y = c[1] + oy, z = c[2] + oz;
  if (x < 0 || x >= g.N[0] || y < 0 || y >= g.N[1] || z < 0 || z >= g.N[2]) return 0.0f;
  return f[(long long)z * g.st[2] + (long long)x * g.st[0] + y];
}

__device__ __forceinline__ bool inside(const Grid3& g, const int c[3], int ox, int oy, int oz) {
  const int x = c[0] + ox, y = c[1] + oy, z = c[2] + oz;
  return x >= 0 && x < g.N[0] && y >= 0 && y < g.N[1] && z >= 0 && z < g.N[2];
}

// physics/fluid.py:52-63 grad(f, gd, fd) on face fd of cell c: backward
// difference when gd == fd, else the central difference along gd linearly
// interpolated onto face fd.
__device__ __forceinline__ float grad(const float* f, const Grid3& g, const int c[3], int gd,
                                      int fd, const float h[3]) {
#pragma clang fp contract(off)
  int e[3] = {0, 0, 0};
  e[gd] = 1;
  if (gd == fd) return (at(f, g, c, 0, 0, 0) - at(f, g, c, -e[0], -e[1], -e[2])) / h[gd];
  const float h2 = 2.0f * h[gd];
  int b[3] = {0, 0, 0};
  b[fd] = -1;
  // central_2 at c and at c - e_fd (0 where that cell is outside the array)
  const float c0 = (at(f, g, c, e[0], e[1], e[2]) - at(f, g, c, -e[0], -e[1], -e[2])) / h2;
  float cm = 0.0f;
  if (inside(g, c, b[0], b[1], b[2]))
    cm = (at(f, g, c, b[0] + e[0], b[1] + e[1], b[2] + e[2]) -
          at(f, g, c, b[0] - e[0], b[1] - e[1], b[2] - e[2])) / h2;
  return 0.5f * (cm + c0);
}

__global__ void __launch_bounds__(256) shear_stress_kernel(const ShearArgs a) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid3 g{{a.N[0], a.N[1], a.N[2]},
                {(long long)a.N[1], 1, (long long)a.N[0] * a.N[1]},
                (long long)a.N[0] * a.N[1] * a.N[2]};
  if (idx >= g.total) return;
  int c[3];
  coords(g, idx, c);
  const float* u[3] = {a.u, a.v, a.w};
  const float third = (float)(1.0 / 3.0);
  float div[3];
#pragma unroll
  for (int fd = 0; fd < 3; ++fd)
    div[fd] = (grad(u[0], g, c, 0, fd, a.h) + grad(u[1], g, c, 1, fd, a.h)) +
              grad(u[2], g, c, 2, fd, a.h);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float s;
      if (i == j) {
        s = grad(u[i], g, c, i, i, a.h) - third * div[i];
      } else {
        s = 0.5f * (grad(u[i], g, c, j, j, a.h) + grad(u[j], g, c, i, j, a.h));
      }
      a.tau[3 * i + j][idx] = ((2.0f * a.rho_f[j][idx]) * a.nu) * s;
    }
  }
}

__global__ void __launch_bounds__(256) heat_flux_kernel(const HeatArgs a) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Grid3 g{{a.N[0], a.N[1], a.N[2]},
                {(long long)a.N[1], 1, (long long)a.N[0] * a.N[1]},
                (long long)a.N[0] * a.N[1] * a.N[2]};
  if (idx >= g.total) return;
  int c[3];
  coords(g, idx, c);
  auto temp = [&](int ox, int oy, int oz) {
    if (!inside(g, c, ox, oy, oz)) return 0.0f;
    return at(a.p, g, c, ox, oy, oz) / (at(a.rho, g, c, ox, oy, oz) * kR);
  };
  auto kappa = [&](int ox, int oy, int oz) {
    if (!inside(g, c, ox, oy, oz)) return 0.0f;
    return (a.inv_pr * kCp) * (a.nu * at(a.rho, g, c, ox, oy, oz));
  };
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int b[3] = {0, 0, 0};
    b[d] = -1;
    const float kf = 0.5f * (kappa(b[0], b[1], b[2]) + kappa(0, 0, 0));
    const float dt = (temp(0, 0, 0) - temp(b[0], b[1], b[2])) / a.h[d];
    a.flux[d][idx] = -kf * dt;
  }
}

}  // namespace

hipError_t launch_shear_stress(const ShearArgs& a, hipStream_t stream) {
  const long long total = (long long)a.N[0] * a.N[1] * a.N[2];
  hipLaunchKernelGGL(shear_stress_kernel, dim3((unsigned)((total + 255) / 256)), 256, 0, stream,
                     a);
  return hipGetLastError();
}

hipError_t launch_heat_flux(const HeatArgs& a, hipStream_t stream) {
  const long long total = (long long)a.N[0] * a.N[1] * a.N[2];
  hipLaunchKernelGGL(heat_flux_kernel, dim3((unsigned)((total + 255) / 256)), 256, 0, stream, a);
  return hipGetLastError();
}

}  // namespace sc
