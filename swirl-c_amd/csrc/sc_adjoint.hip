// Non-templated kernels of the RHS adjoint: clearing the accumulators, the
// viscous transverse-derivative gather, the transposed halo fill and the
// Runge-Kutta stage combination.
#include "sc_adjoint_impl.hpp"
#include "sc_halo.hpp"

namespace sc {

namespace {

constexpr int kZeroMax = 16;
struct PtrList {
  float* p[kZeroMax];
};

__global__ void __launch_bounds__(256) adj_zero_kernel(PtrList l, int np, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    for (int k = 0; k < np; ++k) l.p[k][t] = 0.0f;
  }
}

// u_i enters the central derivative d(u_i)/dx_k of the cells X -+ e_k with
// weight +-1/(2 h_k):  lu_i[X] = sum_k (lG_ik[X - e_k] - lG_ik[X + e_k]) / (2 h_k).
// Cells X cover [h-1, h+n] in every dim (transverse neighbours of the cells
// next to a face).  u_i = (1/rho) * m_i, so lm_i += lu_i / rho and
// lrho -= sum_i lu_i u_i / rho.
__global__ void __launch_bounds__(256) adj_visc_gather_kernel(const AdjArgs A) {
  const SweepArgs& a = A.s;
  const long long ex = a.nx + 2, ey = a.ny + 2, ez = a.nz + 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ex * ey * ez) return;
  const int j = a.h - 1 + (int)(t % ey);
  const int i = a.h - 1 + (int)((t / ey) % ex);
  const int k = a.h - 1 + (int)(t / (ey * ex));
  const long long X = (long long)k * a.sz + (long long)i * a.sx + j;
  const long long st[3] = {a.sx, 1, a.sz};
  float lu[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float* g = A.lG[gidx(c, d)];
      s += (g[X - st[d]] - g[X + st[d]]) * a.vc.inv_two_h[d];
    }
    lu[c] = s;
  }
  const float rho = a.q.p[0][X];
  const float inv = frcp(rho);
  float lrho = 0.0f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float u = inv * a.q.p[1 + c][X];
    A.lq.p[1 + c][X] += lu[c] * inv;
    lrho -= lu[c] * u;
  }
  A.lq.p[0][X] += lrho * inv;
}

// Reverse of halo_dim_body (sc_halo.hip) for one (field, plane cell): the
// forward writes of the h layers are undone in reverse order; a copy
// dst = src becomes l_src += l_dst, l_dst = 0; a DIRICHLET write drops the
// halo adjoint; NEUMANN ghosts all derive from the edge cell.
__global__ void __launch_bounds__(256) halo_adj_dim_kernel(const HaloArgs a, int d) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const int A = a.N[d_a], B = a.N[d_b];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)a.nf * A * B) return;
  const int f = (int)(t / ((long long)A * B));
  const int rem = (int)(t % ((long long)A * B));
  const int ia = rem / B, ib = rem % B;
  float* v = a.q[f];
  const long long base = (long long)ia * a.st[d_a] + (long long)ib * a.st[d_b];
  const long long s = a.st[d];
  const int h = a.h, N = a.N[d], n = a.n[d];
  auto at = [&](int c) -> float& { return v[base + (long long)c * s]; };
  const BcFaceDev& lo = a.bc[f][d][0];
  const BcFaceDev& hi = a.bc[f][d][1];
  if (lo.type == kBcPeriodic || hi.type == kBcPeriodic) {
    for (int i = h - 1; i >= 0; --i) {
      at(h + i) += at(N - h + i);
      at(N - h + i) = 0.0f;
    }
    for (int i = h - 1; i >= 0; --i) {
      at(n + i) += at(i);
      at(i) = 0.0f;
    }
    return;
  }
  for (int side = 1; side >= 0; --side) {
    const BcFaceDev& bc = side == 0 ? lo : hi;
    float sum = 0.0f;
    for (int i = 0; i < h; ++i) {
      const int c = side == 0 ? h - 1 - i : N - h + i;
      sum += at(c);
      at(c) = 0.0f;
    }
    if (bc.type == kBcNeumann) at(side == 0 ? h : N - h - 1) += sum;
  }
}

struct CPtrList {
  const float* p[kMaxFields];
};
struct WPtrList {
  float* p[kMaxFields];
};

// Interior cells: lam_k = c1 * (lam_next + dt * lq), lam0 += c0 * lam_next;
// halo cells of lam_k are set to zero (cotangents live on the interior).
__global__ void __launch_bounds__(256)
rk_adj_combine_kernel(CPtrList ln, CPtrList lq, WPtrList lk, WPtrList l0, int has_l0, int nf,
                      AdjGeom g, float c0, float c1, const float* dtp) {
  const float dt = *dtp;
  const long long n = (long long)g.N[0] * g.N[1] * g.N[2];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const int y = (int)(t % g.N[1]);
    const int x = (int)((t / g.N[1]) % g.N[0]);
    const int z = (int)(t / ((long long)g.N[1] * g.N[0]));
    const bool in = x >= g.h && x < g.h + g.n[0] && y >= g.h && y < g.h + g.n[1] &&
                    z >= g.h && z < g.h + g.n[2];
    for (int f = 0; f < nf; ++f) {
      if (!in) {
        lk.p[f][t] = 0.0f;
        continue;
      }
      const float v = ln.p[f][t];
      if (has_l0) l0.p[f][t] += c0 * v;
      lk.p[f][t] = c1 * (v + dt * lq.p[f][t]);
    }
  }
}

unsigned blocks_for(long long n, long long cap) {
  long long b = (n + 255) / 256;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

hipError_t launch_adj_zero(float* const* p, int np, long long n, hipStream_t stream) {
  for (int k0 = 0; k0 < np; k0 += kZeroMax) {
    PtrList l;
    const int m = np - k0 < kZeroMax ? np - k0 : kZeroMax;
    for (int k = 0; k < kZeroMax; ++k) l.p[k] = k < m ? p[k0 + k] : nullptr;
    hipLaunchKernelGGL(adj_zero_kernel, blocks_for(n, 8192), 256, 0, stream, l, m, n);
  }
  return hipGetLastError();
}

hipError_t launch_adj_visc_gather(const AdjArgs& A, hipStream_t stream) {
  const SweepArgs& a = A.s;
  const long long n = (long long)(a.nx + 2) * (a.ny + 2) * (a.nz + 2);
  hipLaunchKernelGGL(adj_visc_gather_kernel, blocks_for(n, 1ll << 30), 256, 0, stream, A);
  return hipGetLastError();
}

hipError_t launch_halo_adj_dim(const HaloArgs& a, int d, hipStream_t stream) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const long long n = (long long)a.nf * a.N[d_a] * a.N[d_b];
  hipLaunchKernelGGL(halo_adj_dim_kernel, blocks_for(n, 1ll << 30), 256, 0, stream, a, d);
  return hipGetLastError();
}

hipError_t launch_rk_adj_combine(const float* const* lam_next, const float* const* lq,
                                 float* const* lam_k, float* const* lam0, int nf,
                                 const AdjGeom& g, float c0, float c1, const float* dt,
                                 hipStream_t stream) {
  const long long n = (long long)g.N[0] * g.N[1] * g.N[2];
  CPtrList ln, lqq;
  WPtrList lk, l0;
  for (int f = 0; f < kMaxFields; ++f) {
    ln.p[f] = f < nf ? lam_next[f] : nullptr;
    lqq.p[f] = f < nf ? lq[f] : nullptr;
    lk.p[f] = f < nf ? lam_k[f] : nullptr;
    l0.p[f] = (f < nf && lam0 != nullptr) ? lam0[f] : nullptr;
  }
  hipLaunchKernelGGL(rk_adj_combine_kernel, blocks_for(n, 8192), 256, 0, stream, ln, lqq, lk, l0,
                     lam0 != nullptr ? 1 : 0, nf, g, c0, c1, dt);
  return hipGetLastError();
}

}  // namespace sc
