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
// The stage state's fields (read-only).
struct StatePtrs {
  const float* p[kMaxFields];
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

// Number of thread units per plane cell of the transposed fill: unit 0 takes
// the density and every field with its BC on the primitive (those add to the
// density's adjoint, so one thread does them in a fixed order), each other
// field a unit of its own.
__host__ __device__ inline int halo_adj_units(const HaloArgs& a) {
  int u = 1;
  for (int f = 1; f < a.nf; ++f)
    if (!((a.prim_mask >> f) & 1)) ++u;
  return u;
}

// Reverse of halo_dim_body (sc_halo.hip) for one plane cell and one unit of
// fields (halo_adj_units): the forward writes of the h layers are undone in reverse order; a
// copy dst = src becomes l_src += l_dst, l_dst = 0; a DIRICHLET write drops
// the halo adjoint; NEUMANN ghosts all derive from the edge cell.  For a
// field with its BC on the primitive the halo adjoints are primitive
// adjoints (halo_adj_prim_kernel runs first); where the forward read a
// source cell's primitive dnn(q, rho) (a cell not in the halo of an earlier
// dim), the adjoint reaching it is converted back to q and rho (state: the
// filled stage state).
__global__ void __launch_bounds__(256)
halo_adj_dim_kernel(const HaloArgs a, StatePtrs state, int d) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const int A = a.N[d_a], B = a.N[d_b];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int units = halo_adj_units(a);
  if (t >= (long long)units * A * B) return;
  const int unit = (int)(t / ((long long)A * B));
  const int rem = (int)(t % ((long long)A * B));
  const int ia = rem / B, ib = rem % B;
  // Unit > 0: its single (non-primitive-BC) field.
  int only = -1;
  if (unit > 0) {
    int u = 0;
    for (int f = 1; f < a.nf; ++f) {
      if ((a.prim_mask >> f) & 1) continue;
      if (++u == unit) {
        only = f;
        break;
      }
    }
  }
  const long long base = (long long)ia * a.st[d_a] + (long long)ib * a.st[d_b];
  const long long s = a.st[d];
  const int h = a.h, N = a.N[d], n = a.n[d];
  // Sources of this line lie in the halo of an earlier dim (x before y
  // before z) and then hold primitives already.
  bool earlier_halo = false;
  if (d >= 1) {
    const int cx = d_a == 0 ? ia : ib;
    earlier_halo |= (cx < h || cx >= h + a.n[0]);
  }
  if (d == 2) earlier_halo |= (ib < h || ib >= h + a.n[1]);
  for (int f = 0; f < a.nf; ++f) {
    const bool prim_f = (a.prim_mask >> f) & 1;
    if (unit == 0 ? (f != 0 && !prim_f) : f != only) continue;
    float* v = a.q[f];
    const bool conv = prim_f && !earlier_halo;
    auto at = [&](int c) -> float& { return v[base + (long long)c * s]; };
    // l_src += l (a primitive adjoint when conv: dnn(q, rho) at the source).
    auto add_src = [&](int c, float l) {
      const long long idx = base + (long long)c * s;
      if (!conv) {
        v[idx] += l;
        return;
      }
      const float rho = state.p[0][idx];
      if (rho == 0.0f) return;
      const float ir = 1.0f / rho;
      v[idx] += l * ir;
      a.q[0][idx] -= l * state.p[f][idx] * ir * ir;
    };
    const BcFaceDev& lo = a.bc[f][d][0];
    const BcFaceDev& hi = a.bc[f][d][1];
    if (lo.type == kBcNone) continue;
    if (lo.type == kBcPeriodic || hi.type == kBcPeriodic) {
      for (int i = h - 1; i >= 0; --i) {
        const float l = at(N - h + i);
        at(N - h + i) = 0.0f;
        add_src(h + i, l);
      }
      for (int i = h - 1; i >= 0; --i) {
        const float l = at(i);
        at(i) = 0.0f;
        add_src(n + i, l);
      }
      continue;
    }
    for (int side = 1; side >= 0; --side) {
      const BcFaceDev& bc = side == 0 ? lo : hi;
      float sum = 0.0f;
      for (int i = 0; i < h; ++i) {
        const int c = side == 0 ? h - 1 - i : N - h + i;
        sum += at(c);
        at(c) = 0.0f;
      }
      if (bc.type == kBcNeumann) add_src(side == 0 ? h : N - h - 1, sum);
    }
  }
}

// Reverse of halo_prim_finish_kernel: halo cells of primitive-BC fields hold
// q = rho_h * prim_h; their adjoint becomes the primitive adjoint
// l * rho_h, and l * prim_h joins the halo density's adjoint.
__global__ void __launch_bounds__(256) halo_adj_prim_kernel(const HaloArgs a, StatePtrs state) {
  const long long total = (long long)a.N[0] * a.N[1] * a.N[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int y = (int)(t % a.N[1]);
  const int x = (int)((t / a.N[1]) % a.N[0]);
  const int z = (int)(t / ((long long)a.N[1] * a.N[0]));
  const int h = a.h;
  const bool halo = x < h || x >= h + a.n[0] || y < h || y >= h + a.n[1] ||
                    z < h || z >= h + a.n[2];
  if (!halo) return;
  const float r = state.p[0][t];
  for (int f = 1; f < a.nf; ++f) {
    if (!((a.prim_mask >> f) & 1)) continue;
    const float l = a.q[f][t];
    const float prim = r == 0.0f ? 0.0f : state.p[f][t] / r;
    a.q[0][t] += l * prim;
    a.q[f][t] = l * r;
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

hipError_t launch_halo_adj_dim(const HaloArgs& a, const float* const* state, int d,
                               hipStream_t stream) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const long long n = (long long)halo_adj_units(a) * a.N[d_a] * a.N[d_b];
  StatePtrs st;
  for (int f = 0; f < kMaxFields; ++f) st.p[f] = f < a.nf ? state[f] : nullptr;
  hipLaunchKernelGGL(halo_adj_dim_kernel, blocks_for(n, 1ll << 30), 256, 0, stream, a, st, d);
  return hipGetLastError();
}

hipError_t launch_halo_adj_prim(const HaloArgs& a, const float* const* state,
                                hipStream_t stream) {
  const long long n = (long long)a.N[0] * a.N[1] * a.N[2];
  StatePtrs st;
  for (int f = 0; f < kMaxFields; ++f) st.p[f] = f < a.nf ? state[f] : nullptr;
  hipLaunchKernelGGL(halo_adj_prim_kernel, blocks_for(n, 1ll << 30), 256, 0, stream, a, st);
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
