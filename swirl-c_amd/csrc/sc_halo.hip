// Halo fill (cell-average boundary conditions) and the CFL wave-speed
// reduction.
//
// Halo semantics: boundary/boundary.py:34-86,464-539 and swirl_lm's halo fill
// as written out in demo/swirl_c_prototype.ipynb cell 2 (l.24-163): dims are
// processed in the order x, y, z, each over its full halo planes (so edge and
// corner cells follow from the sequence); PERIODIC copies the opposite
// interior layers, DIRICHLET writes a value (scalar, per-layer scalar or
// per-layer plane; layer k -> index k on the low side, n+h+k on the high
// side), NEUMANN continues the neighbour cumulatively (ghost = neighbour -+
// value).  A field whose BC is given on its primitive is filled in primitive
// space (prim = q / rho, divide_no_nan) and multiplied back by the filled rho.
#include "sc_halo.hpp"

namespace sc {

namespace {

struct CFieldPtrs {
  float* p[kMaxFields];
};

// Geometry of one grid for the halo body: field pointers and extents.
struct HaloGeom {
  float* const* q;
  int n[3];
  int N[3];
  long long st[3];
};

// One thread's share of the dim-d fill: thread t of nf * A * B covers one
// (field, plane cell) and all h layers on both sides.  Configuration (BC
// tables, nf, h, prim_mask) comes from a.
__device__ __forceinline__ void halo_dim_body(const HaloArgs& a, const HaloGeom& g, int d,
                                              long long t) {
  // The two other axes in [z][x][y] order.
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const int A = g.N[d_a], B = g.N[d_b];
  if (t >= (long long)a.nf * A * B) return;
  const int f = (int)(t / ((long long)A * B));
  const int rem = (int)(t % ((long long)A * B));
  const int ia = rem / B, ib = rem % B;
  float* v = g.q[f];
  const long long base = (long long)ia * g.st[d_a] + (long long)ib * g.st[d_b];
  const long long s = g.st[d];
  const int h = a.h, N = g.N[d], n = g.n[d];
  const bool prim = (a.prim_mask >> f) & 1;
  // A source cell was already written (and holds a primitive) if it lies in
  // the halo of a dim processed earlier.
  bool earlier_halo = false;
  if (d >= 1) {
    const int cx = d_a == 0 ? ia : ib;  // x coordinate (d = 1: d_b = 0; d = 2: d_a = 0)
    earlier_halo |= (cx < h || cx >= h + g.n[0]);
  }
  if (d == 2) {
    earlier_halo |= (ib < h || ib >= h + g.n[1]);
  }
  const float* rho = g.q[0];
  auto get = [&](int c) -> float {
    const long long idx = base + (long long)c * s;
    const float x = v[idx];
    if (!prim || earlier_halo) return x;
    return dnn_exact(x, rho[idx]);
  };
  auto put = [&](int c, float x) { v[base + (long long)c * s] = x; };

  const BcFaceDev& lo = a.bc[f][d][0];
  const BcFaceDev& hi = a.bc[f][d][1];
  if (lo.type == kBcPeriodic || hi.type == kBcPeriodic) {
    for (int i = 0; i < h; ++i) put(i, get(n + i));
    for (int i = 0; i < h; ++i) put(N - h + i, get(h + i));
    return;
  }
  for (int side = 0; side < 2; ++side) {
    const BcFaceDev& bc = side == 0 ? lo : hi;
    if (bc.type == kBcDirichlet) {
      for (int i = 0; i < h; ++i) {
        float x;
        if (bc.n_layer_values > 0) {
          x = bc.layer_values[i];
        } else if (bc.layer_planes != nullptr) {
          x = bc.layer_planes[((long long)i * A + ia) * B + ib];
        } else {
          x = bc.value;
        }
        put(side == 0 ? i : N - h + i, x);
      }
    } else {  // Neumann
      if (side == 0) {
        float prev = get(h);
        for (int i = 0; i < h; ++i) {
          prev = prev - bc.value;
          put(h - 1 - i, prev);
        }
      } else {
        float prev = get(N - h - 1);
        for (int i = 0; i < h; ++i) {
          prev = prev + bc.value;
          put(N - h + i, prev);
        }
      }
    }
  }
}

__global__ void __launch_bounds__(256) halo_dim_kernel(const HaloArgs a, int d) {
  HaloGeom g;
  g.q = a.q;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = a.n[k];
    g.N[k] = a.N[k];
    g.st[k] = a.st[k];
  }
  halo_dim_body(a, g, d, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256)
batch_halo_dim_kernel(const HaloArgs a, CFieldPtrs q_base, const BatchHaloInst* inst,
                      const BatchHaloTile* tiles, int ntiles, int d) {
  for (int t = (int)blockIdx.x; t < ntiles; t += (int)gridDim.x) {
    const BatchHaloTile tl = tiles[t];
    const BatchHaloInst in = inst[tl.inst];
    float* q[kMaxFields];
    for (int f = 0; f < kMaxFields; ++f)
      q[f] = (f < a.nf && q_base.p[f]) ? q_base.p[f] + in.off : nullptr;
    HaloGeom g;
    g.q = q;
    for (int k = 0; k < 3; ++k) {
      g.n[k] = in.n[k];
      g.N[k] = in.n[k] + 2 * a.h;
    }
    g.st[0] = g.N[1];
    g.st[1] = 1;
    g.st[2] = (long long)g.N[0] * g.N[1];
    halo_dim_body(a, g, d, (long long)tl.blk * blockDim.x + threadIdx.x);
  }
}

__global__ void __launch_bounds__(256) halo_prim_finish_kernel(const HaloArgs a) {
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
  const float r = a.q[0][t];
  for (int f = 1; f < a.nf; ++f) {
    if ((a.prim_mask >> f) & 1) a.q[f][t] = r * a.q[f][t];
  }
}

__global__ void speed_init_kernel(unsigned* bits) { *bits = 0u; }

__global__ void __launch_bounds__(256) max_speed_kernel(const SpeedArgs a, unsigned* bits) {
  const long long total = (long long)a.nx * a.ny * a.nz;
  float m = 0.0f;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int j = a.h + (int)(t % a.ny);
    const int i = a.h + (int)((t / a.ny) % a.nx);
    const int k = a.h + (int)(t / ((long long)a.ny * a.nx));
    const long long idx = (long long)k * a.sz + (long long)i * a.sx + j;
    float q[5], w[7];
#pragma unroll
    for (int f = 0; f < 5; ++f) q[f] = a.q[f][idx];
    const float pe = a.has_pe ? (a.pe_x[i] - a.pe_y[j]) - a.pe_z[k] : 0.0f;
    cons_to_prim<0>(q, pe, w);
    const float c = sqrtf(kGamma * w[kP] / w[kRho]);
    const float s = fmaxf(fmaxf(fabsf(w[kU]) + c, fabsf(w[kV]) + c), fabsf(w[kW]) + c);
    m = fmaxf(m, s);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float part[4];
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) part[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = part[0];
    for (int w = 1; w < (int)(blockDim.x / 64); ++w) b = fmaxf(b, part[w]);
    atomicMax(bits, __float_as_uint(b));
  }
}

__global__ void cfl_dt_kernel(const unsigned* bits, float cfl, float hmin, float* dt) {
  *dt = cfl * hmin / __uint_as_float(*bits);
}

__global__ void batch_speed_init_kernel(unsigned* bits, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) bits[i] = 0u;
}

__global__ void __launch_bounds__(256)
batch_max_speed_kernel(CFieldPtrs q_base, const BatchHaloInst* inst, int h,
                       const BatchHaloTile* tiles, int ntiles, unsigned* bits) {
  __shared__ float part[4];
  for (int tt = (int)blockIdx.x; tt < ntiles; tt += (int)gridDim.x) {
    const BatchHaloTile tl = tiles[tt];
    const BatchHaloInst in = inst[tl.inst];
    const int nx = in.n[0], ny = in.n[1], nz = in.n[2];
    const long long NX = nx + 2 * h, NY = ny + 2 * h;
    const long long total = (long long)nx * ny * nz;
    const long long c0 = (long long)tl.blk * kSpeedChunk;
    float m = 0.0f;
    for (long long t = c0 + threadIdx.x; t < c0 + kSpeedChunk && t < total; t += blockDim.x) {
      const int j = h + (int)(t % ny);
      const int i = h + (int)((t / ny) % nx);
      const int k = h + (int)(t / ((long long)ny * nx));
      const long long idx = in.off + (long long)k * NX * NY + (long long)i * NY + j;
      float q[5], w[7];
#pragma unroll
      for (int f = 0; f < 5; ++f) q[f] = q_base.p[f][idx];
      cons_to_prim<0>(q, 0.0f, w);
      const float c = sqrtf(kGamma * w[kP] / w[kRho]);
      const float s = fmaxf(fmaxf(fabsf(w[kU]) + c, fabsf(w[kV]) + c), fabsf(w[kW]) + c);
      m = fmaxf(m, s);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    const int wave = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) part[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      float b = part[0];
      for (int w = 1; w < (int)(blockDim.x / 64); ++w) b = fmaxf(b, part[w]);
      atomicMax(bits + tl.inst, __float_as_uint(b));
    }
    __syncthreads();
  }
}

__global__ void batch_cfl_dt_kernel(const unsigned* bits, int n, float cfl,
                                    const float* hmin, float* dt) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) dt[i] = cfl * hmin[i] / __uint_as_float(bits[i]);
}

}  // namespace

hipError_t launch_halo_dim(const HaloArgs& a, int d, hipStream_t stream) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const long long n = (long long)a.nf * a.N[d_a] * a.N[d_b];
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(halo_dim_kernel, grid, 256, 0, stream, a, d);
  return hipGetLastError();
}

hipError_t launch_batch_halo_dim(const HaloArgs& cfg, float* const* q_base,
                                 const BatchHaloInst* inst, const BatchHaloTile* tiles,
                                 int ntiles, int grid, int d, hipStream_t stream) {
  CFieldPtrs qb;
  for (int f = 0; f < kMaxFields; ++f) qb.p[f] = f < cfg.nf ? q_base[f] : nullptr;
  hipLaunchKernelGGL(batch_halo_dim_kernel, dim3(grid), 256, 0, stream, cfg, qb, inst, tiles,
                     ntiles, d);
  return hipGetLastError();
}

hipError_t launch_batch_max_speed(float* const* q_base, const BatchHaloInst* inst, int h,
                                  const BatchHaloTile* tiles, int ntiles, int grid,
                                  int n_inst, unsigned* bits, float cfl,
                                  const float* hmin, float* dt, hipStream_t stream) {
  CFieldPtrs qb;
  for (int f = 0; f < kMaxFields; ++f) qb.p[f] = f < 5 ? q_base[f] : nullptr;
  hipLaunchKernelGGL(batch_speed_init_kernel, dim3((n_inst + 255) / 256), 256, 0, stream,
                     bits, n_inst);
  hipLaunchKernelGGL(batch_max_speed_kernel, dim3(grid), 256, 0, stream, qb, inst, h, tiles,
                     ntiles, bits);
  if (dt != nullptr) {
    hipLaunchKernelGGL(batch_cfl_dt_kernel, dim3((n_inst + 255) / 256), 256, 0, stream, bits,
                       n_inst, cfl, hmin, dt);
  }
  return hipGetLastError();
}

hipError_t launch_halo_prim_finish(const HaloArgs& a, hipStream_t stream) {
  const long long n = (long long)a.N[0] * a.N[1] * a.N[2];
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(halo_prim_finish_kernel, grid, 256, 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_max_speed(const SpeedArgs& a, unsigned* max_bits, float cfl,
                            float hmin, float* dt, hipStream_t stream) {
  hipLaunchKernelGGL(speed_init_kernel, 1, 1, 0, stream, max_bits);
  const long long total = (long long)a.nx * a.ny * a.nz;
  long long blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(max_speed_kernel, (unsigned)blocks, 256, 0, stream, a, max_bits);
  if (dt != nullptr) {
    hipLaunchKernelGGL(cfl_dt_kernel, 1, 1, 0, stream, max_bits, cfl, hmin, dt);
  }
  return hipGetLastError();
}

}  // namespace sc
