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

// One halo cell of the dim-D fill: field f, side (0 low, 1 high), layer,
// and its position (ia, ib) in the plane of the two other axes.
struct HaloCell {
  int f, side, layer, ia, ib;
};

// Writes one halo cell.  No cell read here is written by another thread of
// the same dim, so every thread loads its source, if any, then stores once,
// and all of a wave's loads are in flight together.  Configuration (BC
// tables, h, prim_mask) comes from a.
template <int D>
__device__ __forceinline__ void halo_cell_fill(const HaloArgs& a, const HaloGeom& g,
                                               const HaloCell& hc) {
  // The two other axes in [z][x][y] order.
  constexpr int d_a = D == 0 ? 2 : (D == 1 ? 2 : 0);
  constexpr int d_b = D == 0 ? 1 : (D == 1 ? 0 : 1);
  const int A = g.N[d_a], B = g.N[d_b];
  const int h = a.h, N = g.N[D], n = g.n[D];
  const int f = hc.f, side = hc.side, layer = hc.layer, ia = hc.ia, ib = hc.ib;
  float* v = g.q[f];
  const long long base = (long long)ia * g.st[d_a] + (long long)ib * g.st[d_b];
  const long long s = g.st[D];
  const bool prim = (a.prim_mask >> f) & 1;
  // A source cell was already written (and holds a primitive) if it lies in
  // the halo of a dim processed earlier.
  bool earlier_halo = false;
  if (D >= 1) {
    const int cx = d_a == 0 ? ia : ib;  // x coordinate (D = 1: d_b = 0; D = 2: d_a = 0)
    earlier_halo |= (cx < h || cx >= h + g.n[0]);
  }
  if (D == 2) {
    earlier_halo |= (ib < h || ib >= h + g.n[1]);
  }
  const float* rho = g.q[0];
  auto get = [&](int c) -> float {
    const long long idx = base + (long long)c * s;
    const float x = v[idx];
    if (!prim || earlier_halo) return x;
    return dnn_exact(x, rho[idx]);
  };
  const BcFaceDev& lo = a.bc[f][D][0];
  const BcFaceDev& hi = a.bc[f][D][1];
  if (lo.type == kBcNone) return;  // no BC along D: halo cells left alone
  // Layer i of the low side is cell i, of the high side cell N - h + i;
  // Neumann runs outward from the boundary (low side: cell h - 1 - i).
  int dst;
  float x;
  if (lo.type == kBcPeriodic || hi.type == kBcPeriodic) {
    dst = side == 0 ? layer : N - h + layer;
    x = get(side == 0 ? n + layer : h + layer);
  } else {
    const BcFaceDev& bc = side == 0 ? lo : hi;
    if (bc.type == kBcDirichlet) {
      dst = side == 0 ? layer : N - h + layer;
      if (bc.n_layer_values > 0) {
        x = bc.layer_values[layer];
      } else if (bc.layer_planes != nullptr) {
        x = bc.layer_planes[((long long)layer * A + ia) * B + ib];
      } else {
        x = bc.value;
      }
    } else if (side == 0) {  // Neumann: ghost = neighbour - value, cumulatively
      dst = h - 1 - layer;
      x = get(h);
      for (int i = 0; i <= layer; ++i) x = x - bc.value;
    } else {
      dst = N - h + layer;
      x = get(N - h - 1);
      for (int i = 0; i <= layer; ++i) x = x + bc.value;
    }
  }
  v[base + (long long)dst * s] = x;
}

// Rows of one y-fill block: each row end's 2h halo cells are adjacent
// lanes, so a wave's stores cover whole row ends.
__host__ __device__ constexpr int halo_y_rows(int h) { return 256 / (2 * h); }

// Single grid.  D = 0, 2: grid (ceil(B / 256), A, nf * 2h), lanes along ib
// (y, contiguous: coalesced); (f, side, layer) is uniform per block.
// D = 1: grid (ceil(B / rows), A, nf), lanes over (row, side, layer).
template <int D>
__global__ void __launch_bounds__(256) halo_dim_kernel(const HaloArgs a) {
  HaloGeom g;
  g.q = a.q;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = a.n[k];
    g.N[k] = a.N[k];
    g.st[k] = a.st[k];
  }
  constexpr int d_b = D == 0 ? 1 : (D == 1 ? 0 : 1);
  const int B = g.N[d_b];
  const int h = a.h;
  HaloCell hc;
  hc.ia = (int)blockIdx.y;
  if (D == 1) {
    const int rows = halo_y_rows(h);
    const int r = (int)threadIdx.x / (2 * h);
    const int sl = (int)threadIdx.x - r * 2 * h;
    if (r >= rows) return;
    hc.ib = (int)blockIdx.x * rows + r;
    hc.side = sl / h;
    hc.layer = sl - hc.side * h;
    hc.f = (int)blockIdx.z;
  } else {
    hc.ib = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int z = (int)blockIdx.z;
    hc.f = z / (2 * h);
    hc.side = (z / h) & 1;
    hc.layer = z % h;
  }
  if (hc.ib >= B) return;
  halo_cell_fill<D>(a, g, hc);
}

// Every field PERIODIC on all six faces, no primitive-variable BC: the x, y, z
// fill in sequence leaves each halo cell equal to the interior cell at its
// index wrapped along all three dims (edges and corners included), so one
// launch copies that directly.  Threads of one field (blockIdx.y): first the
// full y rows of the z halo planes and of the x halo rows of interior planes,
// then the 2h y-halo cells of each interior row.  Sources are interior cells
// and destinations halo cells, so no thread reads another's output.
__global__ void __launch_bounds__(256)
halo_periodic_kernel(const HaloArgs a, int full_cells, int total) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= total) return;
  const int h = a.h, h2 = 2 * a.h;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int NX = a.N[0], NY = a.N[1];
  int k, i, j;
  if (t < full_cells) {
    const int row = t / NY;
    j = t - row * NY;
    const int zrows = h2 * NX;
    if (row < zrows) {
      const int kk = row / NX;
      i = row - kk * NX;
      k = kk < h ? kk : nz + kk;
    } else {
      const int r = row - zrows;
      const int kk = r / h2;
      const int sl = r - kk * h2;
      k = h + kk;
      i = sl < h ? sl : nx + sl;
    }
  } else {
    const int r = t - full_cells;
    const int row = r / h2;
    const int sl = r - row * h2;
    const int kk = row / nx;
    k = h + kk;
    i = h + (row - kk * nx);
    j = sl < h ? sl : ny + sl;
  }
  auto wrap = [h](int c, int n) { return c < h ? c + n : (c >= h + n ? c - n : c); };
  const long long dst = (long long)k * a.st[2] + (long long)i * a.st[0] + j;
  const long long src =
      (long long)wrap(k, nz) * a.st[2] + (long long)wrap(i, nx) * a.st[0] + wrap(j, ny);
  float* v = a.q[blockIdx.y];
  v[dst] = v[src];
}

// Batched: tile (instance, blk) covers threads blk*256 .. of the instance's
// nf * 2h * A * B halo cells (32-bit: the tile builder checks the count),
// decoded with ib fastest (D = 0, 2) or (side, layer) fastest (D = 1).
template <int D>
__global__ void __launch_bounds__(256)
batch_halo_dim_kernel(const HaloArgs a, CFieldPtrs q_base, const BatchHaloInst* inst,
                      const BatchHaloTile* tiles, int ntiles) {
  constexpr int d_a = D == 0 ? 2 : (D == 1 ? 2 : 0);
  constexpr int d_b = D == 0 ? 1 : (D == 1 ? 0 : 1);
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
    const unsigned A = (unsigned)g.N[d_a], B = (unsigned)g.N[d_b], h = (unsigned)a.h;
    const unsigned id = (unsigned)tl.blk * blockDim.x + threadIdx.x;
    if (id >= (unsigned)a.nf * 2u * h * A * B) continue;
    HaloCell hc;
    unsigned r;
    if (D == 1) {
      hc.layer = (int)(id % h);
      r = id / h;
      hc.side = (int)(r & 1u);
      r >>= 1;
      hc.ib = (int)(r % B);
      r /= B;
      hc.ia = (int)(r % A);
      hc.f = (int)(r / A);
    } else {
      hc.ib = (int)(id % B);
      r = id / B;
      hc.ia = (int)(r % A);
      r /= A;
      hc.layer = (int)(r % h);
      r /= h;
      hc.side = (int)(r & 1u);
      hc.f = (int)(r >> 1);
    }
    halo_cell_fill<D>(a, g, hc);
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

// Waves stride over the interior rows (k, i); the lanes of a wave cover a
// row's y cells 64 at a time (coalesced, no per-cell index division).
__global__ void __launch_bounds__(256) max_speed_kernel(const SpeedArgs a, unsigned* bits) {
  const int rows = a.nx * a.nz;
  const int lane = (int)threadIdx.x & 63;
  const int wave = (int)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  const int nwaves = (int)(gridDim.x * (blockDim.x / 64));
  float m = 0.0f;
  constexpr int U = 4;  // cells per lane per pass, loaded together
  for (int r = wave; r < rows; r += nwaves) {
    const int i = a.h + r % a.nx;
    const int k = a.h + r / a.nx;
    const long long row = (long long)k * a.sz + (long long)i * a.sx;
    for (int j0 = a.h + lane; j0 < a.h + a.ny; j0 += 64 * U) {
      float q[U][5];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + 64 * u;
        const long long idx = row + min(j, a.h + a.ny - 1);
#pragma unroll
        for (int f = 0; f < 5; ++f) q[u][f] = a.q[f][idx];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + 64 * u;
        if (j >= a.h + a.ny) break;
        const float pe = a.has_pe ? (a.pe_x[i] - a.pe_y[j]) - a.pe_z[k] : 0.0f;
        m = fmaxf(m, wave_speed(q[u], pe));
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __shared__ float part[4];
  if (lane == 0) part[threadIdx.x / 64] = m;
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

// dt from the accumulated maximum, then the maximum cleared for the next
// accumulation (one thread: the read precedes the write).
__global__ void cfl_dt_reset_kernel(unsigned* bits, float cfl, float hmin, float* dt) {
  *dt = cfl * hmin / __uint_as_float(*bits);
  *bits = 0u;
}

__global__ void batch_speed_init_kernel(unsigned* bits, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) bits[i] = 0u;
}

__global__ void __launch_bounds__(256)
batch_max_speed_kernel(CFieldPtrs q_base, const BatchHaloInst* inst, int h,
                       const BatchHaloTile* tiles, int ntiles, unsigned* bits,
                       const float* pe_arena) {
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
      float q[5];
#pragma unroll
      for (int f = 0; f < 5; ++f) q[f] = q_base.p[f][idx];
      float pe = 0.0f;
      if (pe_arena != nullptr && in.pe_off >= 0) {
        const float* px = pe_arena + in.pe_off;
        pe = (px[i] - px[NX + j]) - px[NX + NY + k];
      }
      m = fmaxf(m, wave_speed(q, pe));
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

// Per instance: dt from the accumulated maximum, then the maximum cleared for
// the next accumulation (one thread per instance: its read precedes its write).
__global__ void batch_cfl_dt_reset_kernel(unsigned* bits, int n, float cfl, const float* hmin,
                                          float* dt) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) {
    dt[i] = cfl * hmin[i] / __uint_as_float(bits[i]);
    bits[i] = 0u;
  }
}

}  // namespace

hipError_t launch_halo_dim(const HaloArgs& a, int d, hipStream_t stream) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  const unsigned A = (unsigned)a.N[d_a], B = (unsigned)a.N[d_b];
  if (d == 1) {
    const unsigned rows = (unsigned)halo_y_rows(a.h);
    hipLaunchKernelGGL(halo_dim_kernel<1>, dim3((B + rows - 1) / rows, A, (unsigned)a.nf), 256,
                       0, stream, a);
  } else {
    const dim3 grid((B + 255) / 256, A, (unsigned)(a.nf * 2 * a.h));
    if (d == 0) {
      hipLaunchKernelGGL(halo_dim_kernel<0>, grid, 256, 0, stream, a);
    } else {
      hipLaunchKernelGGL(halo_dim_kernel<2>, grid, 256, 0, stream, a);
    }
  }
  return hipGetLastError();
}

long long halo_periodic_threads(const HaloArgs& a) {
  const long long full = ((long long)2 * a.h * a.N[0] + (long long)a.n[2] * 2 * a.h) * a.N[1];
  return full + (long long)a.n[2] * a.n[0] * 2 * a.h;
}

hipError_t launch_halo_periodic(const HaloArgs& a, hipStream_t stream) {
  const int full = (int)(((long long)2 * a.h * a.N[0] + (long long)a.n[2] * 2 * a.h) * a.N[1]);
  const int total = (int)halo_periodic_threads(a);
  hipLaunchKernelGGL(halo_periodic_kernel,
                     dim3((unsigned)(((long long)total + 255) / 256), (unsigned)a.nf), 256, 0,
                     stream, a, full, total);
  return hipGetLastError();
}

hipError_t launch_batch_halo_dim(const HaloArgs& cfg, float* const* q_base,
                                 const BatchHaloInst* inst, const BatchHaloTile* tiles,
                                 int ntiles, int grid, int d, hipStream_t stream) {
  CFieldPtrs qb;
  for (int f = 0; f < kMaxFields; ++f) qb.p[f] = f < cfg.nf ? q_base[f] : nullptr;
  if (d == 0) {
    hipLaunchKernelGGL(batch_halo_dim_kernel<0>, dim3(grid), 256, 0, stream, cfg, qb, inst,
                       tiles, ntiles);
  } else if (d == 1) {
    hipLaunchKernelGGL(batch_halo_dim_kernel<1>, dim3(grid), 256, 0, stream, cfg, qb, inst,
                       tiles, ntiles);
  } else {
    hipLaunchKernelGGL(batch_halo_dim_kernel<2>, dim3(grid), 256, 0, stream, cfg, qb, inst,
                       tiles, ntiles);
  }
  return hipGetLastError();
}

__global__ void __launch_bounds__(256)
batch_halo_prim_finish_kernel(const HaloArgs a, CFieldPtrs q_base, const BatchHaloInst* inst,
                              const BatchHaloTile* tiles, int ntiles) {
  for (int tt = (int)blockIdx.x; tt < ntiles; tt += (int)gridDim.x) {
    const BatchHaloTile tl = tiles[tt];
    const BatchHaloInst in = inst[tl.inst];
    const int h = a.h;
    const long long NX = in.n[0] + 2 * h, NY = in.n[1] + 2 * h, NZ = in.n[2] + 2 * h;
    const long long t = (long long)tl.blk * blockDim.x + threadIdx.x;
    if (t >= NX * NY * NZ) continue;
    const int y = (int)(t % NY);
    const int x = (int)((t / NY) % NX);
    const int z = (int)(t / (NY * NX));
    const bool halo = x < h || x >= h + in.n[0] || y < h || y >= h + in.n[1] ||
                      z < h || z >= h + in.n[2];
    if (!halo) continue;
    const long long idx = in.off + t;
    const float r = q_base.p[0][idx];
    for (int f = 1; f < a.nf; ++f) {
      if ((a.prim_mask >> f) & 1) q_base.p[f][idx] = r * q_base.p[f][idx];
    }
  }
}

hipError_t launch_batch_halo_prim_finish(const HaloArgs& cfg, float* const* q_base,
                                         const BatchHaloInst* inst, const BatchHaloTile* tiles,
                                         int ntiles, int grid, hipStream_t stream) {
  CFieldPtrs qb;
  for (int f = 0; f < kMaxFields; ++f) qb.p[f] = f < cfg.nf ? q_base[f] : nullptr;
  hipLaunchKernelGGL(batch_halo_prim_finish_kernel, dim3(grid), 256, 0, stream, cfg, qb, inst,
                     tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_batch_max_speed(float* const* q_base, const BatchHaloInst* inst, int h,
                                  const BatchHaloTile* tiles, int ntiles, int grid,
                                  int n_inst, unsigned* bits, float cfl,
                                  const float* hmin, float* dt, const float* pe,
                                  hipStream_t stream) {
  CFieldPtrs qb;
  for (int f = 0; f < kMaxFields; ++f) qb.p[f] = f < 5 ? q_base[f] : nullptr;
  hipLaunchKernelGGL(batch_speed_init_kernel, dim3((n_inst + 255) / 256), 256, 0, stream,
                     bits, n_inst);
  hipLaunchKernelGGL(batch_max_speed_kernel, dim3(grid), 256, 0, stream, qb, inst, h, tiles,
                     ntiles, bits, pe);
  if (dt != nullptr) {
    hipLaunchKernelGGL(batch_cfl_dt_kernel, dim3((n_inst + 255) / 256), 256, 0, stream, bits,
                       n_inst, cfl, hmin, dt);
  }
  return hipGetLastError();
}

hipError_t launch_batch_cfl_dt_reset(unsigned* bits, int n_inst, float cfl, const float* hmin,
                                     float* dt, hipStream_t stream) {
  hipLaunchKernelGGL(batch_cfl_dt_reset_kernel, dim3((n_inst + 255) / 256), 256, 0, stream,
                     bits, n_inst, cfl, hmin, dt);
  return hipGetLastError();
}

hipError_t launch_halo_prim_finish(const HaloArgs& a, hipStream_t stream) {
  const long long n = (long long)a.N[0] * a.N[1] * a.N[2];
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(halo_prim_finish_kernel, grid, 256, 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_cfl_dt_reset(unsigned* max_bits, float cfl, float hmin, float* dt,
                               hipStream_t stream) {
  hipLaunchKernelGGL(cfl_dt_reset_kernel, 1, 1, 0, stream, max_bits, cfl, hmin, dt);
  return hipGetLastError();
}

hipError_t launch_max_speed(const SpeedArgs& a, unsigned* max_bits, float cfl,
                            float hmin, float* dt, hipStream_t stream) {
  hipLaunchKernelGGL(speed_init_kernel, 1, 1, 0, stream, max_bits);
  // Up to one wave per interior row, at most 1024 blocks of 4 waves (4 waves
  // per SIMD; each block ends in one atomic on the shared maximum, so more
  // blocks only queue atomics).
  long long blocks = ((long long)a.nx * a.nz + 3) / 4;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(max_speed_kernel, (unsigned)blocks, 256, 0, stream, a, max_bits);
  if (dt != nullptr) {
    hipLaunchKernelGGL(cfl_dt_kernel, 1, 1, 0, stream, max_bits, cfl, hmin, dt);
  }
  return hipGetLastError();
}

}  // namespace sc
