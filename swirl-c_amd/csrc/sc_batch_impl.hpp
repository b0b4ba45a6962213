// Batched sweeps: many independent grids of different sizes advanced by one
// launch per sweep direction over a packed tile list.
//
// Instance i's field f occupies [off_i, off_i + N_i) of arena f (N_i its
// halo-padded cell count); all instances share the physics configuration.
// A tile is one block of the plain launch geometry of one instance; one
// launch covers every tile of every instance (one block per tile; the
// kernels' tile loop also accepts fewer blocks), so nothing is launched per
// instance and small grids leave no idle tail.  Per-instance dt comes from a
// device array, so a whole RK step of the ensemble can be captured in a HIP
// graph and replayed.
#pragma once

#include "sc_sweep_impl.hpp"

namespace sc {

struct BatchInst {
  int nx, ny, nz;
  long long off;  // element offset of the instance in every arena
  long long pe_off;  // its potential tables in BatchArgs::pe (x, y, z), or -1
  float inv_h[3];
  float h[3];
  float inv_two_h[3];
};

struct BatchTile {
  int inst, bx, by, bz;
};

struct BatchArgs {
  SweepArgs proto;  // shared configuration (flags, nu, kfac, RK form, ...)
  CFields q;        // arena bases
  CFields u0;
  Fields r;
  Fields out;
  int nf;
  const BatchInst* inst;
  const BatchTile* tiles;
  int ntiles;
  const float* dt;  // per-instance time steps
  const float* pe;  // potential tables of every instance (gravity), or null
  // Per-instance wave-speed maxima (float bits) the z sweep max-accumulates
  // over the cells it writes, or null.
  unsigned* speed;
};

__device__ __forceinline__ SweepArgs tile_args(const BatchArgs& b, int id, int dir) {
  const BatchInst in = b.inst[id];
  SweepArgs a = b.proto;
#pragma unroll
  for (int f = 0; f < kMaxFields; ++f) {
    const bool on = f < b.nf;
    a.q.p[f] = on && b.q.p[f] ? b.q.p[f] + in.off : nullptr;
    a.u0.p[f] = on && b.u0.p[f] ? b.u0.p[f] + in.off : nullptr;
    a.r.p[f] = on && b.r.p[f] ? b.r.p[f] + in.off : nullptr;
    a.out.p[f] = on && b.out.p[f] ? b.out.p[f] + in.off : nullptr;
  }
  a.nx = in.nx;
  a.ny = in.ny;
  a.nz = in.nz;
  a.NX = in.nx + 2 * a.h;
  a.NY = in.ny + 2 * a.h;
  a.NZ = in.nz + 2 * a.h;
  a.sx = a.NY;
  a.sz = (long long)a.NX * a.NY;
  a.inv_hdir = in.inv_h[dir];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a.vc.inv_h[d] = in.inv_h[d];
    a.vc.h[d] = in.h[d];
    a.vc.inv_two_h[d] = in.inv_two_h[d];
  }
  a.dt = b.dt != nullptr ? b.dt + id : nullptr;
  a.speed_bits = b.speed != nullptr ? b.speed + id : nullptr;
  if (b.pe != nullptr && in.pe_off >= 0) {
    a.has_pe = 1;
    a.pe_x = b.pe + in.pe_off;
    a.pe_y = a.pe_x + a.NX;
    a.pe_z = a.pe_y + a.NY;
  } else {
    a.has_pe = 0;
  }
  return a;
}

// FULL: the build with gravity, primitive-variable BCs and boundary face /
// flux values (decided at run time from the shared configuration); the other
// build leaves that code out.
template <int DIR, int SCHEME, int FLUX, bool VISC, int NS, int PHASE, bool FULL>
__global__ void __launch_bounds__(256) batch_march_kernel(const BatchArgs b) {
  for (int t = (int)blockIdx.x; t < b.ntiles; t += (int)gridDim.x) {
    const BatchTile tl = b.tiles[t];
    const SweepArgs a = tile_args(b, tl.inst, DIR);
    const TileIdx ti{tl.bx, tl.by, tl.bz, (int)threadIdx.x, (int)threadIdx.y,
                     (int)blockDim.y};
    sweep_march_body<DIR, SCHEME, FLUX, VISC, NS, PHASE, false, FULL, !FULL, true>(a, ti);
  }
}

template <int SCHEME, int FLUX, bool VISC, int NS, bool FULL>
__global__ void __launch_bounds__(256) batch_y_kernel(const BatchArgs b) {
  for (int t = (int)blockIdx.x; t < b.ntiles; t += (int)gridDim.x) {
    const BatchTile tl = b.tiles[t];
    const SweepArgs a = tile_args(b, tl.inst, 1);
    const TileIdx ti{tl.bx, tl.by, tl.bz, (int)threadIdx.x, (int)threadIdx.y,
                     (int)blockDim.y};
    sweep_y_body<SCHEME, FLUX, VISC, NS, false, FULL, !FULL>(a, ti, a.seg_y);
  }
}

// Tile lists of one batch (device arrays), one per sweep direction.
struct BatchTiles {
  const BatchTile* t[3];
  int n[3];
  int grid[3];  // blocks launched per direction (one per tile)
};

template <int SCHEME, int FLUX, bool VISC, int NS, bool FULL>
hipError_t launch_batch_sweeps_t(BatchArgs b, const BatchTiles& tiles,
                                 const SweepTuning& tune, hipStream_t stream) {
  {
    BatchArgs bx = b;
    bx.speed = nullptr;  // the CFL maxima come from the last (z) sweep only
    bx.proto.seg_len = tune.seg_x;
    bx.tiles = tiles.t[0];
    bx.ntiles = tiles.n[0];
    hipLaunchKernelGGL((batch_march_kernel<0, SCHEME, FLUX, VISC, NS, 0, FULL>),
                       dim3(tiles.grid[0]), dim3(64, tune.by_x), 0, stream, bx);
  }
  {
    BatchArgs by = b;
    by.speed = nullptr;
    by.proto.seg_y = tune.seg_y;
    by.tiles = tiles.t[1];
    by.ntiles = tiles.n[1];
    hipLaunchKernelGGL((batch_y_kernel<SCHEME, FLUX, VISC, NS, FULL>), dim3(tiles.grid[1]),
                       dim3(64, tune.by_y), 0, stream, by);
  }
  {
    BatchArgs bz = b;
    bz.proto.seg_len = tune.seg_z;
    bz.tiles = tiles.t[2];
    bz.ntiles = tiles.n[2];
    hipLaunchKernelGGL((batch_march_kernel<2, SCHEME, FLUX, VISC, NS, 2, FULL>),
                       dim3(tiles.grid[2]), dim3(64, tune.by_z), 0, stream, bz);
  }
  return hipGetLastError();
}

template <int SCHEME, int FLUX, bool VISC>
hipError_t launch_batch_variant(const BatchArgs& b, const BatchTiles& tiles, int ns, bool full,
                                const SweepTuning& tune, hipStream_t stream) {
  if (ns == 0)
    return full ? launch_batch_sweeps_t<SCHEME, FLUX, VISC, 0, true>(b, tiles, tune, stream)
                : launch_batch_sweeps_t<SCHEME, FLUX, VISC, 0, false>(b, tiles, tune, stream);
  if (ns == 1)
    return full ? launch_batch_sweeps_t<SCHEME, FLUX, VISC, 1, true>(b, tiles, tune, stream)
                : launch_batch_sweeps_t<SCHEME, FLUX, VISC, 1, false>(b, tiles, tune, stream);
  return full ? launch_batch_sweeps_t<SCHEME, FLUX, VISC, 2, true>(b, tiles, tune, stream)
              : launch_batch_sweeps_t<SCHEME, FLUX, VISC, 2, false>(b, tiles, tune, stream);
}

// ns: transported scalars (0..2); full: the configuration has gravity,
// primitive-variable BCs or boundary face / flux values.
template <int SCHEME>
hipError_t launch_batch_sweeps(const BatchArgs& b, const BatchTiles& tiles, int flux,
                               bool visc, int ns, bool full, const SweepTuning& tune,
                               hipStream_t stream);

#define SC_INSTANTIATE_BATCH(SCHEME)                                                      \
  template <>                                                                             \
  hipError_t launch_batch_sweeps<SCHEME>(const BatchArgs& b, const BatchTiles& tiles,     \
                                         int flux, bool visc, int ns, bool full,          \
                                         const SweepTuning& tune, hipStream_t stream) {   \
    if (flux == kHllc)                                                                    \
      return visc ? launch_batch_variant<SCHEME, kHllc, true>(b, tiles, ns, full, tune, stream) \
                  : launch_batch_variant<SCHEME, kHllc, false>(b, tiles, ns, full, tune, stream); \
    return visc ? launch_batch_variant<SCHEME, kHll, true>(b, tiles, ns, full, tune, stream)     \
                : launch_batch_variant<SCHEME, kHll, false>(b, tiles, ns, full, tune, stream);   \
  }

}  // namespace sc
