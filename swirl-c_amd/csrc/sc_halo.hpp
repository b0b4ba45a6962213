// Halo-fill and wave-speed kernels (declarations).
#pragma once

#include <hip/hip_runtime.h>

#include "sc_device.hpp"

namespace sc {

constexpr int kMaxHalo = 4;
constexpr int kBcNone = -1, kBcPeriodic = 0, kBcDirichlet = 1, kBcNeumann = 2;

struct BcFaceDev {
  int type;
  float value;
  int n_layer_values;
  float layer_values[kMaxHalo];
  const float* layer_planes;  // [h][A][B] or null
};

struct HaloArgs {
  float* q[kMaxFields];
  int nf;
  int h;
  int n[3];          // interior extents along x, y, z
  int N[3];          // full extents along x, y, z
  long long st[3];   // element strides along x, y, z
  int prim_mask;     // fields whose BC is on the primitive
  BcFaceDev bc[kMaxFields][3][2];
};

// Threads of the dim-d fill of one grid (extents N along x, y, z): one per
// halo cell (field, side, layer, plane cell).
inline long long halo_dim_threads(int nf, int h, const long long N[3], int d) {
  const int d_a = d == 0 ? 2 : (d == 1 ? 2 : 0);
  const int d_b = d == 0 ? 1 : (d == 1 ? 0 : 1);
  return (long long)nf * 2 * h * N[d_a] * N[d_b];
}

// Applies the BCs along dim d (0 x, 1 y, 2 z) to all fields; must be launched
// for d = 0, 1, 2 in order, then halo_prim_finish if prim_mask != 0.
hipError_t launch_halo_dim(const HaloArgs& a, int d, hipStream_t stream);
hipError_t launch_halo_prim_finish(const HaloArgs& a, hipStream_t stream);

// One-launch fill for grids whose fields are all PERIODIC on every face with
// no primitive-variable BC (same result as the three dim passes); its thread
// count must stay below 2^31.
long long halo_periodic_threads(const HaloArgs& a);
hipError_t launch_halo_periodic(const HaloArgs& a, hipStream_t stream);

// Batched halo fill: `cfg` carries the shared BC tables (its field pointers
// and extents are unused); instance geometry comes from `inst` (nx, ny, nz,
// arena offset) and the tile list (instance, block) of each dim.
struct BatchHaloInst {
  int n[3];
  long long off;
  // Offset of the instance's potential tables in the batch's potential
  // arena (pe_x over NX, then pe_y over NY, then pe_z over NZ), or -1.
  long long pe_off;
};
struct BatchHaloTile {
  int inst, blk;
};
hipError_t launch_batch_halo_dim(const HaloArgs& cfg, float* const* q_base,
                                 const BatchHaloInst* inst, const BatchHaloTile* tiles,
                                 int ntiles, int grid, int d, hipStream_t stream);

// Batched wave speed: tiles (instance, chunk of kSpeedChunk interior cells);
// bits[i] = max speed of instance i, then dt[i] = cfl * hmin[i] / max.
constexpr int kSpeedChunk = 4096;
// pe: the batch's potential arena (instances with pe_off >= 0), or null.
hipError_t launch_batch_max_speed(float* const* q_base, const BatchHaloInst* inst, int h,
                                  const BatchHaloTile* tiles, int ntiles, int grid,
                                  int n_inst, unsigned* bits, float cfl,
                                  const float* hmin, float* dt, const float* pe,
                                  hipStream_t stream);
// Halo cells of fields with a primitive BC hold the primitive after the
// per-dim fills; multiply them by the halo density, instance by instance
// (tiles: one per 256 padded cells of an instance).
hipError_t launch_batch_halo_prim_finish(const HaloArgs& cfg, float* const* q_base,
                                         const BatchHaloInst* inst, const BatchHaloTile* tiles,
                                         int ntiles, int grid, hipStream_t stream);

struct SpeedArgs {
  const float* q[5];
  int nx, ny, nz, h;
  long long sx, sz;
  int has_pe;
  const float* pe_x;
  const float* pe_y;
  const float* pe_z;
};

// dt = cfl * hmin / (max speed held in max_bits), then max_bits = 0.
hipError_t launch_cfl_dt_reset(unsigned* max_bits, float cfl, float hmin, float* dt,
                               hipStream_t stream);

// The same per instance: dt[i] = cfl * hmin[i] / (bits[i] as a float), then
// bits[i] = 0.
hipError_t launch_batch_cfl_dt_reset(unsigned* bits, int n_inst, float cfl, const float* hmin,
                                     float* dt, hipStream_t stream);

// *max_bits = bits of max over interior of max_d(|u_d| + c); then
// *dt = cfl * hmin / max if dt is not null.
hipError_t launch_max_speed(const SpeedArgs& a, unsigned* max_bits, float cfl,
                            float hmin, float* dt, hipStream_t stream);

}  // namespace sc
