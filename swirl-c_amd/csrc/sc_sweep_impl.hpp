// Fused per-direction sweep kernels of the RHS (gfx950).
//
// One launch per direction.  Each launch reads the conserved fields of the
// stage state once, converts them to primitives, reconstructs the face states
// (WENO / MUSCL), solves the Riemann problem (HLL / HLLC), adds the viscous
// and heat flux and takes the flux difference, in the reference's x, y, z
// order (equations/compressible_navier_stokes.py:255-264):
//   x sweep:  R = src - (F_x[i+1] - F_x[i]) / dx            (writes R)
//   y sweep:  R = R - (F_y[j+1] - F_y[j]) / dy               (updates R)
//   z sweep:  R - (F_z[k+1] - F_z[k]) / dz, then either writes R or the
//             Runge-Kutta stage update of the state (RK axpy folded in).
// The RK operands (R, Uk, U0) are read by the z sweep: the marching sweeps
// have memory bandwidth to spare (their arithmetic sets their time), while
// the y sweep, whose windows re-read overlapping positions, is the one the
// extra bytes would slow down.
// Layout [z][x][y], y contiguous.  The x and z sweeps map 64 lanes to 64
// consecutive y cells (coalesced) and march each thread along its pencil
// with a register sliding window; the y sweep maps lanes along y and gets
// the stencil neighbours through DPP lane shifts.
#pragma once

#include "sc_device.hpp"

namespace sc {

struct Fields {
  float* p[kMaxFields];
};
struct CFields {
  const float* p[kMaxFields];
};

// Launch geometry of the sweeps (the tunable knobs, set per direction).
struct SweepTuning {
  int seg_x = 32;   // cells per thread, x sweep
  int by_x = 4;     // waves per block, x sweep
  int seg_z = 32;
  int by_z = 4;
  int by_y = 4;     // waves per block, y sweep
  int xcd_remap = 1;  // 1: XCD-aware block -> tile order (single-grid sweeps)
  int adj_seg = 16;   // adjoint marching passes: owned cells per thread
  int adj_by = 4;     // adjoint marching passes: waves per block
  int adj_by_y = 4;   // adjoint y pass: waves per block
  int seg_y = 16;     // y sweep: z planes per wave
  // Profiling only: when set, each sweep launch records its own start and
  // stop time into these events (hipExtLaunchKernel), adding no packets.
  hipEvent_t ev_start = nullptr;
  hipEvent_t ev_stop = nullptr;
};

// Dirichlet values on the domain's boundary faces of one sweep direction
// (boundary/boundary.py:89-183 intercell fluxes, :230-461 cell faces).
constexpr int kMaxFaceEntries = 8;
// A value is a scalar, or (plane != null) one value per cell of the boundary
// plane, [A][B] over the full extents of the two other axes in [z][x][y]
// order: x faces [z][y], y faces [z][x], z faces [x][y].
struct FaceEntry {
  int kind;   // 1: conserved field (face primitive = value / rho), 2: primitive
  int var;    // index into the reconstructed primitives (1 u .. 6 h, 7+s scalar)
  float val;
  const float* plane;
};
struct BoundaryBc {
  int n_face;                         // face-state entries, applied in order
  FaceEntry face[kMaxFaceEntries];
  int flux_mask;                      // bit 8*t + f: flux type t of field f set
  float flux_val[3][kMaxFields];      // t: 0 convective, 1 diffusive, 2 total
  const float* flux_plane[3][kMaxFields];
};

// Block geometry of one tile: block indices (bx, by, bz) of the plain launch
// and the thread's (tx, ty) with by_dim waves per block.
struct TileIdx {
  int bx, by, bz;
  int tx, ty, by_dim;
};

struct SweepArgs {
  CFields q;    // stage state uk, halos filled
  CFields u0;   // RK base state
  Fields r;     // RHS accumulator
  Fields out;   // RK stage output
  int nx, ny, nz, h;
  int NX, NY, NZ;
  long long sx, sz;  // strides of x and z (y stride is 1)
  float inv_hdir;    // 1 / spacing along the sweep direction
  int seg_len;       // cells per thread (marching sweeps)
  int has_pe;        // gravitational potential present
  const float* pe_x; // pe = (pe_x[i] - pe_y[j]) - pe_z[k]
  const float* pe_y;
  const float* pe_z;
  int grav_src;      // gravity source active
  int grav_mask;     // bit d: momentum d receives the source
  float gdir[3];     // cfg.g components
  int prim_bc_mask;  // fields whose interior is rho*(q/rho)-rounded
  int rk_form;       // 0: write R; SC_RK_* otherwise
  float c0, c1;
  const float* dt;
  ViscConst vc;
  int bbc_dims;             // bit d: boundary face/flux values along dim d
  const BoundaryBc* bbc;    // device table [dim][side]
  int xcd_remap;            // XCD-aware block order (see tile_of_block)
  int seg_y;                // y sweep: z planes per wave
  // Hydrostatic reference state (helper_var p_ref / rho_ref), or null:
  // p_dyn = p - p_ref (zero in halo cells) replaces p in the momentum flux,
  // d_rho = rho - rho_ref replaces rho in the gravity source
  // (compressible_navier_stokes.py:273-335).
  const float* p_ref;
  const float* rho_ref;
  // Last sweep of an RK stage: when set, the largest wave speed of the
  // stage's output cells is max-accumulated here (float bits).
  unsigned* speed_bits;
  // Last sweep of an RK stage on a grid whose halo fill is a pure copy along
  // every dim (periodic: the wrapped interior cell, extent >= 2h; NEUMANN
  // with value 0: the nearest interior cell): also write each output cell's
  // images (the halo cells the fill would copy from it), so out leaves with
  // its halos filled.  Bit d of img_clamp set: dim d clamps (NEUMANN 0),
  // else it wraps (periodic).
  int out_images;
  int img_clamp;
};

// The dispatcher hands consecutive workgroups to the 8 XCDs in turn, and each
// XCD has its own L2.  With xcd_remap the linear block id is permuted so XCD k
// runs the k-th contiguous range of tiles (x-fastest tile order): neighbouring
// tiles, which share halo rows and transverse-neighbour planes, then hit the
// same L2.  The map is a bijection on [0, nblocks).
constexpr int kNumXcd = 8;
__device__ __forceinline__ TileIdx tile_of_block(int remap) {
  int bx = (int)blockIdx.x, by = (int)blockIdx.y, bz = (int)blockIdx.z;
  if (remap) {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int n = gx * gy * (int)gridDim.z;
    const int lin = bx + gx * (by + gy * bz);
    const int q = n / kNumXcd, r = n % kNumXcd;
    const int xcd = lin % kNumXcd, slot = lin / kNumXcd;
    const int t = (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + slot;
    // The divisions are lowered to vector instructions; their results are
    // uniform, and readfirstlane puts them back in scalar registers so the
    // addresses derived from them stay scalar.
    bx = __builtin_amdgcn_readfirstlane(t % gx);
    by = __builtin_amdgcn_readfirstlane((t / gx) % gy);
    bz = __builtin_amdgcn_readfirstlane(t / (gx * gy));
  }
  return TileIdx{bx, by, bz, (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.y};
}

// LEAN builds drop run-time checks from the kernels: kLeanPlain (no
// gravitational potential, no gravity source, no BC given on a primitive) and
// kLeanGrav (potential and gravity source both on, no BC on a primitive);
// LEAN = 0 decides all three at run time.
constexpr int kLeanPlain = 1;
constexpr int kLeanGrav = 2;
template <int LEAN = 0>
__device__ __forceinline__ float pe_at(const SweepArgs& a, int k, int i, int j) {
  if constexpr (LEAN == kLeanPlain) return 0.0f;
  if constexpr (LEAN == kLeanGrav) return (a.pe_x[i] - a.pe_y[j]) - a.pe_z[k];
  return a.has_pe ? (a.pe_x[i] - a.pe_y[j]) - a.pe_z[k] : 0.0f;
}

// Loads the conserved fields of one cell; interior cells of fields with a
// primitive BC get the reference's rho * (q / rho) rounding
// (boundary/boundary.py:516-531).
template <int NF>
__device__ __forceinline__ void load_raw(const SweepArgs& a, long long idx, float* q) {
#pragma unroll
  for (int f = 0; f < NF; ++f) q[f] = a.q.p[f][idx];
}

template <int NF, int LEAN = 0>
__device__ __forceinline__ void prim_bc_round(const SweepArgs& a, bool interior, float* q) {
  if constexpr (LEAN) return;
  if (a.prim_bc_mask && interior) {
#pragma unroll
    for (int f = 1; f < NF; ++f)
      if (a.prim_bc_mask & (1 << f)) q[f] = q[0] * dnn_exact(q[f], q[0]);
  }
}

template <int NF, int LEAN = 0>
__device__ __forceinline__ void load_cons(const SweepArgs& a, long long idx,
                                          bool interior, float* q) {
  load_raw<NF>(a, idx, q);
  prim_bc_round<NF, LEAN>(a, interior, q);
}

// Dynamic pressure of a cell (slot 7 + NS): exchange_halos with zero
// DIRICHLET values on every face leaves 0 in every halo cell.
template <int NS, bool PD>
__device__ __forceinline__ void set_pdyn(const SweepArgs& a, long long idx, bool interior,
                                         float* w) {
  if constexpr (PD) w[7 + NS] = interior ? w[kP] - a.p_ref[idx] : 0.0f;
}

template <int NS, bool PD = false, int LEAN = 0>
__device__ __forceinline__ void load_prim(const SweepArgs& a, long long idx,
                                          bool interior, float pe, float* w) {
  constexpr int NF = 5 + NS;
  float q[NF];
  load_cons<NF, LEAN>(a, idx, interior, q);
  cons_to_prim<NS>(q, pe, w);
  set_pdyn<NS, PD>(a, idx, interior, w);
}

// Velocity components u_c0 and u_c1 at a cell (for transverse derivatives).
template <int LEAN = 0>
__device__ __forceinline__ void load_vel2(const SweepArgs& a, long long idx,
                                          bool interior, int c0, int c1,
                                          float& v0, float& v1) {
  const float rho = a.q.p[0][idx];
  float m0 = a.q.p[1 + c0][idx];
  float m1 = a.q.p[1 + c1][idx];
  if (!LEAN && a.prim_bc_mask && interior) {
    if (a.prim_bc_mask & (1 << (1 + c0))) m0 = rho * dnn_exact(m0, rho);
    if (a.prim_bc_mask & (1 << (1 + c1))) m1 = rho * dnn_exact(m1, rho);
  }
  const float inv = frcp(rho);
  v0 = inv * m0;
  v1 = inv * m1;
}

__device__ __forceinline__ bool in_core(int c, int h, int n) {
  return c >= h && c < h + n;
}

// Central transverse derivatives at cell (k, i, j) for a face normal to D:
// out[t][0] = d(u_t)/dx_t, out[t][1] = d(u_D)/dx_t for the transverse dims t
// in x, y, z order (physics/fluid.py:52-63 with numerics/gradient.py:71-94).
template <int D, int LEAN = 0>
__device__ __forceinline__ void transverse_grads(const SweepArgs& a, int k, int i,
                                                 int j, float (*out)[2]) {
  int t = 0;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) {
    if (dd == D) continue;
    long long sp, sm;
    bool inp, inm;
    const long long c = (long long)k * a.sz + (long long)i * a.sx + j;
    if (dd == 0) {
      sp = c + a.sx;
      sm = c - a.sx;
      inp = in_core(i + 1, a.h, a.nx) && in_core(j, a.h, a.ny) && in_core(k, a.h, a.nz);
      inm = in_core(i - 1, a.h, a.nx) && in_core(j, a.h, a.ny) && in_core(k, a.h, a.nz);
    } else if (dd == 1) {
      sp = c + 1;
      sm = c - 1;
      inp = in_core(i, a.h, a.nx) && in_core(j + 1, a.h, a.ny) && in_core(k, a.h, a.nz);
      inm = in_core(i, a.h, a.nx) && in_core(j - 1, a.h, a.ny) && in_core(k, a.h, a.nz);
    } else {
      sp = c + a.sz;
      sm = c - a.sz;
      inp = in_core(i, a.h, a.nx) && in_core(j, a.h, a.ny) && in_core(k + 1, a.h, a.nz);
      inm = in_core(i, a.h, a.nx) && in_core(j, a.h, a.ny) && in_core(k - 1, a.h, a.nz);
    }
    float ap0, ap1, am0, am1;
    load_vel2<LEAN>(a, sp, inp, dd, D, ap0, ap1);
    load_vel2<LEAN>(a, sm, inm, dd, D, am0, am1);
    const float ith = a.vc.inv_two_h[dd];
    out[t][0] = (ap0 - am0) * ith;
    out[t][1] = (ap1 - am1) * ith;
    ++t;
  }
}

// ---------------------------------------------------------------------------
// Wave-uniform addressing.  Every access of the sweeps is a wave-uniform base
// (a field pointer plus a uniform element offset, held in scalar registers)
// plus the lane's 32-bit byte offset, which global loads and stores take
// directly (the scalar-base form), so no 64-bit vector address arithmetic is
// issued per access.  Two compiler habits defeat that form and are blocked
// here: it hoists the extension of a loop-invariant lane offset out of a loop
// and then adds it to each base in vector registers (lane_off() redefines the
// offset in the block that uses it; this may cost one register move), and it
// re-associates base + uniform offset + lane offset so that the sum of the
// two offsets, shared by several fields, is formed once in vector registers
// (wave_base() passes the base through readfirstlane, which the compiler
// cannot see through and later drops, since its operand is already uniform).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned lane_off(unsigned lb) {
  asm volatile("" : "+v"(lb));
  return lb;
}

// Orders a march iteration's work: the value is taken as redefined here, so
// its computation stays before this point and its uses after it (an empty
// asm statement; volatile ones keep their relative order).  Without it the
// compiler sinks the reconstruction into the flux block and computes the
// transverse derivatives first, which makes every iteration wait for the
// neighbour loads it has just issued.
__device__ __forceinline__ void pin(float& x) { asm volatile("" : "+v"(x)); }

typedef __attribute__((address_space(1))) float global_float;
typedef __attribute__((address_space(1))) char global_char;

// A wave-uniform 64-bit value moved to scalar registers (the compiler may
// otherwise keep a loop-invariant uniform value in vector registers, and the
// addresses formed from it then become vector arithmetic).  v must be the
// same in every lane of the wave.
__device__ __forceinline__ long long uniform_ll(long long v) {
  const unsigned long long u = (unsigned long long)v;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// The base pointer p as a global-memory pointer.  p must be the same in every
// lane of the wave (readfirstlane takes the first active lane's value).
__device__ __forceinline__ global_float* wave_base(const float* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (global_float*)(((unsigned long long)hi << 32) | lo);
}

// Element at byte offset boff of the array starting at `base` (base
// wave-uniform, boff the lane's offset).
__device__ __forceinline__ float ld_plane(const float* base, unsigned boff) {
  return *(const global_float*)((const global_char*)wave_base(base) + boff);
}
__device__ __forceinline__ void st_plane(float* base, unsigned boff, float v) {
  *(global_float*)((global_char*)wave_base(base) + boff) = v;
}
// The same accesses without readfirstlane: the compiler may then form one
// scalar base for accesses whose offsets differ by a constant, so fewer
// scalar base registers are held (for the y sweep, whose many bases would
// otherwise spill scalar registers).
__device__ __forceinline__ float ld_lane(const float* base, unsigned boff) {
  return *(const float*)((const char*)base + boff);
}
__device__ __forceinline__ void st_lane(float* base, unsigned boff, float v) {
  *(float*)((char*)base + boff) = v;
}

// The loads above with a uniform element offset u: cell u + lane.
template <int NF>
__device__ __forceinline__ void load_raw_u(const SweepArgs& a, long long u, unsigned lb0,
                                           float* q) {
  const unsigned lb = lane_off(lb0);
#pragma unroll
  for (int f = 0; f < NF; ++f) q[f] = ld_plane(a.q.p[f] + u, lb);
}

// Potential at (k, i, jb + lane) with k and i uniform.
template <int LEAN = 0>
__device__ __forceinline__ float pe_at_u(const SweepArgs& a, int k, int i, int jb,
                                         unsigned lb0) {
  if constexpr (LEAN == kLeanPlain) return 0.0f;
  if constexpr (LEAN == kLeanGrav)
    return (a.pe_x[i] - ld_plane(a.pe_y + jb, lane_off(lb0))) - a.pe_z[k];
  return a.has_pe ? (a.pe_x[i] - ld_plane(a.pe_y + jb, lane_off(lb0))) - a.pe_z[k] : 0.0f;
}

template <int NS, bool PD>
__device__ __forceinline__ void set_pdyn_u(const SweepArgs& a, long long u, unsigned lb0,
                                           bool interior, float* w) {
  if constexpr (PD) w[7 + NS] = interior ? w[kP] - ld_plane(a.p_ref + u, lane_off(lb0)) : 0.0f;
}

// transverse_grads for a marching sweep's cell (k, i, j) whose lane-0 cell
// sits at uniform element offset u (j = first y cell of the wave + lane), in
// two halves: the loads (rho, u_t, u_D momenta of the four neighbours), issued
// early in a march iteration, and the derivatives, formed once the
// reconstruction has covered the loads' latency.
struct TransRaw {
  float v[2][2][3];  // [transverse dim t][side + / -][rho, m_t, m_D]
};

template <int D>
__device__ __forceinline__ void transverse_load_u(const SweepArgs& a, long long u, unsigned lb0,
                                                  TransRaw& r) {
  const unsigned lb = lane_off(lb0);
  int t = 0;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) {
    if (dd == D) continue;
    const long long st = dd == 0 ? a.sx : (dd == 1 ? 1 : a.sz);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const long long un = s == 0 ? u + st : u - st;
      r.v[t][s][0] = ld_plane(a.q.p[0] + un, lb);
      r.v[t][s][1] = ld_plane(a.q.p[1 + dd] + un, lb);
      r.v[t][s][2] = ld_plane(a.q.p[1 + D] + un, lb);
    }
    ++t;
  }
}

template <int D, int LEAN = 0>
__device__ __forceinline__ void transverse_grads_from(const SweepArgs& a, int k, int i, int j,
                                                      const TransRaw& r, float (*out)[2]) {
  int t = 0;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) {
    if (dd == D) continue;
    float w[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int o = s == 0 ? 1 : -1;
      const bool in = in_core(i + (dd == 0 ? o : 0), a.h, a.nx) &&
                      in_core(j + (dd == 1 ? o : 0), a.h, a.ny) &&
                      in_core(k + (dd == 2 ? o : 0), a.h, a.nz);
      const float rho = r.v[t][s][0];
      float m0 = r.v[t][s][1];
      float m1 = r.v[t][s][2];
      if (!LEAN && a.prim_bc_mask && in) {
        if (a.prim_bc_mask & (1 << (1 + dd))) m0 = rho * dnn_exact(m0, rho);
        if (a.prim_bc_mask & (1 << (1 + D))) m1 = rho * dnn_exact(m1, rho);
      }
      const float inv = frcp(rho);
      w[s][0] = inv * m0;
      w[s][1] = inv * m1;
    }
    const float ith = a.vc.inv_two_h[dd];
    out[t][0] = (w[0][0] - w[1][0]) * ith;
    out[t][1] = (w[0][1] - w[1][1]) * ith;
    ++t;
  }
}

// Gravity source of field f at a cell (physics_models.py:94-143,
// gravity.py:64-102): (rho * G) * g_d on the selected momenta, else 0.
template <int LEAN = 0>
__device__ __forceinline__ float source_of(const SweepArgs& a, int f, float rho) {
  if constexpr (LEAN == kLeanPlain) return 0.0f;
  if ((LEAN == kLeanGrav || a.grav_src) && f >= 1 && f <= 3 &&
      (a.grav_mask & (1 << (f - 1)))) {
    return (rho * kG) * a.gdir[f - 1];
  }
  return 0.0f;
}

// Face-state overwrite of boundary_face_update_1d on one side's reconstructed
// primitives: a conserved value gives primitive = divide_no_nan(value, rho),
// a primitive value is taken as is.
// pidx: the face's index in the boundary plane's [A][B] values.
template <int NS>
__device__ __forceinline__ void apply_face_bc(const BoundaryBc& b, float* W, long long pidx) {
  constexpr int NV = NVar<NS>::value;
  for (int e = 0; e < b.n_face; ++e) {
    const FaceEntry fe = b.face[e];
    const float val = fe.plane != nullptr ? fe.plane[pidx] : fe.val;
    const float v = fe.kind == 1 ? dnn_exact(val, W[kRho]) : val;
#pragma unroll
    for (int k = 1; k < NV; ++k)
      if (k == fe.var) W[k] = v;
  }
}

template <int NF>
__device__ __forceinline__ void apply_flux_bc(const BoundaryBc& b, int t, float* F,
                                              long long pidx) {
#pragma unroll
  for (int f = 0; f < NF; ++f)
    if (b.flux_mask & (1 << (8 * t + f)))
      F[f] = b.flux_plane[t][f] != nullptr ? b.flux_plane[t][f][pidx] : b.flux_val[t][f];
}

// Total flux F of one face: Riemann flux of the face states L, R plus the
// viscous and heat flux.  On a boundary face with Dirichlet values (b not
// null) the face states are overwritten on both sides first, and the
// convective flux, the diffusive flux (viscous runs only) and their sum are
// overwritten in the reference's order (compressible_navier_stokes.py:147-234).
template <int DIR, int FLUX, bool VISC, int NS, bool PD>
__device__ __forceinline__ void face_flux(const BoundaryBc* b, long long pidx, float* L,
                                          float* R, float pe, const float* A, const float* B,
                                          const float (*ta)[2], const float (*tb)[2],
                                          ViscCell ca, ViscCell cb,
                                          const ViscConst& vc, float* F) {
  constexpr int NF = 5 + NS;
  if (b != nullptr) {
    apply_face_bc<NS>(*b, L, pidx);
    apply_face_bc<NS>(*b, R, pidx);
  }
  riemann<FLUX, DIR, NS, PD>(L, R, pe, F);
  if (b != nullptr) apply_flux_bc<NF>(*b, 0, F, pidx);
  if constexpr (VISC) {
    if (b != nullptr) {
      float Fd[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) Fd[f] = 0.0f;
      viscous_flux_cells<DIR, NS>(A, B, ta, tb, ca, cb, vc, Fd);
      apply_flux_bc<NF>(*b, 1, Fd, pidx);
#pragma unroll
      for (int f = 0; f < NF; ++f) F[f] = F[f] + Fd[f];
    } else {
      viscous_flux_cells<DIR, NS>(A, B, ta, tb, ca, cb, vc, F);
    }
  }
  if (b != nullptr) apply_flux_bc<NF>(*b, 2, F, pidx);
}

// ---------------------------------------------------------------------------
// Marching sweep along x (DIR 0) or z (DIR 2).  Block (64, BY): lane -> y,
// threadIdx.y -> the other transverse index, blockIdx.z -> segment of
// seg_len cells along the sweep.  PHASE 0: first sweep (writes R = src - div);
// PHASE 2: last sweep (R - div into R, or the Runge-Kutta stage update of the
// state into out).
// ---------------------------------------------------------------------------

// Operands of the last sweep's output cell: R, and for RK forms 2 and 3 the
// base state u0 (the stage state uk comes from the march's own registers).
template <int NF>
struct OutOps {
  float r[NF], u0[NF];
};

template <int NF>
__device__ __forceinline__ void load_out_ops(const SweepArgs& a, long long u, unsigned lb,
                                             OutOps<NF>& o) {
  {
    const unsigned l = lane_off(lb);
#pragma unroll
    for (int f = 0; f < NF; ++f) o.r[f] = ld_plane(a.r.p[f] + u, l);
  }
  if (a.rk_form == 2 || a.rk_form == 3) {
    const unsigned l = lane_off(lb);
#pragma unroll
    for (int f = 0; f < NF; ++f) o.u0[f] = ld_plane(a.u0.p[f] + u, l);
  }
}

// Stores R - div (rk_form 0) into R, else the RK stage update (rk3.py:69-83,
// rk2.py:67-86): 1 out = uk + L*dt; 2 out = u0*c0 + (uk + L*dt)*c1;
// 3 out = 0.5*(u0 + (uk + dt*L)); one branch on the form for all fields.
// Halo images of the output cell (k, i, j) whose lane-0 cell sits at uniform
// element offset u (k, i wave-uniform, j per lane): the halo cells that the
// sequential fill (x, then y, then z, each over full planes) leaves equal to
// it when every dim's fill is a pure copy.  Along a wrapping dim (periodic,
// extent n >= 2h) the cell at index c has one image, c + n when c < 2h and
// c - n when c >= n; along a clamping dim (NEUMANN with value 0) the first
// interior cell has the h images below it and the last the h above it.  A
// halo cell's value is then the interior cell at its wrapped / clamped index
// in every dim, so the images are every combination of the per-dim images
// (the cell itself included, itself excluded): each halo cell is written
// once, by the lane of the cell it copies.
__device__ __forceinline__ int image_count(int c, int n, int h, bool clamp) {
  if (!clamp) return (c < 2 * h || c >= n) ? 1 : 0;
  return (c == h ? h : 0) + (c == h + n - 1 ? h : 0);
}
// Index shift of the t-th image (0 <= t < image_count).
__device__ __forceinline__ int image_shift(int c, int n, int h, bool clamp, int t) {
  if (!clamp) return c < 2 * h ? n : -n;
  if (c == h && t < h) return -(t + 1);
  return t - (c == h ? h : 0) + 1;
}

// All-periodic grids (extents >= 2h): along each dim a cell within h of the
// low face (c < 2h) has its image n further, one within h of the high face
// (c >= n) n back, and the images are every combination of those shifts.
template <int NF>
__device__ __forceinline__ void store_images_wrap(const SweepArgs& a, long long u,
                                                  unsigned lb, int k, int i, int j,
                                                  const float* res) {
  const int h2 = 2 * a.h;
  const long long zs = k < h2 ? (long long)a.nz * a.sz : (k >= a.nz ? -(long long)a.nz * a.sz : 0);
  const long long xs = i < h2 ? (long long)a.nx * a.sx : (i >= a.nx ? -(long long)a.nx * a.sx : 0);
  const bool ylo = j < h2, yhi = j >= a.ny;
  const long long lane_byte = (long long)lane_off(lb);
#pragma unroll
  for (int zm = 0; zm < 2; ++zm) {
    if (zm && zs == 0) continue;
#pragma unroll
    for (int xm = 0; xm < 2; ++xm) {
      if (xm && xs == 0) continue;
      const long long vb = (u + (zm ? zs : 0) + (xm ? xs : 0)) * 4 + lane_byte;
#pragma unroll
      for (int ym = -1; ym <= 1; ++ym) {
        if (ym == 0 ? !(zm || xm) : !(ym > 0 ? ylo : yhi)) continue;
        const long long b = vb + (long long)ym * a.ny * 4;
#pragma unroll
        for (int f = 0; f < NF; ++f) *(float*)((char*)a.out.p[f] + b) = res[f];
      }
    }
  }
}

// CLAMP: the build that also takes clamping dims (a.img_clamp, the batched
// sweeps); the single-grid sweeps, whose images are periodic only, leave
// that code out.
template <int NF, bool CLAMP>
__device__ __forceinline__ void store_images(const SweepArgs& a, long long u, unsigned lb,
                                             int k, int i, int j, const float* res) {
  if (!CLAMP || a.img_clamp == 0) {
    store_images_wrap<NF>(a, u, lb, k, i, j, res);
    return;
  }
  const int h = a.h;
  const bool cz = (a.img_clamp >> 2) & 1, cx = a.img_clamp & 1, cy = (a.img_clamp >> 1) & 1;
  // Along z and x (wave-uniform) the images are counted; along y (per lane)
  // the low / high side of the lane's cell has 1 (wrap) or h (clamp) images.
  const int nzi = image_count(k, a.nz, h, cz), nxi = image_count(i, a.nx, h, cx);
  const bool ylo = cy ? j == h : j < 2 * h;
  const bool yhi = cy ? j == h + a.ny - 1 : j >= a.ny;
  const int nyi = cy ? h : 1;
  // Rare stores: their addresses are formed per lane in vector registers
  // (the byte offset of cell u + lane plus the image shift), which holds no
  // scalar base per image.
  const long long lane_byte = (long long)lane_off(lb);
  for (int tz = -1; tz < nzi; ++tz) {
    const long long zs = tz < 0 ? 0 : (long long)image_shift(k, a.nz, h, cz, tz) * a.sz;
    for (int tx = -1; tx < nxi; ++tx) {
      const long long xs = tx < 0 ? 0 : (long long)image_shift(i, a.nx, h, cx, tx) * a.sx;
      const long long vb = (u + zs + xs) * 4 + lane_byte;
      if (tz >= 0 || tx >= 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) *(float*)((char*)a.out.p[f] + vb) = res[f];
      }
      if (ylo || yhi) {
        for (int t = 0; t < nyi; ++t) {
          if (ylo) {
            const long long b = vb + (long long)(cy ? -(t + 1) : a.ny) * 4;
#pragma unroll
            for (int f = 0; f < NF; ++f) *(float*)((char*)a.out.p[f] + b) = res[f];
          }
          if (yhi) {
            const long long b = vb + (long long)(cy ? t + 1 : -a.ny) * 4;
#pragma unroll
            for (int f = 0; f < NF; ++f) *(float*)((char*)a.out.p[f] + b) = res[f];
          }
        }
      }
    }
  }
}

// With a.speed_bits set, the output cell's wave speed is max-folded into
// smax (pe: the cell's gravitational potential); with a.out_images set, the
// cell's periodic images are written too.
template <int NF, bool CLAMP>
__device__ __forceinline__ void store_out(const SweepArgs& a, long long u, unsigned lb,
                                          const OutOps<NF>& o, const float* uk,
                                          const float* div, float dt, float pe,
                                          float& smax, int k, int i, int j) {
  float val[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) val[f] = o.r[f] - div[f] * a.inv_hdir;
  if (a.rk_form == 0) {
    const unsigned l = lane_off(lb);
#pragma unroll
    for (int f = 0; f < NF; ++f) st_plane(a.r.p[f] + u, l, val[f]);
    return;
  }
  float res[NF];
  if (a.rk_form == 2) {
#pragma unroll
    for (int f = 0; f < NF; ++f) res[f] = o.u0[f] * a.c0 + (uk[f] + val[f] * dt) * a.c1;
  } else if (a.rk_form == 1) {
#pragma unroll
    for (int f = 0; f < NF; ++f) res[f] = uk[f] + val[f] * dt;
  } else {
#pragma unroll
    for (int f = 0; f < NF; ++f) res[f] = 0.5f * (o.u0[f] + (uk[f] + dt * val[f]));
  }
  {
    const unsigned l = lane_off(lb);
#pragma unroll
    for (int f = 0; f < NF; ++f) st_plane(a.out.p[f] + u, l, res[f]);
  }
  if (a.out_images) store_images<NF, CLAMP>(a, u, lb, k, i, j, res);
  if (a.speed_bits != nullptr) smax = fmaxf(smax, wave_speed(res, pe));
}

// BB: build with the boundary face / flux value handling (only launched when
// the configuration has such values, so other runs carry none of its code).
template <int DIR, int SCHEME, int FLUX, bool VISC, int NS, int PHASE, bool PD = false,
          bool BB = false, int LEAN = 0, bool CLAMP = false>
__device__ __forceinline__ void sweep_march_body(const SweepArgs& a, const TileIdx& ti) {
  constexpr int RAD = Radius<SCHEME>::value;
  constexpr int WN = 2 * RAD + 1;
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  constexpr int NF = 5 + NS;
  // Lanes run along y (64 consecutive cells from jb); the block's waves differ
  // in threadIdx.y only (blocks are 64 wide), so o is wave-uniform.
  const int jb = a.h + ti.bx * 64;
  const int j = jb + ti.tx;
  const int o = a.h + ti.by * ti.by_dim + __builtin_amdgcn_readfirstlane(ti.ty);
  const int n_march = DIR == 0 ? a.nx : a.nz;
  const int n_other = DIR == 0 ? a.nz : a.nx;
  if (j >= a.h + a.ny || o >= a.h + n_other) return;
  const int c_begin = a.h + ti.bz * a.seg_len;
  if (c_begin >= a.h + n_march) return;
  const int c_end = min(c_begin + a.seg_len, a.h + n_march);
  const long long sm = DIR == 0 ? a.sx : a.sz;
  const long long so = DIR == 0 ? a.sz : a.sx;
  // Uniform element offset of the wave's lane-0 cell at march coordinate c,
  // and the lane's byte offset from it.
  const long long ub = uniform_ll((long long)o * so + jb);
  auto uo = [&](int c) { return ub + (long long)c * sm; };
  const unsigned lb = (unsigned)ti.tx * 4u;
  // Cell (k, i, j) of march coordinate c.
  auto kk = [&](int c) { return DIR == 0 ? o : c; };
  auto ii = [&](int c) { return DIR == 0 ? c : o; };

  // Window of primitives, cells c-RAD .. c+RAD at iteration c; the first
  // WN-1 cells are converted here (cons_to_prim rounds the same in every
  // copy, so results do not depend on where a segment starts).
  float W[WN][NV];
  // The last sweep keeps the raw stage values of cells c - 1 .. c + RAD - 1
  // at iteration c: hq[0] is the output cell's uk, read once with the
  // window instead of again from memory.
  constexpr int NH = PHASE == 2 ? RAD + 1 : 1;
  float hq[NH][NF];
#pragma unroll
  for (int w = 0; w < WN - 1; ++w) {
    const int c = c_begin - 1 - RAD + w;
    const bool in = in_core(c, a.h, n_march);
    float q[NF];
    load_raw_u<NF>(a, uo(c), lb, q);
    if constexpr (PHASE == 2) {
      if (w >= RAD - 1) {
#pragma unroll
        for (int f = 0; f < NF; ++f) hq[w - RAD + 1][f] = q[f];
      }
    }
    prim_bc_round<NF, LEAN>(a, in, q);
    cons_to_prim<NS>(q, pe_at_u<LEAN>(a, kk(c), ii(c), jb, lb), W[w]);
    set_pdyn_u<NS, PD>(a, uo(c), lb, in, W[w]);
  }
  float negPrev[NV];
  float Fprev[NF];
  float tPrev[2][2];
  ViscCell vPrev{0.0f, 0.0f};  // heat-flux factors of cell c - 1
#pragma unroll
  for (int v = 0; v < NV; ++v) negPrev[v] = 0.0f;
#pragma unroll
  for (int f = 0; f < NF; ++f) Fprev[f] = 0.0f;
  tPrev[0][0] = tPrev[0][1] = tPrev[1][0] = tPrev[1][1] = 0.0f;

  // Raw conserved values of the cell entering the window, loaded one
  // iteration ahead so the loads overlap the arithmetic of the current cell.
  float qn[NF];
  load_raw_u<NF>(a, uo(c_begin - 1 + RAD), lb, qn);
  // The last sweep's RK time step (a device scalar) and the running maximum
  // of its output cells' wave speeds.
  const float dt = (PHASE == 2 && a.rk_form != 0) ? *a.dt : 0.0f;
  float smax = 0.0f;
  for (int c = c_begin - 1; c <= c_end; ++c) {
    // Loads are issued in the order their values are consumed (the memory
    // counter a wait names is in order): the transverse neighbours (used
    // after the reconstruction), the next window cell (next iteration), and,
    // in the last sweep, the output cell's operands (used at the store).
    TransRaw tr;
    if constexpr (VISC) transverse_load_u<DIR>(a, uo(c), lb, tr);
    float qraw[NF];
    {
      const int cn = c + RAD;
      float qc[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) qc[f] = qn[f];
      if (c < c_end) load_raw_u<NF>(a, uo(cn + 1), lb, qn);
      if constexpr (PHASE == 2) {
        // Enters the history ring at the end of the iteration.
#pragma unroll
        for (int f = 0; f < NF; ++f) qraw[f] = qc[f];
      }
      prim_bc_round<NF, LEAN>(a, in_core(cn, a.h, n_march), qc);
      cons_to_prim<NS>(qc, pe_at_u<LEAN>(a, kk(cn), ii(cn), jb, lb), W[WN - 1]);
      set_pdyn_u<NS, PD>(a, uo(cn), lb, in_core(cn, a.h, n_march), W[WN - 1]);
    }
    float negP[NV], posM[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      float s[WN];
#pragma unroll
      for (int w = 0; w < WN; ++w) s[w] = W[w][v];
      recon<SCHEME>(s, negP[v], posM[v]);
    }
    float tCur[2][2];
    ViscCell vCur{0.0f, 0.0f};
    if constexpr (VISC) {
      // Reconstruction first, then the neighbour values (their loads have
      // had the reconstruction's time to arrive), then the derivatives.
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        pin(negP[v]);
        pin(posM[v]);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int sd = 0; sd < 2; ++sd)
#pragma unroll
          for (int q = 0; q < 3; ++q) pin(tr.v[t][sd][q]);
      transverse_grads_from<DIR, LEAN>(a, kk(c), ii(c), j, tr, tCur);
      vCur = visc_cell(W[RAD], a.vc);
      pin(tCur[0][0]);
      pin(tCur[0][1]);
      pin(tCur[1][0]);
      pin(tCur[1][1]);
    }
    OutOps<NF> oo;
    if constexpr (PHASE == 2) {
      if (c > c_begin) load_out_ops<NF>(a, uo(c - 1), lb, oo);
    }
    if (c >= c_begin) {
      float F[NF];
      const int side = c == a.h ? 0 : (c == a.h + n_march ? 1 : -1);
      const BoundaryBc* bb =
          (BB && ((a.bbc_dims >> DIR) & 1) && side >= 0) ? a.bbc + 2 * DIR + side : nullptr;
      // Boundary plane cell of this face: x faces [z][y], z faces [x][y].
      face_flux<DIR, FLUX, VISC, NS, PD>(bb, (long long)o * a.NY + j, negPrev, posM,
                                         pe_at_u<LEAN>(a, kk(c), ii(c), jb, lb),
                                         W[RAD - 1], W[RAD], tPrev, tCur, vPrev, vCur, a.vc,
                                         F);
      if (c > c_begin) {
        const long long u = uo(c - 1);
        if constexpr (PHASE == 0) {
          const unsigned ls = lane_off(lb);
          // The gravity source takes d_rho = rho - rho_ref with a reference state.
          const float rho =
              PD ? W[RAD - 1][kRho] - ld_plane(a.rho_ref + u, ls) : W[RAD - 1][kRho];
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            st_plane(a.r.p[f] + u, ls,
                     source_of<LEAN>(a, f, rho) - (F[f] - Fprev[f]) * a.inv_hdir);
          }
        } else {
          float div[NF];
#pragma unroll
          for (int f = 0; f < NF; ++f) div[f] = F[f] - Fprev[f];
          store_out<NF, CLAMP>(a, u, lb, oo, hq[0], div, dt,
                        pe_at_u<LEAN>(a, kk(c - 1), ii(c - 1), jb, lb), smax, kk(c - 1),
                        ii(c - 1), j);
        }
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) Fprev[f] = F[f];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) negPrev[v] = negP[v];
    if constexpr (VISC) {
      tPrev[0][0] = tCur[0][0];
      tPrev[0][1] = tCur[0][1];
      tPrev[1][0] = tCur[1][0];
      tPrev[1][1] = tCur[1][1];
      vPrev = vCur;
    }
#pragma unroll
    for (int w = 0; w < WN - 1; ++w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) W[w][v] = W[w + 1][v];
    }
    if constexpr (PHASE == 2) {
#pragma unroll
      for (int m = 0; m < NH - 1; ++m)
#pragma unroll
        for (int f = 0; f < NF; ++f) hq[m][f] = hq[m + 1][f];
#pragma unroll
      for (int f = 0; f < NF; ++f) hq[NH - 1][f] = qraw[f];
    }
  }
  if constexpr (PHASE == 2) {
    // One atomic per wave: the maximum over the wave's active lanes (each
    // lane's maximum is finite or 0: fmaxf drops NaN, as the stand-alone
    // wave-speed kernel does; non-negative floats order as their bits).
    if (a.speed_bits != nullptr && a.rk_form != 0) {
      const unsigned m = __builtin_amdgcn_wave_reduce_max_u32(__float_as_uint(smax), 0);
      if (ti.tx == 0) atomicMax(a.speed_bits, m);  // lane 0 is active if any lane is
    }
  }
}

template <int DIR, int SCHEME, int FLUX, bool VISC, int NS, int PHASE, bool PD = false,
          bool BB = false, int LEAN = 0>
__global__ void __launch_bounds__(256)
sweep_march_kernel(const SweepArgs a) {
  const TileIdx ti = tile_of_block(a.xcd_remap);
  sweep_march_body<DIR, SCHEME, FLUX, VISC, NS, PHASE, PD, BB, LEAN>(a, ti);
}

// ---------------------------------------------------------------------------
// y sweep (contiguous direction, middle).  Block (64, BY): one wave per
// window of 64 consecutive positions of a z plane's run of rows (y_windows);
// the stencil comes from DPP lane shifts (every lane of the wave is active at
// each shift).  Lanes K .. K + OUTW - 1 at interior cells output R - div
// into R.
// ---------------------------------------------------------------------------
template <int SCHEME>
struct YGeom {
  static constexpr int RAD = Radius<SCHEME>::value;
  static constexpr int K = RAD + 1;
  static constexpr int OUTW = 64 - 2 * RAD - 2;
};

// The interior x rows of one z plane are adjacent in memory (row stride NY),
// so the y sweep treats them as one run of nx * NY positions and slides
// 64-lane windows along it with stride OUTW, across row ends: each row's h >= K
// halo cells on either side keep every output's stencil inside its own row.
// Window w's lane K sits at position h + w * OUTW; the last interior cell of
// the plane is position nx * NY - h - 1.  (One window per 58 positions instead
// of ceil(ny / 58) per row: 290 lanes per 256-cell row, not 320.)
__host__ __device__ inline int y_windows(int nx, int ny, int h, int outw) {
  const long long run = (long long)nx * (ny + 2 * h);
  return (int)((run - 2 * h + outw - 1) / outw);
}

// Position of a lane in window `win` of a z plane (see y_windows): its row
// i and column yc, and its byte offset within the plane (below 2^32: a plane
// holds fewer than 2^30 cells, checked at context creation), so loads and
// stores take a wave-uniform plane base and a 32-bit lane offset.
struct YPos {
  int yc, i;
  unsigned boff;
  bool out_lane;
};

template <int SCHEME>
__device__ __forceinline__ YPos y_pos(const SweepArgs& a, int win, int lane) {
  constexpr int K = YGeom<SCHEME>::K;
  constexpr int OUTW = YGeom<SCHEME>::OUTW;
  const int run = a.nx * a.NY;
  const int pos = win * OUTW + a.h - K + lane;
  const int pc = min(pos, run - 1);
  // pc = row * NY + yc: a float estimate of the quotient is off by at most
  // one for row counts below 2^20 (checked at context creation).
  int row = (int)((float)pc * frcp((float)a.NY));
  int yc = pc - row * a.NY;
  if (yc < 0) {
    --row;
    yc += a.NY;
  }
  if (yc >= a.NY) {
    ++row;
    yc -= a.NY;
  }
  YPos p;
  p.yc = yc;
  p.i = a.h + row;
  p.boff = (unsigned)(a.h * (int)a.sx + pc) * 4u;
  p.out_lane = lane >= K && lane < K + OUTW && pos < run && yc >= a.h && yc < a.h + a.ny;
  return p;
}

// Global-memory operands of one lane at one z plane: loaded one plane ahead
// of their use, so a wave's loads for its next plane overlap the arithmetic
// of the current one.
template <int NF, bool VISC, bool PD>
struct YLoad {
  float q[NF];        // stage state (reconstructed)
  float r[NF];        // RHS accumulator
  float tx[2][3];     // rows i + 1, i - 1: rho, rho_u, rho_v
  float pref;         // p_ref (reference-state builds)
};

// Every lane loads (the positions of lanes that do not output are valid cells
// of the plane's interior rows, and plane k may be the first halo plane), so
// no load is predicated.
template <int NF, bool VISC, bool PD>
__device__ __forceinline__ void y_load(const SweepArgs& a, int k, const YPos& p,
                                       YLoad<NF, VISC, PD>& L) {
  const long long pk = (long long)k * a.sz;
  const unsigned b = lane_off(p.boff);
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    L.q[f] = ld_plane(a.q.p[f] + pk, b);
    L.r[f] = ld_lane(a.r.p[f] + pk, b);
  }
  if constexpr (VISC) {
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const long long pn = s == 0 ? pk + a.sx : pk - a.sx;
      L.tx[s][0] = ld_plane(a.q.p[0] + pn, b);
      L.tx[s][1] = ld_plane(a.q.p[1] + pn, b);
      L.tx[s][2] = ld_plane(a.q.p[2] + pn, b);
    }
  }
  if constexpr (PD) L.pref = ld_plane(a.p_ref + pk, b);
}

// Velocity components (u_c0, u_c1) from a loaded (rho, m_c0, m_c1); cells of
// fields with a primitive BC get the reference's rho * (q / rho) rounding.
template <int LEAN>
__device__ __forceinline__ void vel2_of(const SweepArgs& a, float rho, float m0, float m1,
                                        bool interior, int c0, int c1, float& v0, float& v1) {
  if (!LEAN && a.prim_bc_mask && interior) {
    if (a.prim_bc_mask & (1 << (1 + c0))) m0 = rho * dnn_exact(m0, rho);
    if (a.prim_bc_mask & (1 << (1 + c1))) m1 = rho * dnn_exact(m1, rho);
  }
  const float inv = frcp(rho);
  v0 = inv * m0;
  v1 = inv * m1;
}

// One window of the y sweep at plane k from its loaded operands (zp, zm:
// rho, rho_w, rho_v of planes k + 1 and k - 1): R - div_y into R.
template <int SCHEME, int FLUX, bool VISC, int NS, bool PD, bool BB, int LEAN>
__device__ __forceinline__ void y_window(const SweepArgs& a, int k, const YPos& p,
                                         const YLoad<5 + NS, VISC, PD>& L, const float* zp,
                                         const float* zm) {
  constexpr int RAD = YGeom<SCHEME>::RAD;
  constexpr int WN = 2 * RAD + 1;
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  constexpr int NF = 5 + NS;
  const int yc = p.yc, i = p.i;
  const bool interior = in_core(yc, a.h, a.ny);
  float W0[NV];
  {
    float q[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) q[f] = L.q[f];
    prim_bc_round<NF, LEAN>(a, interior, q);
    cons_to_prim<NS>(q, pe_at<LEAN>(a, k, i, yc), W0);
    // Dynamic pressure: zero in every halo cell (Dirichlet-0 helper halos).
    if constexpr (PD) W0[7 + NS] = interior ? W0[kP] - L.pref : 0.0f;
  }
  float negP[NV], posM[NV], Am[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float s[WN];
#pragma unroll
    for (int w = 0; w < WN; ++w) {
      const int off = w - RAD;
      s[w] = lane_shift(W0[v], off);
    }
    Am[v] = s[RAD - 1];
    recon<SCHEME>(s, negP[v], posM[v]);
  }
  float negPrev[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) negPrev[v] = wave_shr1(negP[v]);
  float F[NF];
  float tCur[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, tPrev[2][2];
  ViscCell vCur{0.0f, 0.0f}, vPrev{0.0f, 0.0f};
  if constexpr (VISC) {
    // Central transverse derivatives at (k, i, yc) (physics/fluid.py:52-63
    // with numerics/gradient.py:71-94): [t][0] = d(u_t)/dx_t, [t][1] = dv/dx_t
    // for t = x, z.
    {
      const bool rest = interior && in_core(k, a.h, a.nz);
      float ap0, ap1, am0, am1;
      vel2_of<LEAN>(a, L.tx[0][0], L.tx[0][1], L.tx[0][2], in_core(i + 1, a.h, a.nx) && rest,
                    0, 1, ap0, ap1);
      vel2_of<LEAN>(a, L.tx[1][0], L.tx[1][1], L.tx[1][2], in_core(i - 1, a.h, a.nx) && rest,
                    0, 1, am0, am1);
      const float ith = a.vc.inv_two_h[0];
      tCur[0][0] = (ap0 - am0) * ith;
      tCur[0][1] = (ap1 - am1) * ith;
    }
    {
      const bool rest = in_core(i, a.h, a.nx) && interior;
      float ap0, ap1, am0, am1;
      vel2_of<LEAN>(a, zp[0], zp[1], zp[2], rest && in_core(k + 1, a.h, a.nz), 2, 1, ap0, ap1);
      vel2_of<LEAN>(a, zm[0], zm[1], zm[2], rest && in_core(k - 1, a.h, a.nz), 2, 1, am0, am1);
      const float ith = a.vc.inv_two_h[2];
      tCur[1][0] = (ap0 - am0) * ith;
      tCur[1][1] = (ap1 - am1) * ith;
    }
    vCur = visc_cell(W0, a.vc);
    vPrev = ViscCell{wave_shr1(vCur.T), wave_shr1(vCur.k)};
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    tPrev[t][0] = wave_shr1(tCur[t][0]);
    tPrev[t][1] = wave_shr1(tCur[t][1]);
  }
  const int side = yc == a.h ? 0 : (yc == a.h + a.ny ? 1 : -1);
  const BoundaryBc* bb =
      (BB && ((a.bbc_dims >> 1) & 1) && side >= 0) ? a.bbc + 2 + side : nullptr;
  // Boundary plane cell of this face: y faces [z][x].
  face_flux<1, FLUX, VISC, NS, PD>(bb, (long long)k * a.NX + i, negPrev, posM,
                                   pe_at<LEAN>(a, k, i, yc), Am, W0,
                                   tPrev, tCur, vPrev, vCur, a.vc, F);
  float Fn[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) Fn[f] = wave_shl1(F[f]);
  if (p.out_lane) {
    const long long pk = (long long)k * a.sz;
    const unsigned b = lane_off(p.boff);
#pragma unroll
    for (int f = 0; f < NF; ++f)
      st_lane(a.r.p[f] + pk, b, L.r[f] - (Fn[f] - F[f]) * a.inv_hdir);
  }
}

// A wave holds window (bx, ty) of the launch and marches over z planes
// k0 .. k0 + zseg - 1 (tile bz): the operands of each next plane are loaded
// while the current one is computed (two operand sets used in turn), and the
// z neighbours of the viscous transverse derivatives are the planes before
// and after, already in registers.
template <int SCHEME, int FLUX, bool VISC, int NS, bool PD = false, bool BB = false,
          int LEAN = 0>
__device__ __forceinline__ void sweep_y_body(const SweepArgs& a, const TileIdx& ti, int zseg) {
  constexpr int OUTW = YGeom<SCHEME>::OUTW;
  constexpr int NF = 5 + NS;
  const int lane = ti.tx;
  const int win = ti.bx * ti.by_dim + ti.ty;
  const int k0 = a.h + ti.bz * zseg;
  // uniform per wave
  if (k0 >= a.h + a.nz || win >= y_windows(a.nx, a.ny, a.h, OUTW)) return;
  const int k1 = min(k0 + zseg, a.h + a.nz);
  const YPos p = y_pos<SCHEME>(a, win, lane);
  float zm[3] = {0.0f, 0.0f, 0.0f};
  if constexpr (VISC) {
    const long long pm = (long long)(k0 - 1) * a.sz;
    zm[0] = ld_plane(a.q.p[0] + pm, p.boff);
    zm[1] = ld_plane(a.q.p[3] + pm, p.boff);
    zm[2] = ld_plane(a.q.p[2] + pm, p.boff);
  }
  YLoad<NF, VISC, PD> A, B;
  y_load(a, k0, p, A);
  // Plane k + 1 is always loaded (at the segment's end it supplies only the
  // z neighbour; it is at most the first halo plane).
  for (int k = k0; k < k1; k += 2) {
    y_load(a, k + 1, p, B);
    {
      const float zp[3] = {B.q[0], B.q[3], B.q[2]};
      y_window<SCHEME, FLUX, VISC, NS, PD, BB, LEAN>(a, k, p, A, zp, zm);
      zm[0] = A.q[0];
      zm[1] = A.q[3];
      zm[2] = A.q[2];
    }
    if (k + 1 >= k1) break;
    y_load(a, k + 2, p, A);
    {
      const float zp[3] = {A.q[0], A.q[3], A.q[2]};
      y_window<SCHEME, FLUX, VISC, NS, PD, BB, LEAN>(a, k + 1, p, B, zp, zm);
      zm[0] = B.q[0];
      zm[1] = B.q[3];
      zm[2] = B.q[2];
    }
  }
}

template <int SCHEME, int FLUX, bool VISC, int NS, bool PD = false, bool BB = false,
          int LEAN = 0>
__global__ void __launch_bounds__(256)
sweep_y_kernel(const SweepArgs a) {
  const TileIdx ti = tile_of_block(a.xcd_remap);
  sweep_y_body<SCHEME, FLUX, VISC, NS, PD, BB, LEAN>(a, ti, a.seg_y);
}

// ---------------------------------------------------------------------------
// HLLC branch codes of the interior faces normal to D (parity/debug): one
// thread per face, straightforward stencil loads.
// ---------------------------------------------------------------------------
template <int D, int SCHEME, int NS, bool PD = false>
__global__ void __launch_bounds__(256)
branch_codes_kernel(const SweepArgs a, uint8_t* codes) {
  constexpr int RAD = Radius<SCHEME>::value;
  constexpr int WN = 2 * RAD + 1;
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  const long long nfx = a.nx + (D == 0), nfy = a.ny + (D == 1), nfz = a.nz + (D == 2);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nfx * nfy * nfz) return;
  const int j = a.h + (int)(t % nfy);
  const int i = a.h + (int)((t / nfy) % nfx);
  const int k = a.h + (int)(t / (nfy * nfx));
  const long long s = D == 0 ? a.sx : (D == 1 ? 1 : a.sz);
  const long long c = (long long)k * a.sz + (long long)i * a.sx + j;
  const int pc = D == 0 ? i : (D == 1 ? j : k);
  const int n = D == 0 ? a.nx : (D == 1 ? a.ny : a.nz);
  float Wl[NV], Wr[NV];
  float win[WN + 1][NV];
#pragma unroll
  for (int w = 0; w < WN + 1; ++w) {
    const int off = w - RAD - 1;  // cells c-1-RAD .. c+RAD
    const int kk = k + (D == 2) * off, ii = i + (D == 0) * off, jj = j + (D == 1) * off;
    load_prim<NS, PD>(a, c + off * s, in_core(pc + off, a.h, n) &&
                  in_core(ii, a.h, a.nx) && in_core(jj, a.h, a.ny) && in_core(kk, a.h, a.nz),
                  pe_at(a, kk, ii, jj), win[w]);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float sl[WN], sr[WN], dummy;
#pragma unroll
    for (int w = 0; w < WN; ++w) {
      sl[w] = win[w][v];
      sr[w] = win[w + 1][v];
    }
    recon<SCHEME>(sl, Wl[v], dummy);
    recon<SCHEME>(sr, dummy, Wr[v]);
  }
  const int side = pc == a.h ? 0 : (pc == a.h + n ? 1 : -1);
  if (((a.bbc_dims >> D) & 1) && side >= 0) {
    const long long pidx = D == 0 ? (long long)k * a.NY + j
                                  : (D == 1 ? (long long)k * a.NX + i : (long long)i * a.NY + j);
    apply_face_bc<NS>(a.bbc[2 * D + side], Wl, pidx);
    apply_face_bc<NS>(a.bbc[2 * D + side], Wr, pidx);
  }
  float F[5 + NS];
  codes[c] = (uint8_t)hllc_flux<D, NS, PD>(Wl, Wr, pe_at(a, k, i, j), F);
}

}  // namespace sc
