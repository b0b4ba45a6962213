// Adjoint sweeps: the vector-Jacobian product (VJP) of the RHS,
//   lq = (dR/dq)^T g,
// for a state q (halos filled) and a cotangent g of the RHS on the interior.
//
// The forward RHS is R[c] = src(q[c]) - sum_d (F_d[c+1/2] - F_d[c-1/2]) / h_d
// with every face flux a function of the primitives of the cells of its
// stencil.  Each adjoint pass handles one direction d and is written as a
// gather: a thread owns a run of cells and recomputes every face whose
// stencil reaches them, so each cell's adjoint is summed by one thread, in a
// fixed order (deterministic, no atomics).  Per face:
//   lF = (g[c] - g[c-1]) / h_d                     (flux adjoint)
//   Riemann adjoint (the forward branch and min/max arguments replayed)
//     -> adjoints of the reconstructed left / right face primitives,
//   viscous adjoint -> adjoints of the two cell primitives and of the
//     transverse velocity derivatives of both cells,
//   reconstruction adjoint (WENO smoothness indicators and weights included)
//     -> adjoints of the stencil cells' primitives.
// A cell's primitive adjoints are converted to conserved adjoints once all
// of its faces are done.  Transverse velocity derivatives are central
// differences, so their adjoints are stored per cell (lG) and gathered by a
// separate kernel; the halo fill is then transposed (halo adjoints summed
// into the interior cells they were copied from, z, y, x order).
#pragma once

#include "sc_adjoint_device.hpp"
#include "sc_halo.hpp"
#include "sc_sweep_impl.hpp"

namespace sc {

// lG[gidx(i, k)]: adjoint of the central derivative d(u_i)/dx_k at a cell.
__host__ __device__ constexpr int gidx(int i, int k) { return 3 * i + k; }

struct AdjArgs {
  SweepArgs s;    // the state (halos filled), geometry and physics
  CFields g;      // cotangent of the RHS (interior cells are read)
  Fields lq;      // state adjoint, full grid, accumulated with +=
  float* lG[9];   // velocity-gradient adjoints (viscous), full grid, +=
};

// Conserved adjoint of one cell from its primitive adjoint (and the gravity
// source adjoint of the x pass), accumulated into lq.  PD: lw[7] is the
// adjoint of the cell's dynamic pressure p - p_ref (a constant 0 outside
// the interior along the pass's direction).  The gravity source takes
// rho - rho_ref with a reference state: the same derivative.
template <int NS = 0, bool PD = false, bool LEAN = false>
__device__ __forceinline__ void retire_cell(const AdjArgs& A, long long idx, const float* W,
                                            float pe, const float* lw, bool src,
                                            bool interior) {
  constexpr int NF = 5 + NS, NB = NVar<NS>::value;
  float lq[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) lq[f] = 0.0f;
  if constexpr (PD) {
    float lb[NB];
#pragma unroll
    for (int v = 0; v < NB; ++v) lb[v] = lw[v];
    if (interior) lb[kP] += lw[NB];
    cons_to_prim_adj<NS>(W, pe, lb, lq);
  } else {
    cons_to_prim_adj<NS>(W, pe, lw, lq);
  }
  const SweepArgs& a = A.s;
  if (!LEAN && src && interior && a.grav_src) {
    // R[1+d] += (rho * G) * g_d on the selected momenta.
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (a.grav_mask & (1 << d)) lq[0] += A.g.p[1 + d][idx] * (kG * a.gdir[d]);
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) A.lq.p[f][idx] += lq[f];
}

template <int D>
__device__ __forceinline__ void add_lG(const AdjArgs& A, long long idx, const float (*lt)[2]) {
  int t = 0;
#pragma unroll
  for (int dd = 0; dd < 3; ++dd) {
    if (dd == D) continue;
    A.lG[gidx(dd, dd)][idx] += lt[t][0];
    A.lG[gidx(D, dd)][idx] += lt[t][1];
    ++t;
  }
}

// Boundary face / flux values (the forward's face_flux at a face with a
// BoundaryBc) in reverse: a flux value fixes its component of the
// convective, diffusive or total flux, so that component's adjoint stops
// there; a face value fixes a face-state component (primitive: a constant;
// conserved: value / rho_face, whose derivative reaches the face density).
template <int NF>
__device__ __forceinline__ void flux_bc_adj(const BoundaryBc& b, const float* lF, float* lFc,
                                            float* lFd) {
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const bool tot = (b.flux_mask >> (16 + f)) & 1;
    lFc[f] = (tot || ((b.flux_mask >> f) & 1)) ? 0.0f : lF[f];
    lFd[f] = (tot || ((b.flux_mask >> (8 + f)) & 1)) ? 0.0f : lF[f];
  }
}

// lW: adjoint of the face state after apply_face_bc; on return, of the
// state before it (W, the state before it: its density is never replaced).
template <int NS>
__device__ __forceinline__ void face_bc_adj(const BoundaryBc& b, long long pidx,
                                            const float* W, float* lW) {
  for (int e = b.n_face - 1; e >= 0; --e) {
    const FaceEntry fe = b.face[e];
#pragma unroll
    for (int k = 1; k < NVar<NS>::value; ++k) {
      if (k != fe.var) continue;
      if (fe.kind == 1 && W[kRho] != 0.0f) {
        const float val = fe.plane != nullptr ? fe.plane[pidx] : fe.val;
        lW[kRho] -= lW[k] * (val / (W[kRho] * W[kRho]));
      }
      lW[k] = 0.0f;
    }
  }
}

// Riemann adjoint of one face with the boundary values of bb (or none):
// L, R the reconstructed face states before the face values, lFc the
// convective flux adjoint; lL, lR receive the state adjoints.
template <int FLUX, int D, int NS, bool PD>
__device__ __forceinline__ void face_riemann_adj(const BoundaryBc* bb, long long pidx,
                                                 const float* L, const float* R, float pe,
                                                 const float* lFc, float* lL, float* lR) {
  if (bb == nullptr) {
    riemann_adj<FLUX, D, NS, PD>(L, R, pe, lFc, lL, lR);
    return;
  }
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  float Lm[NV], Rm[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    Lm[v] = L[v];
    Rm[v] = R[v];
  }
  apply_face_bc<NS>(*bb, Lm, pidx);
  apply_face_bc<NS>(*bb, Rm, pidx);
  riemann_adj<FLUX, D, NS, PD>(Lm, Rm, pe, lFc, lL, lR);
  face_bc_adj<NS>(*bb, pidx, L, lL);
  face_bc_adj<NS>(*bb, pidx, R, lR);
}

// ---------------------------------------------------------------------------
// Marching adjoint pass along x (DIR 0) or z (DIR 2).  Block (64, BY): lane ->
// y, threadIdx.y -> the other transverse index, blockIdx.z -> a run of
// seg_len owned cells of the extended range [h-1-RAD, h+n+1+RAD) (every cell
// a face stencil of the direction reaches).  Iteration c handles face c (the
// face between cells c-1 and c) and the reconstruction adjoint of cell c-1.
// ---------------------------------------------------------------------------
// BB: the build with boundary face / flux values; PD: with a reference state
// (the dynamic pressure is an eighth reconstructed variable).
template <int DIR, int SCHEME, int FLUX, bool VISC, int NS, bool BB, bool PD, bool LEAN = false>
__device__ __forceinline__ void adj_march_body(const AdjArgs& A, const TileIdx& ti) {
  constexpr int RAD = Radius<SCHEME>::value;
  constexpr int WN = 2 * RAD + 1;
  constexpr int NW = WN + 1;  // window: cells c-1-RAD .. c+RAD
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  constexpr int NF = 5 + NS;
  const SweepArgs& a = A.s;
  const int j = a.h + ti.bx * 64 + ti.tx;
  const int o = a.h + ti.by * ti.by_dim + ti.ty;
  const int n = DIR == 0 ? a.nx : a.nz;
  const int n_other = DIR == 0 ? a.nz : a.nx;
  if (j >= a.h + a.ny || o >= a.h + n_other) return;
  const int ext_hi = a.h + n + 1 + RAD;
  const int o_lo = a.h - 1 - RAD + ti.bz * a.seg_len;
  if (o_lo >= ext_hi) return;
  const int o_hi = min(o_lo + a.seg_len, ext_hi);
  const int f_hi = a.h + n;  // faces h .. h+n are the interior faces
  const int c_start = max(o_lo - RAD, a.h);
  const int c_last = min(o_hi + RAD, f_hi + 1);
  const int NM = DIR == 0 ? a.NX : a.NZ;
  const long long sm = DIR == 0 ? a.sx : a.sz;
  const long long so = DIR == 0 ? a.sz : a.sx;
  const long long base = (long long)o * so + j;
  auto kk = [&](int c) { return DIR == 0 ? o : c; };
  auto ii = [&](int c) { return DIR == 0 ? c : o; };
  auto interior = [&](int c) { return c >= a.h && c < a.h + n; };
  auto owned = [&](int c) { return c >= o_lo && c < o_hi; };
  auto load_g = [&](int c, float* gv) {
    const bool in = interior(c);
    const long long idx = base + (long long)(in ? c : a.h) * sm;
#pragma unroll
    for (int f = 0; f < NF; ++f) gv[f] = in ? A.g.p[f][idx] : 0.0f;
  };

  // Primitive adjoints of the WN cells c-1-RAD .. c-1+RAD are accumulated
  // in LDS (dynamic, WN * NV floats per thread), a ring of WN slots per
  // thread; a thread touches only its own column, so no barrier is needed.
  HIP_DYNAMIC_SHARED(float, adj_acc_lds)
  const int nthr = 64 * ti.by_dim;
  const int tid = ti.ty * 64 + ti.tx;
  auto acc = [&](int slot, int v) -> float& { return adj_acc_lds[(slot * NV + v) * nthr + tid]; };
  float W[NW][NV];
#pragma unroll
  for (int w = 0; w < WN; ++w) {
    const int c = c_start - 1 - RAD + w;
    load_prim<NS, PD, LEAN>(a, base + (long long)c * sm, in_core(c, a.h, n), pe_at<LEAN>(a, kk(c), ii(c), j),
                 W[w]);
#pragma unroll
    for (int v = 0; v < NV; ++v) acc(w, v) = 0.0f;
  }
  // Slot of cell c-1-RAD (the oldest cell of the window).
  int s0 = 0;
  auto slot = [&](int w) { return s0 + w < WN ? s0 + w : s0 + w - WN; };
  // Left state of face c_start (reconstruction of cell c_start - 1).
  float fpPrev[NV], lRprev[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float s[WN], dummy;
#pragma unroll
    for (int w = 0; w < WN; ++w) s[w] = W[w][v];
    recon_replay<SCHEME>(s, fpPrev[v], dummy);
    lRprev[v] = 0.0f;
  }
  float gPrev[NF];
  load_g(c_start - 1, gPrev);
  float tPrev[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  float ltPrev[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  if constexpr (VISC) transverse_grads<DIR, LEAN>(a, kk(c_start - 1), ii(c_start - 1), j, tPrev);

  for (int c = c_start; c <= c_last; ++c) {
    {
      const int cn = min(c + RAD, NM - 1);
      load_prim<NS, PD, LEAN>(a, base + (long long)cn * sm, in_core(cn, a.h, n),
                   pe_at<LEAN>(a, kk(cn), ii(cn), j), W[NW - 1]);
    }
    const bool face = c <= f_hi;
    // Reconstruction of cell c: right state of face c, left state of c+1.
    float fpC[NV], fmC[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      float s[WN];
#pragma unroll
      for (int w = 0; w < WN; ++w) s[w] = W[w + 1][v];
      recon_replay<SCHEME>(s, fpC[v], fmC[v]);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      pin(fpC[v]);
      pin(fmC[v]);
    }
    float gC[NF];
    load_g(c, gC);
    float lL[NV], lR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) lL[v] = lR[v] = 0.0f;
    float tCur[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
    float ltCur[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
    if (face) {
      float lF[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) lF[f] = (gC[f] - gPrev[f]) * a.inv_hdir;
      const int side = c == a.h ? 0 : (c == a.h + n ? 1 : -1);
      const BoundaryBc* bb =
          (BB && ((a.bbc_dims >> DIR) & 1) && side >= 0) ? a.bbc + 2 * DIR + side : nullptr;
      float lFc[NF], lFd[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) lFc[f] = lFd[f] = lF[f];
      if (bb != nullptr) flux_bc_adj<NF>(*bb, lF, lFc, lFd);
      // Boundary plane cell of this face: x faces [z][y], z faces [x][y].
      face_riemann_adj<FLUX, DIR, NS, PD>(bb, (long long)o * a.NY + j, fpPrev, fmC,
                                  pe_at<LEAN>(a, kk(c), ii(c), j), lFc, lL, lR);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        pin(lL[v]);
        pin(lR[v]);
      }
      if constexpr (VISC) {
        transverse_grads<DIR, LEAN>(a, kk(c), ii(c), j, tCur);
        // Cells c-1 and c sit at window positions RAD and RAD+1.
        float lA[NV], lB[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) lA[v] = lB[v] = 0.0f;
        viscous_adj<DIR>(W[RAD], W[RAD + 1], tPrev, tCur, a.vc, lFd, lA, lB, ltPrev, ltCur);
        const int sa = slot(RAD), sb = slot(RAD + 1);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          acc(sa, v) += lA[v];
          acc(sb, v) += lB[v];
        }
      }
    }
    // Reconstruction adjoint of cell c-1: left state of face c, right state
    // of face c-1.
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      float s[WN], la[WN];
#pragma unroll
      for (int w = 0; w < WN; ++w) {
        s[w] = W[w][v];
        la[w] = 0.0f;
      }
      recon_adj<SCHEME>(s, lL[v], lRprev[v], la);
#pragma unroll
      for (int w = 0; w < WN; ++w) acc(slot(w), v) += la[w];
    }
    if constexpr (VISC) {
      // Cell c-1 has seen both of its faces.
      if (owned(c - 1) && c - 1 <= f_hi) add_lG<DIR>(A, base + (long long)(c - 1) * sm, ltPrev);
    }
    // Cell c-1-RAD is complete; its slot passes to cell c+RAD.
    {
      const int x = c - 1 - RAD;
      float l0[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        l0[v] = acc(s0, v);
        acc(s0, v) = 0.0f;
      }
      if (owned(x))
        retire_cell<NS, PD, LEAN>(A, base + (long long)x * sm, W[0], pe_at<LEAN>(a, kk(x), ii(x), j), l0,
                        DIR == 0,
                    interior(x));
      s0 = s0 + 1 < WN ? s0 + 1 : 0;
    }
#pragma unroll
    for (int w = 0; w < NW - 1; ++w) {
#pragma unroll
      for (int v = 0; v < NV; ++v) W[w][v] = W[w + 1][v];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      fpPrev[v] = fpC[v];
      lRprev[v] = lR[v];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) gPrev[f] = gC[f];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      tPrev[t][0] = tCur[t][0];
      tPrev[t][1] = tCur[t][1];
      ltPrev[t][0] = ltCur[t][0];
      ltPrev[t][1] = ltCur[t][1];
    }
  }
  // Remaining cells of the window: c_last - RAD .. c_last + RAD - 1.
#pragma unroll
  for (int w = 0; w < WN - 1; ++w) {
    const int x = c_last - RAD + w;
    float l0[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) l0[v] = acc(slot(w), v);
    if (owned(x))
      retire_cell<NS, PD, LEAN>(A, base + (long long)x * sm, W[w], pe_at<LEAN>(a, kk(x), ii(x), j), l0,
                      DIR == 0,
                  interior(x));
  }
}

template <int DIR, int SCHEME, int FLUX, bool VISC, int NS, bool BB, bool PD>
__global__ void __launch_bounds__(256) adj_march_kernel(const AdjArgs A) {
  const TileIdx ti{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z,
                   (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.y};
  adj_march_body<DIR, SCHEME, FLUX, VISC, NS, BB, PD>(A, ti);
}

// The LEAN build (no gravity, primitive-variable BCs, boundary values or
// reference state) fits three waves per SIMD (168 VGPRs).
template <int DIR, int SCHEME, int FLUX, bool VISC, int NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
adj_march_lean_kernel(const AdjArgs A) {
  const TileIdx ti{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z,
                   (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.y};
  adj_march_body<DIR, SCHEME, FLUX, VISC, NS, false, false, true>(A, ti);
}

// ---------------------------------------------------------------------------
// y adjoint pass (contiguous direction).  One wave per window of 64
// consecutive positions of a z plane's run of rows (as the forward y sweep):
// lane l computes face l (cells l-1, l) and the reconstruction adjoint of its
// own cell; a cell's adjoint gathers, through DPP lane shifts (every lane
// active at each shift), the contributions of the RAD lanes on either side.
// Lanes KA .. KA + OUTA - 1 write the cells with y in [h-1-RAD, h+ny+1+RAD)
// (every cell a y-face stencil reaches).  Only face and reconstruction lanes
// (y in [h-1, h+ny]) contribute, and all they read lies in their own row
// (RAD <= h-1), so windows may run across row ends.
// ---------------------------------------------------------------------------
template <int SCHEME>
struct YAdjGeom {
  static constexpr int RAD = Radius<SCHEME>::value;
  static constexpr int KA = 2 * RAD + 1;
  static constexpr int OUTA = 62 - 4 * RAD;
};

// Windows of the adjoint y pass over a plane's nx * NY positions: window w's
// lane KA sits at position h-1-RAD + w * OUTA; the last output position is
// (nx-1) * NY + h+ny+RAD.
__host__ __device__ inline int y_adj_windows(int nx, int ny, int h, int rad, int outa) {
  const long long span = (long long)(nx - 1) * (ny + 2 * h) + ny + 2 * rad + 2;
  return (int)((span + outa - 1) / outa);
}

template <int SCHEME, int FLUX, bool VISC, int NS, bool BB, bool PD, bool LEAN = false>
__device__ __forceinline__ void adj_y_body(const AdjArgs& A, const TileIdx& ti) {
  constexpr int RAD = YAdjGeom<SCHEME>::RAD;
  constexpr int KA = YAdjGeom<SCHEME>::KA;
  constexpr int OUTA = YAdjGeom<SCHEME>::OUTA;
  constexpr int WN = 2 * RAD + 1;
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  constexpr int NF = 5 + NS;
  const SweepArgs& a = A.s;
  const int lane = ti.tx;
  const int win = ti.bx * ti.by_dim + ti.ty;
  const int k = a.h + ti.bz;
  // uniform per wave
  if (k >= a.h + a.nz || win >= y_adj_windows(a.nx, a.ny, a.h, RAD, OUTA)) return;
  const int run = a.nx * a.NY;
  const int pos = win * OUTA + (a.h - 1 - RAD) - KA + lane;
  const bool valid = pos >= 0 && pos < run;
  const int pc = min(max(pos, 0), run - 1);
  // pc = r * NY + yc (float quotient estimate, off by at most one for row
  // counts below 2^20, checked at context creation).
  int r = (int)((float)pc * frcp((float)a.NY));
  int yc = pc - r * a.NY;
  if (yc < 0) {
    --r;
    yc += a.NY;
  }
  if (yc >= a.NY) {
    ++r;
    yc -= a.NY;
  }
  const int i = a.h + r;
  const long long idx = (long long)k * a.sz + (long long)a.h * a.sx + pc;
  const bool out_lane = valid && lane >= KA && lane < KA + OUTA && yc >= a.h - 1 - RAD &&
                        yc < a.h + a.ny + 1 + RAD;
  const bool face = valid && yc >= a.h && yc <= a.h + a.ny;
  const bool rcell = valid && yc >= a.h - 1 && yc <= a.h + a.ny;
  const bool inner = valid && yc >= a.h && yc < a.h + a.ny;
  const float pe = pe_at<LEAN>(a, k, i, yc);

  float W0[NV];
  load_prim<NS, PD, LEAN>(a, idx, in_core(yc, a.h, a.ny), pe, W0);
  float S[NV][WN], fp[NV], fm[NV], Am[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
#pragma unroll
    for (int w = 0; w < WN; ++w) {
      const int off = w - RAD;
      S[v][w] = lane_shift(W0[v], off);
    }
    Am[v] = S[v][RAD - 1];
    recon_replay<SCHEME>(S[v], fp[v], fm[v]);
  }
  float gC[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) gC[f] = inner ? A.g.p[f][idx] : 0.0f;
  float L[NV], gP[NF];
#pragma unroll
  for (int v = 0; v < NV; ++v) L[v] = wave_shr1(fp[v]);
#pragma unroll
  for (int f = 0; f < NF; ++f) gP[f] = wave_shr1(gC[f]);
  float lL[NV], lR[NV], lA[NV], lB[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) lL[v] = lR[v] = lA[v] = lB[v] = 0.0f;
  float lta[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, ltb[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
  float tCur[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, tPrev[2][2];
  if constexpr (VISC) transverse_grads<1, LEAN>(a, k, i, yc, tCur);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    tPrev[t][0] = wave_shr1(tCur[t][0]);
    tPrev[t][1] = wave_shr1(tCur[t][1]);
  }
  if (face) {
    float lF[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) lF[f] = (gC[f] - gP[f]) * a.inv_hdir;
    const int side = yc == a.h ? 0 : (yc == a.h + a.ny ? 1 : -1);
    const BoundaryBc* bb =
        (BB && ((a.bbc_dims >> 1) & 1) && side >= 0) ? a.bbc + 2 + side : nullptr;
    float lFc[NF], lFd[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) lFc[f] = lFd[f] = lF[f];
    if (bb != nullptr) flux_bc_adj<NF>(*bb, lF, lFc, lFd);
    // Boundary plane cell of this face: y faces [z][x].
    face_riemann_adj<FLUX, 1, NS, PD>(bb, (long long)k * a.NX + i, L, fm, pe, lFc, lL, lR);
    if constexpr (VISC) viscous_adj<1>(Am, W0, tPrev, tCur, a.vc, lFd, lA, lB, lta, ltb);
  }
  // Reconstruction adjoint of the own cell: left state of face y+1 (lane + 1)
  // and right state of face y.
  float la[NV][WN];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const float lfp = wave_shl1(lL[v]);
#pragma unroll
    for (int w = 0; w < WN; ++w) la[v][w] = 0.0f;
    if (rcell) recon_adj<SCHEME>(S[v], lfp, lR[v], la[v]);
  }
  // Gather: contributions of lanes l - RAD .. l + RAD (reconstruction) and of
  // face l + 1 (viscous, as its cell A).
  float tot[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float s = la[v][RAD] + lB[v];
#pragma unroll
    for (int w = 0; w < WN; ++w) {
      const int off = w - RAD;  // lane l - off holds the contribution at w
      if (off != 0) s += lane_shift(la[v][w], -off);
    }
    tot[v] = s + wave_shl1(lA[v]);
  }
  float lt[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    lt[t][0] = ltb[t][0] + wave_shl1(lta[t][0]);
    lt[t][1] = ltb[t][1] + wave_shl1(lta[t][1]);
  }
  if (out_lane) {
    retire_cell<NS, PD, LEAN>(A, idx, W0, pe, tot, false, inner);
    if constexpr (VISC) {
      if (rcell) add_lG<1>(A, idx, lt);
    }
  }
}

template <int SCHEME, int FLUX, bool VISC, int NS, bool BB, bool PD, bool LEAN = false>
__global__ void __launch_bounds__(256) adj_y_kernel(const AdjArgs A) {
  const TileIdx ti{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z,
                   (int)threadIdx.x, (int)threadIdx.y, (int)blockDim.y};
  adj_y_body<SCHEME, FLUX, VISC, NS, BB, PD, LEAN>(A, ti);
}


// Launches the three adjoint passes (x, y, z) for one scheme.
template <int SCHEME, int FLUX, bool VISC, int NS, bool BB, bool PD, bool LEAN = false>
hipError_t launch_adj_sweeps_t(AdjArgs A, const float h[3], const SweepTuning& tune,
                               hipStream_t stream) {
  static_assert(!LEAN || (!BB && !PD), "LEAN builds carry no boundary values or reference state");
  constexpr int RAD = Radius<SCHEME>::value;
  const SweepArgs& a = A.s;
  const int seg = tune.adj_seg, by = tune.adj_by, by_y = tune.adj_by_y;
  A.s.seg_len = seg;
  const dim3 blk_m(64, by);
  const dim3 blk_y(64, by_y);
  // LDS ring of the marching passes: (2*RAD+1) cells x NV primitives per thread.
  const size_t lds = sizeof(float) * (2 * RAD + 1) * (NVar<NS>::value + (PD ? 1 : 0)) * 64 * by;
  {
    AdjArgs ax = A;
    ax.s.inv_hdir = 1.0f / h[0];
    const int ext = a.nx + 2 * RAD + 2;
    const dim3 grid((a.ny + 63) / 64, (a.nz + by - 1) / by, (ext + seg - 1) / seg);
    if constexpr (LEAN)
      hipLaunchKernelGGL((adj_march_lean_kernel<0, SCHEME, FLUX, VISC, NS>), grid, blk_m, lds,
                         stream, ax);
    else
      hipLaunchKernelGGL((adj_march_kernel<0, SCHEME, FLUX, VISC, NS, BB, PD>), grid, blk_m, lds,
                         stream, ax);
  }
  {
    AdjArgs ay = A;
    ay.s.inv_hdir = 1.0f / h[1];
    const int nwin = y_adj_windows(a.nx, a.ny, a.h, RAD, YAdjGeom<SCHEME>::OUTA);
    const dim3 grid((nwin + by_y - 1) / by_y, 1, a.nz);
    hipLaunchKernelGGL((adj_y_kernel<SCHEME, FLUX, VISC, NS, BB, PD, LEAN>), grid, blk_y, 0,
                       stream, ay);
  }
  {
    AdjArgs az = A;
    az.s.inv_hdir = 1.0f / h[2];
    const int ext = a.nz + 2 * RAD + 2;
    const dim3 grid((a.ny + 63) / 64, (a.nx + by - 1) / by, (ext + seg - 1) / seg);
    if constexpr (LEAN)
      hipLaunchKernelGGL((adj_march_lean_kernel<2, SCHEME, FLUX, VISC, NS>), grid, blk_m, lds,
                         stream, az);
    else
      hipLaunchKernelGGL((adj_march_kernel<2, SCHEME, FLUX, VISC, NS, BB, PD>), grid, blk_m, lds,
                         stream, az);
  }
  return hipGetLastError();
}

// Builds: boundary values (BB) and / or a reference state (PD) for the base
// fields; one or two transported scalars without either.
template <int SCHEME, int FLUX, bool VISC>
hipError_t launch_adj_variant(const AdjArgs& A, const float h[3], int ns,
                              const SweepTuning& tune, hipStream_t stream) {
  const bool bb = A.s.bbc_dims != 0, pd = A.s.p_ref != nullptr;
  const SweepArgs& a = A.s;
  const bool lean = !bb && !pd && a.prim_bc_mask == 0 && !a.has_pe && !a.grav_src;
  if (ns == 1)
    return lean ? launch_adj_sweeps_t<SCHEME, FLUX, VISC, 1, false, false, true>(A, h, tune, stream)
                : launch_adj_sweeps_t<SCHEME, FLUX, VISC, 1, false, false>(A, h, tune, stream);
  if (ns == 2)
    return lean ? launch_adj_sweeps_t<SCHEME, FLUX, VISC, 2, false, false, true>(A, h, tune, stream)
                : launch_adj_sweeps_t<SCHEME, FLUX, VISC, 2, false, false>(A, h, tune, stream);
  if (lean) return launch_adj_sweeps_t<SCHEME, FLUX, VISC, 0, false, false, true>(A, h, tune, stream);
  if (pd)
    return bb ? launch_adj_sweeps_t<SCHEME, FLUX, VISC, 0, true, true>(A, h, tune, stream)
              : launch_adj_sweeps_t<SCHEME, FLUX, VISC, 0, false, true>(A, h, tune, stream);
  return bb ? launch_adj_sweeps_t<SCHEME, FLUX, VISC, 0, true, false>(A, h, tune, stream)
            : launch_adj_sweeps_t<SCHEME, FLUX, VISC, 0, false, false>(A, h, tune, stream);
}

// ns: transported scalars (0..2; boundary values and a reference state with
// ns = 0 only, checked by the caller).
template <int SCHEME>
hipError_t launch_adj_sweeps(const AdjArgs& A, const float h[3], int flux, bool visc, int ns,
                             const SweepTuning& tune, hipStream_t stream);

#define SC_INSTANTIATE_ADJOINT(SCHEME)                                                    \
  template <>                                                                             \
  hipError_t launch_adj_sweeps<SCHEME>(const AdjArgs& A, const float h[3], int flux,      \
                                       bool visc, int ns, const SweepTuning& tune,        \
                                       hipStream_t stream) {                              \
    if (flux == kHllc)                                                                    \
      return visc ? launch_adj_variant<SCHEME, kHllc, true>(A, h, ns, tune, stream)       \
                  : launch_adj_variant<SCHEME, kHllc, false>(A, h, ns, tune, stream);     \
    return visc ? launch_adj_variant<SCHEME, kHll, true>(A, h, ns, tune, stream)          \
                : launch_adj_variant<SCHEME, kHll, false>(A, h, ns, tune, stream);        \
  }

// Non-templated kernels of the adjoint (sc_adjoint.hip).
hipError_t launch_adj_zero(float* const* p, int np, long long n, hipStream_t stream);
hipError_t launch_adj_visc_gather(const AdjArgs& A, hipStream_t stream);
// Transpose of the halo fill along d (lq in a.q, the filled stage state in
// state); call launch_halo_adj_prim first when a.prim_mask is set, then for
// d = 2, 1, 0.
hipError_t launch_halo_adj_dim(const HaloArgs& a, const float* const* state, int d,
                               hipStream_t stream);
hipError_t launch_halo_adj_prim(const HaloArgs& a, const float* const* state,
                                hipStream_t stream);
// Grid extents for the element-wise adjoint kernels.
struct AdjGeom {
  int n[3], N[3], h;
};
// Interior cells: lam_k = c1 * (lam_next + dt * lq), and lam0 += c0 * lam_next
// when lam0 is set; halo cells of lam_k are zeroed.  lam_k may alias lam_next.
hipError_t launch_rk_adj_combine(const float* const* lam_next, const float* const* lq,
                                 float* const* lam_k, float* const* lam0, int nf,
                                 const AdjGeom& g, float c0, float c1, const float* dt,
                                 hipStream_t stream);

}  // namespace sc
