// Device-side physics and numerics of the compressible NS RHS (gfx950).
//
// Every function restates one piece of the reference's per-cell or per-face
// arithmetic, keeping its operand order so fp32 results track the
// reference's fp32 path to a few ulps.  Reference anchors are cited per
// function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sc {

// physics/constant.py:17-41 (computed in double as Python does, rounded once).
constexpr double kGammaD = 1.4;
constexpr double kRD = 8.314 / 0.029;
constexpr double kCvD = kRD / (kGammaD - 1.0);
constexpr double kCpD = kCvD + kRD;
constexpr float kGamma = (float)kGammaD;
constexpr float kGm1 = (float)(kGammaD - 1.0);
constexpr float kR = (float)kRD;
constexpr float kCp = (float)kCpD;
constexpr float kG = 9.81f;
constexpr float kThird = (float)(1.0 / 3.0);
constexpr float kWenoEps = 1e-6f;
constexpr float kSmSnap = 1e-6f;

constexpr int kMaxFields = 7;

// Interpolation scheme ids (match SC_WENO_3 / SC_WENO_5 / SC_MUSCL).
constexpr int kWeno3 = 0;
constexpr int kWeno5 = 1;
constexpr int kMuscl = 2;
// Riemann solver ids (match SC_HLL / SC_HLLC).
constexpr int kHll = 0;
constexpr int kHllc = 1;

// Number of cells on each side of a cell its reconstruction reads.
template <int SCHEME>
struct Radius {
  static constexpr int value = SCHEME == kWeno5 ? 2 : 1;
};

// Reconstructed primitive variables per cell: rho u v w e p h + scalars
// (equations/compressible_navier_stokes.py:107-113,127).
constexpr int kRho = 0, kU = 1, kV = 2, kW = 3, kE = 4, kP = 5, kH = 6;
template <int NS>
struct NVar {
  static constexpr int value = 7 + NS;
};

// Hardware reciprocal and square root (v_rcp_f32 / v_sqrt_f32, 1 ulp).  The
// reference's IEEE divisions become reciprocal-multiplies: every result stays
// within a few ulps of the IEEE value, far inside the 1e-5 parity bound, and
// a division costs 2 instructions instead of about 10.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Value of v held by lane (this lane + off) of the 64-lane wave.  off must
// fold to a constant (small, e.g. a stencil offset in an unrolled loop): each
// step is one DPP move over the whole wave (wave_shr:1 takes lane - 1,
// wave_shl:1 lane + 1), a VALU op with no LDS round trip or wait.  Lanes whose
// source lies outside the wave read 0.
__device__ __forceinline__ float wave_shr1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_shl1(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_shift(float v, int off) {
#pragma unroll
  for (; off < 0; ++off) v = wave_shr1(v);
#pragma unroll
  for (; off > 0; --off) v = wave_shl1(v);
  return v;
}

// tf.math.divide_no_nan / reciprocal_no_nan.
__device__ __forceinline__ float dnn(float x, float y) {
  return y == 0.0f ? 0.0f : x * frcp(y);
}
__device__ __forceinline__ float rnn(float x) {
  return x == 0.0f ? 0.0f : frcp(x);
}
// tf.minimum / tf.maximum: a NaN operand gives NaN (fminf/fmaxf would return
// the other operand), so unphysical states propagate exactly as in the
// reference instead of silently selecting a different wave speed.  gfx950's
// v_minimum3_f32 / v_maximum3_f32 propagate NaN in one instruction.
__device__ __forceinline__ float tf_min(float a, float b) {
  return __builtin_elementwise_minimum(a, b);
}
__device__ __forceinline__ float tf_max(float a, float b) {
  return __builtin_elementwise_maximum(a, b);
}

// divide_no_nan with an IEEE division, for the boundary-condition paths whose
// values must match the reference bit for bit (cheap: halo slabs only).
__device__ __forceinline__ float dnn_exact(float x, float y) {
  return y == 0.0f ? 0.0f : x / y;
}

// Conserved (rho, rho_u, rho_v, rho_w, rho_e, rho_s..) -> primitives
// (rho, u, v, w, e, p, h, s..): common/utils.py:145-183 (phi = (1/rho)*q),
// generic.py:27-88,357-389 (p = ((gamma-1)*rho) * (e - 0.5*(u^2+v^2+w^2) - pe),
// h = e + p/rho).  The multiply-adds are written out and contraction is off,
// so every inlined copy rounds identically: a cell's primitives do not depend
// on which copy (a marching sweep's prologue or its loop) converted it.
template <int NS>
__device__ __forceinline__ void cons_to_prim(const float* q, float pe, float* w) {
#pragma clang fp contract(off)
  const float rho = q[0];
  const float inv = frcp(rho);
  w[kRho] = rho;
  w[kU] = inv * q[1];
  w[kV] = inv * q[2];
  w[kW] = inv * q[3];
  w[kE] = inv * q[4];
#pragma unroll
  for (int s = 0; s < NS; ++s) w[7 + s] = inv * q[5 + s];
  const float ke =
      0.5f * __builtin_fmaf(w[kW], w[kW], __builtin_fmaf(w[kV], w[kV], w[kU] * w[kU]));
  const float eint = (w[kE] - ke) - pe;
  w[kP] = (kGm1 * rho) * eint;
  w[kH] = __builtin_fmaf(w[kP], inv, w[kE]);
}

// Largest signal speed max(|u|, |v|, |w|) + c of one cell's conserved fields
// (c = sqrt(gamma p / rho), IEEE division and square root), the CFL wave
// speed of demo/kelvin_helmholtz_instability.py:66-73 taken per cell.  One
// function for every kernel that computes it, so their maxima agree bit for
// bit.
__device__ __forceinline__ float wave_speed(const float* q, float pe) {
  float w[7];
  cons_to_prim<0>(q, pe, w);
  const float c = sqrtf(kGamma * w[kP] / w[kRho]);
  return fmaxf(fmaxf(fabsf(w[kU]) + c, fabsf(w[kV]) + c), fabsf(w[kW]) + c);
}

// ---------------------------------------------------------------------------
// Face reconstruction of one variable at one cell.  a[] holds the cell values
// at offsets -R..R.  Returns fp = left state of face c+1/2 and fm = right
// state of face c-1/2 (swirl_lm's (neg, pos) at i-1/2 stored at i, for the
// faces on either side of cell c).  WENO follows demo/swirl_c_prototype.ipynb
// cell 4 `weno5` (l.70-211); MUSCL follows `muscl` (l.214-239).
// ---------------------------------------------------------------------------

template <int SCHEME>
__device__ __forceinline__ void recon(const float* a, float& fp, float& fm);

template <>
__device__ __forceinline__ void recon<kWeno5>(const float* a, float& fp, float& fm) {
  const float vm2 = a[0], vm1 = a[1], v0 = a[2], vp1 = a[3], vp2 = a[4];
  const float c16 = (float)(1.0 / 6.0), c13 = (float)(1.0 / 3.0);
  // First and second differences of the stencil.
  const float dm2 = vm1 - vm2, dm1 = v0 - vm1, d0 = vp1 - v0, d1 = vp2 - vp1;
  const float t0 = d1 - d0, t1 = d0 - dm1, t2 = dm1 - dm2;
  // Smoothness indicators beta_r = 13/12 t_r^2 + 1/4 s_r^2 with s_r the
  // one-sided slopes (d1 - 3 d0, d0 + dm1, 3 dm1 - dm2 are 2 s_r).  The
  // nonlinear weights d_r / (eps + beta_r)^2, normalised, do not change when
  // every eps + beta_r is scaled by one factor, so the kernels form
  // E_r = 4 (eps + beta_r) = 13/3 t_r^2 + (2 s_r)^2 + 4 eps directly.
  const float u0 = __builtin_fmaf(-3.0f, d0, d1);
  const float u1 = d0 + dm1;
  const float u2 = __builtin_fmaf(3.0f, dm1, -dm2);
  const float c133 = (float)(13.0 / 3.0), eps4 = 4.0f * kWenoEps;
  const float E0 = __builtin_fmaf(c133 * t0, t0, __builtin_fmaf(u0, u0, eps4));
  const float E1 = __builtin_fmaf(c133 * t1, t1, __builtin_fmaf(u1, u1, eps4));
  const float E2 = __builtin_fmaf(c133 * t2, t2, __builtin_fmaf(u2, u2, eps4));
  // Candidate values of the prototype's c[k][r] stencils from the
  // differences (q = v0 +- d / 2 + t * {1/3, -1/6}; a stencil of equal values
  // gives that value exactly), then the standard weighted average with
  // positive weights, whose rounding keeps a result near zero as small, and
  // of the same sign, as the reference's (the dynamic pressure next to its
  // zero halo cells feeds a square root).
  const float hn0 = v0 + 0.5f * d0, hn1 = v0 + 0.5f * dm1;
  const float hp0 = v0 - 0.5f * d0, hp1 = v0 - 0.5f * dm1;
  const float qn0 = hn0 - c16 * t0, qn1 = hn0 - c16 * t1, qn2 = hn1 + c13 * t2;
  const float qp0 = hp0 + c13 * t0, qp1 = hp1 - c16 * t1, qp2 = hp1 - c16 * t2;
  // alpha_r = d_r / E_r^2 with d = {0.3, 0.6, 0.1} (face c+1/2) and
  // {0.1, 0.6, 0.3} (face c-1/2), both scaled by 10 (again a common factor):
  // one reciprocal per stencil serves both faces.
  const float i0 = frcp(E0 * E0), i1 = frcp(E1 * E1), i2 = frcp(E2 * E2);
  const float a1 = 6.0f * i1, an0 = 3.0f * i0, ap2 = 3.0f * i2;
  const float rn = frcp((an0 + a1) + i2);
  const float rp = frcp((i0 + a1) + ap2);
  fp = ((an0 * qn0 + a1 * qn1) + i2 * qn2) * rn;
  fm = ((i0 * qp0 + a1 * qp1) + ap2 * qp2) * rp;
}

template <>
__device__ __forceinline__ void recon<kWeno3>(const float* a, float& fp, float& fm) {
  const float vm1 = a[0], v0 = a[1], vp1 = a[2];
  const float dp = vp1 - v0, dm = v0 - vm1;
  // eps + beta_r with beta_0 = (v_{+1} - v_0)^2, beta_1 = (v_0 - v_{-1})^2.
  const float e0 = __builtin_fmaf(dp, dp, kWenoEps), e1 = __builtin_fmaf(dm, dm, kWenoEps);
  const float qn0 = 0.5f * v0 + 0.5f * vp1;
  const float qn1 = -0.5f * vm1 + 1.5f * v0;
  const float qp0 = 1.5f * v0 + -0.5f * vp1;
  const float qp1 = 0.5f * vm1 + 0.5f * v0;
  // alpha_r = d_r / (eps + beta_r)^2 with d = {2/3, 1/3} (face c+1/2) and
  // {1/3, 2/3} (face c-1/2), both scaled by 3 (a common factor).
  const float i0 = frcp(e0 * e0), i1 = frcp(e1 * e1);
  const float an0 = 2.0f * i0, ap1 = 2.0f * i1;
  fp = (an0 * qn0 + i1 * qn1) * frcp(an0 + i1);
  fm = (i0 * qp0 + ap1 * qp1) * frcp(i0 + ap1);
}

__device__ __forceinline__ float sgnf(float x) {
  return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f);
}

template <>
__device__ __forceinline__ void recon<kMuscl>(const float* a, float& fp, float& fm) {
  const float vm1 = a[0], v0 = a[1], vp1 = a[2];
  // minmod(central2(v, 2), forward1(v, 1), backward1(v, 1)).
  const float g0 = (vp1 - vm1) * 0.25f;
  const float g1 = vp1 - v0;
  const float g2 = v0 - vm1;
  // sign(sum)/3 is +-1 exactly when all three signs agree (|sum| == 3).
  const float s = (sgnf(g0) + sgnf(g1)) + sgnf(g2);
  const float amin = fminf(fminf(fabsf(g0), fabsf(g1)), fabsf(g2));
  const float m = fabsf(s) == 3.0f ? copysignf(amin, s) : 0.0f;
  fp = v0 + 0.5f * m;
  fm = v0 - 0.5f * m;
}

// ---------------------------------------------------------------------------
// Riemann solvers at one face.  L = neg (left) state, R = pos (right) state,
// both as reconstructed primitives; pe = gravitational potential at the cell
// holding the face.  D = face-normal dimension (0 x, 1 y, 2 z).
// ---------------------------------------------------------------------------

// Physical flux of one side: equations/compressible_navier_stokes.py:338-427,
// with the face conserved momentum m = rho_f * u_f (utils.py:186-216).
// PD: the face state carries the dynamic pressure p - p_ref at index 7 + NS,
// which replaces p in the aligned momentum flux (helper_var 'p_dyn',
// compressible_navier_stokes.py:366-372).
template <int D, int NS, bool PD = false>
__device__ __forceinline__ void physical_flux_inv(const float* s, float is, float* f) {
  const float m = s[kRho] * s[kU + D];
  const float pm = PD ? s[7 + NS] : s[kP];
  f[0] = m;
  f[1] = D == 0 ? m * s[kU] + pm : m * s[kU];
  f[2] = D == 1 ? m * s[kV] + pm : m * s[kV];
  f[3] = D == 2 ? m * s[kW] + pm : m * s[kW];
  f[4] = m * (s[kE] + s[kP] * is);
#pragma unroll
  for (int k = 0; k < NS; ++k) f[5 + k] = m * s[7 + k];
}

// is = 1 / rho of the side (the callers share it with the wave speeds).
template <int D, int NS, bool PD = false>
__device__ __forceinline__ void physical_flux(const float* s, float* f) {
  physical_flux_inv<D, NS, PD>(s, frcp(s[kRho]), f);
}

// Roe / Einfeldt wave-speed estimates: hll.py:178-283, sound speeds
// generic.py:90-159 (opt 'p_rho' per side, opt 'h' for the Roe state).
// iL, iR: 1 / rho of the two sides.
template <int D>
__device__ __forceinline__ void roe_speeds_inv(const float* L, const float* R, float iL,
                                               float iR, float pe, float& sl, float& sr) {
  const float cl = fsqrt(kGamma * L[kP] * iL);
  const float cr = fsqrt(kGamma * R[kP] * iR);
  const float ql = fsqrt(L[kRho]);
  const float qr = fsqrt(R[kRho]);
  const float inv = rnn(ql + qr);
  const float ru = (ql * L[kU] + qr * R[kU]) * inv;
  const float rv = (ql * L[kV] + qr * R[kV]) * inv;
  const float rw = (ql * L[kW] + qr * R[kW]) * inv;
  const float rh = (ql * L[kH] + qr * R[kH]) * inv;
  const float hs = (rh - 0.5f * ((ru * ru + rv * rv) + rw * rw)) - pe;
  const float c = fsqrt(kGm1 * hs);
  const float un = D == 0 ? ru : (D == 1 ? rv : rw);
  sl = tf_min(L[kU + D] - cl, un - c);
  sr = tf_max(R[kU + D] + cr, un + c);
}

template <int D>
__device__ __forceinline__ void roe_speeds(const float* L, const float* R,
                                           float pe, float& sl, float& sr) {
  roe_speeds_inv<D>(L, R, frcp(L[kRho]), frcp(R[kRho]), pe, sl, sr);
}

// HLL flux: hll.py:36-175.
template <int D, int NS, bool PD = false>
__device__ __forceinline__ void hll_flux(const float* L, const float* R, float pe,
                                         float* F) {
  constexpr int NF = 5 + NS;
  float sl, sr;
  const float iL = frcp(L[kRho]), iR = frcp(R[kRho]);
  roe_speeds_inv<D>(L, R, iL, iR, pe, sl, sr);
  const float ds = sr - sl;
  const float rds = ds == 0.0f ? 0.0f : frcp(ds);
  const float t1 = (tf_min(sr, 0.0f) - tf_min(sl, 0.0f)) * rds;
  const float t2 = 1.0f - t1;
  const float t3 = (sr * fabsf(sl) - sl * fabsf(sr)) * (0.5f * rds);
  float fl[NF], fr[NF], ul[NF], ur[NF];
  physical_flux_inv<D, NS, PD>(L, iL, fl);
  physical_flux_inv<D, NS, PD>(R, iR, fr);
  ul[0] = L[kRho];
  ur[0] = R[kRho];
  ul[1] = L[kRho] * L[kU];
  ur[1] = R[kRho] * R[kU];
  ul[2] = L[kRho] * L[kV];
  ur[2] = R[kRho] * R[kV];
  ul[3] = L[kRho] * L[kW];
  ur[3] = R[kRho] * R[kW];
  ul[4] = L[kRho] * L[kE];
  ur[4] = R[kRho] * R[kE];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    ul[5 + k] = L[kRho] * L[7 + k];
    ur[5 + k] = R[kRho] * R[7 + k];
  }
#pragma unroll
  for (int k = 0; k < NF; ++k) F[k] = (t1 * fr[k] + t2 * fl[k]) - t3 * (ur[k] - ul[k]);
}

// HLLC star-state flux of one side: hllc.py:134-192 (Toro eq. 10.38).
template <int D, int NS>
__device__ __forceinline__ void hllc_star(const float* S, float is, float sk, float sm,
                                          const float* f, float* fs) {
  const float rho = S[kRho], vn = S[kU + D];
  const float den = sk - sm;
  const float r = den == 0.0f ? 0.0f : frcp(den);
  const float q = sk * r;
  const float a = rho * (sm - vn);
  fs[0] = f[0] + a * q;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k == D) {
      fs[1 + k] = f[1 + k] + (rho * sk) * (((sk - vn) * r) * sm - vn);
    } else {
      fs[1 + k] = f[1 + k] + (a * S[kU + k]) * q;
    }
  }
  fs[4] = f[4] + ((rho * q) * (sm - vn)) *
                     ((S[kE] + S[kP] * is) + (sk - vn) * sm);
#pragma unroll
  for (int k = 0; k < NS; ++k) fs[5 + k] = f[5 + k] + (a * S[7 + k]) * q;
}

// HLLC flux: hllc.py:33-257 and its wave speeds hllc.py:260-347.  Returns the
// branch code 0 F_L, 1 F*_L, 2 F*_R, 3 F_R (hllc.py:227-243).
// With PD the wave speeds, the contact speed and the star states see the
// dynamic pressure (hllc.py:109-120 swaps p for p_dyn in the primitives they
// use); the physical fluxes keep p in the energy flux.
template <int D, int NS, bool PD = false>
__device__ __forceinline__ int hllc_flux(const float* L0, const float* R0, float pe,
                                         float* F) {
  constexpr int NF = 5 + NS;
  constexpr int NV = NVar<NS>::value + (PD ? 1 : 0);
  float Lp[NV], Rp[NV];
  const float* L = L0;
  const float* R = R0;
  if constexpr (PD) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      Lp[v] = L0[v];
      Rp[v] = R0[v];
    }
    Lp[kP] = L0[7 + NS];
    Rp[kP] = R0[7 + NS];
    L = Lp;
    R = Rp;
  }
  float sl, sr;
  const float iL = frcp(L[kRho]), iR = frcp(R[kRho]);
  roe_speeds_inv<D>(L, R, iL, iR, pe, sl, sr);
  const float vl = L[kU + D], vr = R[kU + D];
  const float rl = L[kRho], rr = R[kRho];
  float sm = dnn(((R[kP] - L[kP]) + (rl * vl) * (sl - vl)) - (rr * vr) * (sr - vr),
                 rl * (sl - vl) - rr * (sr - vr));
  sm = fabsf(sm) < kSmSnap ? 0.0f : sm;
  int br;
  if (sl >= 0.0f) {
    br = 0;
  } else if (sl <= 0.0f && sm >= 0.0f) {
    br = 1;
  } else if (sm <= 0.0f && sr >= 0.0f) {
    br = 2;
  } else {
    br = 3;
  }
  // The flux is built from one side's state (L for branches 0, 1, R for 2,
  // 3), selected per lane, so lanes of a wave on different sides share one
  // evaluation of the physical flux and of the star correction.
  const bool left = br <= 1;
  float S0[NV], Ss[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    S0[v] = left ? L0[v] : R0[v];
    Ss[v] = left ? L[v] : R[v];
  }
  const float sk = left ? sl : sr;
  const float is = left ? iL : iR;
  float f[NF];
  physical_flux_inv<D, NS, PD>(S0, is, f);
  if (br == 1 || br == 2) {
    hllc_star<D, NS>(Ss, is, sk, sm, f, F);
  } else {
#pragma unroll
    for (int k = 0; k < NF; ++k) F[k] = f[k];
  }
  return br;
}

template <int FLUX, int D, int NS, bool PD = false>
__device__ __forceinline__ int riemann(const float* L, const float* R, float pe,
                                       float* F) {
  if constexpr (FLUX == kHllc) {
    return hllc_flux<D, NS, PD>(L, R, pe, F);
  } else {
    hll_flux<D, NS, PD>(L, R, pe, F);
    return 0;
  }
}

// ---------------------------------------------------------------------------
// Viscous + heat flux at one face normal to D:
// equations/compressible_navier_stokes.py:618-707, physics/diffusion.py:33-109,
// physics/fluid.py:31-111, physics/transport/simple.py:44-108.
//   A = cell c-1, B = cell c (primitives); ta/tb = central transverse
//   derivatives at c-1 / c: index [t][0] = d(u_t)/dx_t, [t][1] = d(u_D)/dx_t for
//   the two transverse dims t in x,y,z order (t != D).
// Adds the diffusive flux to F (mass flux unchanged).
// ---------------------------------------------------------------------------
struct ViscConst {
  float nu;        // kinematic viscosity
  float kfac;      // (1/Pr) * cp
  float inv_h[3];      // 1/dx, 1/dy, 1/dz
  float h[3];          // dx, dy, dz
  float inv_two_h[3];  // 1/(2 dx), 1/(2 dy), 1/(2 dz)
};

// Per-cell factors of the heat flux: T = p / (rho R) and kappa / nu-free
// part k = ((1/Pr) cp) (nu rho) (simple.py:104-108); a marching sweep computes
// them once per cell and hands cell c's values to the next face.
struct ViscCell {
  float T, k;
};
__device__ __forceinline__ ViscCell visc_cell(const float* W, const ViscConst& vc) {
  return ViscCell{W[kP] * frcp(W[kRho] * kR), vc.kfac * (vc.nu * W[kRho])};
}

template <int D, int NS>
__device__ __forceinline__ void viscous_flux_cells(const float* A, const float* B,
                                                   const float (*ta)[2],
                                                   const float (*tb)[2], ViscCell ca,
                                                   ViscCell cb, const ViscConst& vc,
                                                   float* F) {
  // The reference's face averages 0.5*(a + b) are carried here as the plain
  // sums a + b, with each factor 1/2 moved to where the sum is used: scaling
  // by a power of two commutes with rounding, so every result below is
  // bit-identical to the direct form (outside the subnormal range) at fewer
  // multiplies.
  const float ih = vc.inv_h[D];
  const float ih2 = 2.0f * ih;
  // Backward (normal) derivatives on the face; g2 = 2 * gn.
  float gn[3], g2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float d = B[kU + i] - A[kU + i];
    if (i == D) gn[i] = d * ih;
    g2[i] = d * ih2;
  }
  // Transverse derivatives interpolated onto the face, times 2: f[c-1] + f[c].
  float st[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    st[t][0] = ta[t][0] + tb[t][0];
    st[t][1] = ta[t][1] + tb[t][1];
  }
  // Twice the divergence on the face, summed in x, y, z order.
  float g3[3];
  {
    int t = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k == D) {
        g3[k] = g2[D];
      } else {
        g3[k] = st[t][0];
        ++t;
      }
    }
  }
  const float div2 = (g3[0] + g3[1]) + g3[2];
  // Strain rate S_iD on the face: gn - div/3 and 0.5*(gn_i + transverse).
  float S[3];
  {
    int t = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i == D) {
        S[i] = gn[D] - (0.5f * kThird) * div2;
      } else {
        S[i] = 0.25f * (g2[i] + st[t][1]);
        ++t;
      }
    }
  }
  // 2 * rho_f * nu with rho_f = 0.5 * (rho[c-1] + rho[c]).
  const float c2 = (A[kRho] + B[kRho]) * vc.nu;
  const float tx = c2 * S[0], ty = c2 * S[1], tz = c2 * S[2];
  // Twice the face velocities.
  const float uf2 = A[kU] + B[kU];
  const float vf2 = A[kV] + B[kV];
  const float wf2 = A[kW] + B[kW];
  // Heat flux -kappa_f * dT/dx, T = p/(rho R), kappa = ((1/Pr)*cp)*(nu*rho),
  // kappa_f = 0.5 * (kappa[c-1] + kappa[c]).
  const float q = -(ca.k + cb.k) * ((cb.T - ca.T) * (0.5f * ih));
  F[1] = F[1] + (-tx);
  F[2] = F[2] + (-ty);
  F[3] = F[3] + (-tz);
  F[4] = F[4] + (q + (-(0.5f * ((uf2 * tx + vf2 * ty) + wf2 * tz))));
}

template <int D, int NS>
__device__ __forceinline__ void viscous_flux(const float* A, const float* B,
                                             const float (*ta)[2],
                                             const float (*tb)[2],
                                             const ViscConst& vc, float* F) {
  viscous_flux_cells<D, NS>(A, B, ta, tb, visc_cell(A, vc), visc_cell(B, vc), vc, F);
}

}  // namespace sc
