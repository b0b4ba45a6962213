// Reverse-mode (adjoint) counterparts of the per-cell and per-face functions
// of sc_device.hpp.  Each *_adj function recomputes the forward values it
// needs and accumulates (+=) the adjoints of its inputs from the adjoints of
// its outputs, for exactly the arithmetic the forward kernels execute (same
// reciprocals, same branch of every min / max / select: the HLLC branch and
// the wave-speed arguments are replayed from the recomputed forward).
#pragma once

#include "sc_device.hpp"

namespace sc {

// Adjoint of cons_to_prim<NS> (common/utils.py:145-183, generic.py:27-88):
// w = (rho, u, v, w, e, p, h, y_s..) from q = (rho, rho u, rho v, rho w,
// rho e, rho y_s..).  Takes the forward primitives W (the velocities, energy
// and scalars are q * (1/rho)).
template <int NS = 0>
__device__ __forceinline__ void cons_to_prim_adj(const float* W, float pe, const float* lw,
                                                 float* lq) {
  const float rho = W[kRho], inv = frcp(rho);
  const float u = W[kU], v = W[kV], w = W[kW], e = W[kE], p = W[kP];
  const float ke = 0.5f * ((u * u + v * v) + w * w);
  const float eint = (e - ke) - pe;
  float lrho = lw[kRho], lu = lw[kU], lv = lw[kV], lww = lw[kW], le = lw[kE];
  float lp = lw[kP];
  const float lh = lw[kH];
  // h = e + p * inv
  le += lh;
  lp += lh * inv;
  float linv = lh * p;
  // p = (g1 * rho) * eint
  lrho += lp * kGm1 * eint;
  const float leint = lp * (kGm1 * rho);
  // eint = (e - ke) - pe
  le += leint;
  const float lke = -leint;
  lu += lke * u;
  lv += lke * v;
  lww += lke * w;
  // phi = inv * q (q = rho * phi up to rounding)
  lq[1] += lu * inv;
  lq[2] += lv * inv;
  lq[3] += lww * inv;
  lq[4] += le * inv;
  linv += (lu * u + lv * v + lww * w + le * e) * rho;
  if constexpr (NS > 0) {
    float ls = 0.0f;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      lq[5 + s] += lw[7 + s] * inv;
      ls += lw[7 + s] * W[7 + s];
    }
    linv += ls * rho;
  }
  lrho += -linv * inv * inv;
  lq[0] += lrho;
}

// ---------------------------------------------------------------------------
// Forward replay of the face reconstruction inside the adjoint passes.  For
// WENO-5 this evaluates the candidates and smoothness indicators directly
// from the cell values (the same function as recon<kWeno5>, rounded
// differently): the difference form of recon<kWeno5> shares its first and
// second differences with recon_adj's, and the compiler then keeps them alive
// across the Riemann adjoint between the two calls (+60 registers, one
// occupancy level less).
// ---------------------------------------------------------------------------
template <int SCHEME>
__device__ __forceinline__ void recon_replay(const float* a, float& fp, float& fm) {
  recon<SCHEME>(a, fp, fm);
}

template <>
__device__ __forceinline__ void recon_replay<kWeno5>(const float* a, float& fp, float& fm) {
  const float vm2 = a[0], vm1 = a[1], v0 = a[2], vp1 = a[3], vp2 = a[4];
  const float c1312 = (float)(13.0 / 12.0);
  float t, s;
  t = v0 - 2.0f * vp1 + vp2;
  s = 3.0f * v0 - 4.0f * vp1 + vp2;
  const float b0 = c1312 * (t * t) + 0.25f * (s * s);
  t = vm1 - 2.0f * v0 + vp1;
  s = vm1 - vp1;
  const float b1 = c1312 * (t * t) + 0.25f * (s * s);
  t = vm2 - 2.0f * vm1 + v0;
  s = vm2 - 4.0f * vm1 + 3.0f * v0;
  const float b2 = c1312 * (t * t) + 0.25f * (s * s);
  const float c13 = (float)(1.0 / 3.0), c56 = (float)(5.0 / 6.0);
  const float c16 = (float)(-1.0 / 6.0), c76 = (float)(-7.0 / 6.0);
  const float c116 = (float)(11.0 / 6.0);
  const float qn0 = (c13 * v0 + c56 * vp1) + c16 * vp2;
  const float qn1 = (c16 * vm1 + c56 * v0) + c13 * vp1;
  const float qn2 = (c13 * vm2 + c76 * vm1) + c116 * v0;
  const float qp0 = (c116 * v0 + c76 * vp1) + c13 * vp2;
  const float qp1 = (c13 * vm1 + c56 * v0) + c16 * vp1;
  const float qp2 = (c16 * vm2 + c56 * vm1) + c13 * v0;
  const float e0 = kWenoEps + b0, e1 = kWenoEps + b1, e2 = kWenoEps + b2;
  const float i0 = frcp(e0 * e0), i1 = frcp(e1 * e1), i2 = frcp(e2 * e2);
  const float an0 = 0.3f * i0, an1 = 0.6f * i1, an2 = 0.1f * i2;
  const float ap0 = 0.1f * i0, ap1 = 0.6f * i1, ap2 = 0.3f * i2;
  const float rn = frcp((an0 + an1) + an2);
  const float rp = frcp((ap0 + ap1) + ap2);
  fp = ((an0 * qn0 + an1 * qn1) + an2 * qn2) * rn;
  fm = ((ap0 * qp0 + ap1 * qp1) + ap2 * qp2) * rp;
}

// ---------------------------------------------------------------------------
// Reconstruction adjoints: given the adjoints lfp (left state of face c+1/2)
// and lfm (right state of face c-1/2), accumulate into la[] (offsets -R..R).
// ---------------------------------------------------------------------------
template <int SCHEME>
__device__ __forceinline__ void recon_adj(const float* a, float lfp, float lfm, float* la);

template <>
__device__ __forceinline__ void recon_adj<kWeno5>(const float* a, float lfp, float lfm,
                                                  float* la) {
  const float vm2 = a[0], vm1 = a[1], v0 = a[2], vp1 = a[3], vp2 = a[4];
  // First differences of the stencil: the smoothness indicators and the
  // differences of the candidate values are formed from them (accurate to
  // their own size on smooth data).
  const float dm2 = vm1 - vm2, dm1 = v0 - vm1, d0 = vp1 - v0, d1 = vp2 - vp1;
  const float c1312 = (float)(13.0 / 12.0);
  const float t0 = d1 - d0, s0 = d1 - 3.0f * d0;
  const float t1 = d0 - dm1, s1 = -(dm1 + d0);
  const float t2 = dm1 - dm2, s2 = 3.0f * dm1 - dm2;
  const float b0 = c1312 * (t0 * t0) + 0.25f * (s0 * s0);
  const float b1 = c1312 * (t1 * t1) + 0.25f * (s1 * s1);
  const float b2 = c1312 * (t2 * t2) + 0.25f * (s2 * s2);
  const float c13 = (float)(1.0 / 3.0), c56 = (float)(5.0 / 6.0);
  const float c16 = (float)(-1.0 / 6.0), c76 = (float)(-7.0 / 6.0);
  const float c116 = (float)(11.0 / 6.0);
  const float e0 = kWenoEps + b0, e1 = kWenoEps + b1, e2 = kWenoEps + b2;
  const float i0 = frcp(e0 * e0), i1 = frcp(e1 * e1), i2 = frcp(e2 * e2);
  const float an0 = 0.3f * i0, an1 = 0.6f * i1, an2 = 0.1f * i2;
  const float ap0 = 0.1f * i0, ap1 = 0.6f * i1, ap2 = 0.3f * i2;
  const float rn = frcp((an0 + an1) + an2);
  const float rp = frcp((ap0 + ap1) + ap2);
  // f = N / S with N = sum a_r q_r, S = sum a_r: df/da_r = (q_r - f) / S,
  // and q_r - f = sum_s w_s (q_r - q_s) (not q_r - f in fp32, which cancels
  // when the candidates nearly agree).
  const float c16p = (float)(1.0 / 6.0);
  const float n01 = c16p * ((2.0f * d0 - dm1) - d1);   // qn0 - qn1
  const float n12 = c13 * ((dm2 - 2.0f * dm1) + d0);   // qn1 - qn2
  const float p01 = c13 * ((dm1 - 2.0f * d0) + d1);    // qp0 - qp1
  const float p12 = c16p * ((2.0f * dm1 - dm2) - d0);  // qp1 - qp2
  const float wn0 = an0 * rn, wn1 = an1 * rn, wn2 = an2 * rn;
  const float wp0 = ap0 * rp, wp1 = ap1 * rp, wp2 = ap2 * rp;
  const float lNn = lfp * rn, lNp = lfm * rp;
  const float lan0 = lNn * ((wn1 + wn2) * n01 + wn2 * n12);
  const float lan1 = lNn * (wn2 * n12 - wn0 * n01);
  const float lan2 = lNn * (-(wn0 * (n01 + n12) + wn1 * n12));
  const float lap0 = lNp * ((wp1 + wp2) * p01 + wp2 * p12);
  const float lap1 = lNp * (wp2 * p12 - wp0 * p01);
  const float lap2 = lNp * (-(wp0 * (p01 + p12) + wp1 * p12));
  const float lqn0 = lNn * an0, lqn1 = lNn * an1, lqn2 = lNn * an2;
  const float lqp0 = lNp * ap0, lqp1 = lNp * ap1, lqp2 = lNp * ap2;
  // alpha = d * i, i = (eps + beta)^-2: d i / d beta = -2 i / (eps + beta),
  // with 1 / (eps + beta) = (eps + beta) * i (no further reciprocal).
  const float li0 = 0.3f * lan0 + 0.1f * lap0;
  const float li1 = 0.6f * (lan1 + lap1);
  const float li2 = 0.1f * lan2 + 0.3f * lap2;
  const float lb0 = -2.0f * li0 * i0 * (e0 * i0);
  const float lb1 = -2.0f * li1 * i1 * (e1 * i1);
  const float lb2 = -2.0f * li2 * i2 * (e2 * i2);
  // beta = c t^2 + 0.25 s^2 (smoothness-indicator derivatives).
  const float lt0 = 2.0f * c1312 * t0 * lb0, ls0 = 0.5f * s0 * lb0;
  const float lt1 = 2.0f * c1312 * t1 * lb1, ls1 = 0.5f * s1 * lb1;
  const float lt2 = 2.0f * c1312 * t2 * lb2, ls2 = 0.5f * s2 * lb2;
  la[2] += lt0 + 3.0f * ls0;
  la[3] += -2.0f * lt0 - 4.0f * ls0;
  la[4] += lt0 + ls0;
  la[1] += lt1 + ls1;
  la[2] += -2.0f * lt1;
  la[3] += lt1 - ls1;
  la[0] += lt2 + ls2;
  la[1] += -2.0f * lt2 - 4.0f * ls2;
  la[2] += lt2 + 3.0f * ls2;
  // Candidate stencils (linear).
  la[2] += c13 * lqn0 + c56 * lqn1 + c116 * lqn2 + c116 * lqp0 + c56 * lqp1 + c13 * lqp2;
  la[3] += c56 * lqn0 + c13 * lqn1 + c76 * lqp0 + c16 * lqp1;
  la[4] += c16 * lqn0 + c13 * lqp0;
  la[1] += c16 * lqn1 + c76 * lqn2 + c13 * lqp1 + c56 * lqp2;
  la[0] += c13 * lqn2 + c16 * lqp2;
}

template <>
__device__ __forceinline__ void recon_adj<kWeno3>(const float* a, float lfp, float lfm,
                                                  float* la) {
  const float vm1 = a[0], v0 = a[1], vp1 = a[2];
  const float t0 = vp1 - v0, t1 = v0 - vm1;
  const float b0 = t0 * t0, b1 = t1 * t1;
  const float qn0 = 0.5f * v0 + 0.5f * vp1;
  const float qn1 = -0.5f * vm1 + 1.5f * v0;
  const float qp0 = 1.5f * v0 + -0.5f * vp1;
  const float qp1 = 0.5f * vm1 + 0.5f * v0;
  const float d0 = (float)(2.0 / 3.0), d1 = (float)(1.0 / 3.0);
  const float e0 = kWenoEps + b0, e1 = kWenoEps + b1;
  const float i0 = frcp(e0 * e0), i1 = frcp(e1 * e1);
  const float an0 = d0 * i0, an1 = d1 * i1;
  const float ap0 = d1 * i0, ap1 = d0 * i1;
  const float rn = frcp(an0 + an1), rp = frcp(ap0 + ap1);
  // df/da_r = (q_r - f) / S with q_r - f from the stencil differences.
  const float n01 = 0.5f * (t0 - t1);  // qn0 - qn1
  const float p01 = 0.5f * (t1 - t0);  // qp0 - qp1
  const float lNn = lfp * rn, lNp = lfm * rp;
  const float lan0 = lNn * (an1 * rn) * n01, lan1 = -lNn * (an0 * rn) * n01;
  const float lap0 = lNp * (ap1 * rp) * p01, lap1 = -lNp * (ap0 * rp) * p01;
  const float lqn0 = lNn * an0, lqn1 = lNn * an1, lqp0 = lNp * ap0, lqp1 = lNp * ap1;
  const float li0 = d0 * lan0 + d1 * lap0, li1 = d1 * lan1 + d0 * lap1;
  const float lb0 = -2.0f * li0 * i0 * (e0 * i0), lb1 = -2.0f * li1 * i1 * (e1 * i1);
  const float lt0 = 2.0f * t0 * lb0, lt1 = 2.0f * t1 * lb1;
  la[2] += lt0;
  la[1] += -lt0 + lt1;
  la[0] += -lt1;
  la[1] += 0.5f * lqn0 + 1.5f * lqn1 + 1.5f * lqp0 + 0.5f * lqp1;
  la[2] += 0.5f * lqn0 - 0.5f * lqp0;
  la[0] += -0.5f * lqn1 + 0.5f * lqp1;
}

template <>
__device__ __forceinline__ void recon_adj<kMuscl>(const float* a, float lfp, float lfm,
                                                  float* la) {
  const float vm1 = a[0], v0 = a[1], vp1 = a[2];
  const float g0 = (vp1 - vm1) * 0.25f, g1 = vp1 - v0, g2 = v0 - vm1;
  const float s = (sgnf(g0) + sgnf(g1)) + sgnf(g2);
  la[1] += lfp + lfm;
  if (fabsf(s) != 3.0f) return;
  // m = the slope of least magnitude (all three share its sign).
  const float lm = 0.5f * (lfp - lfm);
  const float a0 = fabsf(g0), a1 = fabsf(g1), a2 = fabsf(g2);
  if (a0 <= a1 && a0 <= a2) {
    la[2] += 0.25f * lm;
    la[0] -= 0.25f * lm;
  } else if (a1 <= a2) {
    la[2] += lm;
    la[1] -= lm;
  } else {
    la[1] += lm;
    la[0] -= lm;
  }
}

// ---------------------------------------------------------------------------
// Riemann adjoints.  L / R: reconstructed face primitives (rho u v w e p h);
// lF: adjoint of the 5 face fluxes; accumulates into lL, lR.
// ---------------------------------------------------------------------------

// Adjoint of physical_flux<D, NS, PD> (PD: the aligned momentum flux takes
// the dynamic pressure s[7 + NS], the energy flux keeps p; scalar s is
// advected as m * y_s).
template <int D, int NS = 0, bool PD = false>
__device__ __forceinline__ void physical_flux_adj(const float* s, const float* lF, float* ls) {
  const float rho = s[kRho], vn = s[kU + D], ir = frcp(rho);
  const float m = rho * vn;
  float lm = lF[0] + lF[4] * (s[kE] + s[kP] * ir);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lm += lF[1 + k] * s[kU + k];
    ls[kU + k] += lF[1 + k] * m;
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    lm += lF[5 + k] * s[7 + k];
    ls[7 + k] += lF[5 + k] * m;
  }
  if constexpr (PD) {
    ls[7 + NS] += lF[1 + D];
    ls[kP] += lF[4] * m * ir;
  } else {
    ls[kP] += lF[1 + D] + lF[4] * m * ir;
  }
  ls[kE] += lF[4] * m;
  ls[kRho] += -lF[4] * m * s[kP] * ir * ir;
  ls[kRho] += lm * vn;
  ls[kU + D] += lm * rho;
}

// Adjoint of roe_speeds<D> given lsl, lsr.
template <int D>
__device__ __forceinline__ void roe_speeds_adj(const float* L, const float* R, float pe,
                                               float lsl, float lsr, float* lL, float* lR) {
  const float irl = frcp(L[kRho]), irr = frcp(R[kRho]);
  const float al = kGamma * L[kP] * irl, ar = kGamma * R[kP] * irr;
  const float cl = fsqrt(al), cr = fsqrt(ar);
  const float ql = fsqrt(L[kRho]), qr = fsqrt(R[kRho]);
  const float sum = ql + qr;
  const float inv = rnn(sum);
  float rX[4];
  const int idx[4] = {kU, kV, kW, kH};
#pragma unroll
  for (int k = 0; k < 4; ++k) rX[k] = (ql * L[idx[k]] + qr * R[idx[k]]) * inv;
  const float hs = (rX[3] - 0.5f * ((rX[0] * rX[0] + rX[1] * rX[1]) + rX[2] * rX[2])) - pe;
  const float c = fsqrt(kGm1 * hs);
  const float un = rX[D];
  float lcl = 0.0f, lcr = 0.0f, lun = 0.0f, lc = 0.0f;
  // sl = min(vl - cl, un - c)
  if (L[kU + D] - cl <= un - c) {
    lL[kU + D] += lsl;
    lcl -= lsl;
  } else {
    lun += lsl;
    lc -= lsl;
  }
  // sr = max(vr + cr, un + c)
  if (R[kU + D] + cr >= un + c) {
    lR[kU + D] += lsr;
    lcr += lsr;
  } else {
    lun += lsr;
    lc += lsr;
  }
  float lrX[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  lrX[D] += lun;
  // c = sqrt(g1 hs), hs = rh - 0.5 |ru|^2 - pe
  const float lhs = c > 0.0f ? lc * kGm1 * 0.5f * frcp(c) : 0.0f;
  lrX[3] += lhs;
  lrX[0] -= lhs * rX[0];
  lrX[1] -= lhs * rX[1];
  lrX[2] -= lhs * rX[2];
  float lql = 0.0f, lqr = 0.0f, linv = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lL[idx[k]] += lrX[k] * ql * inv;
    lR[idx[k]] += lrX[k] * qr * inv;
    lql += lrX[k] * L[idx[k]] * inv;
    lqr += lrX[k] * R[idx[k]] * inv;
    linv += lrX[k] * (ql * L[idx[k]] + qr * R[idx[k]]);
  }
  const float lsum = sum != 0.0f ? -linv * inv * inv : 0.0f;
  lql += lsum;
  lqr += lsum;
  if (ql > 0.0f) lL[kRho] += lql * 0.5f * frcp(ql);
  if (qr > 0.0f) lR[kRho] += lqr * 0.5f * frcp(qr);
  // c = sqrt(gamma p / rho)
  if (cl > 0.0f) {
    const float la = lcl * 0.5f * frcp(cl);
    lL[kP] += la * kGamma * irl;
    lL[kRho] -= la * al * irl;
  }
  if (cr > 0.0f) {
    const float la = lcr * 0.5f * frcp(cr);
    lR[kP] += la * kGamma * irr;
    lR[kRho] -= la * ar * irr;
  }
}

// Adjoint of hllc_star<D, 0> (flux f enters additively: its adjoint is lfs
// itself and is handled by the caller).
template <int D, int NS = 0>
__device__ __forceinline__ void hllc_star_adj(const float* S, float sk, float sm,
                                              const float* lfs, float* lS, float& lsk,
                                              float& lsm) {
  const float rho = S[kRho], vn = S[kU + D], ir = frcp(rho);
  const float den = sk - sm;
  const float r = den == 0.0f ? 0.0f : frcp(den);
  const float q = sk * r;
  const float a = rho * (sm - vn);
  float la = lfs[0] * q, lq = lfs[0] * a, lr = 0.0f, lvn = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k == D) {
      const float g = ((sk - vn) * r) * sm - vn;
      const float X = rho * sk;
      const float lX = lfs[1 + k] * g, lg = lfs[1 + k] * X;
      lS[kRho] += lX * sk;
      lsk += lX * rho;
      lsk += lg * r * sm;
      lvn -= lg * r * sm;
      lr += lg * (sk - vn) * sm;
      lsm += lg * (sk - vn) * r;
      lvn -= lg;
    } else {
      la += lfs[1 + k] * S[kU + k] * q;
      lS[kU + k] += lfs[1 + k] * a * q;
      lq += lfs[1 + k] * a * S[kU + k];
    }
  }
  // Scalars: fs = f + (a * y) * q.
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    la += lfs[5 + k] * S[7 + k] * q;
    lS[7 + k] += lfs[5 + k] * a * q;
    lq += lfs[5 + k] * a * S[7 + k];
  }
  {
    const float A = (rho * q) * (sm - vn);
    const float B = (S[kE] + S[kP] * ir) + (sk - vn) * sm;
    const float lA = lfs[4] * B, lB = lfs[4] * A;
    lS[kRho] += lA * q * (sm - vn);
    lq += lA * rho * (sm - vn);
    lsm += lA * rho * q;
    lvn -= lA * rho * q;
    lS[kE] += lB;
    lS[kP] += lB * ir;
    lS[kRho] -= lB * S[kP] * ir * ir;
    lsk += lB * sm;
    lvn -= lB * sm;
    lsm += lB * (sk - vn);
  }
  lS[kRho] += la * (sm - vn);
  lsm += la * rho;
  lvn -= la * rho;
  lsk += lq * r;
  lr += lq * sk;
  if (den != 0.0f) {
    const float lden = -lr * r * r;
    lsk += lden;
    lsm -= lden;
  }
  lS[kU + D] += lvn;
}

// PD: L0, R0 carry the dynamic pressure at index 7; the wave speeds, the
// contact speed and the star state use it in place of p (hllc.py:109-120),
// the physical flux both (physical_flux_adj<D, true>).
template <int D, int NS = 0, bool PD = false>
__device__ __forceinline__ void hllc_adj(const float* L0, const float* R0, float pe,
                                         const float* lF, float* lL0, float* lR0) {
  constexpr int NB = NVar<NS>::value;  // primitives without p_dyn
  constexpr int NV = NB + (PD ? 1 : 0);
  constexpr int KD = 7 + NS;           // p_dyn's index
  float Lp[NB], Rp[NB];
  if constexpr (PD) {
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      Lp[v] = L0[v];
      Rp[v] = R0[v];
    }
    Lp[kP] = L0[KD];
    Rp[kP] = R0[KD];
  }
  const float* L = PD ? Lp : L0;
  const float* R = PD ? Rp : R0;
  // Adjoints with respect to L, R (with PD: p stands for p_dyn); merged into
  // lL0, lR0 at the end.
  float lLp[NB], lRp[NB];
  float* lL = lL0;
  float* lR = lR0;
  if constexpr (PD) {
#pragma unroll
    for (int v = 0; v < NB; ++v) lLp[v] = lRp[v] = 0.0f;
    lL = lLp;
    lR = lRp;
  }
  float sl, sr;
  roe_speeds<D>(L, R, pe, sl, sr);
  const float vl = L[kU + D], vr = R[kU + D];
  const float rl = L[kRho], rr = R[kRho];
  const float num = ((R[kP] - L[kP]) + (rl * vl) * (sl - vl)) - (rr * vr) * (sr - vr);
  const float den = rl * (sl - vl) - rr * (sr - vr);
  const float sm_raw = dnn(num, den);
  const bool snapped = fabsf(sm_raw) < kSmSnap;
  const float sm = snapped ? 0.0f : sm_raw;
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
  // One side's flux (L for branches 0, 1, R for 2, 3), selected per lane as
  // in the forward, so the lanes of a wave share one adjoint evaluation.
  const bool left = br <= 1;
  float lsk = 0.0f, lsm = 0.0f;
  if constexpr (PD) {
    // Physical flux of the side's own state (p and p_dyn), star correction
    // of the state with p_dyn for p.
    float S0[NV], Ss[NB], lS0[NV], lSs[NB];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      S0[v] = left ? L0[v] : R0[v];
      lS0[v] = 0.0f;
    }
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      Ss[v] = left ? L[v] : R[v];
      lSs[v] = 0.0f;
    }
    physical_flux_adj<D, NS, true>(S0, lF, lS0);
    if (br == 1 || br == 2)
      hllc_star_adj<D, NS>(Ss, left ? sl : sr, sm, lF, lSs, lsk, lsm);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      lL0[v] += left ? lS0[v] : 0.0f;
      lR0[v] += left ? 0.0f : lS0[v];
    }
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      lL[v] += left ? lSs[v] : 0.0f;
      lR[v] += left ? 0.0f : lSs[v];
    }
  } else {
    float S[NB], lS[NB];
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      S[v] = left ? L[v] : R[v];
      lS[v] = 0.0f;
    }
    physical_flux_adj<D, NS>(S, lF, lS);
    if (br == 1 || br == 2) hllc_star_adj<D, NS>(S, left ? sl : sr, sm, lF, lS, lsk, lsm);
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      lL[v] += left ? lS[v] : 0.0f;
      lR[v] += left ? 0.0f : lS[v];
    }
  }
  float lsl = left ? lsk : 0.0f, lsr = left ? 0.0f : lsk;
  if (!snapped && den != 0.0f && lsm != 0.0f) {
    const float ir = frcp(den);
    const float lnum = lsm * ir;
    const float lden = -lsm * sm_raw * ir;
    lR[kP] += lnum;
    lL[kP] -= lnum;
    lL[kRho] += lnum * vl * (sl - vl);
    lL[kU + D] += lnum * rl * (sl - 2.0f * vl);
    lsl += lnum * rl * vl;
    lR[kRho] -= lnum * vr * (sr - vr);
    lR[kU + D] -= lnum * rr * (sr - 2.0f * vr);
    lsr -= lnum * rr * vr;
    lL[kRho] += lden * (sl - vl);
    lsl += lden * rl;
    lL[kU + D] -= lden * rl;
    lR[kRho] -= lden * (sr - vr);
    lsr -= lden * rr;
    lR[kU + D] += lden * rr;
  }
  roe_speeds_adj<D>(L, R, pe, lsl, lsr, lL, lR);
  if constexpr (PD) {
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      const int dst = v == kP ? KD : v;
      lL0[dst] += lLp[v];
      lR0[dst] += lRp[v];
    }
  }
}

// PD: the physical fluxes take p_dyn (index 7 + NS) in the aligned momentum
// flux; the wave speeds keep p (hll.py:107-115).
template <int D, int NS = 0, bool PD = false>
__device__ __forceinline__ void hll_adj(const float* L, const float* R, float pe,
                                        const float* lF, float* lL, float* lR) {
  constexpr int NF = 5 + NS;
  float sl, sr;
  roe_speeds<D>(L, R, pe, sl, sr);
  const float ds = sr - sl;
  const float rds = ds == 0.0f ? 0.0f : frcp(ds);
  const float n1 = tf_min(sr, 0.0f) - tf_min(sl, 0.0f);
  const float t1 = n1 * rds;
  const float t2 = 1.0f - t1;
  const float n3 = sr * fabsf(sl) - sl * fabsf(sr);
  const float t3 = n3 * (0.5f * rds);
  float fl[NF], fr[NF], ul[NF], ur[NF];
  physical_flux<D, NS, PD>(L, fl);
  physical_flux<D, NS, PD>(R, fr);
  // Conserved face states: u_0 = rho, u_k = rho * phi_k (phi at primitive
  // index k for k = 1..4, 7 + s for scalar s).
  auto pidx = [](int k) { return k < 5 ? k : k + 2; };
  ul[0] = L[kRho];
  ur[0] = R[kRho];
#pragma unroll
  for (int k = 1; k < NF; ++k) {
    ul[k] = L[kRho] * L[pidx(k)];
    ur[k] = R[kRho] * R[pidx(k)];
  }
  float lt1 = 0.0f, lt2 = 0.0f, lt3 = 0.0f, lfl[NF], lfr[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    lt1 += lF[k] * fr[k];
    lt2 += lF[k] * fl[k];
    lt3 -= lF[k] * (ur[k] - ul[k]);
    lfr[k] = lF[k] * t1;
    lfl[k] = lF[k] * t2;
    const float lur = -lF[k] * t3, lul = lF[k] * t3;
    if (k == 0) {
      lR[kRho] += lur;
      lL[kRho] += lul;
    } else {
      const int pk = pidx(k);
      lR[kRho] += lur * R[pk];
      lR[pk] += lur * R[kRho];
      lL[kRho] += lul * L[pk];
      lL[pk] += lul * L[kRho];
    }
  }
  physical_flux_adj<D, NS, PD>(R, lfr, lR);
  physical_flux_adj<D, NS, PD>(L, lfl, lL);
  lt1 -= lt2;
  float lsl = 0.0f, lsr = 0.0f, lrds = lt1 * n1 + lt3 * 0.5f * n3;
  if (sr < 0.0f) lsr += lt1 * rds;
  if (sl < 0.0f) lsl -= lt1 * rds;
  const float ln3 = lt3 * 0.5f * rds;
  lsr += ln3 * (fabsf(sl) - sl * sgnf(sr));
  lsl += ln3 * (sr * sgnf(sl) - fabsf(sr));
  if (ds != 0.0f) {
    const float lds = -lrds * rds * rds;
    lsr += lds;
    lsl -= lds;
  }
  roe_speeds_adj<D>(L, R, pe, lsl, lsr, lL, lR);
}

template <int FLUX, int D, int NS = 0, bool PD = false>
__device__ __forceinline__ void riemann_adj(const float* L, const float* R, float pe,
                                            const float* lF, float* lL, float* lR) {
  if constexpr (FLUX == kHllc) {
    hllc_adj<D, NS, PD>(L, R, pe, lF, lL, lR);
  } else {
    hll_adj<D, NS, PD>(L, R, pe, lF, lL, lR);
  }
}

// Adjoint of viscous_flux<D, 0>.  A, B: prims of cells c-1, c; ta, tb: their
// transverse central derivatives.  Accumulates into lA, lB (rho u v w p) and
// lta, ltb.
template <int D>
__device__ __forceinline__ void viscous_adj(const float* A, const float* B,
                                            const float (*ta)[2], const float (*tb)[2],
                                            const ViscConst& vc, const float* lF, float* lA,
                                            float* lB, float (*lta)[2], float (*ltb)[2]) {
  const float ih = vc.inv_h[D];
  const float gn[3] = {(B[kU] - A[kU]) * ih, (B[kV] - A[kV]) * ih, (B[kW] - A[kW]) * ih};
  float it[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    it[t][0] = 0.5f * (ta[t][0] + tb[t][0]);
    it[t][1] = 0.5f * (ta[t][1] + tb[t][1]);
  }
  float g3[3];
  int tix[3];
  {
    int t = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k == D) {
        g3[k] = gn[D];
        tix[k] = -1;
      } else {
        g3[k] = it[t][0];
        tix[k] = t;
        ++t;
      }
    }
  }
  const float div = (g3[0] + g3[1]) + g3[2];
  float S[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    S[i] = i == D ? gn[D] - kThird * div : 0.5f * (gn[i] + it[tix[i]][1]);
  const float rho_f = 0.5f * (A[kRho] + B[kRho]);
  const float c2 = (2.0f * rho_f) * vc.nu;
  const float tau[3] = {c2 * S[0], c2 * S[1], c2 * S[2]};
  const float uf[3] = {0.5f * (A[kU] + B[kU]), 0.5f * (A[kV] + B[kV]), 0.5f * (A[kW] + B[kW])};
  const float ira = frcp(A[kRho] * kR), irb = frcp(B[kRho] * kR);
  const float Ta = A[kP] * ira, Tb = B[kP] * irb;
  const float ka = vc.kfac * (vc.nu * A[kRho]), kb = vc.kfac * (vc.nu * B[kRho]);
  const float kf = 0.5f * (ka + kb);
  // Reverse.
  float ltau[3], luf[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ltau[i] = -lF[1 + i] - lF[4] * uf[i];
    luf[i] = -lF[4] * tau[i];
  }
  const float lq = lF[4];
  const float lkf = -lq * (Tb - Ta) * ih;
  const float lTb = -lq * kf * ih, lTa = lq * kf * ih;
  lA[kRho] += 0.5f * lkf * vc.kfac * vc.nu;
  lB[kRho] += 0.5f * lkf * vc.kfac * vc.nu;
  lA[kP] += lTa * ira;
  lA[kRho] -= lTa * Ta * frcp(A[kRho]);
  lB[kP] += lTb * irb;
  lB[kRho] -= lTb * Tb * frcp(B[kRho]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lA[kU + i] += 0.5f * luf[i];
    lB[kU + i] += 0.5f * luf[i];
  }
  float lc2 = 0.0f, lS[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lc2 += ltau[i] * S[i];
    lS[i] = ltau[i] * c2;
  }
  const float lrho_f = lc2 * 2.0f * vc.nu;
  lA[kRho] += 0.5f * lrho_f;
  lB[kRho] += 0.5f * lrho_f;
  float lgn[3] = {0.0f, 0.0f, 0.0f}, lit[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}}, ldiv = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i == D) {
      lgn[D] += lS[i];
      ldiv -= kThird * lS[i];
    } else {
      lgn[i] += 0.5f * lS[i];
      lit[tix[i]][1] += 0.5f * lS[i];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k == D) {
      lgn[D] += ldiv;
    } else {
      lit[tix[k]][0] += ldiv;
    }
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      lta[t][c] += 0.5f * lit[t][c];
      ltb[t][c] += 0.5f * lit[t][c];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lB[kU + k] += lgn[k] * ih;
    lA[kU + k] -= lgn[k] * ih;
  }
}

}  // namespace sc
