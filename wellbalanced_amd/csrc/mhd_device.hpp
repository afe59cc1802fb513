// Device math of the ideal-MHD step: MUSCL slopes, the HLLD flux, the CT-contact
// upwind average and the primitive conversion, templated on the arithmetic type.
//
// The double instantiation is the parity path: every expression keeps the tree of
// the reference's scalar kernels (carafe kernels/mhd_sweep/mhd_sweep_pyccel.py,
// ct_contact_pyccel.py, mhd_primitive_pyccel.py, mhd_rate_pyccel.py) and the
// translation unit that instantiates it is built with -ffp-contract=off, so each
// +,-,*,/,sqrt rounds exactly as NumPy's does.  The float instantiation is the
// mixed-precision flux path: reciprocal and square root use the 1-ulp hardware
// instructions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wb {

constexpr int HALO = 2;
constexpr double DEGENERATE = 1.0e-12;  // mhd.py:138, mhd_sweep_pyccel.py:127

template <typename R> struct Arith;

template <> struct Arith<double> {
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double sqrt(double x) { return __builtin_sqrt(x); }
};

template <> struct Arith<float> {
  static __device__ __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
};

template <typename R> __device__ __forceinline__ R absv(R x) { return x < R(0) ? -x : x; }

// Limited slope of a cell from its backward and forward differences
// (mhd_sweep_pyccel.py:25-48): 0 minmod, 1 van Leer, 2 MC, 3 central.
template <typename R>
__device__ __forceinline__ R limited_slope(R a, R b, int limiter) {
  R slope = R(0);
  if (limiter == 3) {
    slope = R(0.5) * (a + b);
  } else if (a * b > R(0)) {
    if (limiter == 0) {
      slope = b;
      if (absv(a) < absv(b)) slope = a;
    } else if (limiter == 1) {
      R product = a * b;
      slope = Arith<R>::div(R(2) * product, a + b);
    } else {
      R central = R(0.5) * (a + b);
      R limit = R(2) * b;
      if (absv(a) < absv(b)) limit = R(2) * a;
      R ac = absv(central), al = absv(limit);
      R size = al < ac ? al : ac;
      if (central > R(0)) slope = size;
      else if (central < R(0)) slope = -size;
    }
  }
  return slope;
}

// Fast magnetosonic speed along the normal, as in the sweep (mhd_sweep_pyccel.py:51-61).
template <typename R>
__device__ __forceinline__ R fast_speed(R rho, R p, R bn, R bt1, R bt2, R gamma) {
  R a2 = Arith<R>::div(gamma * p, rho);
  R b2 = Arith<R>::div(bn * bn + bt1 * bt1 + bt2 * bt2, rho);
  R disc = (a2 + b2) * (a2 + b2) - Arith<R>::div(R(4) * a2 * bn * bn, rho);
  R root = R(0);
  if (disc > R(0)) root = Arith<R>::sqrt(disc);
  return Arith<R>::sqrt(R(0.5) * (a2 + b2 + root));
}

// The CFL kernel's fast speed groups (bn * bn) (mhd_rate_pyccel.py:17-27).
__device__ __forceinline__ double fast_speed_rate(double rho, double p, double bn, double bt1,
                                                  double bt2, double gamma) {
  double a2 = gamma * p / rho;
  double b2 = (bn * bn + bt1 * bt1 + bt2 * bt2) / rho;
  double disc = (a2 + b2) * (a2 + b2) - 4.0 * a2 * (bn * bn) / rho;
  double root = 0.0;
  if (disc > 0.0) root = __builtin_sqrt(disc);
  return __builtin_sqrt(0.5 * (a2 + b2 + root));
}

// Conserved state u[7], flux f[7] and total pressure of a normal-frame state
// w = (rho, vn, vt1, vt2, p, bt1, bt2) (mhd_sweep_pyccel.py:64-98).
template <typename R>
__device__ __forceinline__ R side_state(const R* w, R bn, R gamma, R* u, R* f) {
  R rho = w[0], vn = w[1], vt1 = w[2], vt2 = w[3], p = w[4], bt1 = w[5], bt2 = w[6];
  R magnetic = R(0.5) * (bn * bn + bt1 * bt1 + bt2 * bt2);
  R energy = Arith<R>::div(p, gamma - R(1)) + R(0.5) * rho * (vn * vn + vt1 * vt1 + vt2 * vt2) +
             magnetic;
  R pt = p + magnetic;
  R vb = vn * bn + vt1 * bt1 + vt2 * bt2;
  u[0] = rho;
  u[1] = rho * vn;
  u[2] = rho * vt1;
  u[3] = rho * vt2;
  u[4] = energy;
  u[5] = bt1;
  u[6] = bt2;
  f[0] = rho * vn;
  f[1] = rho * vn * vn + pt - bn * bn;
  f[2] = rho * vn * vt1 - bn * bt1;
  f[3] = rho * vn * vt2 - bn * bt2;
  f[4] = (energy + pt) * vn - bn * vb;
  f[5] = bt1 * vn - bn * vt1;
  f[6] = bt2 * vn - bn * vt2;
  return pt;
}

// HLLD star state us[7] and its tangential velocity/field vt[4] beside the contact
// (mhd_sweep_pyccel.py:101-145).
template <typename R>
__device__ __forceinline__ void star_state(const R* w, R bn, R s, R d, R pt, R sm, R pts,
                                           R energy0, R* us, R* vt) {
  R vn = w[1];
  R rho_s = Arith<R>::div(d, s - sm);
  R den = d * (s - sm) - bn * bn;
  R vt1 = w[2], vt2 = w[3], bt1 = w[5], bt2 = w[6];
  R thr = R(DEGENERATE) * (pt + bn * bn);
  if (absv(den) >= thr) {
    vt1 = w[2] - Arith<R>::div(bn * w[5] * (sm - vn), den);
    vt2 = w[3] - Arith<R>::div(bn * w[6] * (sm - vn), den);
    bt1 = Arith<R>::div(w[5] * (d * (s - vn) - bn * bn), den);
    bt2 = Arith<R>::div(w[6] * (d * (s - vn) - bn * bn), den);
  }
  R vb = vn * bn + w[2] * w[5] + w[3] * w[6];
  R vb_s = sm * bn + vt1 * bt1 + vt2 * bt2;
  R energy = Arith<R>::div((s - vn) * energy0 - pt * vn + pts * sm + bn * (vb - vb_s), s - sm);
  us[0] = rho_s;
  us[1] = rho_s * sm;
  us[2] = rho_s * vt1;
  us[3] = rho_s * vt2;
  us[4] = energy;
  us[5] = bt1;
  us[6] = bt2;
  vt[0] = vt1;
  vt[1] = vt2;
  vt[2] = bt1;
  vt[3] = bt2;
}

// The HLLD flux (Miyoshi and Kusano 2005) of one face between normal-frame states
// L and Rr with normal field bn (mhd_sweep_pyccel.py:148-251).
template <typename R>
__device__ __forceinline__ void hlld(const R* L, const R* Rr, R bn, R gamma, R* flux) {
  R ul[7], fl[7], ur[7], fr[7];
  R ptl = side_state(L, bn, gamma, ul, fl);
  R ptr = side_state(Rr, bn, gamma, ur, fr);
  R rl = L[0], vl = L[1], rr = Rr[0], vr = Rr[1];
  R cfl = fast_speed(rl, L[4], bn, L[5], L[6], gamma);
  R cfr = fast_speed(rr, Rr[4], bn, Rr[5], Rr[6], gamma);
  R cf = cfl > cfr ? cfl : cfr;
  R sl = (vl < vr ? vl : vr) - cf;
  R sr = (vl > vr ? vl : vr) + cf;
  if (sl >= R(0)) {
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fl[c];
    return;
  }
  if (sr < R(0)) {
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fr[c];
    return;
  }
  R dl = (sl - vl) * rl;
  R dr = (sr - vr) * rr;
  R sm = Arith<R>::div(dr * vr - dl * vl - ptr + ptl, dr - dl);
  R pts = Arith<R>::div(dr * ptl - dl * ptr + dl * dr * (vr - vl), dr - dl);
  R usl[7], usr[7], wl[4], wr[4];
  star_state(L, bn, sl, dl, ptl, sm, pts, ul[4], usl, wl);
  star_state(Rr, bn, sr, dr, ptr, sm, pts, ur[4], usr, wr);
  R sql = Arith<R>::sqrt(usl[0]);
  R sqr = Arith<R>::sqrt(usr[0]);
  R sign = R(0);
  if (bn > R(0)) sign = R(1);
  else if (bn < R(0)) sign = R(-1);
  R sum = sql + sqr;
  R vt2_1 = Arith<R>::div(sql * wl[0] + sqr * wr[0] + (wr[2] - wl[2]) * sign, sum);
  R vt2_2 = Arith<R>::div(sql * wl[1] + sqr * wr[1] + (wr[3] - wl[3]) * sign, sum);
  R bt2_1 = Arith<R>::div(sql * wr[2] + sqr * wl[2] + sql * sqr * (wr[0] - wl[0]) * sign, sum);
  R bt2_2 = Arith<R>::div(sql * wr[3] + sqr * wl[3] + sql * sqr * (wr[1] - wl[1]) * sign, sum);
  R vb2 = sm * bn + vt2_1 * bt2_1 + vt2_2 * bt2_2;
  R ssl = sm - Arith<R>::div(absv(bn), sql);
  R ssr = sm + Arith<R>::div(absv(bn), sqr);
  if (ssl >= R(0)) {
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fl[c] + sl * (usl[c] - ul[c]);
  } else if (sm >= R(0)) {
    R vb_s = sm * bn + wl[0] * wl[2] + wl[1] * wl[3];
    R rho_s = usl[0];
    R udd[7] = {rho_s,         rho_s * sm,                               rho_s * vt2_1,
                rho_s * vt2_2, usl[4] + R(-1) * sql * (vb_s - vb2) * sign, bt2_1,
                bt2_2};
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fl[c] + sl * (usl[c] - ul[c]) + ssl * (udd[c] - usl[c]);
  } else if (ssr >= R(0)) {
    R vb_s = sm * bn + wr[0] * wr[2] + wr[1] * wr[3];
    R rho_s = usr[0];
    R udd[7] = {rho_s,         rho_s * sm,                              rho_s * vt2_1,
                rho_s * vt2_2, usr[4] + R(1) * sqr * (vb_s - vb2) * sign, bt2_1,
                bt2_2};
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fr[c] + sr * (usr[c] - ur[c]) + ssr * (udd[c] - usr[c]);
  } else {
#pragma unroll
    for (int c = 0; c < 7; ++c) flux[c] = fr[c] + sr * (usr[c] - ur[c]);
  }
}

// Left and right reconstructed states of the face between cells 1 and 2 of the
// four cells q[c][0..3] along the normal (mhd_sweep_pyccel.py:290-307): limited
// slopes, then first order where a density or pressure would not be positive.
template <typename R>
__device__ __forceinline__ void reconstruct(const R (*q)[4], int limiter, R* L, R* Rr) {
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    R d0 = q[c][1] - q[c][0];
    R d1 = q[c][2] - q[c][1];
    R d2 = q[c][3] - q[c][2];
    R s1 = limited_slope(d0, d1, limiter);
    R s2 = limited_slope(d1, d2, limiter);
    L[c] = q[c][1] + R(0.5) * s1;
    Rr[c] = q[c][2] - R(0.5) * s2;
  }
  if (L[0] <= R(0) || L[4] <= R(0) || Rr[0] <= R(0) || Rr[4] <= R(0)) {
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      L[c] = q[c][1];
      Rr[c] = q[c][2];
    }
  }
}

// ct_contact_pyccel.py:16-23.
__device__ __forceinline__ double upwind(double mass, double positive, double negative) {
  double value = 0.5 * (positive + negative);
  if (mass > 0.0) value = positive;
  else if (mass < 0.0) value = negative;
  return value;
}

// E at one node/edge from the four face fluxes around it and the four cell fields
// (ct_contact_pyccel.py:35-61).  fxb_lo/hi: a-face flux of B_b at b-cells j-1, j;
// gyb_l/r: b-face flux of B_a at a-cells i-1, i; fxr_*, gyr_*: the mass fluxes;
// e00 = cell (i-1, j-1), e10 = (i, j-1), e01 = (i-1, j), e11 = (i, j).
__device__ __forceinline__ double ct_node(double fxb_lo, double fxb_hi, double gyb_l, double gyb_r,
                                          double fxr_lo, double fxr_hi, double gyr_l, double gyr_r,
                                          double e00, double e10, double e01, double e11) {
  double dl2 = upwind(fxr_lo, gyb_l - e00, gyb_r - e10);
  double dr2 = upwind(fxr_hi, gyb_l - e01, gyb_r - e11);
  double dl1 = upwind(gyr_l, -fxb_lo - e00, -fxb_hi - e01);
  double dr1 = upwind(gyr_r, -fxb_lo - e10, -fxb_hi - e11);
  return 0.25 * (gyb_r + gyb_l - fxb_hi - fxb_lo + dl1 + dr1 + dl2 + dr2);
}

__device__ __forceinline__ bool finite(double x) { return __builtin_fabs(x) <= 1.7976931348623157e308; }

}  // namespace wb
