// pair_kernels.hpp -- tiled all-pairs direct-sum kernels for gfx950.
//
// One template covers every O(T*S) sum of the path.  A "pair policy" P
// supplies the per-pair arithmetic; the template supplies the CDNA4 mapping:
//
//  * a wavefront owns 64*R consecutive targets: lane l holds targets
//    t0 + r*64 + l (r < R) in registers, so each source value loaded is used
//    by 64*R pairs;
//  * sources are packed AoS records (8 words) read with wave-uniform
//    addresses, which the compiler turns into scalar (SMEM) loads: the source
//    operands of the pair arithmetic come straight from SGPRs and cost no
//    LDS or VGPR traffic;
//  * sources are walked in tiles of KT; inside a tile each target
//    accumulates plainly in FP32, and the tile partial is folded into a
//    (sum, compensation) pair with an error-free TwoSum (Knuth), so the
//    compensated accumulation costs O(S/KT) per target;
//  * the self pair j == i is excluded by index, exactly, but only tiles whose
//    source-index range overlaps the wave's target range take the checked
//    path (a wave-uniform branch);
//  * a grid of (target tiles) x (source splits) fills the 256 CUs also when T
//    is small; every split writes its FP64 partial per target and a
//    finalize kernel reduces the splits in a fixed order (deterministic).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpm {

// Source records: P::NW words per source (FP32) or doubles (FP64); 8 unless
// a policy needs more.
//   sphere BS   : x y z e w2 idx . .     e = R^2-|x|^2, w2 = -zeta A/(2 pi R)
//   plane  BS   : x y w idx . . . .      w = omega A / (2 pi)
//   PSE         : x y fA A idx . . .     coordinates scaled (see PsePolicy)
//   sphere green: x y z e gr ga idx .    g = -zeta A ln2/(4 pi) (FP32) or -zeta A/(4 pi) (FP64)
//   plane  green: x y g idx . . . .      g = omega A ln2/(4 pi) (FP32) or omega A/(4 pi) (FP64)
//   PSE derivative, interpolation, SWE: see PseDeriv, PseInterp, SwePlane
// Target records: 4 words per target: sphere {x y z e}, plane {x y . .},
// PSE {x y f .}.
constexpr int SREC = 8;
constexpr int TREC = 4;
constexpr int SREC_MAX = 12;

template <typename T>
struct Vec4 {
  T a, b, c, d;
};

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// FP64 reciprocal: the hardware estimate (v_rcp_f64) refined by two
// Newton-Raphson steps r <- r + r (1 - x r), each of which squares the
// relative error, so the result is within an ulp or two of the correctly
// rounded 1 / x for normal x.  Five instructions against about ten for the
// IEEE division sequence; x = 0 gives NaN instead of inf, which the pair
// sums multiply by a zero separation either way.
__device__ __forceinline__ double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// FP64 natural logarithm for the Green's-function sums.  x = m 2^e with m in
// [sqrt(1/2), sqrt(2)); log m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^21/21),
// s = (m - 1) / (m + 1), |s| <= 0.1716, so the first dropped term is below
// 1e-18 relative; e ln 2 uses ln 2 split into a 32-bit head (exact products
// for |e| < 2^21) and its tail.  Within 2 ulp of the correctly rounded log
// for positive normal x (the same steps against the C library's log on
// 1.2e6 arguments from 1e-13 to 7.4); about 30 instructions against ~80 for
// the library's double-double evaluation.  Zero gives a finite value, not
// -inf: the pair sums only meet it in the self pair, which they discard.
__device__ __forceinline__ double log_fp64(double x) {
  constexpr double kLn2Hi = 0.6931471803691238;      // ln 2 to 32 significant bits
  constexpr double kLn2Lo = 1.9082149292705877e-10;  // ln 2 - kLn2Hi
  double m = __builtin_amdgcn_frexp_mant(x);         // [0.5, 1)
  int e = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    --e;
  }
  const double f = m - 1.0;  // exact (Sterbenz)
  const double s = f * rcp_refined(m + 1.0);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double s2 = 2.0 * s;
  const double lm = fma(s2 * z, p, s2);
  const double de = (double)e;
  return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }

// ---------------------------------------------------------------- policies
// A policy splits one pair into three phases so that the kernel can issue the
// phase for all R targets of a lane back to back (independent instruction
// streams hide the latency of the transcendental unit):
//   arg   : differences and the argument of the transcendental (kept in W)
//   trans : the one transcendental per pair (rcp, exp2 or log2)
//   acc   : weight and accumulate into NOUT tile partials
// W holds a policy's per-pair temporaries.

template <typename T>
struct W4 {
  T a, b, c, z;
};

template <typename T, bool USE_E>
struct SphereBS {
  static constexpr int NOUT = 3;
  static constexpr int IDX = 5;
  static constexpr int NW = SREC;
  // acc += s * (x_i - y_j),  s = w2 / (|x_i - y_j|^2 + e_i + e_j); with
  // USE_E == false the points are on the sphere and e_i = e_j = 0.
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    w.a = t.a - s[0];
    w.b = t.b - s[1];
    w.c = t.c - s[2];
    T r2;
    if constexpr (USE_E)
      r2 = fma(w.a, w.a, t.d + s[3]);
    else
      r2 = w.a * w.a;
    r2 = fma(w.b, w.b, r2);
    w.z = fma(w.c, w.c, r2);
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_rcp(w.z);
    else
      return rcp_refined(w.z);
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T f, T a[NOUT]) {
    T q = s[4] * f;
    a[0] = fma(q, w.a, a[0]);
    a[1] = fma(q, w.b, a[1]);
    a[2] = fma(q, w.c, a[2]);
  }
};

template <typename T>
struct PlaneBS {
  static constexpr int NOUT = 2;
  static constexpr int IDX = 3;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    w.a = t.a - s[0];
    w.b = t.b - s[1];
    w.z = fma(w.a, w.a, w.b * w.b);
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_rcp(w.z);
    else
      return rcp_refined(w.z);
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T f, T a[NOUT]) {
    T q = s[2] * f;
    a[0] = fma(q, w.b, a[0]);
    a[1] = fma(q, w.a, a[1]);
  }
};

// PSE Laplacian.  FP32: coordinates are pre-scaled by sqrt(log2 e)/eps so that
// q' = |dx|^2 = log2(e) rho^2, exp(-rho^2) = exp2(-q'), and the polynomial is
// re-expressed in q'.  FP64: coordinates scaled by 1/eps, q = rho^2.
template <typename T, int ORDER>
struct PsePolicy {
  static constexpr int NOUT = 1;
  static constexpr int IDX = 4;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    T dx = t.a - s[0], dy = t.b - s[1];
    w.z = fma(dx, dx, dy * dy);
    w.a = fma(-t.c, s[3], s[2]);  // f_j A_j - f_i A_j
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_exp2(-w.z);
    else
      return exp(-w.z);
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T e, T a[NOUT]) {
    if constexpr (ORDER == 8) {
      // 40 - 40 rho^2 + 10 rho^4 - (2/3) rho^6 in the (scaled) variable q
      constexpr double L = sizeof(T) == 4 ? 1.4426950408889634 : 1.0;
      const T c0 = (T)40.0, c1 = (T)(-40.0 / L), c2 = (T)(10.0 / (L * L)),
              c3 = (T)(-(2.0 / 3.0) / (L * L * L));
      const T q = w.z;
      T p = fma(fma(fma(c3, q, c2), q, c1), q, c0);
      a[0] = fma(p * e, w.a, a[0]);
    } else {
      a[0] = fma(e, w.a, a[0]);
    }
  }
};

// Sphere PSE Laplacian (PSEDirectSum.f90:502-532): K8 of the great-circle
// distance R theta between the directions of x_i and x_j (SphereGeometry.f90:
// atan2(|x_i x x_j|, x_i . x_j) R).  Records hold the unit directions scaled by
// k R (k = sqrt(log2 e) / eps in FP32, 1 / eps in FP64), so that with
// u = |x'_i - x'_j|^2 / (k R)^2 = (c / R)^2 (c the chord) the scaled squared
// distance is q = (k R)^2 theta^2, theta = 2 asin(sqrt(u) / 2).
// SERIES: theta^2 = u g(u), g = sum_n a_n u^(n-1), a_n = 2 / (n^2 binom(2n, n))
// (the Taylor series of 4 asin^2(s / 2)), through n = 7 in FP32 (the caller
// uses it when 7 eps <= 0.56 R, where the first dropped term is below 4e-9
// relative) and n = 10 in FP64 (7 eps <= 0.35 R).  Pairs farther than 7 eps
// get a slightly low theta^2, but their kernel factor exp(-q) is below e^-49.
// Source record: {x', y', z', f A, A, idx, 1/(kR)^2, (kR)^2}; target {x', y', z', f}.
// Scaled squared great-circle distance q = (k R theta)^2 from the squared
// chord c2 = |x'_i - x'_j|^2 of directions scaled by k R; inv = 1/(k R)^2,
// kr2 = (k R)^2.
template <typename T, bool SERIES>
__device__ __forceinline__ T sphere_q(T c2, T inv, T kr2) {
  const T u = c2 * inv;
  if constexpr (SERIES && sizeof(T) == 8) {
    // FP64: through n = 10; the caller uses it when 7 eps <= 0.35 R, where the
    // first dropped term is below 2e-17 relative
    double g = fma(u, 1.0 / 9237800.0, 1.0 / 1969110.0);
    g = fma(u, g, 1.0 / 411840.0);
    g = fma(u, g, 1.0 / 84084.0);
    g = fma(u, g, 1.0 / 16632.0);
    g = fma(u, g, 1.0 / 3150.0);
    g = fma(u, g, 1.0 / 560.0);
    g = fma(u, g, 1.0 / 90.0);
    g = fma(u, g, 1.0 / 12.0);
    g = fma(u, g, 1.0);
    return c2 * g;
  } else if constexpr (SERIES) {
    T g = fma(u, (T)(1.0 / 84084.0), (T)(1.0 / 16632.0));
    g = fma(u, g, (T)(1.0 / 3150.0));
    g = fma(u, g, (T)(1.0 / 560.0));
    g = fma(u, g, (T)(1.0 / 90.0));
    g = fma(u, g, (T)(1.0 / 12.0));
    g = fma(u, g, (T)1.0);
    return c2 * g;
  } else {
    const T th = (T)2.0 * asin(fmin((T)0.5 * sqrt(u), (T)1.0));
    return th * th * kr2;
  }
}

template <typename T, bool SERIES>
struct PseSphere {
  static constexpr int NOUT = 1;
  static constexpr int IDX = 5;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    const T dx = t.a - s[0], dy = t.b - s[1], dz = t.c - s[2];
    const T c2 = fma(dx, dx, fma(dy, dy, dz * dz));
    w.z = sphere_q<T, SERIES>(c2, s[6], s[7]);
    w.a = fma(-t.d, s[4], s[3]);  // f_j A_j - f_i A_j
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) { return PsePolicy<T, 8>::trans(w); }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T e, T a[NOUT]) {
    PsePolicy<T, 8>::acc(s, w, e, a);
  }
};

template <typename T, bool USE_E>
struct SphereGreen {
  static constexpr int NOUT = 2;
  static constexpr int IDX = 6;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    T dx = t.a - s[0], dy = t.b - s[1], dz = t.c - s[2];
    T r2;
    if constexpr (USE_E)
      r2 = fma(dx, dx, t.d + s[3]);
    else
      r2 = dx * dx;
    r2 = fma(dy, dy, r2);
    w.z = fma(dz, dz, r2);
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_log2(w.z);  // log2(2 (R^2 - x.y))
    else
      return log_fp64((T)0.5 * w.z);  // ln(R^2 - x.y)
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T lg, T a[NOUT]) {
    a[0] = fma(lg, s[4], a[0]);
    a[1] = fma(lg, s[5], a[1]);
  }
};

template <typename T>
struct PlaneGreen {
  static constexpr int NOUT = 1;
  static constexpr int IDX = 3;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    T dx = t.a - s[0], dy = t.b - s[1];
    w.z = fma(dx, dx, dy * dy);
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_log2(w.z);
    else
      return log_fp64(w.z);
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T lg, T a[NOUT]) {
    a[0] = fma(lg, s[2], a[0]);
  }
};

// Polynomial parts of the 8th-order PSE kernels in the scaled variable q
// (q = rho^2 in FP64, log2(e) rho^2 in FP32, where exp(-rho^2) = exp2(-q)):
//   derivative kernel (PSEDirectSum.f90:629-634): -20 + 20 r^2 - 5 r^4 + r^6/3
//   delta kernel      (PSEDirectSum.f90:611-616):   4 -  6 r^2 + 2 r^4 - r^6/6
//   Laplacian kernel  (PSEDirectSum.f90:622-627):  40 - 40 r^2 + 10 r^4 - 2 r^6/3
template <typename T>
__device__ __forceinline__ T cubic_in_q(T q, double c0, double c1, double c2, double c3) {
  constexpr double L = sizeof(T) == 4 ? 1.4426950408889634 : 1.0;
  return fma(fma(fma((T)(c3 / (L * L * L)), q, (T)(c2 / (L * L))), q, (T)(c1 / L)), q, (T)c0);
}

template <typename T>
__device__ __forceinline__ T exp_neg_q(T q) {
  if constexpr (sizeof(T) == 4)
    return fast_exp2(-q);
  else
    return exp(-q);
}

// PSE first-derivative sums (PSEDirectSum.f90:180-418, 537-579).  For M source
// fields c_m and the D coordinates of the separation X = x_i - x_j:
//   a[k]             = sum_j X_k K1'(q) A_j           (k < D)
//   a[D + m D + k]   = sum_j X_k K1'(q) c_m,j A_j
// (K1' the unnormalised kernel, coordinates scaled by sigma).  Every
// reference operator is an FP64 epilogue of these sums, since
// sum (c_j + c_i) X K1 A = a[D + m D + k] + c_i a[k].
// Source record: plane {x', y', A, c_0 .. c_{M-1}, idx};
// sphere {x', y', z', A, c_0 .., idx, 1/(kR)^2, (kR)^2} with x' the unit
// direction times k R (the separation is taken as R (d_i - d_j)).
template <typename T, bool SPH, int M, bool SERIES>
struct PseDeriv {
  static constexpr int D = SPH ? 3 : 2;
  static constexpr int NOUT = D * (M + 1);
  static constexpr int IA = D;
  static constexpr int IC = D + 1;
  static constexpr int IDX = D + 1 + M;
  static constexpr int IINV = IDX + 1;
  static constexpr int IKR2 = IDX + 2;
  static constexpr int NW = (SPH ? IKR2 + 1 : IDX + 1) <= SREC ? SREC : SREC_MAX;
  static_assert((SPH ? IKR2 + 1 : IDX + 1) <= SREC_MAX, "record too long");
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    w.a = t.a - s[0];
    w.b = t.b - s[1];
    if constexpr (SPH) {
      w.c = t.c - s[2];
      const T c2 = fma(w.a, w.a, fma(w.b, w.b, w.c * w.c));
      w.z = sphere_q<T, SERIES>(c2, s[IINV], s[IKR2]);
    } else {
      w.z = fma(w.a, w.a, w.b * w.b);
    }
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) { return exp_neg_q(w.z); }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T e, T a[NOUT]) {
    const T g = cubic_in_q(w.z, -20.0, 20.0, -5.0, 1.0 / 3.0) * e * s[IA];
    T X[D];
    X[0] = w.a * g;
    X[1] = w.b * g;
    if constexpr (SPH) X[2] = w.c * g;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      a[k] += X[k];
#pragma unroll
      for (int m = 0; m < M; ++m) a[D + m * D + k] = fma(s[IC + m], X[k], a[D + m * D + k]);
    }
  }
};

// PSE interpolation (PSEDirectSum.f90:121-168): sum_j f_j A_j Kd'(q).
// Source record: plane {x', y', f A, idx}; sphere {x', y', z', f A, idx,
// 1/(kR)^2, (kR)^2}.
template <typename T, bool SPH, bool SERIES>
struct PseInterp {
  static constexpr int D = SPH ? 3 : 2;
  static constexpr int NOUT = 1;
  static constexpr int IFA = D;
  static constexpr int IDX = D + 1;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    const T dx = t.a - s[0], dy = t.b - s[1];
    if constexpr (SPH) {
      const T dz = t.c - s[2];
      w.z = sphere_q<T, SERIES>(fma(dx, dx, fma(dy, dy, dz * dz)), s[IDX + 1], s[IDX + 2]);
    } else {
      w.z = fma(dx, dx, dy * dy);
    }
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) { return exp_neg_q(w.z); }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T e, T a[NOUT]) {
    a[0] = fma(cubic_in_q(w.z, 4.0, -6.0, 2.0, -1.0 / 6.0) * e, s[IFA], a[0]);
  }
};

// Planar shallow-water right-hand side (SWEPlaneSolver.f90:457-538), one pass
// over the pairs for seven sums: velocity (u, v), velocity gradient
// (ux, uy, vx, vy) of the point vortices and sources, and the K8 Laplacian of
// the surface height s = h + bottom.
// Source record {x, y, a, b, s A, A, idx, c}: a = zeta A / (2 pi),
// b = delta A / (2 pi), c = log2(e) / eps^2 (FP32) or 1 / eps^2 (FP64).
// Target record {x, y, s, .}.  With f = 1 / r^2:
//   u  += -Y a f + X b f             v  += X a f + Y b f
//   ux += b f - 2 X (X b - Y a) f^2  uy += -a f - 2 Y (X b - Y a) f^2
//   vx += a f - 2 X (Y b + X a) f^2  vy += b f - 2 Y (Y b + X a) f^2
//   lap += K8'(r^2 c) (s_j - s_i) A_j
template <typename T>
struct SwePlane {
  static constexpr int NOUT = 7;
  static constexpr int IDX = 6;
  static constexpr int NW = SREC;
  __device__ __forceinline__ static void arg(const Vec4<T>& t, const T* s, W4<T>& w) {
    w.a = t.a - s[0];
    w.b = t.b - s[1];
    w.z = fma(w.a, w.a, w.b * w.b);
    w.c = fma(-t.c, s[5], s[4]);  // (s_j - s_i) A_j
  }
  __device__ __forceinline__ static T trans(const W4<T>& w) {
    if constexpr (sizeof(T) == 4)
      return fast_rcp(w.z);
    else
      return rcp_refined(w.z);
  }
  __device__ __forceinline__ static void acc(const T* s, const W4<T>& w, T f, T a[NOUT]) {
    const T X = w.a, Y = w.b;
    const T rot = s[2] * f, pot = s[3] * f;
    a[0] = fma(X, pot, fma(-Y, rot, a[0]));
    a[1] = fma(Y, pot, fma(X, rot, a[1]));
    const T hh = (f + f) * f;
    const T p1 = fma(X, s[3], -Y * s[2]) * hh;
    const T p2 = fma(Y, s[3], X * s[2]) * hh;
    a[2] = fma(-X, p1, a[2] + pot);
    a[3] = fma(-Y, p1, a[3] - rot);
    a[4] = fma(-X, p2, a[4] + rot);
    a[5] = fma(-Y, p2, a[5] + pot);
    const T q = w.z * s[7];
    a[6] = fma(cubic_in_q(q, 40.0, -40.0, 10.0, -2.0 / 3.0) * exp_neg_q(q), w.c, a[6]);
  }
};

// ----------------------------------------------------------- the kernel
struct PairArgs {
  const void* tgt;   // nt target records (TREC words each)
  const void* src;   // ns source records (P::NW words each)
  double* partial;   // [nsplit][NOUT][nt]
  int nt;            // number of targets
  int tgt_idx0;      // particle index of target 0 (self-pair exclusion)
  int ns;            // number of sources
  int chunk;         // sources per split (multiple of kt)
  int nsplit;
  int kt;            // sources per compensated tile
  int n_wtiles;      // wave target tiles = ceil(nt / (64 R))
  // when set, the number of sources is read here and ns is only the bound
  // that sized chunk * nsplit (splits past the count sum nothing)
  const int* ns_dev = nullptr;
};

template <typename T>
__device__ __forceinline__ int rec_index(const T* s, int slot) {
  if constexpr (sizeof(T) == 4)
    return __float_as_int(s[slot]);
  else
    return (int)s[slot];
}

// Error-free (sum, compensation) update with a tile partial p.
template <typename T>
__device__ __forceinline__ void two_sum_acc(T& S, T& C, T p) {
  T t = S + p;
  T bp = t - S;
  T e = (S - (t - bp)) + (p - bp);
  C += e;
  S = t;
}

// One wavefront's work: targets [t_begin, t_begin + 64 R) of the record array
// tg[0, nt) (particle index tgt_idx0 + local) against sources [j_begin, j_end)
// of src; FP64 results go to out[k * out_stride + i] for valid targets i.
template <class P, typename T, int R, bool COMPENSATED>
__device__ __forceinline__ void pair_tile(const Vec4<T>* __restrict__ tg, int nt, int t_begin,
                                          int tgt_idx0, const T* __restrict__ src, int j_begin,
                                          int j_end, int kt, double* __restrict__ out,
                                          int out_stride) {
  constexpr int NO = P::NOUT;
  const int lane = threadIdx.x & 63;
  Vec4<T> t[R];
  int tid[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = t_begin + r * 64 + lane;
    int ic = i < nt ? i : nt - 1;
    t[r] = tg[ic];
    tid[r] = tgt_idx0 + i;
  }
  const int wt_lo = tgt_idx0 + t_begin;
  const int wt_hi = wt_lo + 64 * R - 1;

  T S[R][NO], C[R][NO];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      S[r][k] = (T)0;
      C[r][k] = (T)0;
    }

  for (int jt = j_begin; jt < j_end; jt += kt) {
    const int jt_end = min(j_end, jt + kt);
    const int idx_lo = rec_index(src + (size_t)jt * P::NW, P::IDX);
    const int idx_hi = rec_index(src + (size_t)(jt_end - 1) * P::NW, P::IDX);
    const bool checked = (idx_lo <= wt_hi) && (idx_hi >= wt_lo);
    T p[R][NO];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < NO; ++k) p[r][k] = (T)0;

    if (!checked) {
      // the next source's record is loaded before this one's arithmetic, so
      // the scalar loads' latency overlaps the pair work (the last iteration
      // reloads its own record).  Measured on MI355X against loading each
      // record at its use: PSE-8 at 512^2 15.96 -> 14.45 ms, plane gradient
      // 20.68 -> 18.92 ms, sphere Biot-Savart at icosahedral L7 38.17 -> 37.95 ms.
      T sv[P::NW];
#pragma unroll
      for (int q = 0; q < P::NW; ++q) sv[q] = src[(size_t)jt * P::NW + q];
#pragma unroll 2
      for (int j = jt; j < jt_end; ++j) {
        const int jn = j + 1 < jt_end ? j + 1 : j;
        T sn[P::NW];
#pragma unroll
        for (int q = 0; q < P::NW; ++q) sn[q] = src[(size_t)jn * P::NW + q];
        W4<T> w[R];
        T f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) P::arg(t[r], sv, w[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) f[r] = P::trans(w[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) P::acc(sv, w[r], f[r], p[r]);
#pragma unroll
        for (int q = 0; q < P::NW; ++q) sv[q] = sn[q];
      }
    } else {
      for (int j = jt; j < jt_end; ++j) {
        const T* s = src + (size_t)j * P::NW;
        T sv[P::NW];
#pragma unroll
        for (int q = 0; q < P::NW; ++q) sv[q] = s[q];
        const int sidx = rec_index(sv, P::IDX);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          W4<T> w;
          P::arg(t[r], sv, w);
          T f = P::trans(w);
          T tmp[NO];
#pragma unroll
          for (int k = 0; k < NO; ++k) tmp[k] = (T)0;
          P::acc(sv, w, f, tmp);
          const bool self = (sidx == tid[r]);
#pragma unroll
          for (int k = 0; k < NO; ++k) p[r][k] += self ? (T)0 : tmp[k];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < NO; ++k) {
        if constexpr (COMPENSATED)
          two_sum_acc(S[r][k], C[r][k], p[r][k]);
        else
          S[r][k] += p[r][k];
      }
  }

#pragma unroll
  for (int r = 0; r < R; ++r) {
    int i = t_begin + r * 64 + lane;
    if (i < nt) {
#pragma unroll
      for (int k = 0; k < NO; ++k) out[(size_t)k * out_stride + i] = (double)S[r][k] + (double)C[r][k];
    }
  }
}

template <class P, typename T, int R, bool COMPENSATED>
__global__ __launch_bounds__(256) void pair_sum_kernel(PairArgs a) {
  const int wid = threadIdx.x >> 6;
  const int nblk_t = (a.n_wtiles + 3) >> 2;
  const int split = blockIdx.x / nblk_t;
  const int wtile = (blockIdx.x - split * nblk_t) * 4 + wid;
  if (wtile >= a.n_wtiles) return;  // wave-uniform
  const int ns = a.ns_dev ? *a.ns_dev : a.ns;
  const int j_begin = split * a.chunk;
  const int j_end = min(ns, j_begin + a.chunk);
  pair_tile<P, T, R, COMPENSATED>(reinterpret_cast<const Vec4<T>*>(a.tgt), a.nt, wtile * 64 * R,
                                  a.tgt_idx0, reinterpret_cast<const T*>(a.src), j_begin, j_end, a.kt,
                                  a.partial + (size_t)split * P::NOUT * a.nt, a.nt);
}

// ------------------------------------------------------- batched problems
// Many independent problems in one persistent launch.  Problem k owns target
// records [t_off, t_off + nt) (particle indices t_off + local), source records
// [s_off, s_off + ns) and FP64 partials at part_off ([nsplit][NOUT][nt]).
// A work item is (problem, wave target tile, source split); items are taken
// from a device counter by whole wavefronts until the list is exhausted, so
// every wave reaches the exit.
struct BatchProblem {
  int t_off, nt, s_off, ns;
  int chunk, nsplit;
  long long part_off;
};

struct BatchItem {
  int prob, wtile, split, pad;
};

struct BatchArgs {
  const void* tgt;
  const void* src;
  double* partial;
  const BatchProblem* probs;
  const BatchItem* items;
  int n_items;
  int kt;
  unsigned int* counter;
};

// The buffers are separate __restrict__ kernel parameters rather than fields
// of BatchArgs: the loop stores partials and then reads sources of the next
// item, and only kernel-level noalias lets the compiler keep the source reads
// on the scalar unit (wave-uniform s_load, SGPR operands) across those stores.
// Through a struct, it issued per-lane vector loads of the same addresses.
template <class P, typename T, int R, bool COMPENSATED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void pair_batch_kernel(
    const void* __restrict__ tgt, const void* __restrict__ srcv, double* __restrict__ partial,
    const BatchProblem* __restrict__ probs, const BatchItem* __restrict__ items, int n_items, int kt,
    unsigned int* __restrict__ counter) {
  const int lane = threadIdx.x & 63;
  const Vec4<T>* tg = reinterpret_cast<const Vec4<T>*>(tgt);
  const T* src = reinterpret_cast<const T*>(srcv);
  for (;;) {
    int item = 0;
    if (lane == 0) item = (int)atomicAdd(counter, 1u);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
    if (item >= n_items) break;
    const BatchItem it = items[item];
    const BatchProblem pr = probs[it.prob];
    const int j_begin = pr.s_off + it.split * pr.chunk;
    const int j_end = min(pr.s_off + pr.ns, j_begin + pr.chunk);
    pair_tile<P, T, R, COMPENSATED>(tg + pr.t_off, pr.nt, it.wtile * 64 * R, pr.t_off, src, j_begin, j_end,
                                    kt, partial + pr.part_off + (size_t)it.split * P::NOUT * pr.nt, pr.nt);
  }
}

}  // namespace lpm
