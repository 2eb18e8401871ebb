// lpm_pse_ops.hip -- PSE derivative operators, PSE interpolation and the
// planar shallow-water right-hand side on the shared pair engine.
//
// The derivative kinds compute, in one pass over the pairs, the sums
//   M_k = sum_j X_k K1 A_j  and  S_mk = sum_j X_k K1 c_m,j A_j
// for up to three fields c_m (X = x_i - x_j, k < D coordinates); the FP64
// epilogue forms D[c_m]_k = (S_mk + c_m,i M_k) / (sigma pi eps^4), which is
// the reference's sum of (c_j + c_i) X K1 A_j / eps^4, and from it the
// gradient, second partials, double dot or divergence
// (PSEDirectSum.f90:180-418, 537-579).
#include <math.h>

#include "lpm_internal.hpp"
#include "pair_kernels.hpp"

namespace lpm {

namespace {

template <typename T>
__device__ __forceinline__ T idx_word(int p) {
  if constexpr (sizeof(T) == 4)
    return __int_as_float(p);
  else
    return (T)p;
}

constexpr int kMaxFields = 3;

__host__ __device__ constexpr bool sphere_kind(int kind) {
  return kind == LPM_K_PSE_DERIV_SPHERE || kind == LPM_K_PSE_INTERP_SPHERE;
}

// Source records (see PseDeriv, PseInterp, SwePlane for the layouts).
template <typename T, int KIND>
__global__ __launch_bounds__(256) void ops_pack_sources(PackArgs a, int nw, T* __restrict__ rec) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.count) return;
  const int p = a.list[j];
  T* r = rec + (size_t)j * nw;  // written in place: the field count is a runtime value
  for (int k = 0; k < nw; ++k) r[k] = (T)0;
  const double x = a.x[p], y = a.y[p], A = a.area[p];
  int o = 0;
  double kr = 1.0;
  if constexpr (sphere_kind(KIND)) {
    const double z = a.z[p];
    kr = a.sigma * a.radius;
    const double sc = kr / sqrt(x * x + y * y + z * z);
    r[0] = (T)(x * sc);
    r[1] = (T)(y * sc);
    r[2] = (T)(z * sc);
    o = 3;
  } else if constexpr (KIND == LPM_K_SWE_PLANE) {
    r[0] = (T)x;
    r[1] = (T)y;
    o = 2;
  } else {
    r[0] = (T)(x * a.sigma);
    r[1] = (T)(y * a.sigma);
    o = 2;
  }
  if constexpr (KIND == LPM_K_PSE_DERIV_PLANE || KIND == LPM_K_PSE_DERIV_SPHERE) {
    const double* c[kMaxFields] = {a.q0, a.q1, a.q2};
    r[o] = (T)A;
    for (int m = 0; m < a.m; ++m) r[o + 1 + m] = (T)c[m][p];
    o += 1 + a.m;
  } else if constexpr (KIND == LPM_K_PSE_INTERP_PLANE || KIND == LPM_K_PSE_INTERP_SPHERE) {
    r[o] = (T)(a.q0[p] * A);
    o += 1;
  } else {  // SWE: a, b, (h + bottom) A, A
    constexpr double kTwoPi = 6.28318530717958647692;
    r[2] = (T)(a.q0[p] * A / kTwoPi);
    r[3] = (T)(a.q1[p] * A / kTwoPi);
    r[4] = (T)((a.q2[p] + a.q3[p]) * A);
    r[5] = (T)A;
    o = 6;
  }
  r[o] = idx_word<T>(p);
  if constexpr (sphere_kind(KIND)) {
    r[o + 1] = (T)(1.0 / (kr * kr));
    r[o + 2] = (T)(kr * kr);
  }
  if constexpr (KIND == LPM_K_SWE_PLANE) r[7] = (T)a.aux;
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void ops_pack_targets(PackArgs a, Vec4<T>* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  const int p = a.i0 + i;
  const double x = a.x[p], y = a.y[p];
  Vec4<T> t{(T)0, (T)0, (T)0, (T)0};
  if constexpr (sphere_kind(KIND)) {
    const double z = a.z[p];
    const double sc = a.sigma * a.radius / sqrt(x * x + y * y + z * z);
    t.a = (T)(x * sc);
    t.b = (T)(y * sc);
    t.c = (T)(z * sc);
  } else if constexpr (KIND == LPM_K_SWE_PLANE) {
    t.a = (T)x;
    t.b = (T)y;
    t.c = (T)(a.q2[p] + a.q3[p]);
  } else {
    t.a = (T)(x * a.sigma);
    t.b = (T)(y * a.sigma);
  }
  rec[i] = t;
}

// Epilogues.  NO = number of sums per target, OP the derivative operator.
// OP is a template parameter rather than a runtime switch: with the output
// pointer chosen by a runtime switch (a pointer phi), the gfx950 back end of
// the ROCm 7.2 compiler stored the double dot through the pointer of another
// case (the LLVM IR was correct).
template <int KIND, int NO, int OP>
__global__ __launch_bounds__(256) void ops_finalize(FinArgs f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.nt) return;
  double A[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) A[k] = 0.0;
  for (int s = 0; s < f.nsplit; ++s) {
    const double* ps = f.partial + (size_t)s * NO * f.nt;
#pragma unroll
    for (int k = 0; k < NO; ++k) A[k] += ps[(size_t)k * f.nt + i];
  }
  const int p = f.i0 + i;
  if constexpr (KIND == LPM_K_PSE_INTERP_PLANE || KIND == LPM_K_PSE_INTERP_SPHERE) {
    f.out0[p] = A[0] * f.scale;
  } else if constexpr (KIND == LPM_K_SWE_PLANE) {
    f.out0[p] = A[0];
    f.out1[p] = A[1];
    f.out2[p] = A[2] * A[2] + 2.0 * A[3] * A[4] + A[5] * A[5];
    f.out3[p] = A[6] * f.scale;
  } else {
    constexpr int Dm = KIND == LPM_K_PSE_DERIV_SPHERE ? 3 : 2;
    constexpr int M = NO / Dm - 1;
    const double* cf[kMaxFields] = {f.q0, f.q1, f.q2};
    // G[m][k] = D[c_m]_k = (S_mk + c_m,i M_k) * scale
    double G[M][Dm];
#pragma unroll
    for (int m = 0; m < M; ++m) {
      // sphere double dot: the w row pairs w_j with v_i (PSEDirectSum.f90:401-406)
      const int mi = (KIND == LPM_K_PSE_DERIV_SPHERE && OP == kOpDoubleDot && m == 2) ? 1 : m;
      const double ci = cf[mi][p];
#pragma unroll
      for (int k = 0; k < Dm; ++k) G[m][k] = (A[Dm + m * Dm + k] + ci * A[k]) * f.scale;
    }
    if constexpr (KIND == LPM_K_PSE_DERIV_PLANE && OP == kOpGradient) {
      f.out0[p] = G[0][0];
      f.out1[p] = G[0][1];
    } else if constexpr (KIND == LPM_K_PSE_DERIV_PLANE && OP == kOpSecondPartials) {
      static_assert(M == 2, "second partials need two fields");
      f.out0[p] = G[0][0];
      f.out1[p] = 0.5 * (G[0][1] + G[1][0]);
      f.out2[p] = G[1][1];
    } else if constexpr (KIND == LPM_K_PSE_DERIV_PLANE && OP == kOpDoubleDot) {
      static_assert(M == 2, "the plane double dot needs two fields");
      f.out0[p] = G[0][0] * G[0][0] + 2.0 * G[0][1] * G[1][0] + G[1][1] * G[1][1];
    } else if constexpr (KIND == LPM_K_PSE_DERIV_SPHERE && OP == kOpDoubleDot) {
      static_assert(M == 3, "the sphere double dot needs three fields");
      f.out0[p] = G[0][0] * G[0][0] + G[1][1] * G[1][1] + G[2][2] * G[2][2] +
                  2.0 * (G[0][1] * G[1][0] + G[0][2] * G[2][0] + G[1][2] * G[2][1]);
    } else {
      static_assert(KIND == LPM_K_PSE_DERIV_SPHERE, "unknown operator");
      const double xs[3] = {f.x[p], f.y[p], f.z[p]};
      // SphereProjection (SphereGeometry.f90:578-590): I - x x^T
      auto P = [&](int a, int b) { return (a == b ? 1.0 : 0.0) - xs[a] * xs[b]; };
      if constexpr (OP == kOpGradient) {
        double g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) g[a] = P(a, 0) * G[0][0] + P(a, 1) * G[0][1] + P(a, 2) * G[0][2];
        f.out0[p] = g[0];
        f.out1[p] = g[1];
        f.out2[p] = g[2];
      } else {
        static_assert(OP == kOpDivergence && M == 3, "sphere divergence needs three fields");
        double d = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) d += P(a, b) * G[a][b];
        f.out0[p] = d;
      }
    }
  }
}

using KernelFn = void (*)(PairArgs);

template <class P, typename T>
KernelFn pick_r(int r) {
  constexpr bool C = sizeof(T) == 4;
  switch (r) {
    case 1: return pair_sum_kernel<P, T, 1, C>;
    case 2: return pair_sum_kernel<P, T, 2, C>;
    case 4: return pair_sum_kernel<P, T, 4, C>;
    case 8:
      // the grid is sized for the R the host asked for, so an R without an
      // instance must fail rather than fall back to another R
      if constexpr (C) return pair_sum_kernel<P, T, 8, C>;
      return nullptr;
    default: return nullptr;
  }
}

template <typename T>
KernelFn pick_t(int kind, int r, int m, bool series) {
  switch (kind) {
    case LPM_K_PSE_DERIV_PLANE:
      if (m == 1) return pick_r<PseDeriv<T, false, 1, false>, T>(r);
      if (m == 2) return pick_r<PseDeriv<T, false, 2, false>, T>(r);
      return nullptr;
    case LPM_K_PSE_DERIV_SPHERE:
      if (series) {
        if (m == 1) return pick_r<PseDeriv<T, true, 1, true>, T>(r);
        if (m == 3) return pick_r<PseDeriv<T, true, 3, true>, T>(r);
        return nullptr;
      }
      if (m == 1) return pick_r<PseDeriv<T, true, 1, false>, T>(r);
      if (m == 3) return pick_r<PseDeriv<T, true, 3, false>, T>(r);
      return nullptr;
    case LPM_K_PSE_INTERP_PLANE: return pick_r<PseInterp<T, false, false>, T>(r);
    case LPM_K_PSE_INTERP_SPHERE:
      if (series) return pick_r<PseInterp<T, true, true>, T>(r);
      return pick_r<PseInterp<T, true, false>, T>(r);
    case LPM_K_SWE_PLANE: return pick_r<SwePlane<T>, T>(r);
    default: return nullptr;
  }
}

template <typename T>
void pack_t(int kind, int nw, const PackArgs& as, const PackArgs& at, void* srec, void* trec, int ns, int nt,
            hipStream_t st) {
  dim3 b(256), gs((ns + 255) / 256), gt((nt + 255) / 256);
  PackArgs s = as, t = at;
  s.count = ns;
  t.count = nt;
#define LPM_OPS_PACK(K)                                                                             \
  case K:                                                                                           \
    if (ns > 0) hipLaunchKernelGGL((ops_pack_sources<T, K>), gs, b, 0, st, s, nw, (T*)srec);       \
    if (nt > 0) hipLaunchKernelGGL((ops_pack_targets<T, K>), gt, b, 0, st, t, (Vec4<T>*)trec);     \
    break;
  switch (kind) {
    LPM_OPS_PACK(LPM_K_PSE_DERIV_PLANE)
    LPM_OPS_PACK(LPM_K_PSE_DERIV_SPHERE)
    LPM_OPS_PACK(LPM_K_PSE_INTERP_PLANE)
    LPM_OPS_PACK(LPM_K_PSE_INTERP_SPHERE)
    LPM_OPS_PACK(LPM_K_SWE_PLANE)
  }
#undef LPM_OPS_PACK
}

}  // namespace

bool is_ops_kind(int kind) { return kind >= LPM_K_PSE_DERIV_PLANE && kind <= LPM_K_SWE_PLANE; }

int ops_nout(int kind, int m) {
  switch (kind) {
    case LPM_K_PSE_DERIV_PLANE: return 2 * (m + 1);
    case LPM_K_PSE_DERIV_SPHERE: return 3 * (m + 1);
    case LPM_K_SWE_PLANE: return 7;
    default: return 1;
  }
}

int ops_record_words(int kind, int m) {
  if (kind == LPM_K_PSE_DERIV_SPHERE) return 3 + 1 + m + 3 <= SREC ? SREC : SREC_MAX;
  return SREC;
}

const void* ops_pick_kernel(int kind, int prec, int r, int m, bool series) {
  KernelFn fn = prec == LPM_PREC_FP32_KAHAN ? pick_t<float>(kind, r, m, series) : pick_t<double>(kind, r, m, series);
  return reinterpret_cast<const void*>(fn);
}

void ops_launch_pack(int kind, int prec, const PackArgs& s, const PackArgs& t, void* srec, void* trec, int ns,
                     int nt, hipStream_t st) {
  const int nw = ops_record_words(kind, s.m);
  if (prec == LPM_PREC_FP32_KAHAN)
    pack_t<float>(kind, nw, s, t, srec, trec, ns, nt, st);
  else
    pack_t<double>(kind, nw, s, t, srec, trec, ns, nt, st);
}

int ops_launch_finalize(int kind, int m, const FinArgs& f, hipStream_t st) {
  dim3 b(256), g((f.nt + 255) / 256);
  constexpr int kPlane = LPM_K_PSE_DERIV_PLANE, kSphere = LPM_K_PSE_DERIV_SPHERE;
  if (kind == kPlane && m == 1 && f.op == kOpGradient)
    hipLaunchKernelGGL((ops_finalize<kPlane, 4, kOpGradient>), g, b, 0, st, f);
  else if (kind == kPlane && m == 2 && f.op == kOpSecondPartials)
    hipLaunchKernelGGL((ops_finalize<kPlane, 6, kOpSecondPartials>), g, b, 0, st, f);
  else if (kind == kPlane && m == 2 && f.op == kOpDoubleDot)
    hipLaunchKernelGGL((ops_finalize<kPlane, 6, kOpDoubleDot>), g, b, 0, st, f);
  else if (kind == kSphere && m == 1 && f.op == kOpGradient)
    hipLaunchKernelGGL((ops_finalize<kSphere, 6, kOpGradient>), g, b, 0, st, f);
  else if (kind == kSphere && m == 3 && f.op == kOpDoubleDot)
    hipLaunchKernelGGL((ops_finalize<kSphere, 12, kOpDoubleDot>), g, b, 0, st, f);
  else if (kind == kSphere && m == 3 && f.op == kOpDivergence)
    hipLaunchKernelGGL((ops_finalize<kSphere, 12, kOpDivergence>), g, b, 0, st, f);
  else if (kind == LPM_K_PSE_INTERP_PLANE)
    hipLaunchKernelGGL((ops_finalize<LPM_K_PSE_INTERP_PLANE, 1, 0>), g, b, 0, st, f);
  else if (kind == LPM_K_PSE_INTERP_SPHERE)
    hipLaunchKernelGGL((ops_finalize<LPM_K_PSE_INTERP_SPHERE, 1, 0>), g, b, 0, st, f);
  else if (kind == LPM_K_SWE_PLANE)
    hipLaunchKernelGGL((ops_finalize<LPM_K_SWE_PLANE, 7, 0>), g, b, 0, st, f);
  else
    return LPM_ERR_ARG;  // no epilogue for this kind / field count / operator
  return LPM_OK;
}

}  // namespace lpm
