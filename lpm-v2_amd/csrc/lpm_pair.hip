// lpm_pair.hip -- pack, pair-sum and finalize kernels; the pair-sum driver.
//
// Data flow of one sum (all on the context stream, inputs already in HBM):
//   pack_targets / pack_sources : FP64 SoA particles -> FP32 (or FP64) records
//   pair_sum_kernel             : all pairs, per-split FP64 partials
//   source_totals (green only)  : FP64 sums of the source weights
//   finalize                    : ordered sum over splits + FP64 epilogue
//                                 (cross product, scaling, log offsets)
#include <math.h>

#include <algorithm>
#include <mutex>

#include "lpm_internal.hpp"
#include "pair_kernels.hpp"

namespace lpm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLog2e = 1.4426950408889634074;

int fail(lpm_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(lpm_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) +
                  ") in " + what;
  return fail(ctx, e == hipErrorOutOfMemory ? LPM_ERR_NOMEM : LPM_ERR_HIP, m);
}

// ------------------------------------------------------------------ packing

template <typename T>
__device__ __forceinline__ T idx_word(int p) {
  if constexpr (sizeof(T) == 4)
    return __int_as_float(p);
  else
    return (T)p;
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void pack_sources(PackArgs a, T* __restrict__ rec) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int count = a.count_dev ? *a.count_dev : a.count;
  if (j >= count) return;
  const int p = a.list[j];
  T r[SREC];
#pragma unroll
  for (int k = 0; k < SREC; ++k) r[k] = (T)0;
  const double x = a.x[p], y = a.y[p];
  if constexpr (KIND == LPM_K_SPHERE_BS || KIND == LPM_K_SPHERE_GREEN) {
    const double z = a.z[p];
    const double rad = a.prad ? a.prad[p] : a.radius;
    const double e = rad * rad - (x * x + y * y + z * z);
    r[0] = (T)x;
    r[1] = (T)y;
    r[2] = (T)z;
    r[3] = (T)e;
    if constexpr (KIND == LPM_K_SPHERE_BS) {
      r[4] = (T)(-a.q0[p] * a.area[p] / (2.0 * kPi * rad));
      r[5] = idx_word<T>(p);
    } else {
      const double sc = (sizeof(T) == 4 ? kLn2 : 1.0) / (4.0 * kPi);
      r[4] = (T)(-a.q0[p] * a.area[p] * sc);
      r[5] = (T)(-a.q1[p] * a.area[p] * sc);
      r[6] = idx_word<T>(p);
    }
  } else if constexpr (KIND == LPM_K_PLANE_BS || KIND == LPM_K_PLANE_GREEN) {
    r[0] = (T)x;
    r[1] = (T)y;
    if constexpr (KIND == LPM_K_PLANE_BS)
      r[2] = (T)(a.q0[p] * a.area[p] / (2.0 * kPi));
    else
      r[2] = (T)(a.q0[p] * a.area[p] * (sizeof(T) == 4 ? kLn2 : 1.0) / (4.0 * kPi));
    r[3] = idx_word<T>(p);
  } else if constexpr (KIND == LPM_K_PSE_SPHERE8) {
    // unit direction times k R (k = a.sigma)
    const double z = a.z[p];
    const double kr = a.sigma * a.radius;
    const double sc = kr / sqrt(x * x + y * y + z * z);
    r[0] = (T)(x * sc);
    r[1] = (T)(y * sc);
    r[2] = (T)(z * sc);
    r[3] = (T)(a.q0[p] * a.area[p]);
    r[4] = (T)a.area[p];
    r[5] = idx_word<T>(p);
    r[6] = (T)(1.0 / (kr * kr));
    r[7] = (T)(kr * kr);
  } else {  // plane PSE
    r[0] = (T)(x * a.sigma);
    r[1] = (T)(y * a.sigma);
    r[2] = (T)(a.q0[p] * a.area[p]);
    r[3] = (T)a.area[p];
    r[4] = idx_word<T>(p);
  }
  T* o = rec + (size_t)j * SREC;
#pragma unroll
  for (int k = 0; k < SREC; ++k) o[k] = r[k];
}

template <typename T, int KIND>
__global__ __launch_bounds__(256) void pack_targets(PackArgs a, Vec4<T>* __restrict__ rec) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  const int p = a.i0 + i;
  const double x = a.x[p], y = a.y[p];
  Vec4<T> t{(T)0, (T)0, (T)0, (T)0};
  if constexpr (KIND == LPM_K_SPHERE_BS || KIND == LPM_K_SPHERE_GREEN) {
    const double z = a.z[p];
    const double rad = a.prad ? a.prad[p] : a.radius;
    t.a = (T)x;
    t.b = (T)y;
    t.c = (T)z;
    t.d = (T)(rad * rad - (x * x + y * y + z * z));
  } else if constexpr (KIND == LPM_K_PLANE_BS || KIND == LPM_K_PLANE_GREEN) {
    t.a = (T)x;
    t.b = (T)y;
  } else if constexpr (KIND == LPM_K_PSE_SPHERE8) {
    const double z = a.z[p];
    const double sc = a.sigma * a.radius / sqrt(x * x + y * y + z * z);
    t.a = (T)(x * sc);
    t.b = (T)(y * sc);
    t.c = (T)(z * sc);
    t.d = (T)a.q0[p];
  } else {
    t.a = (T)(x * a.sigma);
    t.b = (T)(y * a.sigma);
    t.c = (T)a.q0[p];
  }
  rec[i] = t;
}

__global__ __launch_bounds__(256) void zero_fill(size_t n, double* __restrict__ p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ----------------------------------------------------- green source totals
// G_k = sum over sources of -q_k A / (4 pi) (sphere) in FP64, one block,
// fixed order: lane-strided partials then a tree in LDS.
__global__ __launch_bounds__(256) void source_totals(int ns, const int* __restrict__ list,
                                                     const double* __restrict__ q0,
                                                     const double* __restrict__ q1,
                                                     const double* __restrict__ area,
                                                     double* __restrict__ G) {
  __shared__ double s0[256], s1[256];
  double a0 = 0.0, a1 = 0.0;
  for (int j = threadIdx.x; j < ns; j += 256) {
    const int p = list[j];
    a0 += -q0[p] * area[p] / (4.0 * kPi);
    a1 += -q1[p] * area[p] / (4.0 * kPi);
  }
  s0[threadIdx.x] = a0;
  s1[threadIdx.x] = a1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    G[0] = s0[0];
    G[1] = s1[0];
  }
}

// ---------------------------------------------------------------- finalize
template <int KIND, int NO>
__global__ __launch_bounds__(256) void finalize_kernel(FinArgs f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
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
  if constexpr (KIND == LPM_K_SPHERE_BS) {
    const double x = f.x[p], y = f.y[p], z = f.z[p];
    f.out0[p] = A[1] * z - A[2] * y;
    f.out1[p] = A[2] * x - A[0] * z;
    f.out2[p] = A[0] * y - A[1] * x;
  } else if constexpr (KIND == LPM_K_PLANE_BS) {
    f.out0[p] = -A[0];
    f.out1[p] = A[1];
  } else if constexpr (KIND == LPM_K_PSE8 || KIND == LPM_K_PSE2 || KIND == LPM_K_PSE_SPHERE8) {
    f.out0[p] = A[0] * f.scale;
  } else if constexpr (KIND == LPM_K_SPHERE_GREEN) {
    if (f.fp32) {
      const double act = f.active[p] ? 1.0 : 0.0;
      const double gr = -f.q0[p] * f.area[p] / (4.0 * kPi) * act;
      const double ga = -f.q1[p] * f.area[p] / (4.0 * kPi) * act;
      f.out0[p] = A[0] - kLn2 * (f.G[0] - gr);
      f.out1[p] = A[1] - kLn2 * (f.G[1] - ga);
    } else {
      f.out0[p] = A[0];
      f.out1[p] = A[1];
    }
  } else {  // plane green
    f.out0[p] = A[0];
  }
}

// ------------------------------------------------------------ dispatching
static int nout_of(int kind) {
  switch (kind) {
    case LPM_K_SPHERE_BS: return 3;
    case LPM_K_PLANE_BS: return 2;
    case LPM_K_SPHERE_GREEN: return 2;
    default: return 1;
  }
}

Tuning default_tuning(int kind, int prec) {
  Tuning t;
  if (prec == LPM_PREC_FP32_KAHAN) {
    // targets per lane from a sweep of R in {1, 2, 4, 8} x KT in {128, 256,
    // 512} on MI355X (icosahedral L7/L8, 512^2 plane, source chunks of
    // ~2,048, each source's record loaded one source ahead of its
    // arithmetic): the sphere Biot-Savart and Green sums run fastest at R = 2
    // (L7: 36.0 / 35.1 ms against 37.7 / 36.1 at R = 4; the L8 solver's sum
    // 515 against 548 ms), the plane Biot-Savart and Green sums at R = 4, the
    // plane PSE sums at R = 2 and the sphere PSE sum (long series distance,
    // more registers) at R = 1; KT other than 256 changes a launch by under 1 %
    switch (kind) {
      case LPM_K_PLANE_BS:
      case LPM_K_PLANE_GREEN: t.r = 4; break;
      case LPM_K_PSE_SPHERE8: t.r = 1; break;
      default: t.r = 2; break;
    }
    t.kt = 256;
  } else {
    // FP64: R = 2, except the PSE Laplacian sums, 1-2 % faster at R = 4
    // (512^2 PSE-2 48.3 against 49.0 ms, sphere PSE-8 at L7 199.8 against
    // 203.1 ms)
    t.r = (kind == LPM_K_PSE2 || kind == LPM_K_PSE8 || kind == LPM_K_PSE_SPHERE8) ? 4 : 2;
    t.kt = 64;
  }
  return t;
}

using KernelFn = void (*)(PairArgs);

template <class P, typename T>
static KernelFn pick_r(int r) {
  constexpr bool C = sizeof(T) == 4;
  switch (r) {
    case 1: return pair_sum_kernel<P, T, 1, C>;
    case 2: return pair_sum_kernel<P, T, 2, C>;
    case 8: return pair_sum_kernel<P, T, 8, C>;
    default: return pair_sum_kernel<P, T, 4, C>;
  }
}

// variant: sphere kinds, targets on the sphere (no e terms); sphere PSE, the
// series form of the distance (FP32 only)
static KernelFn pick_kernel(int kind, int prec, int r, bool variant) {
  const bool on_sphere = variant;
  if (prec == LPM_PREC_FP32_KAHAN) {
    switch (kind) {
      case LPM_K_SPHERE_BS:
        return on_sphere ? pick_r<SphereBS<float, false>, float>(r) : pick_r<SphereBS<float, true>, float>(r);
      case LPM_K_PLANE_BS: return pick_r<PlaneBS<float>, float>(r);
      case LPM_K_PSE8: return pick_r<PsePolicy<float, 8>, float>(r);
      case LPM_K_PSE2: return pick_r<PsePolicy<float, 2>, float>(r);
      case LPM_K_PSE_SPHERE8:
        return variant ? pick_r<PseSphere<float, true>, float>(r) : pick_r<PseSphere<float, false>, float>(r);
      case LPM_K_SPHERE_GREEN:
        return on_sphere ? pick_r<SphereGreen<float, false>, float>(r) : pick_r<SphereGreen<float, true>, float>(r);
      default: return pick_r<PlaneGreen<float>, float>(r);
    }
  }
  switch (kind) {
    case LPM_K_SPHERE_BS:
      return on_sphere ? pick_r<SphereBS<double, false>, double>(r) : pick_r<SphereBS<double, true>, double>(r);
    case LPM_K_PLANE_BS: return pick_r<PlaneBS<double>, double>(r);
    case LPM_K_PSE8: return pick_r<PsePolicy<double, 8>, double>(r);
    case LPM_K_PSE2: return pick_r<PsePolicy<double, 2>, double>(r);
    case LPM_K_PSE_SPHERE8:
      return variant ? pick_r<PseSphere<double, true>, double>(r) : pick_r<PseSphere<double, false>, double>(r);
    case LPM_K_SPHERE_GREEN:
      return on_sphere ? pick_r<SphereGreen<double, false>, double>(r) : pick_r<SphereGreen<double, true>, double>(r);
    default: return pick_r<PlaneGreen<double>, double>(r);
  }
}

template <typename T>
static void launch_pack_t(int kind, const PackArgs& a_src, const PackArgs& a_tgt, void* srec, void* trec,
                          int ns, int nt, hipStream_t st) {
  dim3 b(256);
  dim3 gs((ns + 255) / 256), gt((nt + 255) / 256);
  PackArgs as = a_src, at = a_tgt;
  as.count = ns;
  at.count = nt;
#define LPM_PACK_CASE(K)                                                                    \
  case K:                                                                                   \
    if (ns > 0) hipLaunchKernelGGL((pack_sources<T, K>), gs, b, 0, st, as, (T*)srec);      \
    if (nt > 0) hipLaunchKernelGGL((pack_targets<T, K>), gt, b, 0, st, at, (Vec4<T>*)trec); \
    break;
  switch (kind) {
    LPM_PACK_CASE(LPM_K_SPHERE_BS)
    LPM_PACK_CASE(LPM_K_PLANE_BS)
    LPM_PACK_CASE(LPM_K_PSE8)
    LPM_PACK_CASE(LPM_K_PSE2)
    LPM_PACK_CASE(LPM_K_PSE_SPHERE8)
    LPM_PACK_CASE(LPM_K_SPHERE_GREEN)
    LPM_PACK_CASE(LPM_K_PLANE_GREEN)
  }
#undef LPM_PACK_CASE
}

static void launch_pack2(int kind, int precision, const PackArgs& as, const PackArgs& at, void* srec,
                         void* trec, int ns, int nt, hipStream_t st) {
  if (precision == LPM_PREC_FP32_KAHAN)
    launch_pack_t<float>(kind, as, at, srec, trec, ns, nt, st);
  else
    launch_pack_t<double>(kind, as, at, srec, trec, ns, nt, st);
}

void launch_pack(int kind, int precision, const PackArgs& a, void* srec, void* trec, int ns, int nt,
                 hipStream_t st) {
  launch_pack2(kind, precision, a, a, srec, trec, ns, nt, st);
}

static void launch_finalize(int kind, const FinArgs& f, hipStream_t st) {
  dim3 b(256), g((f.nt + 255) / 256);
  switch (kind) {
    case LPM_K_SPHERE_BS: hipLaunchKernelGGL((finalize_kernel<LPM_K_SPHERE_BS, 3>), g, b, 0, st, f); break;
    case LPM_K_PLANE_BS: hipLaunchKernelGGL((finalize_kernel<LPM_K_PLANE_BS, 2>), g, b, 0, st, f); break;
    case LPM_K_PSE8: hipLaunchKernelGGL((finalize_kernel<LPM_K_PSE8, 1>), g, b, 0, st, f); break;
    case LPM_K_PSE2: hipLaunchKernelGGL((finalize_kernel<LPM_K_PSE2, 1>), g, b, 0, st, f); break;
    case LPM_K_PSE_SPHERE8: hipLaunchKernelGGL((finalize_kernel<LPM_K_PSE_SPHERE8, 1>), g, b, 0, st, f); break;
    case LPM_K_SPHERE_GREEN: hipLaunchKernelGGL((finalize_kernel<LPM_K_SPHERE_GREEN, 2>), g, b, 0, st, f); break;
    default: hipLaunchKernelGGL((finalize_kernel<LPM_K_PLANE_GREEN, 1>), g, b, 0, st, f); break;
  }
}

// Resident 256-thread blocks per CU of a kernel, cached in a fixed table (no
// heap allocation, so the allocation-free entry point may use it; a kernel
// that finds the table full is queried every time).
struct OccEntry {
  const void* fn;
  int nb;
};
constexpr int kOccSlots = 1024;
static OccEntry g_occ[kOccSlots];
static int g_occ_used = 0;
static std::mutex g_occ_mutex;

int blocks_per_cu(const void* kfn) {
  std::lock_guard<std::mutex> lock(g_occ_mutex);
  for (int k = 0; k < g_occ_used; ++k)
    if (g_occ[k].fn == kfn) return g_occ[k].nb;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, 256, 0) != hipSuccess || nb <= 0) nb = 1;
  if (g_occ_used < kOccSlots) g_occ[g_occ_used++] = {kfn, nb};
  return nb;
}

// Source splits.  Short source chunks are faster on MI355X: a sweep of the
// chunk length (1,024 ... 32,768 sources, and the whole-chip-round choice
// below) over the sphere, plane, Green, PSE and SWE kernels, FP32 and FP64,
// at 10^5-2*10^6 targets, put the fastest launches at 1,024-2,048 sources per
// block (7-20 % faster than chunks of 10^4-10^5), from finer dynamic balance
// across the 256 CUs and a source working set that stays in L2.  So: chunks of
// kChunkTarget sources, and at least enough splits that the blocks fill the
// chip in whole rounds of resident blocks (round efficiency B / (rounds *
// slots) >= 0.9, for small problems), within max_split_mem splits (the FP64
// partials of all splits are kept until the ordered reduction).  Problems
// with few target tiles (2*10^5 targets: 192 blocks per split) also want
// kMinRounds rounds of blocks, down to chunks of kChunkMin sources (cubed
// sphere L7: 4.33 ms at 2,048 sources per chunk, 4.23 ms at 1,024).  Smaller
// problems run out of blocks before they run out of rounds: at icosahedral L5
// (30,722 targets, 20,480 sources) chunks of 256 instead of 1,024 took the
// FP64 sum from 0.52 to 0.41 ms and the FP32 sum from 0.233 to 0.214 ms, with
// L6 and the cubed sphere L7 unchanged.
constexpr int kChunkTarget = 2048;
constexpr int kChunkMin = 256;
constexpr int kMinRounds = 12;
constexpr size_t kPartialCap = size_t(2) << 30;  // bytes of FP64 split partials

static int choose_splits(int n_wtiles, int ns, int kt, int slots, int max_split_mem) {
  const int nblk_t = (n_wtiles + 3) / 4;
  const int min_chunk = std::max(kt, 256);
  int max_split = std::max(1, ns / min_chunk);
  max_split = std::min(max_split, std::max(1, max_split_mem));
  int fill = 1;
  double best_eff = -1.0;
  for (int s = 1; s <= max_split; ++s) {
    const double B = (double)nblk_t * s;
    const double rounds = std::ceil(B / slots);
    const double eff = B / (rounds * slots);
    if (eff >= 0.9) {
      fill = s;
      break;
    }
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      fill = s;
    }
  }
  const int by_chunk = (ns + kChunkTarget - 1) / kChunkTarget;
  const long long by_rounds = ((long long)kMinRounds * slots + nblk_t - 1) / std::max(nblk_t, 1);
  const int rounds_split = (int)std::min<long long>(by_rounds, std::max(1, ns / kChunkMin));
  return std::max(1, std::min(max_split, std::max(fill, std::max(by_chunk, rounds_split))));
}

hipEvent_t take_event(lpm_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int collect_timing(lpm_ctx* ctx) {
  if (ctx->pending.empty()) return LPM_OK;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& ep : ctx->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ep.a, ep.b) == hipSuccess) {
      ctx->total_ms[ep.kind] += ms;
      ctx->launches[ep.kind] += 1;
    }
    ctx->event_pool.push_back(ep.a);
    ctx->event_pool.push_back(ep.b);
  }
  ctx->pending.clear();
  return LPM_OK;
}

int run_pair_sum(lpm_ctx* ctx, const PairProblem& pp) {
  const int kind = pp.kind;
  const int prec = pp.precision;
  if (kind < 0 || kind >= LPM_K_COUNT || kind == LPM_K_FAR_FIELD) return fail(ctx, LPM_ERR_ARG, "unknown pair kind");
  if (prec != LPM_PREC_FP64 && prec != LPM_PREC_FP32_KAHAN)
    return fail(ctx, LPM_ERR_ARG, "precision must be LPM_PREC_FP64 or LPM_PREC_FP32_KAHAN");
  const int nt = pp.i1 - pp.i0;
  if (nt <= 0) return LPM_OK;
  const int pi = prec == LPM_PREC_FP32_KAHAN ? 1 : 0;
  Tuning tu = default_tuning(kind, prec);
  const Tuning& ov = ctx->tune[kind][pi];
  if (ov.r) tu.r = ov.r;
  if (ov.kt) tu.kt = ov.kt;
  tu.splits = ov.splits;

  const size_t wsz = prec == LPM_PREC_FP32_KAHAN ? 4 : 8;
  const bool ops = is_ops_kind(kind);
  const int no = ops ? ops_nout(kind, pp.m) : nout_of(kind);
  const int nw = ops ? ops_record_words(kind, pp.m) : SREC;
  if (ops && !ov.r) {
    // targets per lane from a sweep of R in {1, 2, 4, 8} on MI355X (512^2
    // plane, icosahedral L7): in FP32 the sphere operators (series distance)
    // at R = 1, the plane derivative sums and interpolation at R = 2, the
    // 7-sum SWE kernel at R = 1; in FP64 every operator at R = 2 (R = 1 was
    // 3-7 % slower for the sphere gradient and divergence, R = 4 up to 25 %)
    if (prec == LPM_PREC_FP32_KAHAN)
      tu.r = kind == LPM_K_PSE_DERIV_SPHERE || kind == LPM_K_PSE_INTERP_SPHERE || kind == LPM_K_SWE_PLANE ? 1 : 2;
    else
      tu.r = 2;
  }
  const int ns = pp.ns;
  DevBuf& tb = pp.tgt_buf ? *pp.tgt_buf : ctx->tgt;
  DevBuf& sb = pp.src_buf ? *pp.src_buf : ctx->src;
  DevBuf& pb = pp.partial_buf ? *pp.partial_buf : ctx->partial;

  const int per_wave = 64 * tu.r;
  const int n_wtiles = (nt + per_wave - 1) / per_wave;
  // sphere PSE kinds: the series form of the distance when 7 eps <= 0.56 R
  // (FP32, 7 terms) or 7 eps <= 0.35 R (FP64, 10 terms)
  const bool sphere_pse = kind == LPM_K_PSE_SPHERE8 || kind == LPM_K_PSE_DERIV_SPHERE ||
                          kind == LPM_K_PSE_INTERP_SPHERE;
  const double series_reach = prec == LPM_PREC_FP32_KAHAN ? 0.56 : 0.35;
  const bool variant = sphere_pse ? (7.0 * pp.eps <= series_reach * pp.radius) : pp.on_sphere;
  KernelFn fn = ops ? reinterpret_cast<KernelFn>(const_cast<void*>(ops_pick_kernel(kind, prec, tu.r, pp.m, variant)))
                    : pick_kernel(kind, prec, tu.r, variant);
  if (!fn) return fail(ctx, LPM_ERR_ARG, "no pair kernel for this kind, precision and field count");
  if (pp.ns_dev && (ops || kind == LPM_K_SPHERE_GREEN))
    return fail(ctx, LPM_ERR_UNSUPPORTED, "a device-side source count is not supported for this kind");
  int nsplit = 1, chunk = std::max(ns, 1);
  if (ns > 0) {
    const size_t split_bytes = (size_t)no * nt * sizeof(double);
    size_t cap = std::min<size_t>(4096, kPartialCap / split_bytes);
    if (pb.fixed) cap = std::min(cap, pb.bytes / split_bytes);
    if (cap < 1) return fail(ctx, LPM_ERR_NOMEM, "partial buffer too small for one source split");
    const int split_mem = (int)cap;
    int want = tu.splits > 0 ? tu.splits
                             : choose_splits(n_wtiles, ns, tu.kt, blocks_per_cu((const void*)fn) * ctx->num_cus,
                                             split_mem);
    if (pb.fixed) want = std::min(want, split_mem);
    chunk = (ns + want - 1) / want;
    chunk = (chunk + tu.kt - 1) / tu.kt * tu.kt;
    nsplit = (ns + chunk - 1) / chunk;
  }

  LPM_HIP(ctx, tb.ensure((size_t)nt * TREC * wsz));
  LPM_HIP(ctx, sb.ensure((size_t)std::max(ns, 1) * nw * wsz));
  LPM_HIP(ctx, pb.ensure((size_t)nsplit * no * nt * sizeof(double)));

  const bool fp32 = prec == LPM_PREC_FP32_KAHAN;
  PackArgs pa{pp.x, pp.y, pp.z, pp.q0, pp.q1, pp.area, nullptr, pp.src_list, 0, pp.i0, pp.radius, 1.0};
  pa.q2 = pp.q2;
  pa.q3 = pp.q3;
  pa.m = pp.m;
  PackArgs pt = pa;  // targets: the count is nt
  pa.count_dev = pp.ns_dev;
  const bool scaled = kind == LPM_K_PSE8 || kind == LPM_K_PSE2 || kind == LPM_K_PSE_SPHERE8 ||
                      kind == LPM_K_PSE_DERIV_PLANE || kind == LPM_K_PSE_DERIV_SPHERE ||
                      kind == LPM_K_PSE_INTERP_PLANE || kind == LPM_K_PSE_INTERP_SPHERE;
  if (scaled) pa.sigma = (fp32 ? std::sqrt(kLog2e) : 1.0) / pp.eps;
  if (kind == LPM_K_SWE_PLANE) pa.aux = (fp32 ? kLog2e : 1.0) / (pp.eps * pp.eps);
  pt.sigma = pa.sigma;
  pt.aux = pa.aux;
  if (pp.tx) {
    pt.x = pp.tx;
    pt.y = pp.ty;
    pt.z = pp.tz;
    pt.q0 = pp.tq;
  }
  if (ops)
    ops_launch_pack(kind, prec, pa, pt, sb.p, tb.p, ns, nt, ctx->stream);
  else
    launch_pack2(kind, prec, pa, pt, sb.p, tb.p, ns, nt, ctx->stream);
  LPM_HIP(ctx, hipGetLastError());

  if (ns > 0) {
    PairArgs a;
    a.tgt = tb.p;
    a.src = sb.p;
    a.partial = pb.as<double>();
    a.nt = nt;
    // targets at points: an index no source has, so no pair is skipped
    a.tgt_idx0 = pp.tx ? -(1 << 30) : pp.i0;
    a.ns = ns;
    a.chunk = chunk;
    a.nsplit = nsplit;
    a.kt = tu.kt;
    a.n_wtiles = n_wtiles;
    a.ns_dev = pp.ns_dev;
    const int nblk_t = (n_wtiles + 3) / 4;
    dim3 grid((unsigned)(nblk_t * nsplit)), block(256);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing && !pp.ns_dev) {
      e0 = take_event(ctx);
      e1 = take_event(ctx);
      if (e0 && e1) LPM_HIP(ctx, hipEventRecord(e0, ctx->stream));
    }
    hipLaunchKernelGGL(fn, grid, block, 0, ctx->stream, a);
    LPM_HIP(ctx, hipGetLastError());
    if (ctx->timing && e0 && e1) {
      LPM_HIP(ctx, hipEventRecord(e1, ctx->stream));
      ctx->pending.push_back({e0, e1, kind});
    }
  } else {
    // no sources: zero partials by a kernel (a memset node must not enter a captured graph)
    nsplit = 1;
    const size_t nz = (size_t)no * nt;
    hipLaunchKernelGGL(zero_fill, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, ctx->stream, nz, pb.as<double>());
    LPM_HIP(ctx, hipGetLastError());
  }

  FinArgs f;
  f.partial = pb.as<double>();
  f.nt = nt;
  f.i0 = pp.i0;
  f.nsplit = nsplit;
  f.x = pp.x;
  f.y = pp.y;
  f.z = pp.z;
  f.q0 = pp.q0;
  f.q1 = pp.q1;
  f.area = pp.area;
  f.active = pp.active;
  f.G = nullptr;
  f.radius = pp.radius;
  f.scale = 1.0;
  f.fp32 = prec == LPM_PREC_FP32_KAHAN;
  f.out0 = pp.out0;
  f.out1 = pp.out1;
  f.out2 = pp.out2;
  f.q2 = pp.q2;
  f.out3 = pp.out3;
  f.op = pp.op;
  const double e4 = pp.eps * pp.eps * pp.eps * pp.eps;
  if (kind == LPM_K_PSE8 || kind == LPM_K_PSE2 || kind == LPM_K_PSE_SPHERE8)
    f.scale = (kind == LPM_K_PSE2 ? 4.0 : 1.0) / (kPi * e4);
  // derivative sums carry the separation scaled by sigma
  if (kind == LPM_K_PSE_DERIV_PLANE || kind == LPM_K_PSE_DERIV_SPHERE) f.scale = 1.0 / (pa.sigma * kPi * e4);
  if (kind == LPM_K_PSE_INTERP_PLANE || kind == LPM_K_PSE_INTERP_SPHERE) f.scale = 1.0 / (kPi * pp.eps * pp.eps);
  if (kind == LPM_K_SWE_PLANE) f.scale = 1.0 / (kPi * e4);
  if (kind == LPM_K_SPHERE_GREEN && f.fp32) {
    DevBuf& rb = pp.red_buf ? *pp.red_buf : ctx->red;
    LPM_HIP(ctx, rb.ensure(2 * sizeof(double)));
    hipLaunchKernelGGL(source_totals, dim3(1), dim3(256), 0, ctx->stream, ns, pp.src_list, pp.q0,
                       pp.q1, pp.area, rb.as<double>());
    f.G = rb.as<double>();
  }
  if (ops) {
    if (ops_launch_finalize(kind, pp.m, f, ctx->stream) != LPM_OK)
      return fail(ctx, LPM_ERR_ARG, "no epilogue for this operator and field count");
  } else {
    launch_finalize(kind, f, ctx->stream);
  }
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

// ------------------------------------------------------------- diagnostics
// Deterministic two-pass FP64 reduction: fixed grid of 1024 blocks, each a
// strided partial then an LDS tree; the last pass sums the 1024 partials.
constexpr int kRedBlocks = 1024;

__global__ __launch_bounds__(256) void diag_partial(int n, const double* __restrict__ u,
                                                    const double* __restrict__ v,
                                                    const double* __restrict__ w,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ area,
                                                    const int* __restrict__ active,
                                                    double* __restrict__ part) {
  __shared__ double s[3][256];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kRedBlocks * 256) {
    if (!active[i]) continue;
    const double m = u[i] * u[i] + v[i] * v[i] + (w ? w[i] * w[i] : 0.0);
    a0 += m * area[i];
    a1 += q[i] * q[i] * area[i];
    a2 += q[i] * area[i];
  }
  s[0][threadIdx.x] = a0;
  s[1][threadIdx.x] = a1;
  s[2][threadIdx.x] = a2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k)
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) part[c * kRedBlocks + blockIdx.x] = s[c][0];
}

__global__ __launch_bounds__(256) void diag_final(const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double s[3][256];
  for (int c = 0; c < 3; ++c) {
    double a = 0.0;
    for (int k = threadIdx.x; k < kRedBlocks; k += 256) a += part[c * kRedBlocks + k];
    s[c][threadIdx.x] = a;
  }
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k)
      for (int c = 0; c < 3; ++c) s[c][threadIdx.x] += s[c][threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * s[0][0];
    out[1] = 0.5 * s[1][0];
    out[2] = s[2][0];
  }
}

int run_diagnostics(lpm_ctx* ctx, int n, const double* u, const double* v, const double* w,
                    const double* q, const double* area, const int* active, double out[3]) {
  LPM_HIP(ctx, ctx->red.ensure((3 * kRedBlocks + 4) * sizeof(double)));
  double* part = ctx->red.as<double>();
  double* res = part + 3 * kRedBlocks;
  hipLaunchKernelGGL(diag_partial, dim3(kRedBlocks), dim3(256), 0, ctx->stream, n, u, v, w, q,
                     area, active, part);
  hipLaunchKernelGGL(diag_final, dim3(1), dim3(256), 0, ctx->stream, part, res);
  LPM_HIP(ctx, hipGetLastError());
  LPM_HIP(ctx, hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int upload_active_list(lpm_ctx* ctx, const int32_t* active_host, int n, DevBuf& list,
                       DevBuf* flags, int* ns_out) {
  std::vector<int> idx;
  idx.reserve((size_t)n);
  std::vector<int> fl((size_t)n);
  for (int i = 0; i < n; ++i) {
    fl[(size_t)i] = active_host[i] != 0;
    if (fl[(size_t)i]) idx.push_back(i);
  }
  *ns_out = (int)idx.size();
  LPM_HIP(ctx, list.ensure(std::max<size_t>(idx.size(), 1) * sizeof(int)));
  if (!idx.empty())
    LPM_HIP(ctx, hipMemcpyAsync(list.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice,
                                ctx->stream));
  if (flags) {
    LPM_HIP(ctx, flags->ensure((size_t)std::max(n, 1) * sizeof(int)));
    if (n > 0)
      LPM_HIP(ctx, hipMemcpyAsync(flags->p, fl.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice,
                                  ctx->stream));
  }
  // the host vectors die at return: make the copies complete first
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

}  // namespace lpm
