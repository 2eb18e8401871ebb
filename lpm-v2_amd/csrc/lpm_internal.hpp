// lpm_internal.hpp -- context, device buffers and the pair-sum driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lpm_direct.h"

#define LPM_VERSION_NUMBER 10000

namespace lpm {

// Grow-only device buffer owned by a context or solver.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    hipError_t e = hipMalloc(&p, n < 256 ? 256 : n);
    if (e == hipSuccess) bytes = n < 256 ? 256 : n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

struct Tuning {
  int r = 0;       // targets per lane
  int kt = 0;      // sources per compensated tile
  int splits = 0;  // 0 = automatic
};

struct EventPair {
  hipEvent_t a, b;
  int kind;
};

}  // namespace lpm

struct lpm_ctx {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  std::string err;
  // stateless-call staging
  lpm::DevBuf in[8];
  lpm::DevBuf out[3];
  lpm::DevBuf list;
  lpm::DevBuf tgt, src, partial;
  lpm::DevBuf red;  // reductions
  // tuning per kind and precision
  lpm::Tuning tune[LPM_K_COUNT][2];
  // timing
  bool timing = false;
  std::vector<lpm::EventPair> pending;
  std::vector<hipEvent_t> event_pool;
  int64_t launches[LPM_K_COUNT] = {0};
  double total_ms[LPM_K_COUNT] = {0};
};

namespace lpm {

// Error helpers: every failing path records a message on the context.
int fail(lpm_ctx* ctx, int code, const std::string& msg);
int hip_fail(lpm_ctx* ctx, hipError_t e, const char* what);

#define LPM_HIP(ctx, expr)                                   \
  do {                                                       \
    hipError_t _e = (expr);                                  \
    if (_e != hipSuccess) return ::lpm::hip_fail(ctx, _e, #expr); \
  } while (0)

// Description of one all-pairs sum over device-resident FP64 inputs.
// Targets are particles [i0, i1) of the arrays; sources are the particles in
// `src_list` (device int array of ns particle indices, ascending).
struct PairProblem {
  int kind = LPM_K_SPHERE_BS;
  int precision = LPM_PREC_FP32_KAHAN;
  int n = 0;         // particles in the arrays
  int i0 = 0, i1 = 0;
  const double* x = nullptr;
  const double* y = nullptr;
  const double* z = nullptr;
  const double* q0 = nullptr;  // zeta_rel / omega / f
  const double* q1 = nullptr;  // zeta_abs (sphere stream)
  const double* area = nullptr;
  const int* src_list = nullptr;
  int ns = 0;
  const int* active = nullptr;  // device 0/1 flags (sphere stream self term)
  double radius = 1.0;
  bool on_sphere = false;  // all points on the sphere of `radius` (drop the R^2-|x|^2 terms)
  double eps = 1.0;
  double* out0 = nullptr;  // device outputs, indexed by particle (i0..i1-1)
  double* out1 = nullptr;
  double* out2 = nullptr;
  // Optional caller-owned scratch; when null the context's buffers are used.
  DevBuf* tgt_buf = nullptr;
  DevBuf* src_buf = nullptr;
  DevBuf* partial_buf = nullptr;
};

// Enqueue the whole sum (pack, pair kernel, finalize) on ctx->stream.
int run_pair_sum(lpm_ctx* ctx, const PairProblem& pp);

// Record packing (FP64 SoA -> FP32/FP64 pair records).  prad, when not null,
// gives a per-particle sphere radius (batched problems).
struct PackArgs {
  const double *x, *y, *z, *q0, *q1, *area;
  const double* prad;
  const int* list;
  int count;  // records to write
  int i0;     // first target particle (targets)
  double radius, sigma;
};
void launch_pack(int kind, int precision, const PackArgs& a, void* srec, void* trec, int ns, int nt,
                 hipStream_t st);
Tuning default_tuning(int kind, int prec);
int blocks_per_cu(const void* kernel);
hipEvent_t take_event(lpm_ctx* ctx);

// Resolve queued timing events (synchronizes the stream).
int collect_timing(lpm_ctx* ctx);

// FP64 reductions over active particles: {sum |u|^2 A, sum q^2 A, sum q A}.
int run_diagnostics(lpm_ctx* ctx, int n, const double* u, const double* v, const double* w,
                    const double* q, const double* area, const int* active, double out[3]);

// Active-particle list on the device from host flags (ascending indices).
int upload_active_list(lpm_ctx* ctx, const int32_t* active_host, int n, DevBuf& list,
                       DevBuf* flags, int* ns_out);

}  // namespace lpm
