// lpm_internal.hpp -- context, device buffers and the pair-sum driver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/lpm_direct.h"

#define LPM_VERSION_NUMBER 10000

namespace lpm {

// Grow-only device buffer owned by a context or solver.  A `fixed` buffer is
// a view of memory owned elsewhere (a segment of the static scratch array):
// it never allocates or frees, and a request beyond its size fails.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  static DevBuf view(void* ptr, size_t n) {
    DevBuf b;
    b.p = ptr;
    b.bytes = n;
    b.fixed = true;
    return b;
  }
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (fixed) return hipErrorOutOfMemory;
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
    if (p && !fixed) (void)hipFree(p);
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
  lpm::DevBuf in[12];
  lpm::DevBuf out[4];
  lpm::DevBuf list;
  lpm::DevBuf tgt, src, partial;
  lpm::DevBuf red;  // reductions
  // this device's copy of the static scratch array (lpm_api.hip), resolved
  // once at lpm_create; used by the allocation-free entry point only
  void* scratch = nullptr;
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
  // PSE derivative / SWE kinds: more fields and outputs, the epilogue op and
  // the number of derivative fields m.
  const double* q2 = nullptr;
  const double* q3 = nullptr;
  double* out3 = nullptr;
  int op = 0;
  int m = 1;
  // Targets at points other than the particles (interpolation, Laplacian at
  // points): when tx is set, targets are tx/ty/tz[/tq][i0..i1), outputs are
  // indexed by point, and no self pair is excluded.
  const double* tx = nullptr;
  const double* ty = nullptr;
  const double* tz = nullptr;
  const double* tq = nullptr;
  // Optional caller-owned scratch; when null the context's buffers are used.
  DevBuf* tgt_buf = nullptr;
  DevBuf* src_buf = nullptr;
  DevBuf* partial_buf = nullptr;
  DevBuf* red_buf = nullptr;  // FP32 sphere Green: the two FP64 source totals
  // Source count on the device (set by a device-side compaction): ns is then
  // an upper bound that sizes the launches, and the kernels read the count.
  // Such a sum also records no timing events.
  const int* ns_dev = nullptr;
};

// Epilogues of the PSE derivative kinds.
enum PseOp { kOpGradient = 0, kOpSecondPartials = 1, kOpDoubleDot = 2, kOpDivergence = 3 };

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
  const double* q2 = nullptr;
  const double* q3 = nullptr;
  int m = 1;        // derivative fields
  double aux = 0.0; // SWE: log2(e)/eps^2 (FP32) or 1/eps^2 (FP64)
  const int* count_dev = nullptr;  // sources: record count on the device (count is then a bound)
};

// Ordered sum over source splits and the FP64 epilogue of one kind.
struct FinArgs {
  const double* partial;
  int nt, i0, nsplit;
  const double *x, *y, *z, *q0, *q1, *area;
  const int* active;
  const double* G;
  double radius, scale;
  int fp32;
  double *out0, *out1, *out2;
  const double* q2 = nullptr;
  double* out3 = nullptr;
  int op = 0;
};

// The PSE derivative, interpolation and SWE kinds (lpm_pse_ops.hip).
bool is_ops_kind(int kind);
int ops_nout(int kind, int m);
int ops_record_words(int kind, int m);
const void* ops_pick_kernel(int kind, int prec, int r, int m, bool series);
void ops_launch_pack(int kind, int prec, const PackArgs& s, const PackArgs& t, void* srec, void* trec, int ns,
                     int nt, hipStream_t st);
int ops_launch_finalize(int kind, int m, const FinArgs& f, hipStream_t st);
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
