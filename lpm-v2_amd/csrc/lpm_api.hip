// lpm_api.hip -- C ABI: contexts, host-pointer drop-in sums, helpers,
// timing and tuning.  Device-resident solvers live in lpm_solver.hip.
#include <string.h>

#include <string>

#include "lpm_internal.hpp"
#include "pair_kernels.hpp"

using lpm::DevBuf;
using lpm::fail;
using lpm::hip_fail;

// Workspace of the allocation-free sphere Biot-Savart entry point
// (lpm_bve_sphere_velocity_noalloc): the active-source list and its counts,
// the target and source records and the FP64 split partials of one call.
// Each GPU the library runs on holds its own copy for the life of the
// process; host code reaches it only through the address that lpm_create
// resolves per device (ctx->scratch), never through this name.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxN = 131072;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_create_error;

// Active-particle list from device flags, in two passes over blocks of 1024
// particles: per-block counts, then each block's ascending indices at its
// prefix offset (deterministic; the same list the host-side compaction builds).
constexpr int kCompactBlock = 1024;

__global__ __launch_bounds__(kCompactBlock) void count_active_blocks(int n, const int32_t* __restrict__ active,
                                                                     int* __restrict__ block_count) {
  __shared__ int wsum[kCompactBlock / 64];
  const int i = blockIdx.x * kCompactBlock + threadIdx.x;
  const bool a = i < n && active[i] != 0;
  const unsigned long long m = __ballot(a);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < kCompactBlock / 64; ++k) t += wsum[k];
    block_count[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kCompactBlock) void write_active_list(int n, const int32_t* __restrict__ active,
                                                                   const int* __restrict__ block_count,
                                                                   int* __restrict__ list, int* __restrict__ ns) {
  __shared__ int wsum[kCompactBlock / 64];
  __shared__ int base;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    int b = 0;
    for (int k = 0; k < (int)blockIdx.x; ++k) b += block_count[k];
    base = b;
    if (blockIdx.x == gridDim.x - 1) *ns = b + block_count[blockIdx.x];
  }
  const int i = blockIdx.x * kCompactBlock + threadIdx.x;
  const bool a = i < n && active[i] != 0;
  const unsigned long long m = __ballot(a);
  if (lane == 0) wsum[wid] = __popcll(m);
  __syncthreads();
  if (a) {
    int off = base;
    for (int k = 0; k < wid; ++k) off += wsum[k];
    off += __popcll(m & ((1ull << lane) - 1ull));
    list[off] = i;
  }
}

// Device-memory check for the allocation-free entry point: the pointer must
// be device (or managed) memory of this context's device, with `bytes`
// bytes inside its allocation.
int check_device_array(lpm_ctx* ctx, const void* p, size_t bytes, const char* name) {
  const std::string nm = std::string("lpm_bve_sphere_velocity_noalloc: ") + name;
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, LPM_ERR_ARG, nm + " is not device memory");
  }
  if (at.type == hipMemoryTypeManaged) return LPM_OK;
  if (at.type != hipMemoryTypeDevice) return fail(ctx, LPM_ERR_ARG, nm + " is not device memory");
  if (at.device != ctx->device) return fail(ctx, LPM_ERR_ARG, nm + " is memory of another device");
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, LPM_ERR_ARG, nm + ": allocation range unknown");
  }
  if ((const char*)p + bytes > (const char*)base + size)
    return fail(ctx, LPM_ERR_ARG, nm + " is shorter than n elements");
  return LPM_OK;
}

int check_block(lpm_ctx* ctx, int n, int i0, int i1) {
  if (n < 0) return fail(ctx, LPM_ERR_ARG, "n must be >= 0");
  if (i0 < 0 || i1 < i0 || i1 > n)
    return fail(ctx, LPM_ERR_ARG,
                "target block [i0, i1) must satisfy 0 <= i0 <= i1 <= n (got i0=" + std::to_string(i0) +
                    ", i1=" + std::to_string(i1) + ", n=" + std::to_string(n) + ")");
  return LPM_OK;
}

int check_prec(lpm_ctx* ctx, int precision) {
  if (precision != LPM_PREC_FP64 && precision != LPM_PREC_FP32_KAHAN)
    return fail(ctx, LPM_ERR_ARG, "precision must be LPM_PREC_FP64 (0) or LPM_PREC_FP32_KAHAN (1)");
  return LPM_OK;
}

// Upload k host arrays of n doubles into ctx->in[0..k).
int upload(lpm_ctx* ctx, int n, std::initializer_list<const double*> arrays) {
  int k = 0;
  for (const double* a : arrays) {
    DevBuf& b = ctx->in[k++];
    if (!a) continue;
    LPM_HIP(ctx, b.ensure((size_t)std::max(n, 1) * sizeof(double)));
    if (n > 0)
      LPM_HIP(ctx, hipMemcpyAsync(b.p, a, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  return LPM_OK;
}

int prepare_outputs(lpm_ctx* ctx, int n, int nout) {
  for (int k = 0; k < nout; ++k) LPM_HIP(ctx, ctx->out[k].ensure((size_t)std::max(n, 1) * sizeof(double)));
  return LPM_OK;
}

int download(lpm_ctx* ctx, int i0, int i1, std::initializer_list<double*> outs) {
  int k = 0;
  for (double* o : outs) {
    DevBuf& b = ctx->out[k++];
    if (!o || i1 <= i0) continue;
    LPM_HIP(ctx, hipMemcpyAsync(o + i0, b.as<double>() + i0, (size_t)(i1 - i0) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  }
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

// Common body of the PSE operator and SWE entry points.  `inputs` (n doubles
// each) go to ctx->in[0..) in order, `outputs` to ctx->out[0..); the caller's
// `bind` maps them onto the pair problem (kind, fields, op).  Points, when
// given (npts >= 0), are uploaded after the inputs and become the targets.
struct OpsCall {
  const char* name;
  int n, i0, i1, precision;
  double eps, radius;
  const int32_t* active;
  std::initializer_list<const double*> inputs;
  std::initializer_list<double*> outputs;
  int npts = -1;
  std::initializer_list<const double*> points = {};
};

template <class Bind>
int run_ops_call(lpm_ctx* ctx, const OpsCall& c, Bind bind) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  const std::string nm = c.name;
  const bool at_points = c.npts >= 0;
  if (at_points) {
    if (c.n < 0) return fail(ctx, LPM_ERR_ARG, nm + ": n must be >= 0");
  } else if (int rc = check_block(ctx, c.n, c.i0, c.i1)) {
    return rc;
  }
  if (int rc = check_prec(ctx, c.precision)) return rc;
  bool null_arg = !c.active;
  for (const double* a : c.inputs) null_arg = null_arg || !a;
  for (const double* a : c.points) null_arg = null_arg || !a;
  for (double* a : c.outputs) null_arg = null_arg || !a;
  if (null_arg) return fail(ctx, LPM_ERR_ARG, nm + ": null array argument");
  if (!(c.eps > 0.0) || !(c.radius > 0.0)) return fail(ctx, LPM_ERR_ARG, nm + ": eps and radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  const int t0 = at_points ? 0 : c.i0, t1 = at_points ? c.npts : c.i1;
  if (t1 == t0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, c.active, c.n, ctx->list, nullptr, &ns)) return rc;
  int k = 0;
  for (const double* a : c.inputs) {
    DevBuf& b = ctx->in[k++];
    LPM_HIP(ctx, b.ensure((size_t)std::max(c.n, 1) * sizeof(double)));
    if (c.n > 0)
      LPM_HIP(ctx, hipMemcpyAsync(b.p, a, (size_t)c.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  const int k_pts = k;
  for (const double* a : c.points) {
    DevBuf& b = ctx->in[k++];
    LPM_HIP(ctx, b.ensure((size_t)std::max(c.npts, 1) * sizeof(double)));
    LPM_HIP(ctx, hipMemcpyAsync(b.p, a, (size_t)c.npts * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  const int nout_rows = at_points ? c.npts : c.n;
  if (int rc = prepare_outputs(ctx, nout_rows, (int)c.outputs.size())) return rc;
  lpm::PairProblem pp;
  pp.precision = c.precision;
  pp.n = c.n;
  pp.i0 = t0;
  pp.i1 = t1;
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.eps = c.eps;
  pp.radius = c.radius;
  pp.out0 = ctx->out[0].as<double>();
  pp.out1 = ctx->out[1].as<double>();
  pp.out2 = ctx->out[2].as<double>();
  pp.out3 = ctx->out[3].as<double>();
  auto in = [&](int i) { return ctx->in[i].as<double>(); };
  if (at_points) {
    pp.tx = in(k_pts);
    pp.ty = in(k_pts + 1);
    pp.tz = c.points.size() > 2 ? in(k_pts + 2) : nullptr;
    pp.tq = c.points.size() > 3 ? in(k_pts + 3) : nullptr;
  }
  bind(pp, in);
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  // outputs are copied back in the order given
  int o = 0;
  for (double* out : c.outputs) {
    LPM_HIP(ctx, hipMemcpyAsync(out + t0, ctx->out[o++].as<double>() + t0, (size_t)(t1 - t0) * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
  }
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

}  // namespace

extern "C" {

int lpm_version(void) { return LPM_VERSION_NUMBER; }

int lpm_create(lpm_ctx** out, int device) {
  if (!out) {
    g_create_error = "lpm_create: ctx pointer is NULL";
    return LPM_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_error = std::string("lpm_create: no HIP device available (") +
                     (e != hipSuccess ? hipGetErrorString(e) : "0 devices") + ")";
    return LPM_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    g_create_error = "lpm_create: device " + std::to_string(device) + " out of range (" +
                     std::to_string(ndev) + " devices)";
    return LPM_ERR_ARG;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_error = std::string("lpm_create: hipSetDevice failed: ") + hipGetErrorString(e);
    return LPM_ERR_HIP;
  }
  lpm_ctx* ctx = new lpm_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("lpm_create: hipStreamCreate failed: ") + hipGetErrorString(e);
    delete ctx;
    return LPM_ERR_HIP;
  }
  // this device's copy of the static scratch array (device current above)
  e = hipGetSymbolAddress(&ctx->scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ctx->scratch) {
    g_create_error = std::string("lpm_create: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LPM_ERR_HIP;
  }
  g_create_error.clear();
  *out = ctx;
  return LPM_OK;
}

int lpm_destroy(lpm_ctx* ctx) {
  if (!ctx) return LPM_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->in) b.release();
  for (auto& b : ctx->out) b.release();
  ctx->list.release();
  ctx->tgt.release();
  ctx->src.release();
  ctx->partial.release();
  ctx->red.release();
  for (auto& ep : ctx->pending) {
    (void)hipEventDestroy(ep.a);
    (void)hipEventDestroy(ep.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LPM_OK;
}

const char* lpm_last_error(const lpm_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

int lpm_synchronize(lpm_ctx* ctx) {
  if (!ctx) return LPM_ERR_ARG;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_sphere_velocity(lpm_ctx* ctx, double* u, double* v, double* w, const double* x,
                            const double* y, const double* z, const double* relvort,
                            const double* area, double radius, double rotation_rate,
                            const int32_t* active, int n, int i0, int i1, int precision) {
  (void)rotation_rate;  // unused in the reference too (SphereBVESolver.f90:388)
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!u || !v || !w || !x || !y || !z || !relvort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_velocity: null array argument");
  if (!(radius > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_velocity: radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, nullptr, &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, z, relvort, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 3)) return rc;
  lpm::PairProblem pp;
  pp.kind = LPM_K_SPHERE_BS;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.z = ctx->in[2].as<double>();
  pp.q0 = ctx->in[3].as<double>();
  pp.area = ctx->in[4].as<double>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.radius = radius;
  pp.out0 = ctx->out[0].as<double>();
  pp.out1 = ctx->out[1].as<double>();
  pp.out2 = ctx->out[2].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {u, v, w});
}

int lpm_bve_sphere_velocity_noalloc(lpm_ctx* ctx, double* u, double* v, double* w, const double* x,
                                    const double* y, const double* z, const double* relvort,
                                    const double* area, double radius, double rotation_rate,
                                    const int32_t* active, int n, int i0, int i1, int precision) {
  (void)rotation_rate;
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!u || !v || !w || !x || !y || !z || !relvort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_velocity_noalloc: null array argument");
  if (!(radius > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_velocity_noalloc: radius must be > 0");
  if (n > kNoAllocMaxN)
    return fail(ctx, LPM_ERR_ARG,
                "lpm_bve_sphere_velocity_noalloc: n = " + std::to_string(n) + " exceeds the scratch limit of " +
                    std::to_string(kNoAllocMaxN) + " particles");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  const size_t nb = (size_t)n * sizeof(double);
  const void* ins[6] = {x, y, z, relvort, area, active};
  const size_t in_bytes[6] = {nb, nb, nb, nb, nb, (size_t)n * sizeof(int32_t)};
  const char* in_names[6] = {"x", "y", "z", "relvort", "area", "active"};
  for (int k = 0; k < 6; ++k)
    if (int rc = check_device_array(ctx, ins[k], in_bytes[k], in_names[k])) return rc;
  double* outs[3] = {u, v, w};
  const char* out_names[3] = {"u", "v", "w"};
  for (int k = 0; k < 3; ++k)
    if (int rc = check_device_array(ctx, outs[k], nb, out_names[k])) return rc;

  // scratch layout: [block counts | source count] [list] [target records]
  // [source records] [split partials: the rest]
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const int nblk = (n + kCompactBlock - 1) / kCompactBlock;
  const size_t wsz = precision == LPM_PREC_FP32_KAHAN ? 4 : 8;
  const size_t off_list = up((size_t)(nblk + 1) * sizeof(int));
  const size_t off_tgt = off_list + up((size_t)n * sizeof(int));
  const size_t off_src = off_tgt + up((size_t)(i1 - i0) * lpm::TREC * wsz);
  const size_t off_part = off_src + up((size_t)n * lpm::SREC * wsz);
  if (off_part >= kScratchBytes)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_velocity_noalloc: n too large for the scratch array");
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  int* block_count = reinterpret_cast<int*>(base);
  int* ns_dev = block_count + nblk;
  int* list = reinterpret_cast<int*>(base + off_list);
  DevBuf tb = DevBuf::view(base + off_tgt, off_src - off_tgt);
  DevBuf sb = DevBuf::view(base + off_src, off_part - off_src);
  DevBuf pb = DevBuf::view(base + off_part, kScratchBytes - off_part);

  hipLaunchKernelGGL(count_active_blocks, dim3(nblk), dim3(kCompactBlock), 0, ctx->stream, n, active, block_count);
  hipLaunchKernelGGL(write_active_list, dim3(nblk), dim3(kCompactBlock), 0, ctx->stream, n, active, block_count,
                     list, ns_dev);
  LPM_HIP(ctx, hipGetLastError());
  lpm::PairProblem pp;
  pp.kind = LPM_K_SPHERE_BS;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = x;
  pp.y = y;
  pp.z = z;
  pp.q0 = relvort;
  pp.area = area;
  pp.src_list = list;
  pp.ns = n;  // bound; the kernels read the count from ns_dev
  pp.ns_dev = ns_dev;
  pp.radius = radius;
  pp.out0 = u;
  pp.out1 = v;
  pp.out2 = w;
  pp.tgt_buf = &tb;
  pp.src_buf = &sb;
  pp.partial_buf = &pb;
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_plane_velocity(lpm_ctx* ctx, double* u, double* v, const double* x, const double* y,
                       const double* vort, const double* area, const int32_t* active, int n,
                       int i0, int i1, int precision) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!u || !v || !x || !y || !vort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_plane_velocity: null array argument");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, nullptr, &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, vort, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 2)) return rc;
  lpm::PairProblem pp;
  pp.kind = LPM_K_PLANE_BS;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.q0 = ctx->in[2].as<double>();
  pp.area = ctx->in[3].as<double>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.out0 = ctx->out[0].as<double>();
  pp.out1 = ctx->out[1].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {u, v});
}

int lpm_pse_plane_laplacian(lpm_ctx* ctx, double* lap, const double* x, const double* y,
                            const double* f, const double* area, const int32_t* active, int n,
                            int i0, int i1, double eps, int order, int precision) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!lap || !x || !y || !f || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_pse_plane_laplacian: null array argument");
  if (!(eps > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_pse_plane_laplacian: eps must be > 0");
  if (order != 2 && order != 8)
    return fail(ctx, LPM_ERR_ARG, "lpm_pse_plane_laplacian: order must be 2 or 8");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, nullptr, &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, f, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 1)) return rc;
  lpm::PairProblem pp;
  pp.kind = order == 8 ? LPM_K_PSE8 : LPM_K_PSE2;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.q0 = ctx->in[2].as<double>();
  pp.area = ctx->in[3].as<double>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.eps = eps;
  pp.out0 = ctx->out[0].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {lap});
}

int lpm_pse_sphere_laplacian(lpm_ctx* ctx, double* lap, const double* x, const double* y, const double* z,
                             const double* f, const double* area, const int32_t* active, int n, int i0, int i1,
                             double eps, double radius, int precision) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!lap || !x || !y || !z || !f || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_pse_sphere_laplacian: null array argument");
  if (!(eps > 0.0) || !(radius > 0.0))
    return fail(ctx, LPM_ERR_ARG, "lpm_pse_sphere_laplacian: eps and radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, nullptr, &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, z, f, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 1)) return rc;
  lpm::PairProblem pp;
  pp.kind = LPM_K_PSE_SPHERE8;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.z = ctx->in[2].as<double>();
  pp.q0 = ctx->in[3].as<double>();
  pp.area = ctx->in[4].as<double>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.eps = eps;
  pp.radius = radius;
  pp.out0 = ctx->out[0].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {lap});
}

int lpm_pse_sphere_laplacian_at_points(lpm_ctx* ctx, double* lap, const double* px, const double* py,
                                       const double* pz, const double* pf, int nt, const double* x,
                                       const double* y, const double* z, const double* f, const double* area,
                                       const int32_t* active, int n, double eps, double radius, int precision) {
  if (nt < 0) return ctx ? fail(ctx, LPM_ERR_ARG, "lpm_pse_sphere_laplacian_at_points: nt must be >= 0") : LPM_ERR_ARG;
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_sphere_laplacian_at_points", n, 0, 0, precision, eps, radius, active,
                              {x, y, z, f, area}, {lap}, nt, {px, py, pz, pf}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_SPHERE8;
                        pp.x = in(0), pp.y = in(1), pp.z = in(2), pp.q0 = in(3), pp.area = in(4);
                      });
}

int lpm_pse_plane_gradient(lpm_ctx* ctx, double* gx, double* gy, const double* x, const double* y,
                           const double* f, const double* area, const int32_t* active, int n, int i0, int i1,
                           double eps, int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_plane_gradient", n, i0, i1, precision, eps, 1.0, active, {x, y, f, area},
                              {gx, gy}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_PLANE;
                        pp.op = lpm::kOpGradient;
                        pp.m = 1;
                        pp.x = in(0), pp.y = in(1), pp.q0 = in(2), pp.area = in(3);
                      });
}

int lpm_pse_sphere_gradient(lpm_ctx* ctx, double* gx, double* gy, double* gz, const double* x, const double* y,
                            const double* z, const double* f, const double* area, const int32_t* active, int n,
                            int i0, int i1, double eps, double radius, int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_sphere_gradient", n, i0, i1, precision, eps, radius, active,
                              {x, y, z, f, area}, {gx, gy, gz}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_SPHERE;
                        pp.op = lpm::kOpGradient;
                        pp.m = 1;
                        pp.x = in(0), pp.y = in(1), pp.z = in(2), pp.q0 = in(3), pp.area = in(4);
                      });
}

int lpm_pse_plane_second_partials(lpm_ctx* ctx, double* dxx, double* dxy, double* dyy, const double* x,
                                  const double* y, const double* gx, const double* gy, const double* area,
                                  const int32_t* active, int n, int i0, int i1, double eps, int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_plane_second_partials", n, i0, i1, precision, eps, 1.0, active,
                              {x, y, gx, gy, area}, {dxx, dxy, dyy}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_PLANE;
                        pp.op = lpm::kOpSecondPartials;
                        pp.m = 2;
                        pp.x = in(0), pp.y = in(1), pp.q0 = in(2), pp.q1 = in(3), pp.area = in(4);
                      });
}

int lpm_pse_plane_double_dot(lpm_ctx* ctx, double* ddot, const double* x, const double* y, const double* u,
                             const double* v, const double* area, const int32_t* active, int n, int i0, int i1,
                             double eps, int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_plane_double_dot", n, i0, i1, precision, eps, 1.0, active,
                              {x, y, u, v, area}, {ddot}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_PLANE;
                        pp.op = lpm::kOpDoubleDot;
                        pp.m = 2;
                        pp.x = in(0), pp.y = in(1), pp.q0 = in(2), pp.q1 = in(3), pp.area = in(4);
                      });
}

int lpm_pse_sphere_double_dot(lpm_ctx* ctx, double* ddot, const double* x, const double* y, const double* z,
                              const double* u, const double* v, const double* w, const double* area,
                              const int32_t* active, int n, int i0, int i1, double eps, double radius,
                              int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_sphere_double_dot", n, i0, i1, precision, eps, radius, active,
                              {x, y, z, u, v, w, area}, {ddot}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_SPHERE;
                        pp.op = lpm::kOpDoubleDot;
                        pp.m = 3;
                        pp.x = in(0), pp.y = in(1), pp.z = in(2);
                        pp.q0 = in(3), pp.q1 = in(4), pp.q2 = in(5), pp.area = in(6);
                      });
}

int lpm_pse_sphere_divergence(lpm_ctx* ctx, double* div, const double* x, const double* y, const double* z,
                              const double* u, const double* v, const double* w, const double* area,
                              const int32_t* active, int n, int i0, int i1, double eps, double radius,
                              int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_sphere_divergence", n, i0, i1, precision, eps, radius, active,
                              {x, y, z, u, v, w, area}, {div}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_DERIV_SPHERE;
                        pp.op = lpm::kOpDivergence;
                        pp.m = 3;
                        pp.x = in(0), pp.y = in(1), pp.z = in(2);
                        pp.q0 = in(3), pp.q1 = in(4), pp.q2 = in(5), pp.area = in(6);
                      });
}

int lpm_pse_plane_interpolate(lpm_ctx* ctx, double* out, const double* px, const double* py, int nt,
                              const double* x, const double* y, const double* f, const double* area,
                              const int32_t* active, int n, double eps, int precision) {
  if (nt < 0) return ctx ? fail(ctx, LPM_ERR_ARG, "lpm_pse_plane_interpolate: nt must be >= 0") : LPM_ERR_ARG;
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_plane_interpolate", n, 0, 0, precision, eps, 1.0, active, {x, y, f, area},
                              {out}, nt, {px, py}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_INTERP_PLANE;
                        pp.x = in(0), pp.y = in(1), pp.q0 = in(2), pp.area = in(3);
                      });
}

int lpm_pse_sphere_interpolate(lpm_ctx* ctx, double* out, const double* px, const double* py, const double* pz,
                               int nt, const double* x, const double* y, const double* z, const double* f,
                               const double* area, const int32_t* active, int n, double eps, double radius,
                               int precision) {
  if (nt < 0) return ctx ? fail(ctx, LPM_ERR_ARG, "lpm_pse_sphere_interpolate: nt must be >= 0") : LPM_ERR_ARG;
  return run_ops_call(ctx,
                      OpsCall{"lpm_pse_sphere_interpolate", n, 0, 0, precision, eps, radius, active,
                              {x, y, z, f, area}, {out}, nt, {px, py, pz}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_PSE_INTERP_SPHERE;
                        pp.x = in(0), pp.y = in(1), pp.z = in(2), pp.q0 = in(3), pp.area = in(4);
                      });
}

int lpm_swe_plane_rhs(lpm_ctx* ctx, double* u, double* v, double* ddot, double* lapsurf, const double* x,
                      const double* y, const double* vort, const double* div, const double* h,
                      const double* bottom, const double* area, const int32_t* active, int n, int i0, int i1,
                      double eps, int precision) {
  return run_ops_call(ctx,
                      OpsCall{"lpm_swe_plane_rhs", n, i0, i1, precision, eps, 1.0, active,
                              {x, y, vort, div, h, bottom, area}, {u, v, ddot, lapsurf}},
                      [](lpm::PairProblem& pp, auto in) {
                        pp.kind = LPM_K_SWE_PLANE;
                        pp.x = in(0), pp.y = in(1);
                        pp.q0 = in(2), pp.q1 = in(3), pp.q2 = in(4), pp.q3 = in(5), pp.area = in(6);
                      });
}

int lpm_bve_sphere_stream(lpm_ctx* ctx, double* psi_rel, double* psi_abs, const double* x,
                          const double* y, const double* z, const double* relvort,
                          const double* absvort, const double* area, double radius,
                          const int32_t* active, int n, int i0, int i1, int precision) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!psi_rel || !psi_abs || !x || !y || !z || !relvort || !absvort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_stream: null array argument");
  if (!(radius > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_sphere_stream: radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, &ctx->in[7], &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, z, relvort, absvort, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 2)) return rc;
  lpm::PairProblem pp;
  pp.kind = LPM_K_SPHERE_GREEN;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.z = ctx->in[2].as<double>();
  pp.q0 = ctx->in[3].as<double>();
  pp.q1 = ctx->in[4].as<double>();
  pp.area = ctx->in[5].as<double>();
  pp.active = ctx->in[7].as<int>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.radius = radius;
  pp.out0 = ctx->out[0].as<double>();
  pp.out1 = ctx->out[1].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {psi_rel, psi_abs});
}

int lpm_plane_stream(lpm_ctx* ctx, double* psi, const double* x, const double* y,
                     const double* vort, const double* area, const int32_t* active, int n, int i0,
                     int i1, int precision) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (int rc = check_block(ctx, n, i0, i1)) return rc;
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!psi || !x || !y || !vort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_plane_stream: null array argument");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (i1 == i0) return LPM_OK;
  int ns = 0;
  if (int rc = lpm::upload_active_list(ctx, active, n, ctx->list, nullptr, &ns)) return rc;
  if (int rc = upload(ctx, n, {x, y, vort, area})) return rc;
  if (int rc = prepare_outputs(ctx, n, 1)) return rc;
  lpm::PairProblem pp;
  pp.kind = LPM_K_PLANE_GREEN;
  pp.precision = precision;
  pp.n = n;
  pp.i0 = i0;
  pp.i1 = i1;
  pp.x = ctx->in[0].as<double>();
  pp.y = ctx->in[1].as<double>();
  pp.q0 = ctx->in[2].as<double>();
  pp.area = ctx->in[3].as<double>();
  pp.src_list = ctx->list.as<int>();
  pp.ns = ns;
  pp.out0 = ctx->out[0].as<double>();
  if (int rc = lpm::run_pair_sum(ctx, pp)) return rc;
  return download(ctx, i0, i1, {psi});
}

int lpm_pse_eps(double max_edge_length, double power, double* eps) {
  if (!eps || !(max_edge_length > 0.0) || !(power > 0.0)) return LPM_ERR_ARG;
  *eps = pow(max_edge_length, power);
  return LPM_OK;
}

int lpm_load_balance(int n_items, int n_procs, int* index_start, int* index_end,
                     int* message_length) {
  if (n_items < 0 || n_procs <= 0 || !index_start || !index_end || !message_length)
    return LPM_ERR_ARG;
  const int chunk = n_items / n_procs;
  for (int r = 0; r < n_procs; ++r) {
    index_start[r] = r * chunk + 1;
    index_end[r] = (r + 1) * chunk;
  }
  index_end[n_procs - 1] = n_items;
  for (int r = 0; r < n_procs; ++r) message_length[r] = index_end[r] - index_start[r] + 1;
  return LPM_OK;
}

int lpm_diagnostics(lpm_ctx* ctx, const double* u, const double* v, const double* w,
                    const double* vort, const double* area, const int32_t* active, int n,
                    double* ke, double* enstrophy, double* circulation) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->err.clear();
  if (n < 0) return fail(ctx, LPM_ERR_ARG, "lpm_diagnostics: n must be >= 0");
  if (!u || !v || !vort || !area || !active || !ke || !enstrophy || !circulation)
    return fail(ctx, LPM_ERR_ARG, "lpm_diagnostics: null argument");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = upload(ctx, n, {u, v, w, vort, area})) return rc;
  LPM_HIP(ctx, ctx->in[7].ensure((size_t)std::max(n, 1) * sizeof(int)));
  if (n > 0)
    LPM_HIP(ctx, hipMemcpyAsync(ctx->in[7].p, active, (size_t)n * sizeof(int), hipMemcpyHostToDevice,
                                ctx->stream));
  double out[3];
  if (int rc = lpm::run_diagnostics(ctx, n, ctx->in[0].as<double>(), ctx->in[1].as<double>(),
                                    w ? ctx->in[2].as<double>() : nullptr, ctx->in[3].as<double>(),
                                    ctx->in[4].as<double>(), ctx->in[7].as<int>(), out))
    return rc;
  *ke = out[0];
  *enstrophy = out[1];
  *circulation = out[2];
  return LPM_OK;
}

int lpm_set_timing(lpm_ctx* ctx, int enable) {
  if (!ctx) return LPM_ERR_ARG;
  ctx->timing = enable != 0;
  return LPM_OK;
}

int lpm_kernel_stats(lpm_ctx* ctx, int kind, int64_t* launches, double* total_ms) {
  if (!ctx) return LPM_ERR_ARG;
  if (kind < 0 || kind >= LPM_K_COUNT || !launches || !total_ms)
    return fail(ctx, LPM_ERR_ARG, "lpm_kernel_stats: bad kind or null output");
  if (int rc = lpm::collect_timing(ctx)) return rc;
  *launches = ctx->launches[kind];
  *total_ms = ctx->total_ms[kind];
  return LPM_OK;
}

int lpm_reset_kernel_stats(lpm_ctx* ctx) {
  if (!ctx) return LPM_ERR_ARG;
  if (int rc = lpm::collect_timing(ctx)) return rc;
  for (int k = 0; k < LPM_K_COUNT; ++k) {
    ctx->launches[k] = 0;
    ctx->total_ms[k] = 0.0;
  }
  return LPM_OK;
}

int lpm_set_tuning(lpm_ctx* ctx, int kind, int precision, int targets_per_lane, int tile_sources,
                   int splits) {
  if (!ctx) return LPM_ERR_ARG;
  if (kind < 0 || kind >= LPM_K_COUNT) return fail(ctx, LPM_ERR_ARG, "lpm_set_tuning: bad kind");
  if (int rc = check_prec(ctx, precision)) return rc;
  if (!(targets_per_lane == 0 || targets_per_lane == 1 || targets_per_lane == 2 ||
        targets_per_lane == 4 || targets_per_lane == 8))
    return fail(ctx, LPM_ERR_ARG, "lpm_set_tuning: targets_per_lane must be 0, 1, 2, 4 or 8");
  if (tile_sources < 0 || tile_sources % 8 != 0 || tile_sources > 65536)
    return fail(ctx, LPM_ERR_ARG, "lpm_set_tuning: tile_sources must be a multiple of 8 in [0, 65536]");
  if (splits < 0 || splits > 65536) return fail(ctx, LPM_ERR_ARG, "lpm_set_tuning: splits out of range");
  lpm::Tuning& t = ctx->tune[kind][precision == LPM_PREC_FP32_KAHAN ? 1 : 0];
  t.r = targets_per_lane;
  t.kt = tile_sources;
  t.splits = splits;
  return LPM_OK;
}

int lpm_get_tuning(lpm_ctx* ctx, int kind, int precision, int* targets_per_lane, int* tile_sources,
                   int* splits) {
  if (!ctx) return LPM_ERR_ARG;
  if (kind < 0 || kind >= LPM_K_COUNT || !targets_per_lane || !tile_sources || !splits)
    return fail(ctx, LPM_ERR_ARG, "lpm_get_tuning: bad kind or null output");
  if (int rc = check_prec(ctx, precision)) return rc;
  const lpm::Tuning& t = ctx->tune[kind][precision == LPM_PREC_FP32_KAHAN ? 1 : 0];
  *targets_per_lane = t.r;
  *tile_sources = t.kt;
  *splits = t.splits;
  return LPM_OK;
}

}  // extern "C"
