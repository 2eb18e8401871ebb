// lpm_mlp.hip -- far-field surrogate network: parameters, FP32 training
// kernels (forward, backward, Adam) and the FP16-MFMA inference forward.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "lpm_internal.hpp"
#include "lpm_mlp.hpp"
#include "mlp_kernels.hpp"

using lpm::DevBuf;
using lpm::fail;

namespace lpm {

int64_t mlp_param_count(int nin, int width, int nhidden, int nout) {
  int64_t c = (int64_t)width * nin + width;
  for (int l = 1; l < nhidden; ++l) c += (int64_t)width * width + width;
  c += (int64_t)nout * width + nout;
  return c;
}

namespace {

// dZ = dH act'(Z) from the output H alone: tanh' = 1 - H^2;
// ISRU' = (1 + Z^2)^(-3/2) = (1 - H^2)^(3/2)
__global__ __launch_bounds__(256) void act_backward(int n, int act, const float* __restrict__ H,
                                                    float* __restrict__ dH) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = 1.0f - H[i] * H[i];
  dH[i] = dH[i] * (act == kActIsru ? t * sqrtf(fmaxf(t, 0.0f)) : t);
}

// dY = 2 (Y - T) / (B nout); per-block partial loss (fixed order -> deterministic)
__global__ __launch_bounds__(256) void mse_grad(int n, float scale, const float* __restrict__ Y,
                                                const float* __restrict__ T, float* __restrict__ dY,
                                                double* __restrict__ part) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const float r = Y[i] - T[i];
    dY[i] = 2.0f * r * scale;
    a += (double)r * r;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// column sums of a [rows][cols] matrix (bias gradients)
__global__ __launch_bounds__(256) void col_sums(int rows, int cols, const float* __restrict__ X, float* __restrict__ out) {
  const int c = blockIdx.x;
  __shared__ float s[256];
  float a = 0.f;
  for (int r = threadIdx.x; r < rows; r += 256) a += X[(size_t)r * cols + c];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = s[0];
}

__global__ __launch_bounds__(256) void adam_step(int n, float lr, float b1, float b2, float eps, float c1, float c2,
                                                 float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void gather_rows(int n, int cols, const int* __restrict__ idx,
                                                   const float* __restrict__ src, float* __restrict__ dst,
                                                   float scale) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * cols) return;
  const int r = t / cols, c = t % cols;
  dst[t] = src[(size_t)idx[r] * cols + c] * scale;
}

__global__ __launch_bounds__(256) void scale_range(int n, float* __restrict__ p, float s) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] *= s;
}

// sum of squares of n values, one FP64 partial per block (fixed order)
__global__ __launch_bounds__(256) void sumsq_partial(int64_t n, const float* __restrict__ v, double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += (double)v[i] * (double)v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Standalone FP16 inference: features [n][nin] FP32 -> out [n][nout].
template <int ACT>
__global__ __launch_bounds__(256) void mlp_mfma_forward_kernel(int n, int nin, int nout, const float* __restrict__ X,
                                                               const mlp_t* __restrict__ frag,
                                                               const float* __restrict__ bias, float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) mlp_t sfrag[kMlpFragElems];
  __shared__ __attribute__((aligned(16))) float sbias[kMlpBiasElems];
  for (int t = threadIdx.x; t < kMlpFragElems / 8; t += blockDim.x)
    reinterpret_cast<mlpx8_t*>(sfrag)[t] = reinterpret_cast<const mlpx8_t*>(frag)[t];
  for (int t = threadIdx.x; t < kMlpBiasElems; t += blockDim.x) sbias[t] = bias[t];
  __syncthreads();
  constexpr int NCT = 4;  // 64 samples per wave and pass
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int waves = gridDim.x * (blockDim.x >> 6);
  const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int base = wid * 16 * NCT; base < n; base += waves * 16 * NCT) {
    mlpx8_t in8[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int s = base + 16 * ct + (lane & 15);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int f = 8 * g + e;
        in8[ct][e] = (mlp_t)((s < n && f < nin) ? X[(size_t)s * nin + f] : 0.f);
      }
    }
    float out[NCT][4];
    mlp_wave_forward<NCT, ACT>(in8, sfrag, sbias, out);
    if (g == 0) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int s = base + 16 * ct + lane;
        if (s < n)
          for (int v = 0; v < nout && v < 4; ++v) Y[(size_t)s * nout + v] = out[ct][v];
      }
    }
  }
}

// Products with a long K and few output tiles (the weight gradients dZ^T H over
// a minibatch) are split along K when `part` has room (`part_elems` floats).
void gemm(hipStream_t st, bool ta, bool tb, const GemmArgs& g0, float* part = nullptr, size_t part_elems = 0) {
  GemmArgs g = g0;
  g.part = nullptr;
  g.kchunk = 0;
  const int tiles = ((g.N + 63) / 64) * ((g.M + 63) / 64);
  int splits = 1;
  if (part && tiles < 128 && g.K >= 512 && !g.bias && !g.Z && g.beta == 0.f && g.act == 0) {
    splits = std::min((256 + tiles - 1) / tiles, g.K / 256);
    while (splits > 1 && (size_t)splits * g.M * g.N > part_elems) --splits;
  }
  if (splits > 1) {
    g.kchunk = ((g.K + splits - 1) / splits + 15) / 16 * 16;
    splits = (g.K + g.kchunk - 1) / g.kchunk;
    g.part = part;
  }
  dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, splits), blk(256);
  if (!ta && !tb) hipLaunchKernelGGL((gemm_f32<false, false>), grid, blk, 0, st, g);
  else if (!ta && tb) hipLaunchKernelGGL((gemm_f32<false, true>), grid, blk, 0, st, g);
  else if (ta && !tb) hipLaunchKernelGGL((gemm_f32<true, false>), grid, blk, 0, st, g);
  else hipLaunchKernelGGL((gemm_f32<true, true>), grid, blk, 0, st, g);
  if (g.part)
    hipLaunchKernelGGL(gemm_splitk_reduce, dim3((g.M * g.N + 255) / 256), blk, 0, st, splits, g.M, g.N, g.part, g.C,
                       g.ldc);
}

}  // namespace

// ----------------------------------------------------------------- Mlp
Mlp::~Mlp() { release(); }

void Mlp::release() {
  DevBuf* all[] = {&p, &grad, &m, &v, &frag, &bias, &act, &pre, &scratch, &loss_part, &splitk, &xb, &tb, &idx};
  for (DevBuf* b : all) b->release();
}

int Mlp::init(lpm_ctx* c, int nin_, int width_, int nhidden_, int nout_, int act_, uint64_t seed) {
  ctx = c;
  act_kind = act_;
  nin = nin_;
  width = width_;
  nhidden = nhidden_;
  nout = nout_;
  np = mlp_param_count(nin, width, nhidden, nout);
  LPM_HIP(ctx, p.ensure((size_t)np * 4));
  LPM_HIP(ctx, grad.ensure((size_t)np * 4));
  LPM_HIP(ctx, m.ensure((size_t)np * 4));
  LPM_HIP(ctx, v.ensure((size_t)np * 4));
  LPM_HIP(ctx, frag.ensure((size_t)kMlpFragElems * 2));
  LPM_HIP(ctx, bias.ensure((size_t)kMlpBiasElems * 4));
  // Glorot-uniform weights, zero biases (host RNG, seeded)
  std::vector<float> h((size_t)np, 0.f);
  std::mt19937_64 rng(seed);
  int64_t off = 0;
  for (int l = 0; l <= nhidden; ++l) {
    const int di = l == 0 ? nin : width;
    const int dout = l == nhidden ? nout : width;
    const float lim = std::sqrt(6.0f / (float)(di + dout));
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int64_t k = 0; k < (int64_t)di * dout; ++k) h[(size_t)(off + k)] = U(rng);
    off += (int64_t)di * dout + dout;
  }
  LPM_HIP(ctx, hipMemcpyAsync(p.p, h.data(), (size_t)np * 4, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemsetAsync(m.p, 0, (size_t)np * 4, ctx->stream));
  LPM_HIP(ctx, hipMemsetAsync(v.p, 0, (size_t)np * 4, ctx->stream));
  adam_t = 0;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return pack();
}

bool Mlp::mfma_ok() const { return nin <= 32 && width == 128 && nhidden == 3 && nout <= 4; }

int Mlp::pack() {
  if (!mfma_ok()) return LPM_OK;
  hipLaunchKernelGGL(mlp_pack_fragments, dim3(64), dim3(256), 0, ctx->stream, p.as<float>(), nin, nout,
                     frag.as<mlp_t>(), bias.as<float>(), mlp_hidden_scale(act_kind));
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int Mlp::ensure_batch(int n) {
  const size_t per = (size_t)(nhidden + 1) * (size_t)std::max(width, nout);
  LPM_HIP(ctx, act.ensure((size_t)n * per * 4));
  LPM_HIP(ctx, pre.ensure((size_t)n * per * 4));
  LPM_HIP(ctx, scratch.ensure((size_t)n * std::max(width, nout) * 2 * 4));
  LPM_HIP(ctx, loss_part.ensure(1024 * 8));
  // split-K partials of the weight gradients: up to 64 splits of a width x width block
  LPM_HIP(ctx, splitk.ensure((size_t)64 * std::max(width, nin) * std::max(width, nout) * 4));
  return LPM_OK;
}

// activations layout: act + l * n * width holds layer l's output (l = 0..nhidden),
// the last slot (l = nhidden) the network output [n][nout]
int Mlp::forward_f32(int n, const float* X) {
  if (int rc = ensure_batch(n)) return rc;
  const float* in = X;
  int din = nin;
  int64_t off = 0;
  const size_t slot = (size_t)n * std::max(width, nout);
  for (int l = 0; l <= nhidden; ++l) {
    const int dout = l == nhidden ? nout : width;
    GemmArgs g{};
    g.A = in;
    g.B = p.as<float>() + off;
    g.C = act.as<float>() + (size_t)l * slot;
    g.Z = nullptr;
    g.bias = p.as<float>() + off + (int64_t)din * dout;
    g.M = n;
    g.N = dout;
    g.K = din;
    g.lda = din;
    g.ldb = din;
    g.ldc = dout;
    g.alpha = 1.f;
    g.beta = 0.f;
    g.act = l == nhidden ? 0 : act_kind;
    gemm(ctx->stream, false, true, g);
    in = g.C;
    din = dout;
    off += (int64_t)(l == 0 ? nin : width) * dout + dout;
  }
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int Mlp::loss_grad(int n, const float* X, const float* T, double* loss) {
  if (int rc = forward_f32(n, X)) return rc;
  cur_n = n;
  const size_t slot = (size_t)n * std::max(width, nout);
  float* Y = act.as<float>() + (size_t)nhidden * slot;
  float* dcur = scratch.as<float>();
  float* dnext = scratch.as<float>() + slot;
  const int nb = std::min(1024, (n * nout + 255) / 256);
  hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, ctx->stream, n * nout, 1.0f / (float)(n * nout), Y, T,
                     dcur, loss_part.as<double>());
  // layer offsets
  std::vector<int64_t> offs((size_t)nhidden + 2, 0);
  for (int l = 0; l <= nhidden; ++l) {
    const int di = l == 0 ? nin : width, dout = l == nhidden ? nout : width;
    offs[(size_t)l + 1] = offs[(size_t)l] + (int64_t)di * dout + dout;
  }
  for (int l = nhidden; l >= 0; --l) {
    const int di = l == 0 ? nin : width, dout = l == nhidden ? nout : width;
    const float* Hin = l == 0 ? X : act.as<float>() + (size_t)(l - 1) * slot;
    float* gW = grad.as<float>() + offs[(size_t)l];
    float* gb = gW + (int64_t)di * dout;
    // dW = dZ^T H  ([dout][n] x [n][di])
    GemmArgs g{};
    g.A = dcur;
    g.B = Hin;
    g.C = gW;
    g.M = dout;
    g.N = di;
    g.K = n;
    g.lda = dout;
    g.ldb = di;
    g.ldc = di;
    g.alpha = 1.f;
    gemm(ctx->stream, true, false, g, splitk.as<float>(), splitk.bytes / sizeof(float));
    hipLaunchKernelGGL(col_sums, dim3(dout), dim3(256), 0, ctx->stream, n, dout, dcur, gb);
    if (l > 0) {
      // dH = dZ W  ([n][dout] x [dout][di]), then times tanh'
      GemmArgs h{};
      h.A = dcur;
      h.B = p.as<float>() + offs[(size_t)l];
      h.C = dnext;
      h.M = n;
      h.N = di;
      h.K = dout;
      h.lda = dout;
      h.ldb = di;
      h.ldc = di;
      h.alpha = 1.f;
      gemm(ctx->stream, false, false, h);
      hipLaunchKernelGGL(act_backward, dim3((n * di + 255) / 256), dim3(256), 0, ctx->stream, n * di, act_kind,
                         act.as<float>() + (size_t)(l - 1) * slot, dnext);
      std::swap(dcur, dnext);
    }
  }
  LPM_HIP(ctx, hipGetLastError());
  if (loss) {
    std::vector<double> part((size_t)nb);
    LPM_HIP(ctx, hipMemcpyAsync(part.data(), loss_part.p, (size_t)nb * 8, hipMemcpyDeviceToHost, ctx->stream));
    LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (double x : part) s += x;
    *loss = s / (double)(n * nout);
  }
  return LPM_OK;
}

int Mlp::adam(float lr) {
  ++adam_t;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(adam_step, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, (int)np, lr, b1, b2,
                     1e-8f, c1, c2, p.as<float>(), grad.as<float>(), m.as<float>(), v.as<float>());
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int Mlp::scale_output(float s) {
  const int n = nout * width + nout;  // last layer: W[nout][width], b[nout]
  hipLaunchKernelGGL(scale_range, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, p.as<float>() + (np - n), s);
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int Mlp::rms(int64_t n, const float* v, double* out) {
  const int nb = 256;
  LPM_HIP(ctx, loss_part.ensure(nb * sizeof(double)));
  hipLaunchKernelGGL(sumsq_partial, dim3(nb), dim3(256), 0, ctx->stream, n, v, loss_part.as<double>());
  std::vector<double> h(nb);
  LPM_HIP(ctx, hipMemcpyAsync(h.data(), loss_part.p, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double s = 0.0;
  for (double x : h) s += x;
  *out = n > 0 ? std::sqrt(s / (double)n) : 0.0;
  return LPM_OK;
}

int Mlp::train(int n, const float* X, const float* T, int epochs, int batch, float lr, uint64_t seed,
               double* final_loss, std::vector<double>* history, bool cosine, float tscale) {
  if (n <= 0 || batch <= 0) return fail(ctx, LPM_ERR_ARG, "mlp train: empty data or batch");
  batch = std::min(batch, n);
  if (int rc = ensure_batch(batch)) return rc;
  LPM_HIP(ctx, xb.ensure((size_t)batch * nin * 4));
  LPM_HIP(ctx, tb.ensure((size_t)batch * nout * 4));
  LPM_HIP(ctx, idx.ensure((size_t)n * 4));
  std::vector<int> perm((size_t)n);
  std::mt19937_64 rng(seed);
  double last = 0.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int i = 0; i < n; ++i) perm[(size_t)i] = i;
    std::shuffle(perm.begin(), perm.end(), rng);
    LPM_HIP(ctx, hipMemcpyAsync(idx.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    double ep_loss = 0.0;
    int nbatches = 0;
    const float lr_ep = cosine ? (float)(0.5 * lr * (1.0 + std::cos(3.14159265358979323846 * ep / epochs))) : lr;
    for (int b0 = 0; b0 + batch <= n; b0 += batch) {
      hipLaunchKernelGGL(gather_rows, dim3((batch * nin + 255) / 256), dim3(256), 0, ctx->stream, batch, nin,
                         idx.as<int>() + b0, X, xb.as<float>(), 1.0f);
      hipLaunchKernelGGL(gather_rows, dim3((batch * nout + 255) / 256), dim3(256), 0, ctx->stream, batch, nout,
                         idx.as<int>() + b0, T, tb.as<float>(), tscale);
      const bool want = (b0 + 2 * batch > n);  // read the loss of the epoch's last batch only
      double l = 0.0;
      if (int rc = loss_grad(batch, xb.as<float>(), tb.as<float>(), want ? &l : nullptr)) return rc;
      if (int rc = adam(lr_ep)) return rc;
      if (want) {
        ep_loss = l;
        ++nbatches;
      }
    }
    last = ep_loss;
    if (history) history->push_back(ep_loss);
  }
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (int rc = pack()) return rc;
  if (final_loss) *final_loss = last;
  return LPM_OK;
}

int Mlp::forward_mfma(int n, const float* X, float* Y) {
  if (!mfma_ok()) return fail(ctx, LPM_ERR_UNSUPPORTED, "MFMA forward needs nin <= 32, width 128, 3 hidden layers, nout <= 4");
  const int blocks = std::max(1, std::min(2 * ctx->num_cus, (n + 255) / 256));
  if (act_kind == kActIsru)
    hipLaunchKernelGGL(mlp_mfma_forward_kernel<kActIsru>, dim3(blocks), dim3(256), 0, ctx->stream, n, nin, nout, X,
                       frag.as<mlp_t>(), bias.as<float>(), Y);
  else
    hipLaunchKernelGGL(mlp_mfma_forward_kernel<kActTanh>, dim3(blocks), dim3(256), 0, ctx->stream, n, nin, nout, X,
                       frag.as<mlp_t>(), bias.as<float>(), Y);
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

}  // namespace lpm

// ================================================================ C ABI
struct lpm_mlp {
  lpm::Mlp net;
  lpm::DevBuf x, y, t;
};

extern "C" {

int lpm_mlp_new(lpm_ctx* ctx, lpm_mlp** out, int nin, int width, int nhidden, int nout, int activation,
                uint64_t seed) {
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (nin <= 0 || width <= 0 || nhidden <= 0 || nout <= 0 || nhidden > 16)
    return fail(ctx, LPM_ERR_ARG, "lpm_mlp_new: bad dimensions");
  if (activation != LPM_ACT_TANH && activation != LPM_ACT_ISRU)
    return fail(ctx, LPM_ERR_ARG, "lpm_mlp_new: activation must be LPM_ACT_TANH or LPM_ACT_ISRU");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  lpm_mlp* m = new lpm_mlp();
  if (int rc = m->net.init(ctx, nin, width, nhidden, nout, activation, seed)) {
    delete m;
    return rc;
  }
  *out = m;
  return LPM_OK;
}

int lpm_mlp_param_count(lpm_mlp* m, int64_t* n) {
  if (!m || !n) return LPM_ERR_ARG;
  *n = m->net.np;
  return LPM_OK;
}

int lpm_mlp_set_params(lpm_mlp* m, const float* params, int64_t n) {
  if (!m || !params) return LPM_ERR_ARG;
  lpm_ctx* ctx = m->net.ctx;
  if (n != m->net.np) return fail(ctx, LPM_ERR_ARG, "lpm_mlp_set_params: wrong parameter count");
  LPM_HIP(ctx, hipMemcpyAsync(m->net.p.p, params, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = m->net.pack()) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_mlp_get_params(lpm_mlp* m, float* params, int64_t n) {
  if (!m || !params) return LPM_ERR_ARG;
  lpm_ctx* ctx = m->net.ctx;
  if (n != m->net.np) return fail(ctx, LPM_ERR_ARG, "lpm_mlp_get_params: wrong parameter count");
  LPM_HIP(ctx, hipMemcpyAsync(params, m->net.p.p, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_mlp_forward(lpm_mlp* m, int n, const float* in, float* out, int use_mfma) {
  if (!m || !in || !out || n <= 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = m->net.ctx;
  ctx->err.clear();
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, m->x.ensure((size_t)n * m->net.nin * 4));
  LPM_HIP(ctx, m->y.ensure((size_t)n * m->net.nout * 4));
  LPM_HIP(ctx, hipMemcpyAsync(m->x.p, in, (size_t)n * m->net.nin * 4, hipMemcpyHostToDevice, ctx->stream));
  const float* Y = nullptr;
  if (use_mfma) {
    if (int rc = m->net.forward_mfma(n, m->x.as<float>(), m->y.as<float>())) return rc;
    Y = m->y.as<float>();
  } else {
    if (int rc = m->net.forward_f32(n, m->x.as<float>())) return rc;
    Y = m->net.act.as<float>() + (size_t)m->net.nhidden * (size_t)n * std::max(m->net.width, m->net.nout);
  }
  LPM_HIP(ctx, hipMemcpyAsync(out, Y, (size_t)n * m->net.nout * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_mlp_loss_grad(lpm_mlp* m, int n, const float* in, const float* target, float* grad, double* loss) {
  if (!m || !in || !target || !grad || !loss || n <= 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = m->net.ctx;
  ctx->err.clear();
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, m->x.ensure((size_t)n * m->net.nin * 4));
  LPM_HIP(ctx, m->t.ensure((size_t)n * m->net.nout * 4));
  LPM_HIP(ctx, hipMemcpyAsync(m->x.p, in, (size_t)n * m->net.nin * 4, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(m->t.p, target, (size_t)n * m->net.nout * 4, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = m->net.loss_grad(n, m->x.as<float>(), m->t.as<float>(), loss)) return rc;
  LPM_HIP(ctx, hipMemcpyAsync(grad, m->net.grad.p, (size_t)m->net.np * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_mlp_train(lpm_mlp* m, int n, const float* in, const float* target, int epochs, int batch, double lr,
                  uint64_t seed, double* final_loss) {
  if (!m || !in || !target || n <= 0 || epochs < 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = m->net.ctx;
  ctx->err.clear();
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, m->x.ensure((size_t)n * m->net.nin * 4));
  LPM_HIP(ctx, m->t.ensure((size_t)n * m->net.nout * 4));
  LPM_HIP(ctx, hipMemcpyAsync(m->x.p, in, (size_t)n * m->net.nin * 4, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(m->t.p, target, (size_t)n * m->net.nout * 4, hipMemcpyHostToDevice, ctx->stream));
  return m->net.train(n, m->x.as<float>(), m->t.as<float>(), epochs, batch, (float)lr, seed, final_loss, nullptr);
}

int lpm_mlp_delete(lpm_mlp* m) {
  if (!m) return LPM_ERR_ARG;
  (void)hipStreamSynchronize(m->net.ctx->stream);
  m->x.release();
  m->y.release();
  m->t.release();
  delete m;
  return LPM_OK;
}

}  // extern "C"
