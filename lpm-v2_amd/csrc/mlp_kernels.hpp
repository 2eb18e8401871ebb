// mlp_kernels.hpp -- the far-field surrogate network on gfx950.
//
// Network: nin -> width (act) -> ... (nhidden act layers) -> nout, act = tanh or
// ISRU x / sqrt(1 + x^2), parameters per layer W[out][in] (row-major) then
// b[out], FP32 master copy.
//
// Training path (FP32, VALU): tiled GEMMs C = A op(B) with a fused
// bias + activation epilogue, the backward GEMMs, column sums and Adam.
// Inference path (FP16 MFMA, FP32 accumulate), specialised for the
// 13 -> 128 -> 128 -> 128 -> 3 surrogate: activations are kept transposed
// (features = MFMA rows, samples = MFMA columns), so the D registers of one
// layer (lane l: rows 4(l/16)+v of each 16-row tile, column l%16) are, after
// bias + activation + FP16 conversion, exactly the B operand of the next layer once
// the next layer's weight fragments are stored with the matching K order:
// k-slot e of lane group g in k-step ks is feature 32 ks + 16 (e/4) + 4 g + e%4.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpm {

// Inference operands are FP16: the same MFMA rate as BF16 on gfx950 (measured:
// equal launch times) with 3 more mantissa bits, which cut the forward pass's
// error against FP64 from 4.5e-3 to 6e-4 (relative RMS).
typedef _Float16 mlp_t;
typedef mlp_t mlpx8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ floatx4_t mfma16(mlpx8_t a, mlpx8_t b, floatx4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float tanh_fast(float x) {
  // tanh x = 1 - 2 / (exp(2x) + 1), exp via v_exp_f32 (2^y); saturates cleanly
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // 2 log2(e) x
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// Hidden-layer activations: 1 tanh, 2 ISRU x / sqrt(1 + x^2) (one transcendental,
// v_rsq_f32, instead of tanh's two).
constexpr int kActTanh = 1;
constexpr int kActIsru = 2;

template <int ACT>
__device__ __forceinline__ float act_fast(float x) {
  if constexpr (ACT == kActIsru)
    return x * __builtin_amdgcn_rsqf(fmaf(x, x, 1.0f));
  else
    return tanh_fast(x);
}

// Scale of the hidden layers' pre-activations in the MFMA path: with ISRU the
// hidden weights and biases are stored times kIsruScale, so the accumulators
// hold y = x kIsruScale and ISRU(x) = y / sqrt(y^2 + kIsruScale^2), evaluated
// in packed FP16 (v_pk_fma_f16, v_rsq_f16, v_pk_mul_f16) without overflow of
// y^2 for |x| < 16,000.
constexpr float kIsruScale = 1.0f / 64.0f;

// Activation of 8 accumulator values (lo: one 16-row tile's 4, hi: the next
// tile's 4) into the next layer's 8-element B operand.
template <int ACT>
__device__ __forceinline__ mlpx8_t act8(floatx4_t lo, floatx4_t hi) {
  mlpx8_t v;
  if constexpr (ACT == kActIsru) {
    const _Float16 c = (_Float16)(kIsruScale * kIsruScale);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const float a = p < 2 ? lo[2 * p] : hi[2 * p - 4], b = p < 2 ? lo[2 * p + 1] : hi[2 * p - 3];
      const half2_t y = {(_Float16)a, (_Float16)b};
      const half2_t t = y * y + (half2_t){c, c};
      half2_t r;
      r[0] = __builtin_amdgcn_rsqh(t[0]);
      r[1] = __builtin_amdgcn_rsqh(t[1]);
      const half2_t h = y * r;
      v[2 * p] = h[0];
      v[2 * p + 1] = h[1];
    }
    return v;
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (mlp_t)act_fast<ACT>(e < 4 ? lo[e] : hi[e - 4]);
  return v;
}

// Scale the MFMA path applies to the hidden layers' weights and biases.
__host__ __device__ inline float mlp_hidden_scale(int act) {
  return act == kActIsru ? kIsruScale : 1.0f;
}

// ------------------------------------------------------------ FP32 GEMM
// C[M][N] = act(alpha * sum_k A(m, k) B(k, n) + bias[n]) (+ beta C), with
// A(m, k) = TA ? A[k][m] : A[m][k] and B(k, n) = TB ? B[n][k] : B[k][n].
// 64 x 64 tile per 256-thread block, 4 x 4 outputs per thread, K tiles of 16
// staged in LDS.  Z (optional) receives the pre-activation.  Split K: with
// part != nullptr, block z of the grid covers k in [z kchunk, (z+1) kchunk)
// and writes alpha * its partial product to part[z][M][N] (no bias, beta or
// activation); gemm_splitk_reduce then sums the partials in a fixed order.
struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  float* Z;
  const float* bias;
  int M, N, K;
  int lda, ldb, ldc;
  float alpha, beta;
  int act;  // 0 none, kActTanh, kActIsru
  float* part;
  int kchunk;
};

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_f32(GemmArgs g) {
  __shared__ float As[16][64 + 1];
  __shared__ float Bs[16][64 + 1];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kb = g.part ? (int)blockIdx.z * g.kchunk : 0;
  const int ke = g.part ? min(g.K, kb + g.kchunk) : g.K;
  float acc[4][4] = {};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    for (int t = threadIdx.x; t < 16 * 64; t += 256) {
      const int kk = t / 64, mm = t % 64;
      const int m = m0 + mm, k = k0 + kk;
      float a = 0.f;
      if (m < g.M && k < ke) a = TA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
      As[kk][mm] = a;
      const int n = n0 + mm;
      float b = 0.f;
      if (n < g.N && k < ke) b = TB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
      Bs[kk][mm] = b;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  if (g.part) {
    float* P = g.part + (size_t)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + ty * 4 + i;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = n0 + tx * 4 + j;
        if (n < g.N) P[(size_t)m * g.N + n] = g.alpha * acc[i][j];
      }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + ty * 4 + i;
    if (m >= g.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + tx * 4 + j;
      if (n >= g.N) continue;
      float v = g.alpha * acc[i][j];
      if (g.bias) v += g.bias[n];
      if (g.beta != 0.f) v += g.beta * g.C[(size_t)m * g.ldc + n];
      if (g.Z) g.Z[(size_t)m * g.ldc + n] = v;
      if (g.act == kActTanh) v = tanhf(v);
      else if (g.act == kActIsru) v = v / sqrtf(fmaf(v, v, 1.0f));
      g.C[(size_t)m * g.ldc + n] = v;
    }
  }
}

// C[m][n] = sum over splits z = 0, 1, ... of part[z][m][n] (fixed order)
static __global__ __launch_bounds__(256) void gemm_splitk_reduce(int splits, int M, int N, const float* __restrict__ part,
                                                                float* __restrict__ C, int ldc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * N) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(size_t)z * M * N + t];
  C[(size_t)(t / N) * ldc + t % N] = s;
}

// ------------------------------------------------------- FP16 MFMA forward
// Weight fragments, staged once per block in LDS:
//   W1: [8 rt][64 lanes][8]      (K = 32, inputs 0..nin-1, zero padded)
//   W2, W3: [8 rt][4 ks][64][8]  (K order permuted as described above)
//   W4: [4 ks][64][8]            (rows 0..nout-1 of one 16-row tile)
// bias: b1, b2, b3 (128 each), b4 (16), FP32.
constexpr int kW1Frag = 8 * 64 * 8;
constexpr int kWHFrag = 8 * 4 * 64 * 8;
constexpr int kW4Frag = 4 * 64 * 8;
constexpr int kMlpFragElems = kW1Frag + 2 * kWHFrag + kW4Frag;  // FP16 elements
constexpr int kMlpBiasElems = 128 * 3 + 16;

__host__ __device__ inline int hidden_k_feature(int ks, int g, int e) {
  return 32 * ks + 16 * (e / 4) + 4 * g + (e % 4);
}

// Convert FP32 parameters (13 -> 128 x3 -> 3 layout) into fragment order.
static __global__ __launch_bounds__(256) void mlp_pack_fragments(const float* __restrict__ P, int nin, int nout,
                                                         mlp_t* __restrict__ frag,
                                                         float* __restrict__ bias, float hscale) {
  const int W = 128;
  const float* W1 = P;
  const float* b1 = W1 + W * nin;
  const float* W2 = b1 + W;
  const float* b2 = W2 + W * W;
  const float* W3 = b2 + W;
  const float* b3 = W3 + W * W;
  const float* W4 = b3 + W;
  const float* b4 = W4 + nout * W;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < kMlpFragElems; t += gridDim.x * blockDim.x) {
    float v = 0.f;
    if (t < kW1Frag) {
      const int rt = t / 512, l = (t / 8) % 64, e = t % 8;
      const int row = 16 * rt + l % 16, k = 8 * (l / 16) + e;
      v = k < nin ? W1[row * nin + k] : 0.f;
    } else if (t < kW1Frag + 2 * kWHFrag) {
      const int u = (t - kW1Frag) % kWHFrag;
      const float* Wm = (t - kW1Frag) < kWHFrag ? W2 : W3;
      const int rt = u / 2048, ks = (u / 512) % 4, l = (u / 8) % 64, e = u % 8;
      v = Wm[(16 * rt + l % 16) * W + hidden_k_feature(ks, l / 16, e)];
    } else {
      const int u = t - kW1Frag - 2 * kWHFrag;
      const int ks = u / 512, l = (u / 8) % 64, e = u % 8;
      const int row = l % 16;
      v = row < nout ? W4[row * W + hidden_k_feature(ks, l / 16, e)] : 0.f;
    }
    frag[t] = (mlp_t)(t < kW1Frag + 2 * kWHFrag ? v * hscale : v);
  }
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < kMlpBiasElems; t += gridDim.x * blockDim.x) {
    float v;
    if (t < 128) v = b1[t] * hscale;
    else if (t < 256) v = b2[t - 128] * hscale;
    else if (t < 384) v = b3[t - 256] * hscale;
    else v = (t - 384) < nout ? b4[t - 384] : 0.f;
    bias[t] = v;
  }
}

// One wavefront evaluates NCT x 16 samples.  in8[ct] holds this lane's 8 input
// features (features 8 g .. 8 g + 7 of sample 16 ct + l % 16, g = l / 16);
// on return out[ct][0..2] are the outputs of that sample in lanes 0..15
// (lanes 16..63 hold padding rows).  Weights/biases in LDS (fragment order);
// sbias must be 16-byte aligned (it is read as float4 rows).
template <int NCT, int ACT = kActTanh>
__device__ __forceinline__ void mlp_wave_forward(const mlpx8_t in8[NCT], const mlp_t* __restrict__ sfrag,
                                                 const float* __restrict__ sbias, float out[NCT][4]) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const mlpx8_t* F1 = reinterpret_cast<const mlpx8_t*>(sfrag);
  const mlpx8_t* F2 = reinterpret_cast<const mlpx8_t*>(sfrag + kW1Frag);
  const mlpx8_t* F3 = reinterpret_cast<const mlpx8_t*>(sfrag + kW1Frag + kWHFrag);
  const mlpx8_t* F4 = reinterpret_cast<const mlpx8_t*>(sfrag + kW1Frag + 2 * kWHFrag);
  floatx4_t acc[8][NCT];
  mlpx8_t h[4][NCT];  // B operands of the next layer: [ks][ct]

  // The accumulators start at the bias (lane group g holds rows 4 g .. 4 g + 3
  // of each 16-row tile), so the MFMA adds it and the activation pass needs no
  // separate add: a fifth of the vector instructions of a forward pass.
  auto bias4 = [&](const float* b, int rt) -> floatx4_t {
    return *reinterpret_cast<const floatx4_t*>(b + 16 * rt + 4 * g);
  };
  // layer 1: K = 32 (one step)
#pragma unroll
  for (int rt = 0; rt < 8; ++rt) {
    const mlpx8_t a = F1[rt * 64 + lane];
    const floatx4_t z = bias4(sbias, rt);
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[rt][ct] = mfma16(a, in8[ct], z);
  }
  // activation -> next B operands
  auto activate = [&]() {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) h[ks][ct] = act8<ACT>(acc[2 * ks][ct], acc[2 * ks + 1][ct]);
  };
  activate();
  // layers 2, 3: 128 x 128
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    const mlpx8_t* F = layer == 0 ? F2 : F3;
    const float* bl = sbias + 128 * (layer + 1);
#pragma unroll
    for (int rt = 0; rt < 8; ++rt) {
      const floatx4_t z = bias4(bl, rt);
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) acc[rt][ct] = z;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const mlpx8_t a = F[(rt * 4 + ks) * 64 + lane];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[rt][ct] = mfma16(a, h[ks][ct], acc[rt][ct]);
      }
    }
    activate();
  }
  // output layer: one 16-row tile, K = 128
  floatx4_t o[NCT];
  const floatx4_t bo = bias4(sbias + 384, 0);
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) o[ct] = bo;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const mlpx8_t a = F4[ks * 64 + lane];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) o[ct] = mfma16(a, h[ks][ct], o[ct]);
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
    for (int v = 0; v < 4; ++v) out[ct][v] = o[ct][v];
}

}  // namespace lpm
