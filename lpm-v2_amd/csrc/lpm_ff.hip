// lpm_ff.hip -- sphere BVE with a learned far field.
//
// Sources are grouped in coarse cells (the particles of a coarse mesh face).
// For each target cell t and source cell s, with cell centroids c, angular
// radii rho and the angle a(t, s) between their centroid directions:
//   near   a < r_near + rho_t + rho_s : direct FP32-compensated sum over s's sources
//   far    otherwise                  : analytic expansion of s to order P (2..4)
//          in its moments M0..MP of w = zeta A about c_s (P = 4: 35 numbers)
//   band   far and a < r_band + rho_t + rho_s : far + the MLP's correction of the
//          expansion's remainder (FP16 MFMA), trained on direct sums
// Cell membership is Lagrangian (fixed at set-up); centroids, radii and
// moments follow the particles and are recomputed at every evaluation.
//
// Velocity assembly, per target x (on the sphere of radius R):
//   u = V x x,  V = sum_near s_ij (x - y_j) + sum_far G_s(x) / (4 pi R)
// with s_ij = -zeta_j A_j / (4 pi R (R^2 - x.y_j)); the parts of V along x drop
// out of V x x.  G_s(x) = sum_j w_j y_j / (R^2 - x.y_j) expanded about c: with
// phi = 1 / D, D = R^2 - x.c (evaluated as (|x-c|^2 + R^2 - |c|^2) / 2) and
// 1 / (D - x.d) = phi sum_k (phi x.d)^k for y = c + d,
//   G_s = c alpha + beta,  alpha = sum_{k=0..P} phi^(k+1) M_k[x,..,x],
//                          beta  = sum_{k=1..P} phi^k M_k[., x,..,x]
// (M_k[., x^(k-1)] the k-th moment tensor contracted k-1 times with x).
//
// Cells are fixed regions: a particle belongs to the cell whose generator (the
// set-up centroid direction of a coarse mesh face) is nearest, and particles
// are binned anew (stable radix sort by cell) at every evaluation, so cells
// keep their size however the flow strains the particles.
//
// The network works in a frame of its cell: e3 along the centroid c, e1 towards
// the cell's anchor direction (fixed at set-up: its source farthest from c),
// e2 = e3 x e1.  Cells of one mesh level then look alike to it, which fixes the
// orientation of the cell's odd moments that the 10 moments leave open.
// Inputs (13): Q (x - c) / l, M0 / W, Q M1 / (W l), Q M2 Q^T / (W l^2) (6), with
// Q the frame rows, l = 2 R sin(rho / 2) and W = sum |w|.  Output o (3, cell
// frame): the expansion's remainder of G_s / (4 pi R) is Q^T o s_out / r^(P+2)
// with r = |x - c| / l and s_out = W / (4 pi R l^2), a scaling under which the
// remainder (~ W l^(P+1) / |x - c|^(P+3)) is O(1) at every distance in the band.
#include <math.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "lpm_internal.hpp"
#include "lpm_mlp.hpp"
#include "mlp_kernels.hpp"
#include "pair_kernels.hpp"

using lpm::DevBuf;
using lpm::fail;

namespace {

constexpr double kPi = 3.14159265358979323846;
// targets per lane in the near / far kernels.  Wave tiles are per target cell
// (~1,536 targets at L8 with 1,280 cells), so the last tile of a cell is
// partly empty: 96 % of lanes are busy at 2 targets per lane against 92 % at
// 4, and the expansion-mode evaluation at L8 took 23.6 ms against 25.1 ms.
constexpr int kR = 2;
constexpr int kKT = 256;     // sources per compensated tile
constexpr int kNMom = 35;    // moments M0..M4: 1 + 3 + 6 + 10 + 15
constexpr int kCellRec = 40; // floats per far-field cell record
constexpr int kMlpRec = 32;  // floats per MLP cell record
constexpr int kMlpNct = 2;   // 16-target column tiles per network pass
constexpr int kMlpTile = 16 * kMlpNct;  // targets per network work item
constexpr int kLocR = 2;  // particles per lane in the binning kernel
// The near and far kernels split their work (near lists, far record lists) in
// up to kSplitMax parts, enough for about kSplitRounds rounds of resident waves.
constexpr double kSplitRounds = 6.0;
constexpr int kSplitMax = 8;

// Symmetric moment tensors are stored by their sorted index tuples in
// lexicographic order: M2 {00 01 02 11 12 22}, M3 {000 001 002 011 012 022 111
// 112 122 222}, M4 {0000 0001 ... 2222}; moment k starts at kMomOff[k].
constexpr int kMomOff[5] = {0, 1, 4, 10, 20};
__host__ __device__ constexpr int sym2(int a, int b) {
  return a <= b ? a * (5 - a) / 2 + b : b * (5 - b) / 2 + a;
}
__host__ __device__ constexpr int sym3(int a, int b, int c) {
  // sort a <= b <= c
  return a > b ? sym3(b, a, c) : (b > c ? sym3(a, c, b) : (a == 0 ? 0 : (a == 1 ? 6 : 9)) + sym2(b, c) - sym2(a, a));
}
__host__ __device__ constexpr int sym4(int a, int b, int c, int d) {
  return a > b ? sym4(b, a, c, d)
               : (b > c ? sym4(a, c, b, d)
                        : (c > d ? sym4(a, b, d, c)
                                 : (a == 0 ? 0 : (a == 1 ? 10 : 14)) + sym3(b, c, d) - sym3(a, a, a)));
}
// unique index tuples of M2, M3, M4 and their multiplicities in a full contraction
constexpr int kPair[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
constexpr int kTri[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                             {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
constexpr int kQuad[15][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2},
                              {0, 0, 2, 2}, {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 2, 2},
                              {1, 1, 1, 1}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}, {2, 2, 2, 2}};
constexpr float kPairMul[6] = {1, 2, 2, 1, 2, 1};
constexpr float kTriMul[10] = {1, 3, 3, 3, 6, 3, 1, 3, 3, 1};

// Position in m[] of M_k[i, rest] for each free index i and unique tuple `rest`
// of the remaining k-1 indices, tabulated at compile time so that the unrolled
// contractions index the moments with constants.
struct MomIndex {
  int b2[3][3];
  int b3[3][6];
  int b4[3][10];
};
constexpr MomIndex make_mom_index() {
  MomIndex t{};
  for (int k = 0; k < 3; ++k) {
    for (int j = 0; j < 3; ++j) t.b2[k][j] = kMomOff[2] + sym2(k, j);
    for (int p = 0; p < 6; ++p) t.b3[k][p] = kMomOff[3] + sym3(k, kPair[p][0], kPair[p][1]);
    for (int q = 0; q < 10; ++q) t.b4[k][q] = kMomOff[4] + sym4(k, kTri[q][0], kTri[q][1], kTri[q][2]);
  }
  return t;
}
constexpr MomIndex kMomIdx = make_mom_index();

// Expansion of G = sum_j w_j y_j / (R^2 - x.y_j) to order P about c from the
// moments m[35] (T = float: the far kernel, m already divided by 4 pi R;
// T = double: training targets).  D = (|x-c|^2 + e_c) / 2, e_c = R^2 - |c|^2.
__device__ __forceinline__ float recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double recip(double v) { return 1.0 / v; }

template <int P, typename T>
__device__ __forceinline__ void expansion(const T c[3], T e_c, const T* m, const T x[3], T G[3]) {
  const T dx = x[0] - c[0], dy = x[1] - c[1], dz = x[2] - c[2];
  const T D = T(0.5) * (dx * dx + dy * dy + dz * dz + e_c);
  const T phi = recip(D);
  // b_k = M_k[., x^(k-1)], a_k = x . b_k
  T b[5][3] = {};
  T a[5] = {};
  a[0] = m[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) b[1][k] = m[1 + k];
  if (P >= 2) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T s = 0;
#pragma unroll
      for (int j = 0; j < 3; ++j) s += m[kMomIdx.b2[k][j]] * x[j];
      b[2][k] = s;
    }
  }
  if (P >= 3) {
    T q[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) q[p] = T(kPairMul[p]) * x[kPair[p][0]] * x[kPair[p][1]];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T s = 0;
#pragma unroll
      for (int p = 0; p < 6; ++p) s += m[kMomIdx.b3[k][p]] * q[p];
      b[3][k] = s;
    }
  }
  if (P >= 4) {
    T q[10];
#pragma unroll
    for (int t = 0; t < 10; ++t) q[t] = T(kTriMul[t]) * x[kTri[t][0]] * x[kTri[t][1]] * x[kTri[t][2]];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      T s = 0;
#pragma unroll
      for (int t = 0; t < 10; ++t) s += m[kMomIdx.b4[k][t]] * q[t];
      b[4][k] = s;
    }
  }
#pragma unroll
  for (int k = 1; k <= P; ++k) a[k] = x[0] * b[k][0] + x[1] * b[k][1] + x[2] * b[k][2];
  // Horner in phi
  T al = a[P];
#pragma unroll
  for (int k = P - 1; k >= 0; --k) al = a[k] + phi * al;
  al = al * phi;
  T be[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    T s = b[P][i];
#pragma unroll
    for (int k = P - 1; k >= 1; --k) s = b[k][i] + phi * s;
    be[i] = s * phi;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) G[i] = c[i] * al + be[i];
}

// The far kernel's FP32 form of expansion<P>, accumulated into p.  Two
// identities shorten it:
//   a_k = x . b_k (k >= 1), so alpha = phi (M_0 + x . beta) needs no a_k, and
//   phi = 2 psi with psi = 1 / (|x-c|^2 + e_c),
// so with the record's moments scaled by 2^k (M'_k = 2^k M_k / (4 pi R)) and
// B_k = M'_k[., x^(k-1)]:
//   beta = psi h,  h = B_1 + psi (B_2 + ... + psi B_P)
//   alpha = psi (M'_0 + psi (2x) . h)
//   p += c alpha + psi h.
// cv = {c (3), e_c, M'_0..M'_P}; x2 = 2x.  84 vector instructions per target
// and cell at P = 4 against 102 for expansion<4, float> and the sum into p.
template <int P>
__device__ __forceinline__ void far_expansion_acc(const float* cv, const float x[3], const float x2[3], float p[3]) {
  const float dx = x[0] - cv[0], dy = x[1] - cv[1], dz = x[2] - cv[2];
  const float psi = recip(fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, cv[3]))));
  const float* m = cv + 4;
  float B[5][3] = {};
#pragma unroll
  for (int k = 0; k < 3; ++k) B[1][k] = m[1 + k];
  if (P >= 2) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      B[2][k] = fmaf(m[kMomIdx.b2[k][2]], x[2], fmaf(m[kMomIdx.b2[k][1]], x[1], m[kMomIdx.b2[k][0]] * x[0]));
  }
  if (P >= 3) {
    float q[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) q[i] = kPairMul[i] * x[kPair[i][0]] * x[kPair[i][1]];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float s = m[kMomIdx.b3[k][0]] * q[0];
#pragma unroll
      for (int i = 1; i < 6; ++i) s = fmaf(m[kMomIdx.b3[k][i]], q[i], s);
      B[3][k] = s;
    }
  }
  if (P >= 4) {
    float q[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) q[i] = kTriMul[i] * x[kTri[i][0]] * x[kTri[i][1]] * x[kTri[i][2]];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float s = m[kMomIdx.b4[k][0]] * q[0];
#pragma unroll
      for (int i = 1; i < 10; ++i) s = fmaf(m[kMomIdx.b4[k][i]], q[i], s);
      B[4][k] = s;
    }
  }
  float h[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s = B[P][i];
#pragma unroll
    for (int k = P - 1; k >= 1; --k) s = fmaf(psi, s, B[k][i]);
    h[i] = s;
  }
  const float xh = fmaf(x2[0], h[0], fmaf(x2[1], h[1], x2[2] * h[2]));
  const float alpha = psi * fmaf(psi, xh, m[0]);
#pragma unroll
  for (int i = 0; i < 3; ++i) p[i] = fmaf(cv[i], alpha, fmaf(psi, h[i], p[i]));
}

#pragma clang fp contract(off)

__device__ __forceinline__ void project(double& x, double& y, double& z, double radius) {
  const double s = radius / sqrt(x * x + y * y + z * z);
  x = x * s;
  y = y * s;
  z = z * s;
}

// ---------------------------------------------------------------- binning
// Cell of each particle: the generator with the largest dot product (lowest
// index on ties); sort keys for targets (all particles) and sources (inactive
// particles get key ncell and sort last).
__global__ __launch_bounds__(256) void ff_locate(int n, int ncell, const lpm::Vec4<float>* __restrict__ gen,
                                                 const double* __restrict__ x, const double* __restrict__ y,
                                                 const double* __restrict__ z, const int* __restrict__ active,
                                                 int* __restrict__ cellof, int* __restrict__ key_t,
                                                 int* __restrict__ key_s, int* __restrict__ iota) {
  // kLocR particles per lane; generators in groups of 8, so that one wait
  // covers 8 scalar loads
  const int i0 = blockIdx.x * blockDim.x * kLocR + threadIdx.x;
  if (i0 >= n) return;
  float px[kLocR], py[kLocR], pz[kLocR], best[kLocR];
  int bs[kLocR];
#pragma unroll
  for (int r = 0; r < kLocR; ++r) {
    const int i = min(i0 + r * (int)blockDim.x, n - 1);
    px[r] = (float)x[i];
    py[r] = (float)y[i];
    pz[r] = (float)z[i];
    best[r] = -3.f;
    bs[r] = 0;
  }
  auto visit = [&](const lpm::Vec4<float>& g, int s) {
#pragma unroll
    for (int r = 0; r < kLocR; ++r) {
      const float d = fmaf(g.a, px[r], fmaf(g.b, py[r], g.c * pz[r]));
      if (d > best[r]) {  // strict: the lowest index wins a tie
        best[r] = d;
        bs[r] = s;
      }
    }
  };
  int s = 0;
  for (; s + 8 <= ncell; s += 8) {
    lpm::Vec4<float> g[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) g[e] = gen[s + e];
#pragma unroll
    for (int e = 0; e < 8; ++e) visit(g[e], s + e);
  }
  for (; s < ncell; ++s) visit(gen[s], s);
#pragma unroll
  for (int r = 0; r < kLocR; ++r) {
    const int i = i0 + r * (int)blockDim.x;
    if (i < n) {
      cellof[i] = bs[r];
      key_t[i] = bs[r];
      key_s[i] = active[i] ? bs[r] : ncell;
      iota[i] = i;
    }
  }
}

__device__ __forceinline__ int lower_bound_dev(const int* __restrict__ a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Start of every cell (and the end, index ncell) in the sorted key arrays.
__global__ __launch_bounds__(256) void ff_bounds(int n, int ncell, const int* __restrict__ tkey,
                                                 const int* __restrict__ skey, int* __restrict__ tgt_start,
                                                 int* __restrict__ src_start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell) return;
  tgt_start[c] = lower_bound_dev(tkey, n, c);
  src_start[c] = lower_bound_dev(skey, n, c);
}

// Work items (cell, tile) for tiles of tileA and of tileB targets, in cell
// order; counts[0], counts[1] = their numbers.  One block of 1024 threads.
__global__ __launch_bounds__(1024) void ff_make_items(int ncell, const int* __restrict__ tgt_start, int tileA,
                                                      int tileB, int2* __restrict__ items_a,
                                                      int2* __restrict__ items_b, int* __restrict__ counts) {
  __shared__ int sa[1024], sb[1024];
  const int tid = threadIdx.x;
  const int per = (ncell + 1023) / 1024;
  const int c0 = min(ncell, tid * per), c1 = min(ncell, c0 + per);
  int na = 0, nb = 0;
  for (int c = c0; c < c1; ++c) {
    const int cnt = tgt_start[c + 1] - tgt_start[c];
    na += (cnt + tileA - 1) / tileA;
    nb += (cnt + tileB - 1) / tileB;
  }
  sa[tid] = na;
  sb[tid] = nb;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const int va = tid >= off ? sa[tid - off] : 0, vb = tid >= off ? sb[tid - off] : 0;
    __syncthreads();
    sa[tid] += va;
    sb[tid] += vb;
    __syncthreads();
  }
  int oa = sa[tid] - na, ob = sb[tid] - nb;
  for (int c = c0; c < c1; ++c) {
    const int cnt = tgt_start[c + 1] - tgt_start[c];
    for (int j = 0; j * tileA < cnt; ++j) items_a[oa++] = make_int2(c, j);
    for (int j = 0; j * tileB < cnt; ++j) items_b[ob++] = make_int2(c, j);
  }
  if (tid == 1023) {
    counts[0] = sa[1023];
    counts[1] = sb[1023];
  }
}

// ---------------------------------------------------------------- packing
// Targets in cell order: FP32 {x, y, z, 0}.
__global__ __launch_bounds__(256) void ff_pack_targets(int n, const int* __restrict__ perm, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       lpm::Vec4<float>* __restrict__ rec) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int p = perm[k];
  rec[k] = lpm::Vec4<float>{(float)x[p], (float)y[p], (float)z[p], 0.f};
}

// Per cell: FP64 centroids/radii of its sources and of its targets, the 10
// moments of w = zeta A about the source centroid, sum |w|; writes the far
// record (moments / (4 pi R)), the MLP record and FP64 geometry.
struct GeomArgs {
  int ncell;
  const int *src_perm, *src_start, *tgt_perm, *tgt_start;
  const double *x, *y, *z, *q, *area;
  double radius;
  float* far_rec;      // [ncell][40]: c(3), e_c, M0..M4 (35) / (4 pi R) with M0 scaled by 2 and Mk by 2^k, rho
  float* mlp_rec;      // [ncell][32]: c(3), 1/l, Q(9), s_out, pad(2), features 3..12, pad(6)
  const double* adir;  // [ncell][3] unit anchor direction of each cell's frame
  double* geo;         // [ncell][8]: src dir (3), src rho, tgt dir (3), tgt rho
  double* cen64;       // [ncell][3] FP64 source centroids
  double* mom64;       // [ncell][kNMom]
};

__device__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ double block_max(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Area-weighted centroid c of the sources [s0, s1) (cell order), their
// moments M0..M4 of w = zeta A about c, sum |w| and angular radius about c's
// direction.  Every thread of the block gets the block's values.
struct SourceMoments {
  double cx, cy, cz, ux, uy, uz, rho, wabs;
  double m[kNMom];
};

__device__ void source_moments(const int* __restrict__ src_perm, const double* __restrict__ x,
                               const double* __restrict__ y, const double* __restrict__ z,
                               const double* __restrict__ q, const double* __restrict__ area, int s0, int s1,
                               double* sh, SourceMoments& r) {
  double ax = 0, ay = 0, az = 0, aa = 0;
  for (int j = s0 + threadIdx.x; j < s1; j += 256) {
    const int p = src_perm[j];
    const double w = area[p];
    ax += w * x[p];
    ay += w * y[p];
    az += w * z[p];
    aa += w;
  }
  ax = block_sum(ax, sh);
  ay = block_sum(ay, sh);
  az = block_sum(az, sh);
  aa = block_sum(aa, sh);
  const double cx = aa > 0 ? ax / aa : 0.0, cy = aa > 0 ? ay / aa : 0.0, cz = aa > 0 ? az / aa : 1.0;
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  const double ux = cx / cn, uy = cy / cn, uz = cz / cn;
  double* m = r.m;
  for (int k = 0; k < kNMom; ++k) m[k] = 0.0;
  double wabs = 0.0, rho = 0.0;
  for (int j = s0 + threadIdx.x; j < s1; j += 256) {
    const int p = src_perm[j];
    const double w = q[p] * area[p];
    const double d[3] = {x[p] - cx, y[p] - cy, z[p] - cz};
    m[0] += w;
#pragma unroll
    for (int i = 0; i < 3; ++i) m[1 + i] += w * d[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) m[kMomOff[2] + i] += w * d[kPair[i][0]] * d[kPair[i][1]];
#pragma unroll
    for (int i = 0; i < 10; ++i) m[kMomOff[3] + i] += w * d[kTri[i][0]] * d[kTri[i][1]] * d[kTri[i][2]];
#pragma unroll
    for (int i = 0; i < 15; ++i)
      m[kMomOff[4] + i] += w * d[kQuad[i][0]] * d[kQuad[i][1]] * d[kQuad[i][2]] * d[kQuad[i][3]];
    wabs += fabs(w);
    const double rr = sqrt(x[p] * x[p] + y[p] * y[p] + z[p] * z[p]);
    const double cosang = (x[p] * ux + y[p] * uy + z[p] * uz) / rr;
    rho = fmax(rho, acos(fmin(1.0, fmax(-1.0, cosang))));
  }
#pragma unroll
  for (int k = 0; k < kNMom; ++k) m[k] = block_sum(m[k], sh);
  r.wabs = block_sum(wabs, sh);
  r.rho = block_max(rho, sh);
  r.cx = cx;
  r.cy = cy;
  r.cz = cz;
  r.ux = ux;
  r.uy = uy;
  r.uz = uz;
}

// The FP32 far record of far_expansion_acc: c, e_c = R^2 - |c|^2, the moments
// / (4 pi R) with M_0 and M_1 scaled by 2, M_2 by 4, M_3 by 8, M_4 by 16, rho.
__device__ void write_far_record(float* fr, const SourceMoments& sm, double R) {
  const double inv = 1.0 / (4.0 * kPi * R);
  fr[0] = (float)sm.cx;
  fr[1] = (float)sm.cy;
  fr[2] = (float)sm.cz;
  fr[3] = (float)(R * R - (sm.cx * sm.cx + sm.cy * sm.cy + sm.cz * sm.cz));
  for (int k = 0; k < kNMom; ++k)
    fr[4 + k] = (float)(sm.m[k] * inv * (k < 4 ? 2.0 : k < 10 ? 4.0 : k < 20 ? 8.0 : 16.0));
  fr[4 + kNMom] = (float)sm.rho;
}

__global__ __launch_bounds__(256) void ff_cell_geometry(GeomArgs a) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  const int s0 = a.src_start[c], s1 = a.src_start[c + 1];
  const int t0 = a.tgt_start[c], t1 = a.tgt_start[c + 1];
  SourceMoments sm;
  source_moments(a.src_perm, a.x, a.y, a.z, a.q, a.area, s0, s1, sh, sm);
  const double cx = sm.cx, cy = sm.cy, cz = sm.cz, ux = sm.ux, uy = sm.uy, uz = sm.uz;
  const double* m = sm.m;
  const double wabs = sm.wabs, rho = sm.rho;
  // target cell direction and radius
  double tx = 0, ty = 0, tz = 0;
  for (int k = t0 + threadIdx.x; k < t1; k += 256) {
    const int p = a.tgt_perm[k];
    tx += a.x[p];
    ty += a.y[p];
    tz += a.z[p];
  }
  tx = block_sum(tx, sh);
  ty = block_sum(ty, sh);
  tz = block_sum(tz, sh);
  double tn = sqrt(tx * tx + ty * ty + tz * tz);
  if (!(tn > 0)) {
    tx = ux;
    ty = uy;
    tz = uz;
    tn = 1.0;
  }
  const double vx = tx / tn, vy = ty / tn, vz = tz / tn;
  double trho = 0.0;
  for (int k = t0 + threadIdx.x; k < t1; k += 256) {
    const int p = a.tgt_perm[k];
    const double r = sqrt(a.x[p] * a.x[p] + a.y[p] * a.y[p] + a.z[p] * a.z[p]);
    const double cosang = (a.x[p] * vx + a.y[p] * vy + a.z[p] * vz) / r;
    trho = fmax(trho, acos(fmin(1.0, fmax(-1.0, cosang))));
  }
  trho = block_max(trho, sh);
  if (threadIdx.x == 0) {
    const double R = a.radius;
    const double inv = 1.0 / (4.0 * kPi * R);
    write_far_record(a.far_rec + (size_t)c * kCellRec, sm, R);
    const double l = fmax(2.0 * R * sin(0.5 * rho), 1e-12);
    const double W = wabs + 1e-300;
    // cell frame: e3 = c/|c|, e1 = tangential part of (R anchor - c), e2 = e3 x e1
    double Q[3][3];
    Q[2][0] = ux;
    Q[2][1] = uy;
    Q[2][2] = uz;
    {
      const double* ad = a.adir + (size_t)c * 3;
      double ex = R * ad[0] - cx, ey = R * ad[1] - cy, ez = R * ad[2] - cz;
      const double dn = ex * ux + ey * uy + ez * uz;
      ex -= dn * ux;
      ey -= dn * uy;
      ez -= dn * uz;
      double en = sqrt(ex * ex + ey * ey + ez * ez);
      if (!(en > 1e-300)) {  // one-source cell: any tangent direction
        if (fabs(ux) < 0.9) {
          ex = 0.0; ey = -uz; ez = uy;
        } else {
          ex = uz; ey = 0.0; ez = -ux;
        }
        en = sqrt(ex * ex + ey * ey + ez * ez);
      }
      Q[0][0] = ex / en;
      Q[0][1] = ey / en;
      Q[0][2] = ez / en;
    }
    Q[1][0] = Q[2][1] * Q[0][2] - Q[2][2] * Q[0][1];
    Q[1][1] = Q[2][2] * Q[0][0] - Q[2][0] * Q[0][2];
    Q[1][2] = Q[2][0] * Q[0][1] - Q[2][1] * Q[0][0];
    const double S[3][3] = {{m[4], m[5], m[6]}, {m[5], m[7], m[8]}, {m[6], m[8], m[9]}};
    double QS[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) QS[i][j] = Q[i][0] * S[0][j] + Q[i][1] * S[1][j] + Q[i][2] * S[2][j];
    float* mr = a.mlp_rec + (size_t)c * kMlpRec;
    for (int k = 0; k < kMlpRec; ++k) mr[k] = 0.f;
    mr[0] = (float)cx;
    mr[1] = (float)cy;
    mr[2] = (float)cz;
    mr[3] = (float)(1.0 / l);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) mr[4 + 3 * i + j] = (float)Q[i][j];
    mr[13] = (float)(W * inv / (l * l));  // output scale s_out
    mr[16] = (float)(m[0] / W);
    for (int i = 0; i < 3; ++i) mr[17 + i] = (float)((Q[i][0] * m[1] + Q[i][1] * m[2] + Q[i][2] * m[3]) / (W * l));
    const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int k = 0; k < 6; ++k) {
      const int i = pr[k][0], j = pr[k][1];
      mr[20 + k] = (float)((QS[i][0] * Q[j][0] + QS[i][1] * Q[j][1] + QS[i][2] * Q[j][2]) / (W * l * l));
    }
    double* g = a.geo + (size_t)c * 8;
    g[0] = ux;
    g[1] = uy;
    g[2] = uz;
    g[3] = rho;
    g[4] = vx;
    g[5] = vy;
    g[6] = vz;
    g[7] = trho;
    a.cen64[(size_t)c * 3 + 0] = cx;
    a.cen64[(size_t)c * 3 + 1] = cy;
    a.cen64[(size_t)c * 3 + 2] = cz;
    for (int k = 0; k < kNMom; ++k) a.mom64[(size_t)c * kNMom + k] = m[k];
  }
}

// Coarse level: parent p is the four cells 4p .. 4p+3, whose sources are
// contiguous in cell order.  Its far record (after the cells' records) and its
// geometry {direction of c (3), rho}.
struct ParentArgs {
  const int *src_perm, *src_start;
  const double *x, *y, *z, *q, *area;
  double radius;
  float* far_rec;  // [npar][kCellRec]
  double* pgeo;    // [npar][4]
};

__global__ __launch_bounds__(256) void ff_parent_geometry(ParentArgs a) {
  __shared__ double sh[256];
  const int p = blockIdx.x;
  SourceMoments sm;
  source_moments(a.src_perm, a.x, a.y, a.z, a.q, a.area, a.src_start[4 * p], a.src_start[4 * p + 4], sh, sm);
  if (threadIdx.x == 0) {
    write_far_record(a.far_rec + (size_t)p * kCellRec, sm, a.radius);
    double* g = a.pgeo + (size_t)p * 4;
    g[0] = sm.ux;
    g[1] = sm.uy;
    g[2] = sm.uz;
    g[3] = sm.rho;
  }
}

// Class of source cell s for target cell t (0 far, 1 band, 2 near) from
// their FP64 geometry rows (geo: src dir (3), src rho, tgt dir (3), tgt rho).
__device__ __forceinline__ int cell_class(const double* gt, const double* gs, double r_near, double r_band) {
  const double cosang = gt[4] * gs[0] + gt[5] * gs[1] + gt[6] * gs[2];
  const double ang = acos(fmin(1.0, fmax(-1.0, cosang)));
  const double pad = gt[7] + gs[3];
  return ang < r_near + pad ? 2 : (ang < r_band + pad ? 1 : 0);
}

// Per target cell t: class of every source cell (0 far, 1 band, 2 near) and
// compact lists of near and band cells.
struct ClassArgs {
  int ncell;
  const double* geo;
  double r_near, r_band;
  int* near_list;      // [ncell][ncell]
  int* near_count;
  int* band_list;      // [ncell][ncell]
  int* band_count;
};

__global__ __launch_bounds__(256) void ff_classify(ClassArgs a) {
  __shared__ int nn, nb;
  const int t = blockIdx.x;
  if (threadIdx.x == 0) {
    nn = 0;
    nb = 0;
  }
  __syncthreads();
  const double* gt = a.geo + (size_t)t * 8;
  for (int base = 0; base < a.ncell; base += 256) {
    const int s = base + threadIdx.x;
    int cls = -1;
    if (s < a.ncell) cls = cell_class(gt, a.geo + (size_t)s * 8, a.r_near, a.r_band);
    // ordered compaction: one wave at a time keeps lists sorted by s
    for (int w = 0; w < 4; ++w) {
      if ((int)(threadIdx.x >> 6) == w) {
        const unsigned long long mn = __ballot(cls == 2), mb = __ballot(cls == 1);
        const int lane = threadIdx.x & 63;
        const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
        if (cls == 2) a.near_list[(size_t)t * a.ncell + nn + __popcll(mn & below)] = s;
        if (cls == 1) a.band_list[(size_t)t * a.ncell + nb + __popcll(mb & below)] = s;
        if (lane == 0) {
          nn += __popcll(mn);
          nb += __popcll(mb);
        }
      }
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) {
    a.near_count[t] = nn;
    a.band_count[t] = nb;
  }
}

// Per target cell t: the far kernel's list of expansion records, in order.
// Cells are visited in groups (4 siblings with the coarse level, else 1); a
// group of 4 far cells whose parent p is far enough, angle >= r_par + rho_t +
// rho_p, contributes the parent's record ncell + p, else every one of its
// cells that is not near contributes its own.
struct FarListArgs {
  int ncell, group;  // group 4 (coarse level) or 1
  const double* geo;
  const double* pgeo;
  double r_near, r_band, r_par;
  int* far_list;   // [ncell][ncell]
  int* far_count;
};

__global__ __launch_bounds__(256) void ff_far_list(FarListArgs a) {
  __shared__ int sc[256];
  __shared__ int base;
  const int t = blockIdx.x;
  const double* gt = a.geo + (size_t)t * 8;
  if (threadIdx.x == 0) base = 0;
  const int ng = a.ncell / a.group;
  for (int g0 = 0; g0 < ng; g0 += 256) {
    const int g = g0 + threadIdx.x;
    int ent[4];
    int n = 0;
    if (g < ng) {
      int cls[4];
      bool all_far = true;
      for (int k = 0; k < a.group; ++k) {
        cls[k] = cell_class(gt, a.geo + (size_t)(g * a.group + k) * 8, a.r_near, a.r_band);
        all_far = all_far && cls[k] == 0;
      }
      bool parent = false;
      if (a.group == 4 && all_far) {
        const double* gp = a.pgeo + (size_t)g * 4;
        const double cosang = gt[4] * gp[0] + gt[5] * gp[1] + gt[6] * gp[2];
        parent = acos(fmin(1.0, fmax(-1.0, cosang))) >= a.r_par + gt[7] + gp[3];
      }
      if (parent) {
        ent[n++] = a.ncell + g;
      } else {
        for (int k = 0; k < a.group; ++k)
          if (cls[k] != 2) ent[n++] = g * a.group + k;
      }
    }
    // ordered placement: inclusive scan of the counts over the block
    sc[threadIdx.x] = n;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
      const int v = threadIdx.x >= (unsigned)d ? sc[threadIdx.x - d] : 0;
      __syncthreads();
      sc[threadIdx.x] += v;
      __syncthreads();
    }
    int* out = a.far_list + (size_t)t * a.ncell + base + sc[threadIdx.x] - n;
    for (int i = 0; i < n; ++i) out[i] = ent[i];
    __syncthreads();
    if (threadIdx.x == 255) base += sc[255];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.far_count[t] = base;
}

// --------------------------------------------------------- near field
struct NearArgs {
  const lpm::Vec4<float>* tgt;  // cell-ordered targets
  const int* tgt_perm;          // particle index of each ordered target
  const int* tgt_start;
  const float* src;             // cell-ordered source records (SREC)
  const int* src_start;
  const int* near_list;
  const int* near_count;
  const int2* items;            // (target cell, wave tile)
  const int* n_items;  // device count
  int ncell;
  double* out;                  // [split][3][nt] FP64, cell order
  int nt;
  int nsplit;                   // near lists are split in nsplit parts (grid y)
};

__global__ __launch_bounds__(256) void ff_near_kernel(NearArgs a) {
  using P = lpm::SphereBS<float, false>;
  // the wave's item index is wave-uniform: readfirstlane puts it (and every
  // index and record address derived from it) in scalar registers
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (item >= *a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int2 it = a.items[item];
  const int t = it.x;
  const int k0 = a.tgt_start[t] + it.y * 64 * kR;
  const int k_end = a.tgt_start[t + 1];
  lpm::Vec4<float> tv[kR];
  int tid[kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    const int kc = k < k_end ? k : k_end - 1;
    tv[r] = a.tgt[kc];
    tid[r] = k < k_end ? a.tgt_perm[kc] : -1;
  }
  float S[kR][3], C[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[r][c] = C[r][c] = 0.f;
  // this split's share of the near list
  const int nn = a.near_count[t];
  const int split = blockIdx.y;
  const int qb = (int)((long long)nn * split / a.nsplit), qe = (int)((long long)nn * (split + 1) / a.nsplit);
  double* out = a.out + (size_t)split * 3 * a.nt;
  // particle ids of the wave's targets: ascending inside a cell (stable binning)
  const int id_lo = a.tgt_perm[k0], id_hi = a.tgt_perm[min(k0 + 64 * kR, k_end) - 1];
  for (int q = qb; q < qe; ++q) {
    const int s = __builtin_amdgcn_readfirstlane(a.near_list[(size_t)t * a.ncell + q]);
    const int j0 = a.src_start[s], j1 = a.src_start[s + 1];
    for (int jt = j0; jt < j1; jt += kKT) {
      const int je = min(j1, jt + kKT);
      // a target's own source lives in its own cell, in the source tile whose
      // (ascending) particle ids span the target's
      const bool checked = s == t && __float_as_int(a.src[(size_t)jt * lpm::SREC + P::IDX]) <= id_hi &&
                           __float_as_int(a.src[(size_t)(je - 1) * lpm::SREC + P::IDX]) >= id_lo;
      float p[kR][3];
#pragma unroll
      for (int r = 0; r < kR; ++r) p[r][0] = p[r][1] = p[r][2] = 0.f;
      if (!checked) {
        // each source's record is loaded one source ahead of its arithmetic, as
        // in the plain pair kernel
        float sv[lpm::SREC];
#pragma unroll
        for (int e = 0; e < lpm::SREC; ++e) sv[e] = a.src[(size_t)jt * lpm::SREC + e];
#pragma unroll 2
        for (int j = jt; j < je; ++j) {
          const int jn = j + 1 < je ? j + 1 : j;
          float sn[lpm::SREC];
#pragma unroll
          for (int e = 0; e < lpm::SREC; ++e) sn[e] = a.src[(size_t)jn * lpm::SREC + e];
          lpm::W4<float> w[kR];
          float f[kR];
#pragma unroll
          for (int r = 0; r < kR; ++r) P::arg(tv[r], sv, w[r]);
#pragma unroll
          for (int r = 0; r < kR; ++r) f[r] = P::trans(w[r]);
#pragma unroll
          for (int r = 0; r < kR; ++r) P::acc(sv, w[r], f[r], p[r]);
#pragma unroll
          for (int e = 0; e < lpm::SREC; ++e) sv[e] = sn[e];
        }
      } else {
        for (int j = jt; j < je; ++j) {
          const float* sr = a.src + (size_t)j * lpm::SREC;
          float sv[lpm::SREC];
#pragma unroll
          for (int e = 0; e < lpm::SREC; ++e) sv[e] = sr[e];
          const int sidx = __float_as_int(sv[P::IDX]);
#pragma unroll
          for (int r = 0; r < kR; ++r) {
            lpm::W4<float> w;
            P::arg(tv[r], sv, w);
            const float f = P::trans(w);
            float tmp[3] = {0.f, 0.f, 0.f};
            P::acc(sv, w, f, tmp);
            const bool self = sidx == tid[r];
#pragma unroll
            for (int c = 0; c < 3; ++c) p[r][c] += self ? 0.f : tmp[c];
          }
        }
      }
#pragma unroll
      for (int r = 0; r < kR; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) lpm::two_sum_acc(S[r][c], C[r][c], p[r][c]);
    }
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    if (k < k_end)
#pragma unroll
      for (int c = 0; c < 3; ++c) out[(size_t)c * a.nt + k] = (double)S[r][c] + (double)C[r][c];
  }
}

// ------------------------------------------------------- far (expansion)
// Every record of a target cell's far list (cells not near it, or their
// parents) by its expansion.  The list is split in gridDim.y contiguous
// ranges, one FP64 partial per range, so that the launch has several rounds
// of waves and its last round is short.
struct FarArgs {
  const lpm::Vec4<float>* tgt;
  const int* tgt_start;
  const float* cell;          // far records: cells, then parents
  const int* far_list;        // [ncell][ncell] record indices per target cell
  const int* far_count;
  const int2* items;
  const int* n_items;  // device count
  int ncell, nt;
  double* part;               // [gridDim.y][3][nt] this launch's partials
};

template <int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void ff_far_kernel(FarArgs a) {
  // the wave's item index is wave-uniform: readfirstlane puts it (and every
  // index and record address derived from it) in scalar registers
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (item >= *a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int2 it = a.items[item];
  const int t = it.x;
  const int k0 = a.tgt_start[t] + it.y * 64 * kR;
  const int k_end = a.tgt_start[t + 1];
  float X[kR][3], X2[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    const lpm::Vec4<float> v = a.tgt[k < k_end ? k : k_end - 1];
    X[r][0] = v.a;
    X[r][1] = v.b;
    X[r][2] = v.c;
#pragma unroll
    for (int c = 0; c < 3; ++c) X2[r][c] = 2.f * X[r][c];
  }
  float S[kR][3], C[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[r][c] = C[r][c] = 0.f;
  const int* list = a.far_list + (size_t)t * a.ncell;
  const int cnt = a.far_count[t];
  const int q_lo = (int)((long long)cnt * blockIdx.y / gridDim.y);
  const int q_hi = (int)((long long)cnt * (blockIdx.y + 1) / gridDim.y);
  // one compensated block per 64 records of this part of the list
  for (int b0 = q_lo; b0 < q_hi; b0 += 64) {
    const int b1 = min(q_hi, b0 + 64);
    float p[kR][3];
#pragma unroll
    for (int r = 0; r < kR; ++r) p[r][0] = p[r][1] = p[r][2] = 0.f;
    for (int qi = b0; qi < b1; ++qi) {
      const int s = list[qi];
      const float* cr = a.cell + (size_t)s * kCellRec;
      constexpr int NV = 4 + kMomOff[P] + (P + 1) * (P + 2) / 2;  // c, e_c, M'0..M'P
      float cv[NV];
#pragma unroll
      for (int e = 0; e < NV; ++e) cv[e] = cr[e];
#pragma unroll
      for (int r = 0; r < kR; ++r) far_expansion_acc<P>(cv, X[r], X2[r], p[r]);
    }
#pragma unroll
    for (int r = 0; r < kR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) lpm::two_sum_acc(S[r][c], C[r][c], p[r][c]);
  }
  double* out = a.part + (size_t)blockIdx.y * 3 * a.nt;
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    if (k < k_end)
#pragma unroll
      for (int c = 0; c < 3; ++c) out[(size_t)c * a.nt + k] = (double)S[r][c] + (double)C[r][c];
  }
}

// V = the near partials, then the far partials, summed in this fixed order.
__global__ __launch_bounds__(256) void ff_sum_parts(int nt, int nparts, const double* __restrict__ part,
                                                    double* __restrict__ V) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 3 * nt) return;
  double v = part[k];
  for (int sp = 1; sp < nparts; ++sp) v += part[(size_t)sp * 3 * nt + k];
  V[k] = v;
}

// --------------------------------------------------- MLP correction (MFMA)
// Target position in the cell frame, Q (x - c) / l.
__device__ __forceinline__ void mlp_local(const float* mr, float x, float y, float z, float loc[3]) {
  const float dx = (x - mr[0]) * mr[3], dy = (y - mr[1]) * mr[3], dz = (z - mr[2]) * mr[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) loc[i] = fmaf(mr[4 + 3 * i], dx, fmaf(mr[5 + 3 * i], dy, mr[6 + 3 * i] * dz));
}

// Input f of a sample: 0..2 the local position, 3..12 the cell's moment features.
__device__ __forceinline__ float mlp_feature(const float* mr, const float loc[3], int f) {
  if (f < 3) return loc[f];
  if (f < 13) return mr[16 + (f - 3)];
  return 0.f;
}

// r^(P+2) for squared radius r2 and expansion order P in 2..4
__device__ __forceinline__ float out_pow(float r2, int order) {
  const float r4 = r2 * r2;
  return order == 2 ? r4 : (order == 3 ? r4 * sqrtf(r2) : r4 * r2);
}

// s_out / r^(P+2) for the local position
__device__ __forceinline__ float mlp_out_scale(const float* mr, const float loc[3], int order) {
  const float r2 = fmaf(loc[0], loc[0], fmaf(loc[1], loc[1], loc[2] * loc[2]));
  return mr[13] * __builtin_amdgcn_rcpf(out_pow(r2, order));
}

struct MlpCorrArgs {
  const lpm::Vec4<float>* tgt;
  const int* tgt_start;
  const float* mlp_rec;
  const int* band_list;
  const int* band_count;
  const int2* items;  // (target cell, 64-target tile)
  const int* n_items;  // device count
  int ncell, nt, order;
  const lpm::mlp_t* frag;
  const float* bias;
  double* out;        // [3][nt] accumulated into
};

// 512 threads: 8 waves (2 per SIMD) share one LDS copy of the weights, which
// alone takes half a CU's LDS.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void ff_mlp_kernel(MlpCorrArgs a) {
  __shared__ __attribute__((aligned(16))) lpm::mlp_t sfrag[lpm::kMlpFragElems];
  __shared__ __attribute__((aligned(16))) float sbias[lpm::kMlpBiasElems];
  for (int t = threadIdx.x; t < lpm::kMlpFragElems / 8; t += blockDim.x)
    reinterpret_cast<lpm::mlpx8_t*>(sfrag)[t] = reinterpret_cast<const lpm::mlpx8_t*>(a.frag)[t];
  for (int t = threadIdx.x; t < lpm::kMlpBiasElems; t += blockDim.x) sbias[t] = a.bias[t];
  __syncthreads();
  constexpr int NCT = kMlpNct;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int waves = gridDim.x * (blockDim.x >> 6);
  const int n_items = *a.n_items;
  const int wave0 = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  for (int item = wave0; item < n_items; item += waves) {
    const int2 it = a.items[item];
    const int t = it.x;
    const int k0 = a.tgt_start[t] + it.y * kMlpTile;
    const int k_end = a.tgt_start[t + 1];
    float xs[NCT][3];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int k = min(k0 + 16 * ct + (lane & 15), k_end - 1);
      const lpm::Vec4<float> v = a.tgt[k];
      xs[ct][0] = v.a;
      xs[ct][1] = v.b;
      xs[ct][2] = v.c;
    }
    float acc[NCT][3];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[ct][0] = acc[ct][1] = acc[ct][2] = 0.f;
    const int nb = a.band_count[t];
    for (int q = 0; q < nb; ++q) {
      const int s = __builtin_amdgcn_readfirstlane(a.band_list[(size_t)t * a.ncell + q]);
      const float* mr = a.mlp_rec + (size_t)s * kMlpRec;
      // the cell's 10 moment features (scalar loads); lane group g holds inputs
      // 8g .. 8g+7: g = 0 the local position and features 3..7, g = 1 features
      // 8..12, g = 2, 3 the zero padding -- selected without branches
      float mv[10];
#pragma unroll
      for (int k = 0; k < 10; ++k) mv[k] = mr[16 + k];
      lpm::mlpx8_t in8[NCT];
      float loc[NCT][3];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        mlp_local(mr, xs[ct][0], xs[ct][1], xs[ct][2], loc[ct]);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float f0 = e < 3 ? loc[ct][e] : mv[e - 3];
          const float f1 = e < 5 ? mv[5 + e] : 0.f;
          in8[ct][e] = (lpm::mlp_t)(g == 0 ? f0 : (g == 1 ? f1 : 0.f));
        }
      }
      float o[NCT][4];
      lpm::mlp_wave_forward<NCT, lpm::kActIsru>(in8, sfrag, sbias, o);
      // lanes 0..15 hold the outputs: scale, rotate to world axes (Q^T o), accumulate
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const float so = mlp_out_scale(mr, loc[ct], a.order);
        const float o0 = o[ct][0] * so, o1 = o[ct][1] * so, o2 = o[ct][2] * so;
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[ct][c] = fmaf(mr[4 + c], o0, fmaf(mr[7 + c], o1, fmaf(mr[10 + c], o2, acc[ct][c])));
      }
    }
    if (g == 0) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int k = k0 + 16 * ct + lane;
        if (k < k_end)
#pragma unroll
          for (int c = 0; c < 3; ++c) a.out[(size_t)c * a.nt + k] += (double)acc[ct][c];
      }
    }
  }
}

// ---------------------------------------------------------------- epilogue
__global__ __launch_bounds__(256) void ff_epilogue(int nt, const int* __restrict__ perm, const double* __restrict__ V,
                                                   const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, double* __restrict__ u,
                                                   double* __restrict__ v, double* __restrict__ w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt) return;
  const int p = perm[k];
  const double A0 = V[k], A1 = V[(size_t)nt + k], A2 = V[(size_t)2 * nt + k];
  const double px = x[p], py = y[p], pz = z[p];
  u[p] = A1 * pz - A2 * py;
  v[p] = A2 * px - A0 * pz;
  w[p] = A0 * py - A1 * px;
}

// --------------------------------------------------- training samples
// Sample q: a random target (particle) and a random band cell of its cell.
// Inputs: the 13 MLP features; target: Q (G_exact - G_expansion) / (4 pi R) r^(P+2) / s_out,
// the remainder in the cell frame and scaled, in FP64 (G_exact over the cell's
// sources, with w = zeta A).  ids receives (particle, cell) of each sample.
struct SampleArgs {
  int n_samples;
  unsigned long long seed;
  int nt, ncell, order;
  const int *tgt_perm, *tgt_cell, *src_perm, *src_start, *band_list, *band_count;
  const double *x, *y, *z, *q, *area;
  const double *cen64, *mom64;
  const float* mlp_rec;
  double radius;
  float* X;  // [n][13]
  float* Y;  // [n][3]
  int* ids;  // [n][2]: target particle, source cell (-1, -1 for an invalid sample)
  int draws = 1;  // band cells drawn per sample; the one nearest the target is kept
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ void expansion_fp64(int order, const double* c, const double* m, double R, const double x[3],
                               double G[3]) {
  const double e_c = R * R - (c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (order == 2) expansion<2, double>(c, e_c, m, x, G);
  else if (order == 3) expansion<3, double>(c, e_c, m, x, G);
  else expansion<4, double>(c, e_c, m, x, G);
}

__global__ __launch_bounds__(256) void ff_make_samples(SampleArgs a) {
  const int qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= a.n_samples) return;
  unsigned long long h = mix64(a.seed ^ mix64((unsigned long long)qi));
  int ok = 0, pid = -1, sid = -1;
  float feat[13];
  float tgt3[3] = {0.f, 0.f, 0.f};
  for (int attempt = 0; attempt < 8 && !ok; ++attempt) {
    const int k = (int)(h % (unsigned long long)a.nt);
    h = mix64(h);
    const int t = a.tgt_cell[k];
    const int nb = a.band_count[t];
    if (nb == 0) continue;
    const int p = a.tgt_perm[k];
    const double x = a.x[p], y = a.y[p], z = a.z[p];
    // band cell: the nearest (centroid distance) of `draws` uniform draws, so
    // that the cells whose remainders dominate the error are sampled more
    int s = -1;
    double best = 0.0;
    for (int d = 0; d < a.draws; ++d) {
      const int c = a.band_list[(size_t)t * a.ncell + (int)(h % (unsigned long long)nb)];
      h = mix64(h);
      const double* cc = a.cen64 + (size_t)c * 3;
      const double d2 = (x - cc[0]) * (x - cc[0]) + (y - cc[1]) * (y - cc[1]) + (z - cc[2]) * (z - cc[2]);
      if (s < 0 || d2 < best) {
        s = c;
        best = d2;
      }
    }
    const double R = a.radius;
    double Ge[3] = {0.0, 0.0, 0.0};
    for (int j = a.src_start[s]; j < a.src_start[s + 1]; ++j) {
      const int pj = a.src_perm[j];
      const double dx = x - a.x[pj], dy = y - a.y[pj], dz = z - a.z[pj];
      const double D = 0.5 * (dx * dx + dy * dy + dz * dz +
                              (R * R - (a.x[pj] * a.x[pj] + a.y[pj] * a.y[pj] + a.z[pj] * a.z[pj])));
      const double w = a.q[pj] * a.area[pj] / D;
      Ge[0] += w * a.x[pj];
      Ge[1] += w * a.y[pj];
      Ge[2] += w * a.z[pj];
    }
    double Gm[3];
    const double* c64 = a.cen64 + (size_t)s * 3;
    const double xv[3] = {x, y, z};
    expansion_fp64(a.order, c64, a.mom64 + (size_t)s * kNMom, R, xv, Gm);
    const float* mr = a.mlp_rec + (size_t)s * kMlpRec;
    // local position in FP64 for the scaling, FP32 (as at inference) for the inputs
    const double il = (double)mr[3];
    double r2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double li = ((double)mr[4 + 3 * i] * (x - c64[0]) + (double)mr[5 + 3 * i] * (y - c64[1]) +
                         (double)mr[6 + 3 * i] * (z - c64[2])) * il;
      r2 += li * li;
    }
    const double rp = a.order == 2 ? r2 * r2 : (a.order == 3 ? r2 * r2 * sqrt(r2) : r2 * r2 * r2);
    const double inv = rp / (4.0 * kPi * R) / (double)mr[13];
    for (int i = 0; i < 3; ++i)
      tgt3[i] = (float)(((double)mr[4 + 3 * i] * (Ge[0] - Gm[0]) + (double)mr[5 + 3 * i] * (Ge[1] - Gm[1]) +
                         (double)mr[6 + 3 * i] * (Ge[2] - Gm[2])) * inv);
    float loc[3];
    mlp_local(mr, (float)x, (float)y, (float)z, loc);
    for (int f = 0; f < 13; ++f) feat[f] = mlp_feature(mr, loc, f);
    ok = 1;
    pid = p;
    sid = s;
  }
  for (int f = 0; f < 13; ++f) a.X[(size_t)qi * 13 + f] = ok ? feat[f] : 0.f;
  for (int c = 0; c < 3; ++c) a.Y[(size_t)qi * 3 + c] = tgt3[c];
  a.ids[(size_t)qi * 2] = pid;
  a.ids[(size_t)qi * 2 + 1] = sid;
}

// ---------------------------------------------------------- RK4 algebra
__global__ __launch_bounds__(256) void ff_stage(int n, double dt, double omega, double radius, double h, int first,
                                                int make_input, const double* __restrict__ x,
                                                const double* __restrict__ y, const double* __restrict__ z,
                                                const double* __restrict__ q, const double* __restrict__ u,
                                                const double* __restrict__ v, const double* __restrict__ w,
                                                double* kx, double* ky, double* kz, double* kq, double* xi,
                                                double* yi, double* zi, double* qi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double vx = first ? u[i] : kx[i], vy = first ? v[i] : ky[i], vz = first ? w[i] : kz[i];
  const double dq = -dt * 2.0 * omega * vz / radius;
  const double a = dt * vx, b = dt * vy, c = dt * vz;
  kx[i] = a;
  ky[i] = b;
  kz[i] = c;
  kq[i] = dq;
  if (make_input) {
    double px = x[i] + h * a, py = y[i] + h * b, pz = z[i] + h * c;
    project(px, py, pz, radius);
    xi[i] = px;
    yi[i] = py;
    zi[i] = pz;
    qi[i] = q[i] + h * dq;
  }
}

__global__ __launch_bounds__(256) void ff_combine(int n, double radius, double* x, double* y, double* z, double* q,
                                                  const double* const* __restrict__ k) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* st[4] = {x, y, z, q};
  double v[4];
  for (int c = 0; c < 4; ++c)
    v[c] = st[c][i] + k[c * 4 + 0][i] / 6.0 + k[c * 4 + 1][i] / 3.0 + k[c * 4 + 2][i] / 3.0 + k[c * 4 + 3][i] / 6.0;
  project(v[0], v[1], v[2], radius);
  for (int c = 0; c < 4; ++c) st[c][i] = v[c];
}

__global__ __launch_bounds__(256) void ff_project(int n, double radius, double* x, double* y, double* z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i], b = y[i], c = z[i];
  project(a, b, c, radius);
  x[i] = a;
  y[i] = b;
  z[i] = c;
}

// max_i |a_i - b_i| and max_i |b_i| over 3-vectors (two-pass, deterministic)
__global__ __launch_bounds__(256) void ff_err_partial(int n, const double* __restrict__ au, const double* __restrict__ av,
                                                      const double* __restrict__ aw, const double* __restrict__ bu,
                                                      const double* __restrict__ bv, const double* __restrict__ bw,
                                                      double* __restrict__ part) {
  __shared__ double s0[256], s1[256];
  double d = 0.0, m = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double ex = au[i] - bu[i], ey = av[i] - bv[i], ez = aw[i] - bw[i];
    d = fmax(d, sqrt(ex * ex + ey * ey + ez * ez));
    m = fmax(m, sqrt(bu[i] * bu[i] + bv[i] * bv[i] + bw[i] * bw[i]));
  }
  s0[threadIdx.x] = d;
  s1[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + k]);
      s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

}  // namespace

// ====================================================================== object
struct lpm_bve_ff {
  lpm_ctx* ctx = nullptr;
  int n = 0, ns = 0, ncell = 0;
  double radius = 1.0, omega = 0.0, r_near = 0.1, r_band = 0.45;
  int order = 4;  // expansion order of far cells
  int draws = 1;  // training samples: band cells drawn per sample (nearest kept)
  DevBuf x, y, z, q, u, v, w, area;
  DevBuf k[4][4];
  DevBuf xi, yi, zi, qi;
  DevBuf pu, pv, pw;  // plain-sum velocity (judge)
  // binning: cell-ordered particle lists rebuilt at every evaluation
  DevBuf src_perm, src_key, src_start, tgt_perm, tgt_cell, tgt_start;
  DevBuf key_t, key_s, iota, cellof, active, sort_tmp, item_count;
  DevBuf gen, adir;  // cell generator directions (float4), frame anchor directions (FP64)
  int end_bit = 0;
  size_t sort_bytes = 0;
  DevBuf trec, srec, far_rec, mlp_rec, geo, cen64, mom64;
  DevBuf near_list, near_count, band_list, band_count;
  DevBuf far_list, far_count;  // expansion records per target cell (cells or parents)
  DevBuf pgeo;                 // coarse level: parent geometry [ncell / 4][4]
  int npar = 0;                // parents in use (0: no coarse level)
  double r_par = 0.0;          // a parent replaces its 4 cells beyond this radius
  DevBuf items, mlp_items;
  DevBuf V;
  DevBuf Vpart;        // FP64 partials [near_split + far_split][3][n]: near, then far
  int near_split = 1;  // near lists split in this many parts (fills the chip)
  int far_split = 1;   // far cells split in this many ranges
  DevBuf list_plain;  // ascending active list for the plain path
  DevBuf tgt_plain, src_plain, partial_plain;
  DevBuf ptrs, err_part;
  DevBuf sx, sy, sid;  // collected training set: inputs, targets, (particle, cell)
  int n_items = 0, n_mlp_items = 0;
  int64_t n_collected = 0, cap_collected = 0;
  lpm::Mlp net;
  bool trained = false;
  double ys = 0.0;  // RMS of the training targets (0: not fixed yet)
  int online_n = 0, online_epochs = 1, online_batch = 4096, online_rounds = 0;
  double online_from = 0.0;  // fine-tune only at judged errors above online_from * tol
  double online_lr = 2e-4;
  bool loaded = false;  // false after a failed particle reload
};

namespace {

int ff_upload(lpm_ctx* ctx, DevBuf& b, const void* host, size_t bytes) {
  LPM_HIP(ctx, b.ensure(bytes));
  if (host && bytes) LPM_HIP(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  return LPM_OK;
}

// Bin the particles at (x, y, z) into cells: cell-ordered target and source
// lists (stable: ascending particle index inside a cell), cell starts, items.
int ff_bin(lpm_bve_ff* f, const double* x, const double* y, const double* z) {
  lpm_ctx* ctx = f->ctx;
  const int n = f->n, nc = f->ncell;
  hipLaunchKernelGGL(ff_locate, dim3((n + 256 * kLocR - 1) / (256 * kLocR)), dim3(256), 0, ctx->stream, n, nc,
                     f->gen.as<lpm::Vec4<float>>(), x, y, z, f->active.as<int>(), f->cellof.as<int>(),
                     f->key_t.as<int>(), f->key_s.as<int>(), f->iota.as<int>());
  size_t bytes = f->sort_bytes;
  LPM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(f->sort_tmp.p, bytes, f->key_t.as<int>(), f->tgt_cell.as<int>(),
                                                  f->iota.as<int>(), f->tgt_perm.as<int>(), n, 0, f->end_bit,
                                                  ctx->stream));
  bytes = f->sort_bytes;
  LPM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(f->sort_tmp.p, bytes, f->key_s.as<int>(), f->src_key.as<int>(),
                                                  f->iota.as<int>(), f->src_perm.as<int>(), n, 0, f->end_bit,
                                                  ctx->stream));
  hipLaunchKernelGGL(ff_bounds, dim3((nc + 1 + 255) / 256), dim3(256), 0, ctx->stream, n, nc, f->tgt_cell.as<int>(),
                     f->src_key.as<int>(), f->tgt_start.as<int>(), f->src_start.as<int>());
  hipLaunchKernelGGL(ff_make_items, dim3(1), dim3(1024), 0, ctx->stream, nc, f->tgt_start.as<int>(), 64 * kR, kMlpTile,
                     f->items.as<int2>(), f->mlp_items.as<int2>(), f->item_count.as<int>());
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

// velocity of the current/stage state by `mode`: 0 plain, 1 expansion, 2 expansion + MLP
int ff_eval(lpm_bve_ff* f, int mode, const double* x, const double* y, const double* z, const double* q, double* u,
            double* v, double* w) {
  lpm_ctx* ctx = f->ctx;
  if (mode == 0) {
    lpm::PairProblem pp;
    pp.kind = LPM_K_SPHERE_BS;
    pp.precision = LPM_PREC_FP32_KAHAN;
    pp.n = f->n;
    pp.i0 = 0;
    pp.i1 = f->n;
    pp.x = x;
    pp.y = y;
    pp.z = z;
    pp.q0 = q;
    pp.area = f->area.as<double>();
    pp.src_list = f->list_plain.as<int>();
    pp.ns = f->ns;
    pp.radius = f->radius;
    pp.on_sphere = true;
    pp.out0 = u;
    pp.out1 = v;
    pp.out2 = w;
    pp.tgt_buf = &f->tgt_plain;
    pp.src_buf = &f->src_plain;
    pp.partial_buf = &f->partial_plain;
    return lpm::run_pair_sum(ctx, pp);
  }
  const int nc = f->ncell;
  dim3 b256(256);
  if (int rc = ff_bin(f, x, y, z)) return rc;
  hipLaunchKernelGGL(ff_pack_targets, dim3((f->n + 255) / 256), b256, 0, ctx->stream, f->n, f->tgt_perm.as<int>(), x,
                     y, z, f->trec.as<lpm::Vec4<float>>());
  lpm::PackArgs pa{x, y, z, q, nullptr, f->area.as<double>(), nullptr, f->src_perm.as<int>(), 0, 0, f->radius, 1.0};
  lpm::launch_pack(LPM_K_SPHERE_BS, LPM_PREC_FP32_KAHAN, pa, f->srec.p, nullptr, f->ns, 0, ctx->stream);
  GeomArgs ga{nc, f->src_perm.as<int>(), f->src_start.as<int>(), f->tgt_perm.as<int>(), f->tgt_start.as<int>(),
              x, y, z, q, f->area.as<double>(), f->radius, f->far_rec.as<float>(), f->mlp_rec.as<float>(),
              f->adir.as<double>(), f->geo.as<double>(), f->cen64.as<double>(), f->mom64.as<double>()};
  hipLaunchKernelGGL(ff_cell_geometry, dim3(nc), b256, 0, ctx->stream, ga);
  if (f->npar) {
    ParentArgs pa2{f->src_perm.as<int>(), f->src_start.as<int>(), x, y, z, q, f->area.as<double>(), f->radius,
                   f->far_rec.as<float>() + (size_t)nc * kCellRec, f->pgeo.as<double>()};
    hipLaunchKernelGGL(ff_parent_geometry, dim3(f->npar), b256, 0, ctx->stream, pa2);
  }
  ClassArgs ca{nc, f->geo.as<double>(), f->r_near, f->r_band, f->near_list.as<int>(),
               f->near_count.as<int>(), f->band_list.as<int>(), f->band_count.as<int>()};
  hipLaunchKernelGGL(ff_classify, dim3(nc), b256, 0, ctx->stream, ca);
  FarListArgs la{nc, f->npar ? 4 : 1, f->geo.as<double>(), f->pgeo.as<double>(), f->r_near, f->r_band, f->r_par,
                 f->far_list.as<int>(), f->far_count.as<int>()};
  hipLaunchKernelGGL(ff_far_list, dim3(nc), b256, 0, ctx->stream, la);
  NearArgs na{f->trec.as<lpm::Vec4<float>>(), f->tgt_perm.as<int>(), f->tgt_start.as<int>(), f->srec.as<float>(),
              f->src_start.as<int>(), f->near_list.as<int>(), f->near_count.as<int>(), f->items.as<int2>(),
              f->item_count.as<int>(), nc, f->Vpart.as<double>(), f->n, f->near_split};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing) {
    e0 = lpm::take_event(ctx);
    e1 = lpm::take_event(ctx);
    if (e0 && e1) LPM_HIP(ctx, hipEventRecord(e0, ctx->stream));
  }
  hipLaunchKernelGGL(ff_near_kernel, dim3((f->n_items + 3) / 4, f->near_split), b256, 0, ctx->stream, na);
  FarArgs fa{f->trec.as<lpm::Vec4<float>>(), f->tgt_start.as<int>(), f->far_rec.as<float>(), f->far_list.as<int>(),
             f->far_count.as<int>(), f->items.as<int2>(), f->item_count.as<int>(), nc, f->n,
             f->Vpart.as<double>() + (size_t)f->near_split * 3 * f->n};
  const dim3 gfar((f->n_items + 3) / 4, f->far_split);
  if (f->order == 2) hipLaunchKernelGGL(ff_far_kernel<2>, gfar, b256, 0, ctx->stream, fa);
  else if (f->order == 3) hipLaunchKernelGGL(ff_far_kernel<3>, gfar, b256, 0, ctx->stream, fa);
  else hipLaunchKernelGGL(ff_far_kernel<4>, gfar, b256, 0, ctx->stream, fa);
  hipLaunchKernelGGL(ff_sum_parts, dim3((3 * f->n + 255) / 256), b256, 0, ctx->stream, f->n,
                     f->near_split + f->far_split, f->Vpart.as<double>(), f->V.as<double>());
  if (mode == 2 && f->trained) {
    MlpCorrArgs ma{f->trec.as<lpm::Vec4<float>>(), f->tgt_start.as<int>(), f->mlp_rec.as<float>(),
                   f->band_list.as<int>(), f->band_count.as<int>(), f->mlp_items.as<int2>(), f->item_count.as<int>() + 1, nc, f->n,
                   f->order,
                   f->net.frag.as<lpm::mlp_t>(), f->net.bias.as<float>(), f->V.as<double>()};
    hipLaunchKernelGGL(ff_mlp_kernel, dim3(ctx->num_cus), dim3(512), 0, ctx->stream, ma);
  }
  if (ctx->timing && e0 && e1) {
    LPM_HIP(ctx, hipEventRecord(e1, ctx->stream));
    ctx->pending.push_back({e0, e1, LPM_K_FAR_FIELD});
  }
  hipLaunchKernelGGL(ff_epilogue, dim3((f->n + 255) / 256), b256, 0, ctx->stream, f->n, f->tgt_perm.as<int>(),
                     f->V.as<double>(), x, y, z, u, v, w);
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int ff_step(lpm_bve_ff* f, double dt, int mode) {
  lpm_ctx* ctx = f->ctx;
  const int n = f->n;
  dim3 g((n + 255) / 256), b(256);
  auto K = [&](int s, int c) { return f->k[s][c].as<double>(); };
  hipLaunchKernelGGL(ff_stage, g, b, 0, ctx->stream, n, dt, f->omega, f->radius, 0.5, 1, 1, f->x.as<double>(),
                     f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->u.as<double>(), f->v.as<double>(),
                     f->w.as<double>(), K(0, 0), K(0, 1), K(0, 2), K(0, 3), f->xi.as<double>(), f->yi.as<double>(),
                     f->zi.as<double>(), f->qi.as<double>());
  for (int s = 1; s < 4; ++s) {
    if (int rc = ff_eval(f, mode, f->xi.as<double>(), f->yi.as<double>(), f->zi.as<double>(), f->qi.as<double>(),
                         K(s, 0), K(s, 1), K(s, 2)))
      return rc;
    hipLaunchKernelGGL(ff_stage, g, b, 0, ctx->stream, n, dt, f->omega, f->radius, s == 1 ? 0.5 : 1.0, 0, s < 3 ? 1 : 0,
                       f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), nullptr, nullptr,
                       nullptr, K(s, 0), K(s, 1), K(s, 2), K(s, 3), f->xi.as<double>(), f->yi.as<double>(),
                       f->zi.as<double>(), f->qi.as<double>());
  }
  hipLaunchKernelGGL(ff_combine, g, b, 0, ctx->stream, n, f->radius, f->x.as<double>(), f->y.as<double>(),
                     f->z.as<double>(), f->q.as<double>(), (const double* const*)f->ptrs.p);
  LPM_HIP(ctx, hipGetLastError());
  return ff_eval(f, mode, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->u.as<double>(),
                 f->v.as<double>(), f->w.as<double>());
}

int ff_error(lpm_bve_ff* f, const double* au, const double* av, const double* aw, const double* bu, const double* bv,
             const double* bw, double* err) {
  lpm_ctx* ctx = f->ctx;
  const int nb = 256;
  LPM_HIP(ctx, f->err_part.ensure(2 * nb * sizeof(double)));
  hipLaunchKernelGGL(ff_err_partial, dim3(nb), dim3(256), 0, ctx->stream, f->n, au, av, aw, bu, bv, bw,
                     f->err_part.as<double>());
  std::vector<double> h(2 * nb);
  LPM_HIP(ctx, hipMemcpyAsync(h.data(), f->err_part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double d = 0.0, m = 0.0;
  for (int i = 0; i < nb; ++i) {
    d = std::max(d, h[2 * i]);
    m = std::max(m, h[2 * i + 1]);
  }
  *err = m > 0 ? d / m : d;
  return LPM_OK;
}

// Load a particle set (at creation, and after a remesh): everything sized by
// the particle count -- active list, work-item bounds, near-list splits, the
// radix-sort workspace and the particle arrays (grown, never shrunk) -- then
// the plain velocity of the set (as BVESolver New) and its binning.  The
// cells, their frames, the network and the collected samples are kept.
int ff_load(lpm_bve_ff* f, int n, const double* x, const double* y, const double* z, const double* relvort,
            const double* area, const int32_t* active) {
  lpm_ctx* ctx = f->ctx;
  const int ncell = f->ncell;
  f->loaded = false;
  f->n = n;
  std::vector<int> plain;
  std::vector<int32_t> act((size_t)n);
  for (int i = 0; i < n; ++i) {
    act[(size_t)i] = active[i] ? 1 : 0;
    if (active[i]) plain.push_back(i);
  }
  f->ns = (int)plain.size();
  // work-item upper bounds: per cell ceil(count / tile) <= count / tile + 1
  f->n_items = ncell + n / (64 * kR) + 1;
  f->n_mlp_items = ncell + n / kMlpTile + 1;
  // near-list and far-cell splits: about kSplitRounds rounds of each kernel's
  // resident waves
  {
    const double waves = (double)n / (64 * kR) + 0.5 * ncell;
    auto split_for = [&](const void* kernel) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb <= 0) nb = 1;
      const double slots = (double)nb * 4 * ctx->num_cus;
      return std::min(kSplitMax, std::max(1, (int)std::ceil(kSplitRounds * slots / std::max(waves, 1.0))));
    };
    f->near_split = split_for(reinterpret_cast<const void*>(ff_near_kernel));
    f->far_split = split_for(f->order == 2   ? reinterpret_cast<const void*>(ff_far_kernel<2>)
                             : f->order == 3 ? reinterpret_cast<const void*>(ff_far_kernel<3>)
                                             : reinterpret_cast<const void*>(ff_far_kernel<4>));
  }
  const int end_bit = 32 - __builtin_clz((unsigned)ncell);  // keys 0..ncell
  size_t sort_bytes = 0;
  const hipError_t sort_query =
      hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int*)nullptr, (int*)nullptr, (const int*)nullptr,
                                         (int*)nullptr, n, 0, end_bit, ctx->stream);
  f->end_bit = end_bit;
  f->sort_bytes = sort_bytes;
  const size_t nb = (size_t)n * sizeof(double);
  const size_t ni = (size_t)n * sizeof(int);
  int rc = sort_query == hipSuccess ? LPM_OK
                                    : fail(ctx, LPM_ERR_HIP, "lpm_bve_ff_new: radix-sort workspace query failed");
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
    return rc == LPM_OK;
  };
  chk(ff_upload(ctx, f->x, x, nb)) && chk(ff_upload(ctx, f->y, y, nb)) && chk(ff_upload(ctx, f->z, z, nb)) &&
      chk(ff_upload(ctx, f->q, relvort, nb)) && chk(ff_upload(ctx, f->area, area, nb)) &&
      chk(ff_upload(ctx, f->u, nullptr, nb)) && chk(ff_upload(ctx, f->v, nullptr, nb)) &&
      chk(ff_upload(ctx, f->w, nullptr, nb)) && chk(ff_upload(ctx, f->pu, nullptr, nb)) &&
      chk(ff_upload(ctx, f->pv, nullptr, nb)) && chk(ff_upload(ctx, f->pw, nullptr, nb)) &&
      chk(ff_upload(ctx, f->xi, nullptr, nb)) && chk(ff_upload(ctx, f->yi, nullptr, nb)) &&
      chk(ff_upload(ctx, f->zi, nullptr, nb)) && chk(ff_upload(ctx, f->qi, nullptr, nb)) &&
      chk(ff_upload(ctx, f->src_perm, nullptr, ni)) && chk(ff_upload(ctx, f->src_key, nullptr, ni)) &&
      chk(ff_upload(ctx, f->src_start, nullptr, ((size_t)ncell + 1) * 4)) &&
      chk(ff_upload(ctx, f->tgt_perm, nullptr, ni)) && chk(ff_upload(ctx, f->tgt_cell, nullptr, ni)) &&
      chk(ff_upload(ctx, f->tgt_start, nullptr, ((size_t)ncell + 1) * 4)) &&
      chk(ff_upload(ctx, f->key_t, nullptr, ni)) && chk(ff_upload(ctx, f->key_s, nullptr, ni)) &&
      chk(ff_upload(ctx, f->iota, nullptr, ni)) && chk(ff_upload(ctx, f->cellof, nullptr, ni)) &&
      chk(ff_upload(ctx, f->active, act.data(), ni)) &&
      chk(ff_upload(ctx, f->sort_tmp, nullptr, std::max<size_t>(sort_bytes, 16))) &&
      chk(ff_upload(ctx, f->item_count, nullptr, 2 * 4)) &&
      chk(ff_upload(ctx, f->list_plain, plain.data(), std::max<size_t>(plain.size(), 1) * 4)) &&
      chk(ff_upload(ctx, f->items, nullptr, (size_t)f->n_items * sizeof(int2))) &&
      chk(ff_upload(ctx, f->mlp_items, nullptr, (size_t)f->n_mlp_items * sizeof(int2))) &&
      chk(ff_upload(ctx, f->trec, nullptr, (size_t)n * 16)) &&
      chk(ff_upload(ctx, f->srec, nullptr, std::max<size_t>((size_t)f->ns, 1) * lpm::SREC * 4)) &&
      chk(ff_upload(ctx, f->far_rec, nullptr, (size_t)(ncell + ncell / 4) * kCellRec * 4)) &&
      chk(ff_upload(ctx, f->mlp_rec, nullptr, (size_t)ncell * kMlpRec * 4)) &&
      chk(ff_upload(ctx, f->geo, nullptr, (size_t)ncell * 8 * 8)) &&
      chk(ff_upload(ctx, f->cen64, nullptr, (size_t)ncell * 3 * 8)) &&
      chk(ff_upload(ctx, f->mom64, nullptr, (size_t)ncell * kNMom * 8)) &&
      chk(ff_upload(ctx, f->far_list, nullptr, (size_t)ncell * ncell * 4)) &&
      chk(ff_upload(ctx, f->far_count, nullptr, (size_t)ncell * 4)) &&
      chk(ff_upload(ctx, f->near_list, nullptr, (size_t)ncell * ncell * 4)) &&
      chk(ff_upload(ctx, f->near_count, nullptr, (size_t)ncell * 4)) &&
      chk(ff_upload(ctx, f->band_list, nullptr, (size_t)ncell * ncell * 4)) &&
      chk(ff_upload(ctx, f->band_count, nullptr, (size_t)ncell * 4)) && chk(ff_upload(ctx, f->V, nullptr, 3 * nb)) &&
      chk(ff_upload(ctx, f->Vpart, nullptr, (size_t)(f->near_split + f->far_split) * 3 * nb));
  for (int s = 0; s < 4 && rc == LPM_OK; ++s)
    for (int c = 0; c < 4 && rc == LPM_OK; ++c) chk(ff_upload(ctx, f->k[s][c], nullptr, nb));
  if (rc == LPM_OK) {
    std::vector<double*> tab(16);
    for (int c = 0; c < 4; ++c)
      for (int s = 0; s < 4; ++s) tab[(size_t)(c * 4 + s)] = f->k[s][c].as<double>();
    chk(ff_upload(ctx, f->ptrs, tab.data(), tab.size() * sizeof(double*)));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  if (rc == LPM_OK) {
    // start on the sphere, with the plain velocity stored (as BVESolver New)
    hipLaunchKernelGGL(ff_project, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, f->radius,
                       f->x.as<double>(), f->y.as<double>(), f->z.as<double>());
    chk(ff_eval(f, 0, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->u.as<double>(),
                f->v.as<double>(), f->w.as<double>()));
    // bin once so that the cell queries are defined before any evaluation
    if (rc == LPM_OK) chk(ff_bin(f, f->x.as<double>(), f->y.as<double>(), f->z.as<double>()));
  }
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  f->loaded = rc == LPM_OK;
  return rc;
}

int ff_usable(lpm_bve_ff* f) {
  if (f->loaded) return LPM_OK;
  return fail(f->ctx, LPM_ERR_STATE, "far-field particles were not loaded (a particle reload failed)");
}

}  // namespace

extern "C" {

int lpm_bve_ff_new(lpm_ctx* ctx, lpm_bve_ff** out, int n, const double* x, const double* y, const double* z,
                   const double* relvort, const double* area, const int32_t* active, const int32_t* cell, int ncell,
                   double radius, double rotation_rate, double near_radius, double band_radius, int flags) {
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (n <= 0 || ncell <= 0 || !x || !y || !z || !relvort || !area || !active || !cell)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_new: bad arguments");
  if (!(radius > 0.0) || !(near_radius >= 0.0) || !(band_radius >= near_radius))
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_new: need radius > 0 and 0 <= near_radius <= band_radius");
  const int order = (flags & LPM_FF_ORDER_MASK) == 0 ? 4 : (flags & LPM_FF_ORDER_MASK);
  if (order < 2 || order > 4 || (flags & ~(LPM_FF_ORDER_MASK | LPM_FF_SAMPLE_DRAWS_MASK)))
    return fail(ctx, LPM_ERR_ARG,
                "lpm_bve_ff_new: flags hold the expansion order 2, 3 or 4 (0: 4) and the sample draws only");
  const int draws = std::max(1, (flags & LPM_FF_SAMPLE_DRAWS_MASK) >> 3);
  for (int i = 0; i < n; ++i)
    if (cell[i] < 0 || cell[i] >= ncell) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_new: cell id out of range");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  lpm_bve_ff* f = new lpm_bve_ff();
  f->ctx = ctx;
  f->n = n;
  f->ncell = ncell;
  f->radius = radius;
  f->omega = rotation_rate;
  f->r_near = near_radius;
  f->r_band = band_radius;
  f->order = order;
  f->draws = draws;
  // Cells are fixed regions of the sphere: each is the set of points nearest to
  // its generator, the direction of the area-weighted centroid of the given
  // cell's active sources at set-up (of all its particles if it has none).  Its
  // frame anchor is the direction of its active source farthest from that
  // centroid (lowest particle index on ties).  Particles are binned anew at
  // every evaluation, so cells never stretch with the flow.
  std::vector<lpm::Vec4<float>> gen((size_t)ncell);
  std::vector<double> adir((size_t)ncell * 3, 0.0);
  {
    std::vector<double> cs((size_t)ncell * 4, 0.0), ca((size_t)ncell * 4, 0.0);
    for (int i = 0; i < n; ++i) {
      double* c = &ca[(size_t)cell[i] * 4];
      c[0] += x[i];
      c[1] += y[i];
      c[2] += z[i];
      c[3] += 1.0;
      if (active[i]) {
        double* s = &cs[(size_t)cell[i] * 4];
        s[0] += area[i] * x[i];
        s[1] += area[i] * y[i];
        s[2] += area[i] * z[i];
        s[3] += area[i];
      }
    }
    std::vector<double> best((size_t)ncell, -1.0);
    for (int i = 0; i < n; ++i)
      if (active[i]) {
        const double* c = &cs[(size_t)cell[i] * 4];
        const double inv = c[3] > 0 ? 1.0 / c[3] : 0.0;
        const double dx = x[i] - c[0] * inv, dy = y[i] - c[1] * inv, dz = z[i] - c[2] * inv;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > best[(size_t)cell[i]]) {
          best[(size_t)cell[i]] = d2;
          const double r = std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
          adir[(size_t)cell[i] * 3 + 0] = x[i] / r;
          adir[(size_t)cell[i] * 3 + 1] = y[i] / r;
          adir[(size_t)cell[i] * 3 + 2] = z[i] / r;
        }
      }
    for (int c = 0; c < ncell; ++c) {
      const double* s = cs[(size_t)c * 4 + 3] > 0 ? &cs[(size_t)c * 4] : &ca[(size_t)c * 4];
      const double r = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
      if (r > 0) gen[(size_t)c] = lpm::Vec4<float>{(float)(s[0] / r), (float)(s[1] / r), (float)(s[2] / r), 0.f};
      else gen[(size_t)c] = lpm::Vec4<float>{0.f, 0.f, 0.f, 0.f};  // empty cell: never nearest
    }
  }
  int rc = LPM_OK;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
    return rc == LPM_OK;
  };
  chk(ff_upload(ctx, f->gen, gen.data(), (size_t)ncell * sizeof(lpm::Vec4<float>))) &&
      chk(ff_upload(ctx, f->adir, adir.data(), adir.size() * 8)) && chk(f->net.init(ctx, 13, 128, 3, 3, LPM_ACT_ISRU, 12345)) &&
      chk(ff_load(f, n, x, y, z, relvort, area, active));
  if (rc != LPM_OK) {
    lpm_bve_ff_delete(f);
    return rc;
  }
  *out = f;
  return lpm::collect_timing(ctx);
}

int lpm_bve_ff_velocity(lpm_bve_ff* f, int mode, double* u, double* v, double* w) {
  if (!f || !u || !v || !w) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  if (mode < 0 || mode > 2) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_velocity: mode must be 0, 1 or 2");
  if (mode == 2 && !f->trained) return fail(ctx, LPM_ERR_STATE, "lpm_bve_ff_velocity: the network is not trained");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = ff_eval(f, mode, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(),
                       f->pu.as<double>(), f->pv.as<double>(), f->pw.as<double>()))
    return rc;
  const size_t nb = (size_t)f->n * sizeof(double);
  LPM_HIP(ctx, hipMemcpyAsync(u, f->pu.p, nb, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(v, f->pv.p, nb, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(w, f->pw.p, nb, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_ff_collect(lpm_bve_ff* f, int n_samples, uint64_t seed, int64_t* total) {
  if (!f || n_samples <= 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  // refresh geometry, moments and lists at the current state (expansion mode)
  if (int rc = ff_eval(f, 1, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(),
                       f->pu.as<double>(), f->pv.as<double>(), f->pw.as<double>()))
    return rc;
  const int64_t need = f->n_collected + n_samples;
  if (need > f->cap_collected) {
    const int64_t cap = std::max<int64_t>(need, 2 * f->cap_collected);
    DevBuf nx, ny, ni;
    LPM_HIP(ctx, nx.ensure((size_t)cap * 13 * 4));
    LPM_HIP(ctx, ny.ensure((size_t)cap * 3 * 4));
    LPM_HIP(ctx, ni.ensure((size_t)cap * 2 * 4));
    if (f->n_collected > 0) {
      LPM_HIP(ctx, hipMemcpyAsync(nx.p, f->sx.p, (size_t)f->n_collected * 13 * 4, hipMemcpyDeviceToDevice, ctx->stream));
      LPM_HIP(ctx, hipMemcpyAsync(ny.p, f->sy.p, (size_t)f->n_collected * 3 * 4, hipMemcpyDeviceToDevice, ctx->stream));
      LPM_HIP(ctx, hipMemcpyAsync(ni.p, f->sid.p, (size_t)f->n_collected * 2 * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    f->sx.release();
    f->sy.release();
    f->sid.release();
    f->sx = nx;
    f->sy = ny;
    f->sid = ni;
    nx.p = ny.p = ni.p = nullptr;
    f->cap_collected = cap;
  }
  SampleArgs sa{n_samples, (unsigned long long)seed, f->n, f->ncell, f->order, f->tgt_perm.as<int>(), f->tgt_cell.as<int>(),
                f->src_perm.as<int>(), f->src_start.as<int>(), f->band_list.as<int>(), f->band_count.as<int>(),
                f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->area.as<double>(),
                f->cen64.as<double>(), f->mom64.as<double>(), f->mlp_rec.as<float>(), f->radius,
                f->sx.as<float>() + (size_t)f->n_collected * 13, f->sy.as<float>() + (size_t)f->n_collected * 3,
                f->sid.as<int>() + (size_t)f->n_collected * 2};
  sa.draws = f->draws;
  hipLaunchKernelGGL(ff_make_samples, dim3((n_samples + 255) / 256), dim3(256), 0, ctx->stream, sa);
  LPM_HIP(ctx, hipGetLastError());
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  f->n_collected += n_samples;
  if (total) *total = f->n_collected;
  return LPM_OK;
}

int lpm_bve_ff_train(lpm_bve_ff* f, int epochs, int batch, double lr, uint64_t seed, double* final_loss) {
  if (!f || epochs < 0 || batch <= 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (f->n_collected <= 0) return fail(ctx, LPM_ERR_STATE, "lpm_bve_ff_train: no samples collected");
  if (f->n_collected > INT32_MAX) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_train: too many samples");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  // targets are trained at unit RMS (scale fixed at the first training); the
  // stored parameters keep the scale folded into the output layer
  if (!(f->ys > 0.0)) {
    double r = 0.0;
    if (int rc = f->net.rms(f->n_collected * 3, f->sy.as<float>(), &r)) return rc;
    f->ys = r > 0.0 ? r : 1.0;
  }
  if (f->trained)
    if (int rc = f->net.scale_output((float)(1.0 / f->ys))) return rc;
  if (int rc = f->net.train((int)f->n_collected, f->sx.as<float>(), f->sy.as<float>(), epochs, batch, (float)lr, seed,
                            final_loss, nullptr, true, (float)(1.0 / f->ys)))
    return rc;
  if (int rc = f->net.scale_output((float)f->ys)) return rc;
  if (int rc = f->net.pack()) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  f->trained = true;
  return LPM_OK;
}

int lpm_bve_ff_set_online(lpm_bve_ff* f, int n_samples, int epochs, int batch, double lr) {
  if (!f || n_samples < 0 || epochs < 0 || batch <= 0 || !(lr >= 0.0)) return LPM_ERR_ARG;
  f->online_n = n_samples;
  f->online_epochs = epochs;
  f->online_batch = batch;
  f->online_lr = lr;
  return LPM_OK;
}

int lpm_bve_ff_set_online_from(lpm_bve_ff* f, double fraction) {
  if (!f || !(fraction >= 0.0 && fraction <= 1.0)) return LPM_ERR_ARG;
  f->online_from = fraction;
  return LPM_OK;
}

int lpm_bve_ff_set_params(lpm_bve_ff* f, const float* params, int64_t count) {
  if (!f || !params) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  if (count != f->net.np) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_set_params: wrong parameter count");
  LPM_HIP(ctx, hipMemcpyAsync(f->net.p.p, params, (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = f->net.pack()) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  f->trained = true;
  return LPM_OK;
}

int lpm_bve_ff_get_params(lpm_bve_ff* f, float* params, int64_t count) {
  if (!f || !params) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  if (count != f->net.np) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_get_params: wrong parameter count");
  LPM_HIP(ctx, hipMemcpyAsync(params, f->net.p.p, (size_t)count * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_samples(lpm_bve_ff* f, int64_t offset, int64_t count, float* in, float* target, int32_t* ids) {
  if (!f || !in || !target || offset < 0 || count < 0 || offset + count > f->n_collected) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  LPM_HIP(ctx, hipMemcpyAsync(in, f->sx.as<float>() + (size_t)offset * 13, (size_t)count * 13 * 4,
                              hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(target, f->sy.as<float>() + (size_t)offset * 3, (size_t)count * 3 * 4,
                              hipMemcpyDeviceToHost, ctx->stream));
  if (ids)
    LPM_HIP(ctx, hipMemcpyAsync(ids, f->sid.as<int>() + (size_t)offset * 2, (size_t)count * 2 * 4,
                                hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_steps(lpm_bve_ff* f, double dt, int nsteps, int mode, int judge_every, double tol, int* mode_out,
                     double* last_error, int* judged) {
  if (!f || nsteps < 0) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  if (mode < 0 || mode > 2) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_steps: mode must be 0, 1 or 2");
  if (mode == 2 && !f->trained) return fail(ctx, LPM_ERR_STATE, "lpm_bve_ff_steps: the network is not trained");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  int cur = mode, nj = 0;
  double err = -1.0;
  for (int s = 1; s <= nsteps; ++s) {
    if (int rc = ff_step(f, dt, cur)) return rc;
    if (cur != 0 && judge_every > 0 && s % judge_every == 0) {
      // the plain direct sum judges the stored velocity; above tol fall back for good
      if (int rc = ff_eval(f, 0, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(),
                           f->pu.as<double>(), f->pv.as<double>(), f->pw.as<double>()))
        return rc;
      if (int rc = ff_error(f, f->u.as<double>(), f->v.as<double>(), f->w.as<double>(), f->pu.as<double>(),
                            f->pv.as<double>(), f->pw.as<double>(), &err))
        return rc;
      ++nj;
      if (!(err <= tol)) {
        cur = 0;
        LPM_HIP(ctx, hipMemcpyAsync(f->u.p, f->pu.p, (size_t)f->n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        LPM_HIP(ctx, hipMemcpyAsync(f->v.p, f->pv.p, (size_t)f->n * 8, hipMemcpyDeviceToDevice, ctx->stream));
        LPM_HIP(ctx, hipMemcpyAsync(f->w.p, f->pw.p, (size_t)f->n * 8, hipMemcpyDeviceToDevice, ctx->stream));
      } else if (cur == 2 && f->online_n > 0 && err > f->online_from * tol) {
        // fine-tune on direct sums of the judged state for the steps that follow
        f->n_collected = 0;
        if (int rc = lpm_bve_ff_collect(f, f->online_n, 0x5eedull + (uint64_t)(++f->online_rounds), nullptr))
          return rc;
        if (f->online_epochs > 0)
          if (int rc = lpm_bve_ff_train(f, f->online_epochs, f->online_batch, f->online_lr,
                                        (uint64_t)f->online_rounds, nullptr))
            return rc;
      }
    }
  }
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (mode_out) *mode_out = cur;
  if (last_error) *last_error = err;
  if (judged) *judged = nj;
  return lpm::collect_timing(ctx);
}

int lpm_bve_ff_get_state(lpm_bve_ff* f, double* x, double* y, double* z, double* relvort, double* u, double* v,
                         double* w) {
  if (!f) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  const DevBuf* src[7] = {&f->x, &f->y, &f->z, &f->q, &f->u, &f->v, &f->w};
  double* dst[7] = {x, y, z, relvort, u, v, w};
  for (int k = 0; k < 7; ++k)
    if (dst[k])
      LPM_HIP(ctx, hipMemcpyAsync(dst[k], src[k]->p, (size_t)f->n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_set_particles(lpm_bve_ff* f, int n, const double* x, const double* y, const double* z,
                             const double* relvort, const double* area, const int32_t* active) {
  if (!f) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (n <= 0 || !x || !y || !z || !relvort || !area || !active)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_set_particles: n must be > 0 and the arrays non-null");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = ff_load(f, n, x, y, z, relvort, area, active)) return rc;
  return lpm::collect_timing(ctx);
}

int lpm_bve_ff_set_state(lpm_bve_ff* f, const double* x, const double* y, const double* z, const double* relvort) {
  if (!f || !x || !y || !z || !relvort) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  const size_t nb = (size_t)f->n * sizeof(double);
  LPM_HIP(ctx, hipMemcpyAsync(f->x.p, x, nb, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(f->y.p, y, nb, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(f->z.p, z, nb, hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(f->q.p, relvort, nb, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(ff_project, dim3((f->n + 255) / 256), dim3(256), 0, ctx->stream, f->n, f->radius,
                     f->x.as<double>(), f->y.as<double>(), f->z.as<double>());
  if (int rc = ff_eval(f, 0, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(),
                       f->u.as<double>(), f->v.as<double>(), f->w.as<double>()))
    return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_ff_particle_cells(lpm_bve_ff* f, int32_t* cell) {
  if (!f || !cell) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, hipMemcpyAsync(cell, f->cellof.p, (size_t)f->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_cells(lpm_bve_ff* f, int* near_cells, int* band_cells) {
  if (!f || !near_cells || !band_cells) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, hipMemcpyAsync(near_cells, f->near_count.p, (size_t)f->ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(band_cells, f->band_count.p, (size_t)f->ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_set_parents(lpm_bve_ff* f, double parent_radius) {
  if (!f) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (!(parent_radius >= 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_set_parents: parent_radius must be >= 0");
  if (parent_radius > 0.0 && f->ncell % 4 != 0)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_set_parents: the coarse level needs a cell count divisible by 4");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (parent_radius > 0.0) LPM_HIP(ctx, f->pgeo.ensure((size_t)(f->ncell / 4) * 4 * sizeof(double)));
  f->r_par = parent_radius;
  f->npar = parent_radius > 0.0 ? f->ncell / 4 : 0;
  return LPM_OK;
}

int lpm_bve_ff_far_counts(lpm_bve_ff* f, int* far_records) {
  if (!f || !far_records) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  ctx->err.clear();
  if (int rc = ff_usable(f)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  LPM_HIP(ctx, hipMemcpyAsync(far_records, f->far_count.p, (size_t)f->ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_delete(lpm_bve_ff* f) {
  if (!f) return LPM_ERR_ARG;
  (void)hipStreamSynchronize(f->ctx->stream);
  DevBuf* all[] = {&f->x, &f->y, &f->z, &f->q, &f->u, &f->v, &f->w, &f->area, &f->xi, &f->yi, &f->zi, &f->qi,
                   &f->pu, &f->pv, &f->pw, &f->src_perm, &f->src_start, &f->tgt_perm, &f->tgt_start, &f->tgt_cell,
                   &f->trec, &f->srec, &f->far_rec, &f->mlp_rec, &f->geo, &f->cen64, &f->mom64, &f->far_list, &f->far_count, &f->pgeo,
                   &f->near_list, &f->near_count, &f->band_list, &f->band_count, &f->items, &f->mlp_items, &f->V,
                   &f->list_plain, &f->tgt_plain, &f->src_plain, &f->partial_plain, &f->ptrs, &f->err_part, &f->sx,
                   &f->sy, &f->sid, &f->src_key, &f->key_t, &f->key_s, &f->iota, &f->cellof, &f->active,
                   &f->sort_tmp, &f->item_count, &f->gen, &f->adir, &f->Vpart};
  for (DevBuf* b : all) b->release();
  for (auto& row : f->k)
    for (auto& b : row) b.release();
  delete f;
  return LPM_OK;
}

}  // extern "C"
