// lpm_ff.hip -- sphere BVE with a learned far field.
//
// Sources are grouped in coarse cells (the particles of a coarse mesh face).
// For each target cell t and source cell s, with cell centroids c, angular
// radii rho and the angle a(t, s) between their centroid directions:
//   near   a < r_near + rho_t + rho_s : direct FP32-compensated sum over s's sources
//   far    otherwise                  : analytic expansion of s to 2nd order in
//          its 10 moments (M0, M1, M2 of zeta A about c_s)
//   band   far and a < r_band + rho_t + rho_s : far + the MLP's correction of the
//          expansion's remainder (BF16 MFMA), trained on direct sums
// Cell membership is Lagrangian (fixed at set-up); centroids, radii and
// moments follow the particles and are recomputed at every evaluation.
//
// The network works in a frame of its cell: e3 along the centroid c, e1 towards
// the cell's anchor source (the source farthest from c at set-up, fixed), e2 =
// e3 x e1.  Cells of one mesh level then look alike to it, which fixes the
// orientation of the cell's third moment that the 10 moments leave open.
// Inputs (13): Q (x - c) / l, M0 / W, Q M1 / (W l), Q M2 Q^T / (W l^2) (6), with
// Q the frame rows, l = 2 R sin(rho / 2) and W = sum |w|.  Output o (3, cell
// frame): the expansion's remainder of G_s / (4 pi R) is Q^T o s_out / r^4 with
// r = |x - c| / l and s_out = W / (4 pi R l^2), a scaling under which the
// remainder (~ W l^3 / |x - c|^5) is O(1) at every distance in the band.
//
// Velocity assembly, per target x (on the sphere of radius R):
//   u = V x x,  V = sum_near s_ij (x - y_j) + sum_far G_s(x) / (4 pi R)
// with s_ij = -zeta_j A_j / (4 pi R (R^2 - x.y_j)) and, for the expansion,
// G_s(x) = c alpha + M1 / D + (M2 x) / D^2,
// alpha = M0 / D + x.M1 / D^2 + x^T M2 x / D^3, D = R^2 - x.c
// (sum_j w_j y_j / (R^2 - x.y_j) to second order about c; the parts of V
// along x drop out of V x x).  D is evaluated as (|x-c|^2 + R^2 - |c|^2) / 2.
#include <math.h>

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
constexpr int kR = 4;        // targets per lane in the near / far kernels
constexpr int kKT = 256;     // sources per compensated tile
constexpr int kCellRec = 16; // floats per far-field cell record
constexpr int kMlpRec = 32;  // floats per MLP cell record
constexpr int kOutPow = 4;   // remainder scaling power r^kOutPow

#pragma clang fp contract(off)

__device__ __forceinline__ void project(double& x, double& y, double& z, double radius) {
  const double s = radius / sqrt(x * x + y * y + z * z);
  x = x * s;
  y = y * s;
  z = z * s;
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
  float* far_rec;      // [ncell][16]: c(3), e_c, M0', M1'(3), M2'(6: xx xy xz yy yz zz), pad
  float* mlp_rec;      // [ncell][32]: c(3), 1/l, Q(9), s_out, pad(2), features 3..12, pad(6)
  const int* anchor;   // [ncell] anchor source (particle index) of each cell's frame
  double* geo;         // [ncell][8]: src dir (3), src rho, tgt dir (3), tgt rho
  double* cen64;       // [ncell][3] FP64 source centroids
  double* mom64;       // [ncell][10]
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

__global__ __launch_bounds__(256) void ff_cell_geometry(GeomArgs a) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  const int s0 = a.src_start[c], s1 = a.src_start[c + 1];
  const int t0 = a.tgt_start[c], t1 = a.tgt_start[c + 1];
  // area-weighted source centroid
  double ax = 0, ay = 0, az = 0, aa = 0;
  for (int j = s0 + threadIdx.x; j < s1; j += 256) {
    const int p = a.src_perm[j];
    const double w = a.area[p];
    ax += w * a.x[p];
    ay += w * a.y[p];
    az += w * a.z[p];
    aa += w;
  }
  ax = block_sum(ax, sh);
  ay = block_sum(ay, sh);
  az = block_sum(az, sh);
  aa = block_sum(aa, sh);
  const double cx = aa > 0 ? ax / aa : 0.0, cy = aa > 0 ? ay / aa : 0.0, cz = aa > 0 ? az / aa : 1.0;
  const double cn = sqrt(cx * cx + cy * cy + cz * cz);
  const double ux = cx / cn, uy = cy / cn, uz = cz / cn;
  // moments about c, radius
  double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  double wabs = 0.0, rho = 0.0;
  for (int j = s0 + threadIdx.x; j < s1; j += 256) {
    const int p = a.src_perm[j];
    const double w = a.q[p] * a.area[p];
    const double dx = a.x[p] - cx, dy = a.y[p] - cy, dz = a.z[p] - cz;
    m[0] += w;
    m[1] += w * dx;
    m[2] += w * dy;
    m[3] += w * dz;
    m[4] += w * dx * dx;
    m[5] += w * dx * dy;
    m[6] += w * dx * dz;
    m[7] += w * dy * dy;
    m[8] += w * dy * dz;
    m[9] += w * dz * dz;
    wabs += fabs(w);
    const double r = sqrt(a.x[p] * a.x[p] + a.y[p] * a.y[p] + a.z[p] * a.z[p]);
    const double cosang = (a.x[p] * ux + a.y[p] * uy + a.z[p] * uz) / r;
    rho = fmax(rho, acos(fmin(1.0, fmax(-1.0, cosang))));
  }
  for (int k = 0; k < 10; ++k) m[k] = block_sum(m[k], sh);
  wabs = block_sum(wabs, sh);
  rho = block_max(rho, sh);
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
    float* fr = a.far_rec + (size_t)c * kCellRec;
    fr[0] = (float)cx;
    fr[1] = (float)cy;
    fr[2] = (float)cz;
    fr[3] = (float)(R * R - (cx * cx + cy * cy + cz * cz));
    for (int k = 0; k < 10; ++k) fr[4 + k] = (float)(m[k] * inv);
    fr[14] = (float)rho;
    fr[15] = 0.f;
    const double l = fmax(2.0 * R * sin(0.5 * rho), 1e-12);
    const double W = wabs + 1e-300;
    // cell frame: e3 = c/|c|, e1 = tangential part of (anchor - c), e2 = e3 x e1
    double Q[3][3];
    Q[2][0] = ux;
    Q[2][1] = uy;
    Q[2][2] = uz;
    {
      const int pa = a.anchor[c];
      double ex = a.x[pa] - cx, ey = a.y[pa] - cy, ez = a.z[pa] - cz;
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
    for (int k = 0; k < 10; ++k) a.mom64[(size_t)c * 10 + k] = m[k];
  }
}

// Per target cell t: class of every source cell (0 far, 1 band, 2 near) and
// compact lists of near and band cells.
struct ClassArgs {
  int ncell;
  const double* geo;
  double r_near, r_band;
  unsigned char* cls;  // [ncell][ncell]
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
    if (s < a.ncell) {
      const double* gs = a.geo + (size_t)s * 8;
      const double cosang = gt[4] * gs[0] + gt[5] * gs[1] + gt[6] * gs[2];
      const double ang = acos(fmin(1.0, fmax(-1.0, cosang)));
      const double pad = gt[7] + gs[3];
      cls = ang < a.r_near + pad ? 2 : (ang < a.r_band + pad ? 1 : 0);
      a.cls[(size_t)t * a.ncell + s] = (unsigned char)cls;
    }
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
  int n_items, ncell;
  double* out;                  // [3][nt] FP64, cell order
  int nt;
};

__global__ __launch_bounds__(256) void ff_near_kernel(NearArgs a) {
  using P = lpm::SphereBS<float, false>;
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= a.n_items) return;
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
  const int nn = a.near_count[t];
  for (int q = 0; q < nn; ++q) {
    const int s = a.near_list[(size_t)t * a.ncell + q];
    const int j0 = a.src_start[s], j1 = a.src_start[s + 1];
    const bool checked = (s == t);  // a target's own source lives in its own cell
    for (int jt = j0; jt < j1; jt += kKT) {
      const int je = min(j1, jt + kKT);
      float p[kR][3];
#pragma unroll
      for (int r = 0; r < kR; ++r) p[r][0] = p[r][1] = p[r][2] = 0.f;
      if (!checked) {
        for (int j = jt; j < je; ++j) {
          const float* sr = a.src + (size_t)j * lpm::SREC;
          float sv[lpm::SREC];
#pragma unroll
          for (int e = 0; e < lpm::SREC; ++e) sv[e] = sr[e];
          lpm::W4<float> w[kR];
          float f[kR];
#pragma unroll
          for (int r = 0; r < kR; ++r) P::arg(tv[r], sv, w[r]);
#pragma unroll
          for (int r = 0; r < kR; ++r) f[r] = P::trans(w[r]);
#pragma unroll
          for (int r = 0; r < kR; ++r) P::acc(sv, w[r], f[r], p[r]);
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
      for (int c = 0; c < 3; ++c) a.out[(size_t)c * a.nt + k] = (double)S[r][c] + (double)C[r][c];
  }
}

// ------------------------------------------------------- far (expansion)
struct FarArgs {
  const lpm::Vec4<float>* tgt;
  const int* tgt_start;
  const float* cell;          // far records
  const unsigned char* cls;
  const int2* items;
  int n_items, ncell, nt;
  double* out;                // [3][nt], accumulated into (+=)
};

__global__ __launch_bounds__(256) void ff_far_kernel(FarArgs a) {
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= a.n_items) return;
  const int lane = threadIdx.x & 63;
  const int2 it = a.items[item];
  const int t = it.x;
  const int k0 = a.tgt_start[t] + it.y * 64 * kR;
  const int k_end = a.tgt_start[t + 1];
  float X[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    const lpm::Vec4<float> v = a.tgt[k < k_end ? k : k_end - 1];
    X[r][0] = v.a;
    X[r][1] = v.b;
    X[r][2] = v.c;
  }
  float S[kR][3], C[kR][3];
#pragma unroll
  for (int r = 0; r < kR; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[r][c] = C[r][c] = 0.f;
  const unsigned char* row = a.cls + (size_t)t * a.ncell;
  for (int s0 = 0; s0 < a.ncell; s0 += 64) {
    const int s1 = min(a.ncell, s0 + 64);
    float p[kR][3];
#pragma unroll
    for (int r = 0; r < kR; ++r) p[r][0] = p[r][1] = p[r][2] = 0.f;
    for (int s = s0; s < s1; ++s) {
      if (row[s] == 2) continue;  // near: summed directly (wave-uniform branch)
      const float* cr = a.cell + (size_t)s * kCellRec;
      float cv[14];
#pragma unroll
      for (int e = 0; e < 14; ++e) cv[e] = cr[e];
#pragma unroll
      for (int r = 0; r < kR; ++r) {
        const float dx = X[r][0] - cv[0], dy = X[r][1] - cv[1], dz = X[r][2] - cv[2];
        const float D = 0.5f * fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, cv[3])));
        const float iD = __builtin_amdgcn_rcpf(D);
        // M2 x
        const float mx = fmaf(cv[8], X[r][0], fmaf(cv[9], X[r][1], cv[10] * X[r][2]));
        const float my = fmaf(cv[9], X[r][0], fmaf(cv[11], X[r][1], cv[12] * X[r][2]));
        const float mz = fmaf(cv[10], X[r][0], fmaf(cv[12], X[r][1], cv[13] * X[r][2]));
        const float xm1 = fmaf(cv[5], X[r][0], fmaf(cv[6], X[r][1], cv[7] * X[r][2]));
        const float xm2x = fmaf(mx, X[r][0], fmaf(my, X[r][1], mz * X[r][2]));
        const float alpha = iD * fmaf(iD, fmaf(iD, xm2x, xm1), cv[4]);
        const float iD2 = iD * iD;
        p[r][0] += fmaf(cv[0], alpha, fmaf(cv[5], iD, mx * iD2));
        p[r][1] += fmaf(cv[1], alpha, fmaf(cv[6], iD, my * iD2));
        p[r][2] += fmaf(cv[2], alpha, fmaf(cv[7], iD, mz * iD2));
      }
    }
#pragma unroll
    for (int r = 0; r < kR; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) lpm::two_sum_acc(S[r][c], C[r][c], p[r][c]);
  }
#pragma unroll
  for (int r = 0; r < kR; ++r) {
    const int k = k0 + 64 * r + lane;
    if (k < k_end)
#pragma unroll
      for (int c = 0; c < 3; ++c) a.out[(size_t)c * a.nt + k] += (double)S[r][c] + (double)C[r][c];
  }
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

// s_out / r^kOutPow for the local position
__device__ __forceinline__ float mlp_out_scale(const float* mr, const float loc[3]) {
  const float r2 = fmaf(loc[0], loc[0], fmaf(loc[1], loc[1], loc[2] * loc[2]));
  return mr[13] * __builtin_amdgcn_rcpf(r2 * r2);
}

struct MlpCorrArgs {
  const lpm::Vec4<float>* tgt;
  const int* tgt_start;
  const float* mlp_rec;
  const int* band_list;
  const int* band_count;
  const int2* items;  // (target cell, 64-target tile)
  int n_items, ncell, nt;
  const __bf16* frag;
  const float* bias;
  double* out;        // [3][nt] accumulated into
};

__global__ __launch_bounds__(256) void ff_mlp_kernel(MlpCorrArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 sfrag[lpm::kMlpFragElems];
  __shared__ float sbias[lpm::kMlpBiasElems];
  for (int t = threadIdx.x; t < lpm::kMlpFragElems / 8; t += blockDim.x)
    reinterpret_cast<lpm::bf16x8_t*>(sfrag)[t] = reinterpret_cast<const lpm::bf16x8_t*>(a.frag)[t];
  for (int t = threadIdx.x; t < lpm::kMlpBiasElems; t += blockDim.x) sbias[t] = a.bias[t];
  __syncthreads();
  constexpr int NCT = 4;
  const int lane = threadIdx.x & 63, g = lane >> 4;
  const int waves = gridDim.x * 4;
  for (int item = blockIdx.x * 4 + (threadIdx.x >> 6); item < a.n_items; item += waves) {
    const int2 it = a.items[item];
    const int t = it.x;
    const int k0 = a.tgt_start[t] + it.y * 64;
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
      const int s = a.band_list[(size_t)t * a.ncell + q];
      const float* mr = a.mlp_rec + (size_t)s * kMlpRec;
      lpm::bf16x8_t in8[NCT];
      float loc[NCT][3];
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        mlp_local(mr, xs[ct][0], xs[ct][1], xs[ct][2], loc[ct]);
#pragma unroll
        for (int e = 0; e < 8; ++e) in8[ct][e] = (__bf16)mlp_feature(mr, loc[ct], 8 * g + e);
      }
      float o[NCT][4];
      lpm::mlp_wave_forward<NCT>(in8, sfrag, sbias, o);
      // lanes 0..15 hold the outputs: scale, rotate to world axes (Q^T o), accumulate
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const float so = mlp_out_scale(mr, loc[ct]);
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
// Inputs: the 13 MLP features; target: Q (G_exact - G_expansion) / (4 pi R) r^4 / s_out,
// the remainder in the cell frame and scaled, in FP64 (G_exact over the cell's
// sources, with w = zeta A).  ids receives (particle, cell) of each sample.
struct SampleArgs {
  int n_samples;
  unsigned long long seed;
  int nt, ncell;
  const int *tgt_perm, *tgt_cell, *src_perm, *src_start, *band_list, *band_count;
  const double *x, *y, *z, *q, *area;
  const double *cen64, *mom64;
  const float* mlp_rec;
  double radius;
  float* X;  // [n][13]
  float* Y;  // [n][3]
  int* ids;  // [n][2]: target particle, source cell (-1, -1 for an invalid sample)
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ void expansion_G(const double* c, const double* m, double R, double x, double y, double z, double G[3]) {
  const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
  const double D = 0.5 * (dx * dx + dy * dy + dz * dz + (R * R - (c[0] * c[0] + c[1] * c[1] + c[2] * c[2])));
  const double iD = 1.0 / D;
  const double mx = m[4] * x + m[5] * y + m[6] * z;
  const double my = m[5] * x + m[7] * y + m[8] * z;
  const double mz = m[6] * x + m[8] * y + m[9] * z;
  const double xm1 = m[1] * x + m[2] * y + m[3] * z;
  const double xm2x = mx * x + my * y + mz * z;
  const double alpha = m[0] * iD + xm1 * iD * iD + xm2x * iD * iD * iD;
  G[0] = c[0] * alpha + m[1] * iD + mx * iD * iD;
  G[1] = c[1] * alpha + m[2] * iD + my * iD * iD;
  G[2] = c[2] * alpha + m[3] * iD + mz * iD * iD;
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
    const int s = a.band_list[(size_t)t * a.ncell + (int)(h % (unsigned long long)nb)];
    h = mix64(h);
    const int p = a.tgt_perm[k];
    const double x = a.x[p], y = a.y[p], z = a.z[p];
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
    expansion_G(c64, a.mom64 + (size_t)s * 10, R, x, y, z, Gm);
    const float* mr = a.mlp_rec + (size_t)s * kMlpRec;
    // local position in FP64 for the scaling, FP32 (as at inference) for the inputs
    const double il = (double)mr[3];
    double r2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double li = ((double)mr[4 + 3 * i] * (x - c64[0]) + (double)mr[5 + 3 * i] * (y - c64[1]) +
                         (double)mr[6 + 3 * i] * (z - c64[2])) * il;
      r2 += li * li;
    }
    const double inv = r2 * r2 / (4.0 * kPi * R) / (double)mr[13];
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
  DevBuf x, y, z, q, u, v, w, area;
  DevBuf k[4][4];
  DevBuf xi, yi, zi, qi;
  DevBuf pu, pv, pw;  // plain-sum velocity (judge)
  DevBuf src_perm, src_start, tgt_perm, tgt_start, tgt_cell, anchor;
  DevBuf trec, srec, far_rec, mlp_rec, geo, cen64, mom64;
  DevBuf cls, near_list, near_count, band_list, band_count;
  DevBuf items, mlp_items;
  DevBuf V;
  DevBuf list_plain;  // ascending active list for the plain path
  DevBuf tgt_plain, src_plain, partial_plain;
  DevBuf ptrs, err_part;
  DevBuf sx, sy, sid;  // collected training set: inputs, targets, (particle, cell)
  int n_items = 0, n_mlp_items = 0;
  int64_t n_collected = 0, cap_collected = 0;
  lpm::Mlp net;
  bool trained = false;
};

namespace {

int ff_upload(lpm_ctx* ctx, DevBuf& b, const void* host, size_t bytes) {
  LPM_HIP(ctx, b.ensure(bytes));
  if (host && bytes) LPM_HIP(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
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
  hipLaunchKernelGGL(ff_pack_targets, dim3((f->n + 255) / 256), b256, 0, ctx->stream, f->n, f->tgt_perm.as<int>(), x,
                     y, z, f->trec.as<lpm::Vec4<float>>());
  lpm::PackArgs pa{x, y, z, q, nullptr, f->area.as<double>(), nullptr, f->src_perm.as<int>(), 0, 0, f->radius, 1.0};
  lpm::launch_pack(LPM_K_SPHERE_BS, LPM_PREC_FP32_KAHAN, pa, f->srec.p, nullptr, f->ns, 0, ctx->stream);
  GeomArgs ga{nc, f->src_perm.as<int>(), f->src_start.as<int>(), f->tgt_perm.as<int>(), f->tgt_start.as<int>(),
              x, y, z, q, f->area.as<double>(), f->radius, f->far_rec.as<float>(), f->mlp_rec.as<float>(),
              f->anchor.as<int>(), f->geo.as<double>(), f->cen64.as<double>(), f->mom64.as<double>()};
  hipLaunchKernelGGL(ff_cell_geometry, dim3(nc), b256, 0, ctx->stream, ga);
  ClassArgs ca{nc, f->geo.as<double>(), f->r_near, f->r_band, f->cls.as<unsigned char>(), f->near_list.as<int>(),
               f->near_count.as<int>(), f->band_list.as<int>(), f->band_count.as<int>()};
  hipLaunchKernelGGL(ff_classify, dim3(nc), b256, 0, ctx->stream, ca);
  NearArgs na{f->trec.as<lpm::Vec4<float>>(), f->tgt_perm.as<int>(), f->tgt_start.as<int>(), f->srec.as<float>(),
              f->src_start.as<int>(), f->near_list.as<int>(), f->near_count.as<int>(), f->items.as<int2>(),
              f->n_items, nc, f->V.as<double>(), f->n};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing) {
    e0 = lpm::take_event(ctx);
    e1 = lpm::take_event(ctx);
    if (e0 && e1) LPM_HIP(ctx, hipEventRecord(e0, ctx->stream));
  }
  hipLaunchKernelGGL(ff_near_kernel, dim3((f->n_items + 3) / 4), b256, 0, ctx->stream, na);
  FarArgs fa{f->trec.as<lpm::Vec4<float>>(), f->tgt_start.as<int>(), f->far_rec.as<float>(), f->cls.as<unsigned char>(),
             f->items.as<int2>(), f->n_items, nc, f->n, f->V.as<double>()};
  hipLaunchKernelGGL(ff_far_kernel, dim3((f->n_items + 3) / 4), b256, 0, ctx->stream, fa);
  if (mode == 2 && f->trained) {
    MlpCorrArgs ma{f->trec.as<lpm::Vec4<float>>(), f->tgt_start.as<int>(), f->mlp_rec.as<float>(),
                   f->band_list.as<int>(), f->band_count.as<int>(), f->mlp_items.as<int2>(), f->n_mlp_items, nc, f->n,
                   f->net.frag.as<__bf16>(), f->net.bias.as<float>(), f->V.as<double>()};
    hipLaunchKernelGGL(ff_mlp_kernel, dim3(2 * ctx->num_cus), b256, 0, ctx->stream, ma);
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

}  // namespace

extern "C" {

int lpm_bve_ff_new(lpm_ctx* ctx, lpm_bve_ff** out, int n, const double* x, const double* y, const double* z,
                   const double* relvort, const double* area, const int32_t* active, const int32_t* cell, int ncell,
                   double radius, double rotation_rate, double near_radius, double band_radius, int flags) {
  (void)flags;
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (n <= 0 || ncell <= 0 || !x || !y || !z || !relvort || !area || !active || !cell)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_new: bad arguments");
  if (!(radius > 0.0) || !(near_radius >= 0.0) || !(band_radius >= near_radius))
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_ff_new: need radius > 0 and 0 <= near_radius <= band_radius");
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
  // cell orderings (stable: ascending particle index inside a cell)
  std::vector<int> tcount((size_t)ncell + 1, 0), scount((size_t)ncell + 1, 0);
  std::vector<int> plain;
  for (int i = 0; i < n; ++i) {
    tcount[(size_t)cell[i] + 1]++;
    if (active[i]) {
      scount[(size_t)cell[i] + 1]++;
      plain.push_back(i);
    }
  }
  f->ns = (int)plain.size();
  std::partial_sum(tcount.begin(), tcount.end(), tcount.begin());
  std::partial_sum(scount.begin(), scount.end(), scount.begin());
  std::vector<int> tperm((size_t)n), sperm((size_t)f->ns), tpos(tcount.begin(), tcount.end() - 1),
      spos(scount.begin(), scount.end() - 1), tcell((size_t)n);
  for (int i = 0; i < n; ++i) {
    const int c = cell[i];
    tcell[(size_t)tpos[(size_t)c]] = c;
    tperm[(size_t)tpos[(size_t)c]++] = i;
    if (active[i]) sperm[(size_t)spos[(size_t)c]++] = i;
  }
  // frame anchor of each cell: its active source farthest from the area-weighted
  // centroid at set-up (lowest particle index on ties), then fixed
  std::vector<int> anchor((size_t)ncell, -1);
  {
    std::vector<double> cen((size_t)ncell * 4, 0.0);
    for (int i = 0; i < n; ++i)
      if (active[i]) {
        double* c = &cen[(size_t)cell[i] * 4];
        c[0] += area[i] * x[i];
        c[1] += area[i] * y[i];
        c[2] += area[i] * z[i];
        c[3] += area[i];
      }
    std::vector<double> best((size_t)ncell, -1.0);
    for (int i = 0; i < n; ++i)
      if (active[i]) {
        const double* c = &cen[(size_t)cell[i] * 4];
        const double inv = c[3] > 0 ? 1.0 / c[3] : 0.0;
        const double dx = x[i] - c[0] * inv, dy = y[i] - c[1] * inv, dz = z[i] - c[2] * inv;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > best[(size_t)cell[i]]) {
          best[(size_t)cell[i]] = d2;
          anchor[(size_t)cell[i]] = i;
        }
      }
    for (int c = 0; c < ncell; ++c)
      if (anchor[(size_t)c] < 0) anchor[(size_t)c] = 0;  // no sources: the frame is never used
  }
  std::vector<int2> items, mitems;
  for (int c = 0; c < ncell; ++c) {
    const int cnt = tcount[(size_t)c + 1] - tcount[(size_t)c];
    for (int w = 0; w * 64 * kR < cnt; ++w) items.push_back(make_int2(c, w));
    for (int w = 0; w * 64 < cnt; ++w) mitems.push_back(make_int2(c, w));
  }
  f->n_items = (int)items.size();
  f->n_mlp_items = (int)mitems.size();
  const size_t nb = (size_t)n * sizeof(double);
  int rc = LPM_OK;
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
      chk(ff_upload(ctx, f->src_perm, sperm.data(), std::max<size_t>(sperm.size(), 1) * 4)) &&
      chk(ff_upload(ctx, f->src_start, scount.data(), scount.size() * 4)) &&
      chk(ff_upload(ctx, f->tgt_perm, tperm.data(), tperm.size() * 4)) &&
      chk(ff_upload(ctx, f->tgt_start, tcount.data(), tcount.size() * 4)) &&
      chk(ff_upload(ctx, f->tgt_cell, tcell.data(), tcell.size() * 4)) &&
      chk(ff_upload(ctx, f->anchor, anchor.data(), anchor.size() * 4)) &&
      chk(ff_upload(ctx, f->list_plain, plain.data(), std::max<size_t>(plain.size(), 1) * 4)) &&
      chk(ff_upload(ctx, f->items, items.data(), std::max<size_t>(items.size(), 1) * sizeof(int2))) &&
      chk(ff_upload(ctx, f->mlp_items, mitems.data(), std::max<size_t>(mitems.size(), 1) * sizeof(int2))) &&
      chk(ff_upload(ctx, f->trec, nullptr, (size_t)n * 16)) &&
      chk(ff_upload(ctx, f->srec, nullptr, std::max<size_t>((size_t)f->ns, 1) * lpm::SREC * 4)) &&
      chk(ff_upload(ctx, f->far_rec, nullptr, (size_t)ncell * kCellRec * 4)) &&
      chk(ff_upload(ctx, f->mlp_rec, nullptr, (size_t)ncell * kMlpRec * 4)) &&
      chk(ff_upload(ctx, f->geo, nullptr, (size_t)ncell * 8 * 8)) &&
      chk(ff_upload(ctx, f->cen64, nullptr, (size_t)ncell * 3 * 8)) &&
      chk(ff_upload(ctx, f->mom64, nullptr, (size_t)ncell * 10 * 8)) &&
      chk(ff_upload(ctx, f->cls, nullptr, (size_t)ncell * ncell)) &&
      chk(ff_upload(ctx, f->near_list, nullptr, (size_t)ncell * ncell * 4)) &&
      chk(ff_upload(ctx, f->near_count, nullptr, (size_t)ncell * 4)) &&
      chk(ff_upload(ctx, f->band_list, nullptr, (size_t)ncell * ncell * 4)) &&
      chk(ff_upload(ctx, f->band_count, nullptr, (size_t)ncell * 4)) && chk(ff_upload(ctx, f->V, nullptr, 3 * nb));
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
  if (rc == LPM_OK) chk(f->net.init(ctx, 13, 128, 3, 3, 12345));
  if (rc == LPM_OK) {
    // start on the sphere, with the plain velocity stored (as BVESolver New)
    hipLaunchKernelGGL(ff_project, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, radius, f->x.as<double>(),
                       f->y.as<double>(), f->z.as<double>());
    chk(ff_eval(f, 0, f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->u.as<double>(),
                f->v.as<double>(), f->w.as<double>()));
  }
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
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
  SampleArgs sa{n_samples, (unsigned long long)seed, f->n, f->ncell, f->tgt_perm.as<int>(), f->tgt_cell.as<int>(),
                f->src_perm.as<int>(), f->src_start.as<int>(), f->band_list.as<int>(), f->band_count.as<int>(),
                f->x.as<double>(), f->y.as<double>(), f->z.as<double>(), f->q.as<double>(), f->area.as<double>(),
                f->cen64.as<double>(), f->mom64.as<double>(), f->mlp_rec.as<float>(), f->radius,
                f->sx.as<float>() + (size_t)f->n_collected * 13, f->sy.as<float>() + (size_t)f->n_collected * 3,
                f->sid.as<int>() + (size_t)f->n_collected * 2};
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
  if (int rc = f->net.train((int)f->n_collected, f->sx.as<float>(), f->sy.as<float>(), epochs, batch, (float)lr, seed,
                            final_loss, nullptr, true))
    return rc;
  f->trained = true;
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
  const DevBuf* src[7] = {&f->x, &f->y, &f->z, &f->q, &f->u, &f->v, &f->w};
  double* dst[7] = {x, y, z, relvort, u, v, w};
  for (int k = 0; k < 7; ++k)
    if (dst[k])
      LPM_HIP(ctx, hipMemcpyAsync(dst[k], src[k]->p, (size_t)f->n * 8, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_cells(lpm_bve_ff* f, int* near_cells, int* band_cells) {
  if (!f || !near_cells || !band_cells) return LPM_ERR_ARG;
  lpm_ctx* ctx = f->ctx;
  LPM_HIP(ctx, hipMemcpyAsync(near_cells, f->near_count.p, (size_t)f->ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipMemcpyAsync(band_cells, f->band_count.p, (size_t)f->ncell * 4, hipMemcpyDeviceToHost, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_ff_delete(lpm_bve_ff* f) {
  if (!f) return LPM_ERR_ARG;
  (void)hipStreamSynchronize(f->ctx->stream);
  DevBuf* all[] = {&f->x, &f->y, &f->z, &f->q, &f->u, &f->v, &f->w, &f->area, &f->xi, &f->yi, &f->zi, &f->qi,
                   &f->pu, &f->pv, &f->pw, &f->src_perm, &f->src_start, &f->tgt_perm, &f->tgt_start, &f->tgt_cell,
                   &f->trec, &f->srec, &f->far_rec, &f->mlp_rec, &f->geo, &f->cen64, &f->mom64, &f->cls,
                   &f->near_list, &f->near_count, &f->band_list, &f->band_count, &f->items, &f->mlp_items, &f->V,
                   &f->list_plain, &f->tgt_plain, &f->src_plain, &f->partial_plain, &f->ptrs, &f->err_part, &f->sx,
                   &f->sy, &f->sid, &f->anchor};
  for (DevBuf* b : all) b->release();
  for (auto& row : f->k)
    for (auto& b : row) b.release();
  delete f;
  return LPM_OK;
}

}  // extern "C"
