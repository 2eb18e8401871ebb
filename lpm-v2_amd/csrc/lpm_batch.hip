// lpm_batch.hip -- many independent sphere BVE problems of uneven size in one
// persistent launch per velocity evaluation, replayed as one hipGraph per RK4
// stage.
//
// Layout: all problems' particles are concatenated (CSR offsets); each
// particle carries its problem's radius and rotation rate.  A velocity
// evaluation is: reset the work counter (a one-thread kernel, so the graph
// holds no memset node), pack FP32/FP64 records for every particle and every
// active source, one persistent pair_batch_kernel that drains a work list of
// (problem, wave target tile, source split) items, and a batched finalize.
#include <math.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "lpm_internal.hpp"
#include "pair_kernels.hpp"

using lpm::DevBuf;
using lpm::fail;

namespace {

#pragma clang fp contract(off)

__device__ __forceinline__ void project(double& x, double& y, double& z, double radius) {
  const double s = radius / sqrt(x * x + y * y + z * z);
  x = x * s;
  y = y * s;
  z = z * s;
}

__global__ void reset_counter(unsigned int* c) {
  if (threadIdx.x == 0) c[0] = 0u;
}

struct StageArgs {
  int n;
  double dt;
  double h;
  int make_input, proj, first;
  const double *prad, *pomg;
  const double *x, *y, *z, *q;
  const double *u, *v, *w;  // stage 1 only
  double *kx, *ky, *kz, *kq;
  double *xi, *yi, *zi, *qi;
};

// Stage algebra of timestepPrivate (SphereBVESolver.f90:238-313) with the
// per-particle radius and rotation rate of the particle's problem.
__global__ __launch_bounds__(256) void batch_stage(StageArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double R = a.prad[i], om = a.pomg[i];
  double vx, vy, vz;
  if (a.first) {
    vx = a.u[i];
    vy = a.v[i];
    vz = a.w[i];
  } else {
    vx = a.kx[i];
    vy = a.ky[i];
    vz = a.kz[i];
  }
  const double q = -a.dt * 2.0 * om * vz / R;
  const double ka = a.dt * vx, kb = a.dt * vy, kc = a.dt * vz;
  a.kx[i] = ka;
  a.ky[i] = kb;
  a.kz[i] = kc;
  a.kq[i] = q;
  if (a.make_input) {
    double px = a.x[i] + a.h * ka, py = a.y[i] + a.h * kb, pz = a.z[i] + a.h * kc;
    if (a.proj) project(px, py, pz, R);
    a.xi[i] = px;
    a.yi[i] = py;
    a.zi[i] = pz;
    a.qi[i] = a.q[i] + a.h * q;
  }
}

struct CombineArgs {
  int n, proj;
  const double* prad;
  double *x, *y, *z, *q;
  const double* k[4][4];  // [component][stage]
};

__global__ __launch_bounds__(256) void batch_combine(CombineArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double* st[4] = {a.x, a.y, a.z, a.q};
  double v[4];
  for (int c = 0; c < 4; ++c)
    v[c] = st[c][i] + a.k[c][0][i] / 6.0 + a.k[c][1][i] / 3.0 + a.k[c][2][i] / 3.0 + a.k[c][3][i] / 6.0;
  if (a.proj) project(v[0], v[1], v[2], a.prad[i]);
  for (int c = 0; c < 4; ++c) st[c][i] = v[c];
}

struct BatchFinArgs {
  int n;
  const int* pinst;
  const lpm::BatchProblem* probs;
  const double* partial;
  const double *x, *y, *z;
  double *u, *v, *w;
};

__global__ __launch_bounds__(256) void batch_finalize(BatchFinArgs f) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= f.n) return;
  const lpm::BatchProblem pr = f.probs[f.pinst[p]];
  const int i = p - pr.t_off;
  double A0 = 0.0, A1 = 0.0, A2 = 0.0;
  for (int s = 0; s < pr.nsplit; ++s) {
    const double* ps = f.partial + pr.part_off + (size_t)s * 3 * pr.nt;
    A0 += ps[i];
    A1 += ps[(size_t)pr.nt + i];
    A2 += ps[(size_t)2 * pr.nt + i];
  }
  const double x = f.x[p], y = f.y[p], z = f.z[p];
  f.u[p] = A1 * z - A2 * y;
  f.v[p] = A2 * x - A0 * z;
  f.w[p] = A0 * y - A1 * x;
}

using BatchFn = void (*)(const void*, const void*, double*, const lpm::BatchProblem*, const lpm::BatchItem*, int,
                         int, unsigned int*);

template <typename T, bool USE_E>
BatchFn pick_batch_r(int r) {
  constexpr bool C = sizeof(T) == 4;
  using P = lpm::SphereBS<T, USE_E>;
  switch (r) {
    case 2: return lpm::pair_batch_kernel<P, T, 2, C>;
    case 8: return lpm::pair_batch_kernel<P, T, 8, C>;
    default: return lpm::pair_batch_kernel<P, T, 4, C>;
  }
}

BatchFn pick_batch(int prec, bool on_sphere, int r) {
  if (prec == LPM_PREC_FP32_KAHAN)
    return on_sphere ? pick_batch_r<float, false>(r) : pick_batch_r<float, true>(r);
  return on_sphere ? pick_batch_r<double, false>(r) : pick_batch_r<double, true>(r);
}

}  // namespace

struct lpm_bve_batch {
  lpm_ctx* ctx = nullptr;
  int n_inst = 0, n = 0, ns = 0, precision = 0, flags = 0;
  int r = 4, kt = 256, chunk = 16384;
  std::vector<int64_t> off;
  std::vector<int> src_host;  // global active indices
  DevBuf x, y, z, q, u, v, w, area, prad, pomg, pinst, list;
  DevBuf k[4][4];  // [stage][x, y, z, q]
  DevBuf xi, yi, zi, qi;
  DevBuf tgt, src, partial, probs, items, counter;
  int n_items = 0;
  int grid = 0;
  hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraph_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  double graph_dt = 0.0;
  int64_t pairs_per_eval = 0;
};

namespace {

void drop_graphs(lpm_bve_batch* b) {
  for (int g = 0; g < 4; ++g) {
    if (b->exec[g]) (void)hipGraphExecDestroy(b->exec[g]);
    if (b->graph[g]) (void)hipGraphDestroy(b->graph[g]);
    b->exec[g] = nullptr;
    b->graph[g] = nullptr;
  }
}

// (Re)build the problem table and work list for the current tile shape.
int plan(lpm_bve_batch* b) {
  lpm_ctx* ctx = b->ctx;
  const int per_wave = 64 * b->r;
  const size_t wsz = b->precision == LPM_PREC_FP32_KAHAN ? 4 : 8;
  std::vector<lpm::BatchProblem> probs((size_t)b->n_inst);
  struct It {
    lpm::BatchItem it;
    int cost;
  };
  std::vector<It> items;
  long long part = 0;
  int s_off = 0;
  for (int k = 0; k < b->n_inst; ++k) {
    lpm::BatchProblem& p = probs[(size_t)k];
    p.t_off = (int)b->off[(size_t)k];
    p.nt = (int)(b->off[(size_t)k + 1] - b->off[(size_t)k]);
    int nsrc = 0;
    for (int j = s_off; j < b->ns && b->src_host[(size_t)j] < b->off[(size_t)k + 1]; ++j) ++nsrc;
    p.s_off = s_off;
    p.ns = nsrc;
    s_off += nsrc;
    int chunk = std::max(b->kt, (b->chunk + b->kt - 1) / b->kt * b->kt);
    p.nsplit = nsrc > 0 ? (nsrc + chunk - 1) / chunk : 0;
    p.chunk = chunk;
    p.part_off = part;
    part += (long long)std::max(p.nsplit, 1) * 3 * p.nt;
    const int nw = (p.nt + per_wave - 1) / per_wave;
    // split-major: the waves that take consecutive items read the same source
    // chunk, so it stays in L2 while they run
    for (int s = 0; s < p.nsplit; ++s)
      for (int t = 0; t < nw; ++t) {
        const int cost = std::min(chunk, nsrc - s * chunk);
        items.push_back({{k, t, s, 0}, cost});
      }
  }
  // longest work first, so the queue's tail holds the short items
  std::stable_sort(items.begin(), items.end(), [](const It& a, const It& c) { return a.cost > c.cost; });
  std::vector<lpm::BatchItem> flat(items.size());
  for (size_t i = 0; i < items.size(); ++i) flat[i] = items[i].it;
  b->n_items = (int)flat.size();
  LPM_HIP(ctx, b->probs.ensure(probs.size() * sizeof(lpm::BatchProblem)));
  LPM_HIP(ctx, b->items.ensure(std::max<size_t>(flat.size(), 1) * sizeof(lpm::BatchItem)));
  LPM_HIP(ctx, b->partial.ensure((size_t)std::max<long long>(part, 1) * sizeof(double)));
  LPM_HIP(ctx, b->tgt.ensure((size_t)b->n * lpm::TREC * wsz));
  LPM_HIP(ctx, b->src.ensure((size_t)std::max(b->ns, 1) * lpm::SREC * wsz));
  LPM_HIP(ctx, b->counter.ensure(256));
  LPM_HIP(ctx, hipMemcpyAsync(b->probs.p, probs.data(), probs.size() * sizeof(lpm::BatchProblem),
                              hipMemcpyHostToDevice, ctx->stream));
  if (!flat.empty())
    LPM_HIP(ctx, hipMemcpyAsync(b->items.p, flat.data(), flat.size() * sizeof(lpm::BatchItem),
                                hipMemcpyHostToDevice, ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const bool on_sphere = (b->flags & LPM_SOLVER_PROJECT_STAGES) != 0;
  BatchFn fn = pick_batch(b->precision, on_sphere, b->r);
  b->grid = lpm::blocks_per_cu((const void*)fn) * ctx->num_cus;
  drop_graphs(b);
  return LPM_OK;
}

// One velocity evaluation of every problem at (x, y, z, q) into (u, v, w).
int eval_velocity(lpm_bve_batch* b, const double* x, const double* y, const double* z, const double* q,
                  double* u, double* v, double* w) {
  lpm_ctx* ctx = b->ctx;
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, ctx->stream, b->counter.as<unsigned int>());
  lpm::PackArgs pa{x, y, z, q, nullptr, b->area.as<double>(), b->prad.as<double>(), b->list.as<int>(),
                   0, 0, 1.0, 1.0};
  lpm::launch_pack(LPM_K_SPHERE_BS, b->precision, pa, b->src.p, b->tgt.p, b->ns, b->n, ctx->stream);
  const bool on_sphere = (b->flags & LPM_SOLVER_PROJECT_STAGES) != 0;
  lpm::BatchArgs ba;
  ba.tgt = b->tgt.p;
  ba.src = b->src.p;
  ba.partial = b->partial.as<double>();
  ba.probs = b->probs.as<lpm::BatchProblem>();
  ba.items = b->items.as<lpm::BatchItem>();
  ba.n_items = b->n_items;
  ba.kt = b->kt;
  ba.counter = b->counter.as<unsigned int>();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing) {
    e0 = lpm::take_event(ctx);
    e1 = lpm::take_event(ctx);
    if (e0 && e1) LPM_HIP(ctx, hipEventRecord(e0, ctx->stream));
  }
  if (b->n_items > 0)
    hipLaunchKernelGGL(pick_batch(b->precision, on_sphere, b->r), dim3((unsigned)b->grid), dim3(256), 0,
                       ctx->stream, ba.tgt, ba.src, ba.partial, ba.probs, ba.items, ba.n_items, ba.kt, ba.counter);
  if (ctx->timing && e0 && e1) {
    LPM_HIP(ctx, hipEventRecord(e1, ctx->stream));
    ctx->pending.push_back({e0, e1, LPM_K_SPHERE_BS});
  }
  BatchFinArgs f{b->n, b->pinst.as<int>(), b->probs.as<lpm::BatchProblem>(), b->partial.as<double>(), x, y, z,
                 u, v, w};
  hipLaunchKernelGGL(batch_finalize, dim3((b->n + 255) / 256), dim3(256), 0, ctx->stream, f);
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

// RK4 stage g (0..3) of a step; g = 3 ends with the new velocity.
int enqueue_stage(lpm_bve_batch* b, int g, double dt) {
  lpm_ctx* ctx = b->ctx;
  const int n = b->n;
  dim3 blk(256), grd((n + 255) / 256);
  const int proj = (b->flags & LPM_SOLVER_PROJECT_STAGES) ? 1 : 0;
  auto K = [&](int s, int c) { return b->k[s][c].as<double>(); };
  StageArgs a;
  a.n = n;
  a.dt = dt;
  a.proj = proj;
  a.prad = b->prad.as<double>();
  a.pomg = b->pomg.as<double>();
  a.x = b->x.as<double>();
  a.y = b->y.as<double>();
  a.z = b->z.as<double>();
  a.q = b->q.as<double>();
  a.u = b->u.as<double>();
  a.v = b->v.as<double>();
  a.w = b->w.as<double>();
  a.xi = b->xi.as<double>();
  a.yi = b->yi.as<double>();
  a.zi = b->zi.as<double>();
  a.qi = b->qi.as<double>();
  if (g == 0) {  // stage 1 algebra from the stored velocity
    a.first = 1;
    a.make_input = 1;
    a.h = 0.5;
    a.kx = K(0, 0);
    a.ky = K(0, 1);
    a.kz = K(0, 2);
    a.kq = K(0, 3);
    hipLaunchKernelGGL(batch_stage, grd, blk, 0, ctx->stream, a);
  }
  if (g <= 2) {  // stage g+2 velocity at the stage input, then its algebra
    const int s = g + 1;
    if (int rc = eval_velocity(b, b->xi.as<double>(), b->yi.as<double>(), b->zi.as<double>(), b->qi.as<double>(),
                               K(s, 0), K(s, 1), K(s, 2)))
      return rc;
    a.first = 0;
    a.make_input = s < 3 ? 1 : 0;
    a.h = s == 1 ? 0.5 : 1.0;
    a.kx = K(s, 0);
    a.ky = K(s, 1);
    a.kz = K(s, 2);
    a.kq = K(s, 3);
    hipLaunchKernelGGL(batch_stage, grd, blk, 0, ctx->stream, a);
  }
  if (g == 2) {
    CombineArgs c;
    c.n = n;
    c.proj = proj;
    c.prad = b->prad.as<double>();
    c.x = b->x.as<double>();
    c.y = b->y.as<double>();
    c.z = b->z.as<double>();
    c.q = b->q.as<double>();
    for (int comp = 0; comp < 4; ++comp)
      for (int s = 0; s < 4; ++s) c.k[comp][s] = K(s, comp);
    hipLaunchKernelGGL(batch_combine, grd, blk, 0, ctx->stream, c);
  }
  if (g == 3)
    if (int rc = eval_velocity(b, b->x.as<double>(), b->y.as<double>(), b->z.as<double>(), b->q.as<double>(),
                               b->u.as<double>(), b->v.as<double>(), b->w.as<double>()))
      return rc;
  LPM_HIP(ctx, hipGetLastError());
  return LPM_OK;
}

int batch_step(lpm_bve_batch* b, double dt) {
  lpm_ctx* ctx = b->ctx;
  const bool use_graph = (b->flags & LPM_SOLVER_USE_GRAPH) && !ctx->timing;
  if (!use_graph) {
    for (int g = 0; g < 4; ++g)
      if (int rc = enqueue_stage(b, g, dt)) return rc;
    return LPM_OK;
  }
  if (!b->exec[0] || b->graph_dt != dt) {
    drop_graphs(b);
    for (int g = 0; g < 4; ++g) {
      LPM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_stage(b, g, dt);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        drop_graphs(b);
        return rc;
      }
      if (e != hipSuccess) return lpm::hip_fail(ctx, e, "hipStreamEndCapture");
      b->graph[g] = gr;
      LPM_HIP(ctx, hipGraphInstantiate(&b->exec[g], gr, nullptr, nullptr, 0));
    }
    b->graph_dt = dt;
  }
  for (int g = 0; g < 4; ++g) LPM_HIP(ctx, hipGraphLaunch(b->exec[g], ctx->stream));
  return LPM_OK;
}

int upload(lpm_ctx* ctx, DevBuf& b, const void* host, size_t bytes) {
  LPM_HIP(ctx, b.ensure(bytes));
  if (host && bytes) LPM_HIP(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  return LPM_OK;
}

}  // namespace

extern "C" {

int lpm_bve_batch_new(lpm_ctx* ctx, lpm_bve_batch** out, int n_inst, const int64_t* offsets, const double* x,
                      const double* y, const double* z, const double* relvort, const double* area,
                      const int32_t* active, const double* radius, const double* rotation_rate, int precision,
                      int flags) {
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (n_inst <= 0 || !offsets) return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: need n_inst > 0 and offsets");
  if (offsets[0] != 0) return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: offsets[0] must be 0");
  for (int k = 0; k < n_inst; ++k)
    if (offsets[k + 1] <= offsets[k])
      return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: offsets must increase (empty problem " + std::to_string(k) + ")");
  if (offsets[n_inst] > (int64_t)INT32_MAX / 2) return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: too many particles");
  if (!x || !y || !z || !relvort || !area || !active || !radius || !rotation_rate)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: null array argument");
  if (precision != LPM_PREC_FP64 && precision != LPM_PREC_FP32_KAHAN)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: bad precision");
  for (int k = 0; k < n_inst; ++k)
    if (!(radius[k] > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_new: radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  lpm_bve_batch* b = new lpm_bve_batch();
  b->ctx = ctx;
  b->n_inst = n_inst;
  b->n = (int)offsets[n_inst];
  b->precision = precision;
  b->flags = flags;
  if (precision == LPM_PREC_FP64) {
    b->r = 2;
    b->kt = 64;
  }
  b->off.assign(offsets, offsets + n_inst + 1);
  const int n = b->n;
  std::vector<double> prad((size_t)n), pomg((size_t)n);
  std::vector<int> pinst((size_t)n);
  for (int k = 0; k < n_inst; ++k)
    for (int64_t p = offsets[k]; p < offsets[k + 1]; ++p) {
      prad[(size_t)p] = radius[k];
      pomg[(size_t)p] = rotation_rate[k];
      pinst[(size_t)p] = k;
    }
  for (int p = 0; p < n; ++p)
    if (active[p]) b->src_host.push_back(p);
  b->ns = (int)b->src_host.size();
  const size_t nb = (size_t)n * sizeof(double);
  int rc = LPM_OK;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
    return rc == LPM_OK;
  };
  chk(upload(ctx, b->x, x, nb)) && chk(upload(ctx, b->y, y, nb)) && chk(upload(ctx, b->z, z, nb)) &&
      chk(upload(ctx, b->q, relvort, nb)) && chk(upload(ctx, b->area, area, nb)) &&
      chk(upload(ctx, b->u, nullptr, nb)) && chk(upload(ctx, b->v, nullptr, nb)) &&
      chk(upload(ctx, b->w, nullptr, nb)) && chk(upload(ctx, b->prad, prad.data(), nb)) &&
      chk(upload(ctx, b->pomg, pomg.data(), nb)) && chk(upload(ctx, b->pinst, pinst.data(), (size_t)n * 4)) &&
      chk(upload(ctx, b->list, b->src_host.data(), std::max<size_t>(b->src_host.size(), 1) * 4)) &&
      chk(upload(ctx, b->xi, nullptr, nb)) && chk(upload(ctx, b->yi, nullptr, nb)) &&
      chk(upload(ctx, b->zi, nullptr, nb)) && chk(upload(ctx, b->qi, nullptr, nb));
  for (int s = 0; s < 4 && rc == LPM_OK; ++s)
    for (int c = 0; c < 4 && rc == LPM_OK; ++c) chk(upload(ctx, b->k[s][c], nullptr, nb));
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  if (rc == LPM_OK) chk(plan(b));
  int64_t pairs = 0;
  for (int k = 0; k < n_inst && rc == LPM_OK; ++k) {
    int64_t nt = offsets[k + 1] - offsets[k], na = 0;
    for (int64_t p = offsets[k]; p < offsets[k + 1]; ++p) na += active[p] != 0;
    pairs += nt * na - na;
  }
  b->pairs_per_eval = pairs;
  if (rc == LPM_OK && (flags & LPM_SOLVER_PROJECT_STAGES)) {
    // start on the sphere: reuse the combine kernel with zero increments
    CombineArgs c{};
    c.n = n;
    c.proj = 1;
    c.prad = b->prad.as<double>();
    c.x = b->x.as<double>();
    c.y = b->y.as<double>();
    c.z = b->z.as<double>();
    c.q = b->q.as<double>();
    for (int comp = 0; comp < 4; ++comp)
      for (int s = 0; s < 4; ++s) c.k[comp][s] = b->k[s][comp].as<double>();
    for (int s = 0; s < 4 && rc == LPM_OK; ++s)
      for (int comp = 0; comp < 4 && rc == LPM_OK; ++comp) {
        hipError_t e = hipMemsetAsync(b->k[s][comp].p, 0, nb, ctx->stream);
        if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipMemsetAsync");
      }
    if (rc == LPM_OK) hipLaunchKernelGGL(batch_combine, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, c);
  }
  if (rc == LPM_OK)
    chk(eval_velocity(b, b->x.as<double>(), b->y.as<double>(), b->z.as<double>(), b->q.as<double>(),
                      b->u.as<double>(), b->v.as<double>(), b->w.as<double>()));
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  if (rc != LPM_OK) {
    lpm_bve_batch_delete(b);
    return rc;
  }
  *out = b;
  return lpm::collect_timing(ctx);
}

int lpm_bve_batch_steps(lpm_bve_batch* b, double dt, int nsteps) {
  if (!b) return LPM_ERR_ARG;
  lpm_ctx* ctx = b->ctx;
  ctx->err.clear();
  if (nsteps < 0) return fail(ctx, LPM_ERR_ARG, "nsteps must be >= 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  for (int s = 0; s < nsteps; ++s)
    if (int rc = batch_step(b, dt)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_batch_get_state(lpm_bve_batch* b, double* x, double* y, double* z, double* relvort, double* u,
                            double* v, double* w) {
  if (!b) return LPM_ERR_ARG;
  lpm_ctx* ctx = b->ctx;
  ctx->err.clear();
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  const DevBuf* src[7] = {&b->x, &b->y, &b->z, &b->q, &b->u, &b->v, &b->w};
  double* dst[7] = {x, y, z, relvort, u, v, w};
  for (int k = 0; k < 7; ++k)
    if (dst[k])
      LPM_HIP(ctx, hipMemcpyAsync(dst[k], src[k]->p, (size_t)b->n * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_batch_set_tuning(lpm_bve_batch* b, int targets_per_lane, int tile_sources, int chunk_sources) {
  if (!b) return LPM_ERR_ARG;
  lpm_ctx* ctx = b->ctx;
  ctx->err.clear();
  if (!(targets_per_lane == 2 || targets_per_lane == 4 || targets_per_lane == 8))
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_set_tuning: targets_per_lane must be 2, 4 or 8");
  if (tile_sources < 8 || tile_sources % 8 || tile_sources > 65536)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_set_tuning: tile_sources must be a multiple of 8 in [8, 65536]");
  if (chunk_sources < tile_sources) return fail(ctx, LPM_ERR_ARG, "lpm_bve_batch_set_tuning: chunk < tile");
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  b->r = targets_per_lane;
  b->kt = tile_sources;
  b->chunk = chunk_sources;
  return plan(b);
}

int lpm_bve_batch_get_tuning(lpm_bve_batch* b, int* targets_per_lane, int* tile_sources, int* chunk_sources,
                             int* n_items) {
  if (!b || !targets_per_lane || !tile_sources || !chunk_sources || !n_items) return LPM_ERR_ARG;
  *targets_per_lane = b->r;
  *tile_sources = b->kt;
  *chunk_sources = b->chunk;
  *n_items = b->n_items;
  return LPM_OK;
}

int lpm_bve_batch_pairs(lpm_bve_batch* b, int64_t* pairs_per_eval) {
  if (!b || !pairs_per_eval) return LPM_ERR_ARG;
  *pairs_per_eval = b->pairs_per_eval;
  return LPM_OK;
}

int lpm_bve_batch_autotune(lpm_bve_batch* b, int* targets_per_lane, int* tile_sources, int* chunk_sources,
                           double* best_ms) {
  if (!b) return LPM_ERR_ARG;
  lpm_ctx* ctx = b->ctx;
  ctx->err.clear();
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  // Time one velocity evaluation per candidate tile shape (state untouched:
  // the evaluation writes the stage-2 slot, which the next step overwrites).
  static const int rs[3] = {2, 4, 8};
  static const int kts[2] = {128, 256};
  static const int chunks[3] = {4096, 16384, 65536};
  int best_r = b->r, best_kt = b->kt, best_chunk = b->chunk;
  float best = 1e30f;
  hipEvent_t e0, e1;
  LPM_HIP(ctx, hipEventCreate(&e0));
  LPM_HIP(ctx, hipEventCreate(&e1));
  const bool was_timing = ctx->timing;
  ctx->timing = false;
  int rc = LPM_OK;
  for (int r : rs)
    for (int kt : kts)
      for (int ch : chunks) {
        if (b->precision == LPM_PREC_FP64 && r == 8) continue;
        b->r = r;
        b->kt = kt;
        b->chunk = ch;
        if ((rc = plan(b))) break;
        for (int rep = 0; rep < 2 && rc == LPM_OK; ++rep) {  // first run warms caches
          hipError_t e = hipEventRecord(e0, ctx->stream);
          if (e == hipSuccess)
            rc = eval_velocity(b, b->x.as<double>(), b->y.as<double>(), b->z.as<double>(), b->q.as<double>(),
                               b->k[1][0].as<double>(), b->k[1][1].as<double>(), b->k[1][2].as<double>());
          if (rc == LPM_OK) e = hipEventRecord(e1, ctx->stream);
          if (rc == LPM_OK && e == hipSuccess) e = hipEventSynchronize(e1);
          if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "autotune timing");
          float ms = 0.f;
          if (rc == LPM_OK && rep == 1 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms < best) {
            best = ms;
            best_r = r;
            best_kt = kt;
            best_chunk = ch;
          }
        }
      }
  ctx->timing = was_timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  b->r = best_r;
  b->kt = best_kt;
  b->chunk = best_chunk;
  if ((rc = plan(b))) return rc;
  if (targets_per_lane) *targets_per_lane = best_r;
  if (tile_sources) *tile_sources = best_kt;
  if (chunk_sources) *chunk_sources = best_chunk;
  if (best_ms) *best_ms = best;
  return LPM_OK;
}

int lpm_bve_batch_delete(lpm_bve_batch* b) {
  if (!b) return LPM_ERR_ARG;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graphs(b);
  DevBuf* all[] = {&b->x, &b->y, &b->z, &b->q, &b->u, &b->v, &b->w, &b->area, &b->prad, &b->pomg,
                   &b->pinst, &b->list, &b->xi, &b->yi, &b->zi, &b->qi, &b->tgt, &b->src, &b->partial,
                   &b->probs, &b->items, &b->counter};
  for (DevBuf* d : all) d->release();
  for (auto& row : b->k)
    for (auto& d : row) d.release();
  delete b;
  return LPM_OK;
}

}  // extern "C"
