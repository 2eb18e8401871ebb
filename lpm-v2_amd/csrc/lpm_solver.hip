// lpm_solver.hip -- device-resident RK4 solvers (sphere BVE, planar
// incompressible).  State lives in HBM for the solver's lifetime; a step is
// 4 pair sums plus FP64 stage algebra, all on the context stream, optionally
// replayed as one captured hipGraph per step.
//
// Stage algebra restates timestepPrivate (SphereBVESolver.f90:219-353;
// PlaneIncompressibleSolver.f90:171-259) operation for operation in FP64
// with contraction off, so given equal velocities the update is bitwise the
// reference's.
#include <string>
#include <vector>

#include "lpm_internal.hpp"

using lpm::DevBuf;
using lpm::fail;

namespace {

#pragma clang fp contract(off)

// x <- R x / |x| in FP64 (used with LPM_SOLVER_PROJECT_STAGES).
__device__ __forceinline__ void project(double& x, double& y, double& z, double radius) {
  const double s = radius / sqrt(x * x + y * y + z * z);
  x = x * s;
  y = y * s;
  z = z * s;
}

// Stage 1 (SphereBVESolver.f90:238-244): k1 = dt u, dzeta1 = -dt 2 Omega w / R,
// and the stage-2 input x + 0.5 k1 (:249-255).
__global__ __launch_bounds__(256) void bve_stage1(int n, double dt, double omega, double radius,
                                                  int proj, const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ z,
                                                  const double* __restrict__ zeta,
                                                  const double* __restrict__ u,
                                                  const double* __restrict__ v,
                                                  const double* __restrict__ w, double* kx,
                                                  double* ky, double* kz, double* kq, double* xi,
                                                  double* yi, double* zi, double* qi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = dt * u[i], b = dt * v[i], c = dt * w[i];
  const double q = -dt * 2.0 * omega * w[i] / radius;
  kx[i] = a;
  ky[i] = b;
  kz[i] = c;
  kq[i] = q;
  double px = x[i] + 0.5 * a, py = y[i] + 0.5 * b, pz = z[i] + 0.5 * c;
  if (proj) project(px, py, pz, radius);
  xi[i] = px;
  yi[i] = py;
  zi[i] = pz;
  qi[i] = zeta[i] + 0.5 * q;
}

// After a stage's velocity sum (:261-267): dzeta = -dt 2 Omega w / R, k = dt u,
// then the next stage input x + h k (h = 0.5 or 1; :272-301).
__global__ __launch_bounds__(256) void bve_stage_k(int n, double dt, double omega, double radius,
                                                   double h, int make_input, int proj,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ zeta, double* kx,
                                                   double* ky, double* kz, double* kq, double* xi,
                                                   double* yi, double* zi, double* qi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double q = -dt * 2.0 * omega * kz[i] / radius;
  const double a = dt * kx[i], b = dt * ky[i], c = dt * kz[i];
  kq[i] = q;
  kx[i] = a;
  ky[i] = b;
  kz[i] = c;
  if (make_input) {
    double px = x[i] + h * a, py = y[i] + h * b, pz = z[i] + h * c;
    if (proj) project(px, py, pz, radius);
    xi[i] = px;
    yi[i] = py;
    zi[i] = pz;
    qi[i] = zeta[i] + h * q;
  }
}

// RK4 combine (:319-329): x += k1/6 + k2/3 + k3/3 + k4/6, left to right;
// with proj the first three components (a sphere position) are projected.
__global__ __launch_bounds__(256) void rk4_combine(int n, int ncomp, int proj, double radius,
                                                   double* const* __restrict__ st,
                                                   const double* const* __restrict__ k) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[4];
  for (int c = 0; c < ncomp; ++c) {
    const double* s = st[c];
    v[c] = s[i] + k[c * 4 + 0][i] / 6.0 + k[c * 4 + 1][i] / 3.0 + k[c * 4 + 2][i] / 3.0 +
           k[c * 4 + 3][i] / 6.0;
  }
  if (proj) project(v[0], v[1], v[2], radius);
  for (int c = 0; c < ncomp; ++c) st[c][i] = v[c];
}

__global__ __launch_bounds__(256) void project_positions(int n, double radius, double* x, double* y,
                                                         double* z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a = x[i], b = y[i], c = z[i];
  project(a, b, c, radius);
  x[i] = a;
  y[i] = b;
  z[i] = c;
}

// Planar stage 1 (PlaneIncompressibleSolver.f90:183-194).
__global__ __launch_bounds__(256) void plane_stage1(int n, double dt, const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ v, double* kx,
                                                    double* ky, double* xi, double* yi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = dt * u[i], b = dt * v[i];
  kx[i] = a;
  ky[i] = b;
  xi[i] = x[i] + 0.5 * a;
  yi[i] = y[i] + 0.5 * b;
}

__global__ __launch_bounds__(256) void plane_stage_k(int n, double dt, double h, int make_input,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ y, double* kx,
                                                     double* ky, double* xi, double* yi) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = dt * kx[i], b = dt * ky[i];
  kx[i] = a;
  ky[i] = b;
  if (make_input) {
    xi[i] = x[i] + h * a;
    yi[i] = y[i] + h * b;
  }
}

int upload_array(lpm_ctx* ctx, DevBuf& b, const void* host, size_t bytes) {
  LPM_HIP(ctx, b.ensure(bytes));
  if (host && bytes)
    LPM_HIP(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  return LPM_OK;
}

int download_array(lpm_ctx* ctx, double* host, const DevBuf& b, int n) {
  if (!host || n <= 0) return LPM_OK;
  LPM_HIP(ctx, hipMemcpyAsync(host, b.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  return LPM_OK;
}

}  // namespace

// =================================================================== sphere
struct lpm_bve_solver {
  lpm_ctx* ctx = nullptr;
  int n = 0, ns = 0, precision = 0, flags = 0;
  double radius = 1.0, omega = 0.0;
  DevBuf x, y, z, zeta, absv, u, v, w, area, active, list;
  DevBuf k[4][4];  // [stage][x, y, z, zeta]
  DevBuf xi, yi, zi, qi;
  DevBuf psi_rel, psi_abs;
  DevBuf tgt, src, partial;
  DevBuf tot;   // FP32 stream functions: FP64 source totals
  DevBuf ptrs;  // device pointer tables for rk4_combine
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0.0;
  // addresses of the workspaces a captured step uses, as captured; an on-demand
  // stream-function call may regrow them, and then the step is captured again
  void* graph_bufs[6] = {};
  bool has_abs = false;
  bool loaded = false;  // false after a failed particle reload
};

namespace {

int bve_velocity(lpm_bve_solver* s, const double* x, const double* y, const double* z,
                 const double* q, double* u, double* v, double* w) {
  lpm::PairProblem pp;
  pp.kind = LPM_K_SPHERE_BS;
  pp.precision = s->precision;
  pp.n = s->n;
  pp.i0 = 0;
  pp.i1 = s->n;
  pp.x = x;
  pp.y = y;
  pp.z = z;
  pp.q0 = q;
  pp.area = s->area.as<double>();
  pp.src_list = s->list.as<int>();
  pp.ns = s->ns;
  pp.radius = s->radius;
  pp.on_sphere = (s->flags & LPM_SOLVER_PROJECT_STAGES) != 0;
  pp.out0 = u;
  pp.out1 = v;
  pp.out2 = w;
  pp.tgt_buf = &s->tgt;
  pp.src_buf = &s->src;
  pp.partial_buf = &s->partial;
  return lpm::run_pair_sum(s->ctx, pp);
}

int bve_stream(lpm_bve_solver* s) {
  lpm_ctx* ctx = s->ctx;
  if (!s->has_abs) return fail(ctx, LPM_ERR_STATE, "stream functions need absvort at lpm_bve_solver_new");
  LPM_HIP(ctx, s->psi_rel.ensure((size_t)s->n * sizeof(double)));
  LPM_HIP(ctx, s->psi_abs.ensure((size_t)s->n * sizeof(double)));
  lpm::PairProblem pp;
  pp.kind = LPM_K_SPHERE_GREEN;
  pp.precision = s->precision;
  pp.n = s->n;
  pp.i0 = 0;
  pp.i1 = s->n;
  pp.x = s->x.as<double>();
  pp.y = s->y.as<double>();
  pp.z = s->z.as<double>();
  pp.q0 = s->zeta.as<double>();
  pp.q1 = s->absv.as<double>();
  pp.area = s->area.as<double>();
  pp.active = s->active.as<int>();
  pp.src_list = s->list.as<int>();
  pp.ns = s->ns;
  pp.radius = s->radius;
  pp.on_sphere = (s->flags & LPM_SOLVER_PROJECT_STAGES) != 0;
  pp.out0 = s->psi_rel.as<double>();
  pp.out1 = s->psi_abs.as<double>();
  pp.tgt_buf = &s->tgt;
  pp.src_buf = &s->src;
  pp.partial_buf = &s->partial;
  pp.red_buf = &s->tot;
  return lpm::run_pair_sum(ctx, pp);
}

// Enqueue one RK4 step (no synchronization).
int bve_enqueue_step(lpm_bve_solver* s, double dt) {
  lpm_ctx* ctx = s->ctx;
  const int n = s->n;
  dim3 b(256), g((n + 255) / 256);
  double* K[4][4];
  for (int a = 0; a < 4; ++a)
    for (int c = 0; c < 4; ++c) K[a][c] = s->k[a][c].as<double>();
  const int proj = (s->flags & LPM_SOLVER_PROJECT_STAGES) ? 1 : 0;
  hipLaunchKernelGGL(bve_stage1, g, b, 0, ctx->stream, n, dt, s->omega, s->radius, proj, s->x.as<double>(),
                     s->y.as<double>(), s->z.as<double>(), s->zeta.as<double>(), s->u.as<double>(),
                     s->v.as<double>(), s->w.as<double>(), K[0][0], K[0][1], K[0][2], K[0][3],
                     s->xi.as<double>(), s->yi.as<double>(), s->zi.as<double>(), s->qi.as<double>());
  static const double h_next[4] = {0.0, 0.5, 1.0, 0.0};
  for (int st = 1; st < 4; ++st) {
    if (int rc = bve_velocity(s, s->xi.as<double>(), s->yi.as<double>(), s->zi.as<double>(),
                              s->qi.as<double>(), K[st][0], K[st][1], K[st][2]))
      return rc;
    hipLaunchKernelGGL(bve_stage_k, g, b, 0, ctx->stream, n, dt, s->omega, s->radius, h_next[st],
                       st < 3 ? 1 : 0, proj, s->x.as<double>(), s->y.as<double>(), s->z.as<double>(),
                       s->zeta.as<double>(), K[st][0], K[st][1], K[st][2], K[st][3],
                       s->xi.as<double>(), s->yi.as<double>(), s->zi.as<double>(), s->qi.as<double>());
  }
  hipLaunchKernelGGL(rk4_combine, g, b, 0, ctx->stream, n, 4, proj, s->radius, (double* const*)s->ptrs.p,
                     (const double* const*)((double**)s->ptrs.p + 4));
  LPM_HIP(ctx, hipGetLastError());
  if (int rc = bve_velocity(s, s->x.as<double>(), s->y.as<double>(), s->z.as<double>(),
                            s->zeta.as<double>(), s->u.as<double>(), s->v.as<double>(), s->w.as<double>()))
    return rc;
  if ((s->flags & LPM_SOLVER_STREAM_EACH_STEP) && s->has_abs)
    if (int rc = bve_stream(s)) return rc;
  return LPM_OK;
}

// The workspaces a step grows on demand (the state and stage arrays change
// only in bve_load, which drops the graph).
void bve_work_bufs(const lpm_bve_solver* s, void* out[6]) {
  const DevBuf* b[6] = {&s->tgt, &s->src, &s->partial, &s->psi_rel, &s->psi_abs, &s->tot};
  for (int k = 0; k < 6; ++k) out[k] = b[k]->p;
}

template <class Solver, class BufsFn>
bool graph_bufs_unchanged(const Solver* s, BufsFn bufs) {
  void* now[6];
  bufs(s, now);
  for (int k = 0; k < 6; ++k)
    if (now[k] != s->graph_bufs[k]) return false;
  return true;
}

int bve_step(lpm_bve_solver* s, double dt) {
  lpm_ctx* ctx = s->ctx;
  const bool use_graph = (s->flags & LPM_SOLVER_USE_GRAPH) && !ctx->timing;
  if (!use_graph) return bve_enqueue_step(s, dt);
  if (s->exec && s->graph_dt == dt && graph_bufs_unchanged(s, bve_work_bufs)) {
    LPM_HIP(ctx, hipGraphLaunch(s->exec, ctx->stream));
    return LPM_OK;
  }
  if (s->exec) {
    (void)hipGraphExecDestroy(s->exec);
    s->exec = nullptr;
  }
  if (s->graph) {
    (void)hipGraphDestroy(s->graph);
    s->graph = nullptr;
  }
  // eager first step: allocates workspaces and fills the occupancy cache
  if (int rc = bve_enqueue_step(s, dt)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  LPM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = bve_enqueue_step(s, dt);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return lpm::hip_fail(ctx, e, "hipStreamEndCapture");
  s->graph = gr;
  LPM_HIP(ctx, hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
  s->graph_dt = dt;
  bve_work_bufs(s, s->graph_bufs);
  return LPM_OK;
}

// A captured step holds the sizes and buffer addresses of its capture.
template <class Solver>
void drop_graph(Solver* s) {
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->exec = nullptr;
  s->graph = nullptr;
}

// Check the particle arrays of lpm_bve_solver_new / _set_particles.
int bve_check_particles(lpm_ctx* ctx, const char* who, int n, const double* x, const double* y,
                        const double* z, const double* relvort, const double* area, const int32_t* active,
                        const double* u, const double* v, const double* w) {
  const std::string nm = who;
  if (n <= 0) return fail(ctx, LPM_ERR_ARG, nm + ": n must be > 0");
  if (!x || !y || !z || !relvort || !area || !active) return fail(ctx, LPM_ERR_ARG, nm + ": null array argument");
  if ((u || v || w) && !(u && v && w)) return fail(ctx, LPM_ERR_ARG, nm + ": give all of u, v, w or none");
  return LPM_OK;
}

// Load a particle set (at creation, and after a remesh): state arrays,
// stage workspaces (grown, never shrunk), the active-source list, the
// pointer tables of the RK4 combine, and the velocity when none is given
// (SetVelocityOnMesh after a remesh, SsrfpackRemesh.f90:552).
int bve_load(lpm_bve_solver* s, int n, const double* x, const double* y, const double* z, const double* relvort,
             const double* absvort, const double* area, const int32_t* active, const double* u, const double* v,
             const double* w) {
  lpm_ctx* ctx = s->ctx;
  drop_graph(s);
  s->loaded = false;
  s->n = n;
  s->has_abs = absvort != nullptr;
  const size_t nb = (size_t)n * sizeof(double);
  int rc = LPM_OK;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
    return rc == LPM_OK;
  };
  chk(upload_array(ctx, s->x, x, nb)) && chk(upload_array(ctx, s->y, y, nb)) &&
      chk(upload_array(ctx, s->z, z, nb)) && chk(upload_array(ctx, s->zeta, relvort, nb)) &&
      chk(upload_array(ctx, s->area, area, nb)) && chk(upload_array(ctx, s->u, u, nb)) &&
      chk(upload_array(ctx, s->v, v, nb)) && chk(upload_array(ctx, s->w, w, nb)) &&
      chk(upload_array(ctx, s->absv, absvort, absvort ? nb : 8));
  for (int a = 0; a < 4 && rc == LPM_OK; ++a)
    for (int c = 0; c < 4 && rc == LPM_OK; ++c) chk(upload_array(ctx, s->k[a][c], nullptr, nb));
  if (rc == LPM_OK)
    chk(upload_array(ctx, s->xi, nullptr, nb)) && chk(upload_array(ctx, s->yi, nullptr, nb)) &&
        chk(upload_array(ctx, s->zi, nullptr, nb)) && chk(upload_array(ctx, s->qi, nullptr, nb));
  if (rc == LPM_OK) chk(lpm::upload_active_list(ctx, active, n, s->list, &s->active, &s->ns));
  if (rc == LPM_OK) {
    // pointer tables for rk4_combine: 4 state arrays, then k[stage] per component
    std::vector<double*> tab(4 + 16);
    tab[0] = s->x.as<double>();
    tab[1] = s->y.as<double>();
    tab[2] = s->z.as<double>();
    tab[3] = s->zeta.as<double>();
    for (int c = 0; c < 4; ++c)
      for (int a = 0; a < 4; ++a) tab[4 + c * 4 + a] = s->k[a][c].as<double>();
    chk(upload_array(ctx, s->ptrs, tab.data(), tab.size() * sizeof(double*)));
    if (rc == LPM_OK) {
      hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
    }
  }
  if (rc == LPM_OK && (s->flags & LPM_SOLVER_PROJECT_STAGES)) {
    hipLaunchKernelGGL(project_positions, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, s->radius,
                       s->x.as<double>(), s->y.as<double>(), s->z.as<double>());
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "project_positions");
  }
  if (rc == LPM_OK && !u)
    chk(bve_velocity(s, s->x.as<double>(), s->y.as<double>(), s->z.as<double>(), s->zeta.as<double>(),
                     s->u.as<double>(), s->v.as<double>(), s->w.as<double>()));
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  s->loaded = rc == LPM_OK;
  return rc;
}

int bve_usable(lpm_bve_solver* s) {
  if (s->loaded) return LPM_OK;
  return fail(s->ctx, LPM_ERR_STATE, "solver particles were not loaded (a particle reload failed)");
}

}  // namespace

extern "C" {

int lpm_bve_solver_new(lpm_ctx* ctx, lpm_bve_solver** out, int n, const double* x, const double* y,
                       const double* z, const double* relvort, const double* absvort,
                       const double* area, const int32_t* active, const double* u, const double* v,
                       const double* w, double radius, double rotation_rate, int precision,
                       int flags) {
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (int rc = bve_check_particles(ctx, "lpm_bve_solver_new", n, x, y, z, relvort, area, active, u, v, w))
    return rc;
  if (precision != LPM_PREC_FP64 && precision != LPM_PREC_FP32_KAHAN)
    return fail(ctx, LPM_ERR_ARG, "lpm_bve_solver_new: bad precision");
  if (!(radius > 0.0)) return fail(ctx, LPM_ERR_ARG, "lpm_bve_solver_new: radius must be > 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  lpm_bve_solver* s = new lpm_bve_solver();
  s->ctx = ctx;
  s->precision = precision;
  s->flags = flags;
  s->radius = radius;
  s->omega = rotation_rate;
  if (int rc = bve_load(s, n, x, y, z, relvort, absvort, area, active, u, v, w)) {
    lpm_bve_solver_delete(s);
    return rc;
  }
  *out = s;
  return lpm::collect_timing(ctx);
}

int lpm_bve_solver_set_particles(lpm_bve_solver* s, int n, const double* x, const double* y, const double* z,
                                 const double* relvort, const double* absvort, const double* area,
                                 const int32_t* active, const double* u, const double* v, const double* w) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = bve_check_particles(ctx, "lpm_bve_solver_set_particles", n, x, y, z, relvort, area, active, u, v,
                                   w))
    return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = bve_load(s, n, x, y, z, relvort, absvort, area, active, u, v, w)) return rc;
  return lpm::collect_timing(ctx);
}

int lpm_bve_solver_timestep(lpm_bve_solver* s, double dt) { return lpm_bve_solver_steps(s, dt, 1); }

int lpm_bve_solver_steps(lpm_bve_solver* s, double dt, int nsteps) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = bve_usable(s)) return rc;
  if (nsteps < 0) return fail(ctx, LPM_ERR_ARG, "nsteps must be >= 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  for (int k = 0; k < nsteps; ++k)
    if (int rc = bve_step(s, dt)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_solver_get_state(lpm_bve_solver* s, double* x, double* y, double* z, double* relvort,
                             double* u, double* v, double* w) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = bve_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  int rc = LPM_OK;
  const DevBuf* src[7] = {&s->x, &s->y, &s->z, &s->zeta, &s->u, &s->v, &s->w};
  double* dst[7] = {x, y, z, relvort, u, v, w};
  for (int k = 0; k < 7 && rc == LPM_OK; ++k) rc = download_array(ctx, dst[k], *src[k], s->n);
  if (rc) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_bve_solver_stream(lpm_bve_solver* s, double* psi_rel, double* psi_abs) {
  if (!s || !psi_rel || !psi_abs) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = bve_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = bve_stream(s)) return rc;
  if (int rc = download_array(ctx, psi_rel, s->psi_rel, s->n)) return rc;
  if (int rc = download_array(ctx, psi_abs, s->psi_abs, s->n)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_bve_solver_diagnostics(lpm_bve_solver* s, double* ke, double* enstrophy,
                               double* circulation) {
  if (!s || !ke || !enstrophy || !circulation) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = bve_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  double out[3];
  if (int rc = lpm::run_diagnostics(ctx, s->n, s->u.as<double>(), s->v.as<double>(), s->w.as<double>(),
                                    s->zeta.as<double>(), s->area.as<double>(), s->active.as<int>(), out))
    return rc;
  *ke = out[0];
  *enstrophy = out[1];
  *circulation = out[2];
  return LPM_OK;
}

int lpm_bve_solver_delete(lpm_bve_solver* s) {
  if (!s) return LPM_ERR_ARG;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  DevBuf* all[] = {&s->x, &s->y, &s->z, &s->zeta, &s->absv, &s->u, &s->v, &s->w, &s->area,
                   &s->active, &s->list, &s->xi, &s->yi, &s->zi, &s->qi, &s->psi_rel, &s->psi_abs,
                   &s->tgt, &s->src, &s->partial, &s->tot, &s->ptrs};
  for (DevBuf* b : all) b->release();
  for (auto& row : s->k)
    for (auto& b : row) b.release();
  delete s;
  return LPM_OK;
}

}  // extern "C"

// ==================================================================== plane
struct lpm_plane_solver {
  lpm_ctx* ctx = nullptr;
  int n = 0, ns = 0, precision = 0, flags = 0;
  DevBuf x, y, vort, u, v, area, active, list;
  DevBuf k[4][2];
  DevBuf xi, yi;
  DevBuf psi;
  DevBuf tgt, src, partial;
  DevBuf ptrs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0.0;
  void* graph_bufs[6] = {};  // workspace addresses as captured (see lpm_bve_solver)
  bool loaded = false;  // false after a failed particle reload
};

namespace {

int plane_velocity(lpm_plane_solver* s, const double* x, const double* y, double* u, double* v) {
  lpm::PairProblem pp;
  pp.kind = LPM_K_PLANE_BS;
  pp.precision = s->precision;
  pp.n = s->n;
  pp.i0 = 0;
  pp.i1 = s->n;
  pp.x = x;
  pp.y = y;
  pp.q0 = s->vort.as<double>();
  pp.area = s->area.as<double>();
  pp.src_list = s->list.as<int>();
  pp.ns = s->ns;
  pp.out0 = u;
  pp.out1 = v;
  pp.tgt_buf = &s->tgt;
  pp.src_buf = &s->src;
  pp.partial_buf = &s->partial;
  return lpm::run_pair_sum(s->ctx, pp);
}

int plane_stream(lpm_plane_solver* s) {
  LPM_HIP(s->ctx, s->psi.ensure((size_t)s->n * sizeof(double)));
  lpm::PairProblem pp;
  pp.kind = LPM_K_PLANE_GREEN;
  pp.precision = s->precision;
  pp.n = s->n;
  pp.i0 = 0;
  pp.i1 = s->n;
  pp.x = s->x.as<double>();
  pp.y = s->y.as<double>();
  pp.q0 = s->vort.as<double>();
  pp.area = s->area.as<double>();
  pp.src_list = s->list.as<int>();
  pp.ns = s->ns;
  pp.out0 = s->psi.as<double>();
  pp.tgt_buf = &s->tgt;
  pp.src_buf = &s->src;
  pp.partial_buf = &s->partial;
  return lpm::run_pair_sum(s->ctx, pp);
}

int plane_enqueue_step(lpm_plane_solver* s, double dt) {
  lpm_ctx* ctx = s->ctx;
  const int n = s->n;
  dim3 b(256), g((n + 255) / 256);
  hipLaunchKernelGGL(plane_stage1, g, b, 0, ctx->stream, n, dt, s->x.as<double>(), s->y.as<double>(),
                     s->u.as<double>(), s->v.as<double>(), s->k[0][0].as<double>(),
                     s->k[0][1].as<double>(), s->xi.as<double>(), s->yi.as<double>());
  static const double h_next[4] = {0.0, 0.5, 1.0, 0.0};
  for (int st = 1; st < 4; ++st) {
    if (int rc = plane_velocity(s, s->xi.as<double>(), s->yi.as<double>(), s->k[st][0].as<double>(),
                                s->k[st][1].as<double>()))
      return rc;
    hipLaunchKernelGGL(plane_stage_k, g, b, 0, ctx->stream, n, dt, h_next[st], st < 3 ? 1 : 0,
                       s->x.as<double>(), s->y.as<double>(), s->k[st][0].as<double>(),
                       s->k[st][1].as<double>(), s->xi.as<double>(), s->yi.as<double>());
  }
  hipLaunchKernelGGL(rk4_combine, g, b, 0, ctx->stream, n, 2, 0, 1.0, (double* const*)s->ptrs.p,
                     (const double* const*)((double**)s->ptrs.p + 2));
  LPM_HIP(ctx, hipGetLastError());
  if (int rc = plane_velocity(s, s->x.as<double>(), s->y.as<double>(), s->u.as<double>(), s->v.as<double>()))
    return rc;
  if (s->flags & LPM_SOLVER_STREAM_EACH_STEP)
    if (int rc = plane_stream(s)) return rc;
  return LPM_OK;
}

void plane_work_bufs(const lpm_plane_solver* s, void* out[6]) {
  const DevBuf* b[4] = {&s->tgt, &s->src, &s->partial, &s->psi};
  for (int k = 0; k < 4; ++k) out[k] = b[k]->p;
  out[4] = out[5] = nullptr;
}

int plane_step(lpm_plane_solver* s, double dt) {
  lpm_ctx* ctx = s->ctx;
  const bool use_graph = (s->flags & LPM_SOLVER_USE_GRAPH) && !ctx->timing;
  if (!use_graph) return plane_enqueue_step(s, dt);
  if (s->exec && s->graph_dt == dt && graph_bufs_unchanged(s, plane_work_bufs)) {
    LPM_HIP(ctx, hipGraphLaunch(s->exec, ctx->stream));
    return LPM_OK;
  }
  if (s->exec) {
    (void)hipGraphExecDestroy(s->exec);
    s->exec = nullptr;
  }
  if (s->graph) {
    (void)hipGraphDestroy(s->graph);
    s->graph = nullptr;
  }
  if (int rc = plane_enqueue_step(s, dt)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  LPM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = plane_enqueue_step(s, dt);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &gr);
  if (rc) {
    if (gr) (void)hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return lpm::hip_fail(ctx, e, "hipStreamEndCapture");
  s->graph = gr;
  LPM_HIP(ctx, hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
  s->graph_dt = dt;
  plane_work_bufs(s, s->graph_bufs);
  return LPM_OK;
}

int plane_check_particles(lpm_ctx* ctx, const char* who, int n, const double* x, const double* y,
                          const double* vort, const double* area, const int32_t* active, const double* u,
                          const double* v) {
  const std::string nm = who;
  if (n <= 0) return fail(ctx, LPM_ERR_ARG, nm + ": n must be > 0");
  if (!x || !y || !vort || !area || !active) return fail(ctx, LPM_ERR_ARG, nm + ": null array argument");
  if ((u || v) && !(u && v)) return fail(ctx, LPM_ERR_ARG, nm + ": give both u and v or none");
  return LPM_OK;
}

int plane_usable(lpm_plane_solver* s) {
  if (s->loaded) return LPM_OK;
  return fail(s->ctx, LPM_ERR_STATE, "solver particles were not loaded (a particle reload failed)");
}

// Load a particle set (at creation, and after a remesh); see bve_load.
int plane_load(lpm_plane_solver* s, int n, const double* x, const double* y, const double* vort,
               const double* area, const int32_t* active, const double* u, const double* v) {
  lpm_ctx* ctx = s->ctx;
  drop_graph(s);
  s->loaded = false;
  s->n = n;
  const size_t nb = (size_t)n * sizeof(double);
  int rc = LPM_OK;
  auto chk = [&](int r) {
    if (r && !rc) rc = r;
    return rc == LPM_OK;
  };
  chk(upload_array(ctx, s->x, x, nb)) && chk(upload_array(ctx, s->y, y, nb)) &&
      chk(upload_array(ctx, s->vort, vort, nb)) && chk(upload_array(ctx, s->area, area, nb)) &&
      chk(upload_array(ctx, s->u, u, nb)) && chk(upload_array(ctx, s->v, v, nb)) &&
      chk(upload_array(ctx, s->xi, nullptr, nb)) && chk(upload_array(ctx, s->yi, nullptr, nb));
  for (int a = 0; a < 4 && rc == LPM_OK; ++a)
    for (int c = 0; c < 2 && rc == LPM_OK; ++c) chk(upload_array(ctx, s->k[a][c], nullptr, nb));
  if (rc == LPM_OK) chk(lpm::upload_active_list(ctx, active, n, s->list, &s->active, &s->ns));
  if (rc == LPM_OK) {
    std::vector<double*> tab(2 + 8);
    tab[0] = s->x.as<double>();
    tab[1] = s->y.as<double>();
    for (int c = 0; c < 2; ++c)
      for (int a = 0; a < 4; ++a) tab[2 + c * 4 + a] = s->k[a][c].as<double>();
    chk(upload_array(ctx, s->ptrs, tab.data(), tab.size() * sizeof(double*)));
    if (rc == LPM_OK) {
      hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
    }
  }
  if (rc == LPM_OK && !u)
    chk(plane_velocity(s, s->x.as<double>(), s->y.as<double>(), s->u.as<double>(), s->v.as<double>()));
  if (rc == LPM_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = lpm::hip_fail(ctx, e, "hipStreamSynchronize");
  }
  s->loaded = rc == LPM_OK;
  return rc;
}

}  // namespace

extern "C" {

int lpm_plane_solver_new(lpm_ctx* ctx, lpm_plane_solver** out, int n, const double* x,
                         const double* y, const double* vort, const double* area,
                         const int32_t* active, const double* u, const double* v, int precision,
                         int flags) {
  if (!ctx || !out) return LPM_ERR_ARG;
  ctx->err.clear();
  *out = nullptr;
  if (int rc = plane_check_particles(ctx, "lpm_plane_solver_new", n, x, y, vort, area, active, u, v)) return rc;
  if (precision != LPM_PREC_FP64 && precision != LPM_PREC_FP32_KAHAN)
    return fail(ctx, LPM_ERR_ARG, "lpm_plane_solver_new: bad precision");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  lpm_plane_solver* s = new lpm_plane_solver();
  s->ctx = ctx;
  s->precision = precision;
  s->flags = flags;
  if (int rc = plane_load(s, n, x, y, vort, area, active, u, v)) {
    lpm_plane_solver_delete(s);
    return rc;
  }
  *out = s;
  return lpm::collect_timing(ctx);
}

int lpm_plane_solver_set_particles(lpm_plane_solver* s, int n, const double* x, const double* y,
                                   const double* vort, const double* area, const int32_t* active,
                                   const double* u, const double* v) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = plane_check_particles(ctx, "lpm_plane_solver_set_particles", n, x, y, vort, area, active, u, v))
    return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = plane_load(s, n, x, y, vort, area, active, u, v)) return rc;
  return lpm::collect_timing(ctx);
}

int lpm_plane_solver_timestep(lpm_plane_solver* s, double dt) { return lpm_plane_solver_steps(s, dt, 1); }

int lpm_plane_solver_steps(lpm_plane_solver* s, double dt, int nsteps) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = plane_usable(s)) return rc;
  if (nsteps < 0) return fail(ctx, LPM_ERR_ARG, "nsteps must be >= 0");
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  for (int k = 0; k < nsteps; ++k)
    if (int rc = plane_step(s, dt)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_plane_solver_get_state(lpm_plane_solver* s, double* x, double* y, double* u, double* v) {
  if (!s) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = plane_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  const DevBuf* src[4] = {&s->x, &s->y, &s->u, &s->v};
  double* dst[4] = {x, y, u, v};
  for (int k = 0; k < 4; ++k)
    if (int rc = download_array(ctx, dst[k], *src[k], s->n)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LPM_OK;
}

int lpm_plane_solver_stream(lpm_plane_solver* s, double* psi) {
  if (!s || !psi) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = plane_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = plane_stream(s)) return rc;
  if (int rc = download_array(ctx, psi, s->psi, s->n)) return rc;
  LPM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return lpm::collect_timing(ctx);
}

int lpm_plane_solver_diagnostics(lpm_plane_solver* s, double* ke, double* enstrophy,
                                 double* circulation) {
  if (!s || !ke || !enstrophy || !circulation) return LPM_ERR_ARG;
  lpm_ctx* ctx = s->ctx;
  ctx->err.clear();
  if (int rc = plane_usable(s)) return rc;
  LPM_HIP(ctx, hipSetDevice(ctx->device));
  double out[3];
  if (int rc = lpm::run_diagnostics(ctx, s->n, s->u.as<double>(), s->v.as<double>(), nullptr,
                                    s->vort.as<double>(), s->area.as<double>(), s->active.as<int>(), out))
    return rc;
  *ke = out[0];
  *enstrophy = out[1];
  *circulation = out[2];
  return LPM_OK;
}

int lpm_plane_solver_delete(lpm_plane_solver* s) {
  if (!s) return LPM_ERR_ARG;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  DevBuf* all[] = {&s->x, &s->y, &s->vort, &s->u, &s->v, &s->area, &s->active, &s->list,
                   &s->xi, &s->yi, &s->psi, &s->tgt, &s->src, &s->partial, &s->ptrs};
  for (DevBuf* b : all) b->release();
  for (auto& row : s->k)
    for (auto& b : row) b.release();
  delete s;
  return LPM_OK;
}

}  // extern "C"
