/*
 * lpm_oracle.c -- FP64 CPU restatement of the lpm-v2 particle direct sums.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed
 * CPU baseline.  The product library never links or calls it.
 *
 * Every loop below restates the reference's FP64 loop with its own
 * expression order: target loop over the caller's block, source loops split
 * at i (self pair skipped by index), `active` mask branch, denominators in
 * the reference's R^2 - x.y form.  Reference files are cited per function.
 * OpenMP splits the target block into contiguous static chunks, the same
 * decomposition MPISetup.f90:132-146 uses across ranks.
 *
 * lpm_oracle_mutation != 0 plants a known mistake; the oracle pin tests use
 * it to show that each pin catches the class of mistake it is meant to.
 */
#include "lpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int lpm_oracle_mutation = 0;

int lpm_oracle_set_mutation(int m) {
  int old = lpm_oracle_mutation;
  lpm_oracle_mutation = m;
  return old;
}

int lpm_oracle_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

/* Mutation ids (planted mistakes). */
enum {
  MUT_NONE = 0,
  MUT_BVE_DROP_4PI = 1,      /* dropped 4*pi in the strength               */
  MUT_BVE_SIGN = 2,          /* wrong sign of the strength                 */
  MUT_BVE_TRANSPOSE = 3,     /* cross product operands transposed          */
  MUT_BVE_SKIP_NEXT = 4,     /* second j loop starts at i+2 (index slip)   */
  MUT_PSE_SIGN = 5,          /* (f_i - f_j) instead of (f_j - f_i)         */
  MUT_PSE_DROP_EPS2 = 6,     /* one 1/eps^2 factor dropped                 */
  MUT_PSE_K8_COEF = 7,       /* 10 rho^4 -> 12 rho^4 in K8                 */
  MUT_PLANE_SIGN_V = 8,      /* wrong sign on the v component              */
  MUT_RK4_WEIGHTS = 9,       /* stage weights 1/6,1/3,1/3,1/6 -> 1/4 each  */
  MUT_STREAM_DROP_4PI = 10,  /* Green's kernel missing 1/(4 pi)            */
  MUT_BVE_ZETA_SOURCE = 11,  /* d(zeta)/dt uses u instead of w             */
  MUT_PSE_K2_COEF = 12,      /* 4/pi -> 2/pi in the 2nd-order kernel       */
  MUT_BVE_RADIUS = 13,       /* strength divided by R^2 instead of R        */
  MUT_PSE_GRAD_SIGN = 14,    /* (x_j - x_i) instead of (x_i - x_j)          */
  MUT_PSE_K1_COEF = 15,      /* 5 rho^4 -> 4 rho^4 in the derivative kernel */
  MUT_PSE_DELTA_COEF = 16,   /* 2 rho^4 -> 3 rho^4 in the delta kernel      */
  MUT_PSE_NO_PROJ = 17,      /* sphere gradient/divergence not projected    */
  MUT_DDOT_CROSS = 18,       /* 2 uy vx -> uy vx in the double dot          */
  MUT_DDOT_W_ROW = 19,       /* sphere double dot w row with w_i, not v_i   */
  MUT_PARTIALS_SYM = 20,     /* mixed partial dxy instead of (dxy + dyx)/2  */
  MUT_SWE_POT_SIGN = 21,     /* potential (divergence) velocity sign        */
  MUT_SWE_DENOM2 = 22,       /* velocity-gradient denominator 2 pi r^4      */
  MUT_SWE_LAP_EPS = 23,      /* surface Laplacian missing the outer 1/eps^2 */
};

/* --------------------------------------------------------------------------
 * Sphere Biot-Savart velocity.
 * Restates BVESphereVelocity, SphereBVESolver.f90:377-430 (loop :395-420),
 * identical arithmetic in setVelocityFromVorticity, SphereBVE.f90:489-531.
 *   u_i = sum_{j != i, active j} (x_i x x_j) s_ij,
 *   s_ij = -zeta_j A_j / (4 pi R (R^2 - x_i . x_j))
 * ------------------------------------------------------------------------ */
static inline void bve_pair(double xi, double yi, double zi, double xj, double yj, double zj,
                            double zeta, double area, double radius, double *u, double *v,
                            double *w) {
  double fourpi = 4.0 * M_PI;
  if (lpm_oracle_mutation == MUT_BVE_DROP_4PI) fourpi = 1.0;
  double rfac = radius;
  if (lpm_oracle_mutation == MUT_BVE_RADIUS) rfac = radius * radius;
  double strength =
      -zeta * area / (fourpi * rfac * (radius * radius - xi * xj - yi * yj - zi * zj));
  if (lpm_oracle_mutation == MUT_BVE_SIGN) strength = -strength;
  if (lpm_oracle_mutation == MUT_BVE_TRANSPOSE) {
    *u += (yj * zi - zj * yi) * strength;
    *v += (zj * xi - xj * zi) * strength;
    *w += (xj * yi - yj * xi) * strength;
  } else {
    *u += (yi * zj - zi * yj) * strength;
    *v += (zi * xj - xi * zj) * strength;
    *w += (xi * yj - yi * xj) * strength;
  }
}

void lpm_oracle_bve_sphere_velocity(int n, const double *x, const double *y, const double *z,
                                    const double *relvort, const double *area, const int *active,
                                    double radius, int i0, int i1, double *u, double *v,
                                    double *w) {
  const int skip = (lpm_oracle_mutation == MUT_BVE_SKIP_NEXT) ? 2 : 1;
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double ui = 0.0, vi = 0.0, wi = 0.0;
    for (int j = 0; j < i; ++j)
      if (active[j]) bve_pair(x[i], y[i], z[i], x[j], y[j], z[j], relvort[j], area[j], radius, &ui, &vi, &wi);
    for (int j = i + skip; j < n; ++j)
      if (active[j]) bve_pair(x[i], y[i], z[i], x[j], y[j], z[j], relvort[j], area[j], radius, &ui, &vi, &wi);
    u[i] = ui;
    v[i] = vi;
    w[i] = wi;
  }
}

/* --------------------------------------------------------------------------
 * Planar Biot-Savart velocity (singular point-vortex kernel).
 * Restates planarIncompressibleVelocity, PlaneIncompressibleSolver.f90:278-316
 * (loop :291-308); same arithmetic as PlanarIncompressible.f90:426-466.
 * ------------------------------------------------------------------------ */
static inline void plane_pair(double xi, double yi, double xj, double yj, double vort,
                              double area, double *u, double *v) {
  double strength = vort * area / (2.0 * M_PI * ((xi - xj) * (xi - xj) + (yi - yj) * (yi - yj)));
  *u = *u - (yi - yj) * strength;
  if (lpm_oracle_mutation == MUT_PLANE_SIGN_V)
    *v = *v - (xi - xj) * strength;
  else
    *v = *v + (xi - xj) * strength;
}

void lpm_oracle_plane_velocity(int n, const double *x, const double *y, const double *vort,
                               const double *area, const int *active, int i0, int i1, double *u,
                               double *v) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double ui = 0.0, vi = 0.0;
    for (int j = 0; j < i; ++j)
      if (active[j]) plane_pair(x[i], y[i], x[j], y[j], vort[j], area[j], &ui, &vi);
    for (int j = i + 1; j < n; ++j)
      if (active[j]) plane_pair(x[i], y[i], x[j], y[j], vort[j], area[j], &ui, &vi);
    u[i] = ui;
    v[i] = vi;
  }
}

/* --------------------------------------------------------------------------
 * PSE kernels.
 * order 8: bivariateLaplacianKernel8, PSEDirectSum.f90:622-627
 *   K8(rho) = (40 - 40 rho^2 + 10 rho^4 - 2 rho^6 / 3) exp(-rho^2) / pi
 * order 2: the 2nd-order Gaussian Laplacian kernel (no reference routine;
 *   same normalisation as K8, second moment int eta(|x|) x_1^2 dx = 2)
 *   K2(rho) = (4 / pi) exp(-rho^2)
 * ------------------------------------------------------------------------ */
double lpm_oracle_pse_kernel(int order, double rho) {
  double r2 = rho * rho;
  if (order == 2) {
    double c = (lpm_oracle_mutation == MUT_PSE_K2_COEF) ? 2.0 : 4.0;
    return c * exp(-r2) / M_PI;
  }
  double c4 = (lpm_oracle_mutation == MUT_PSE_K8_COEF) ? 12.0 : 10.0;
  return (40.0 - 40.0 * r2 + c4 * r2 * r2 - 2.0 * r2 * r2 * r2 / 3.0) * exp(-r2) / M_PI;
}

/* PSEPlaneLaplacianAtParticles, PSEDirectSum.f90:467-500: j = i included (its
 * term is 0), distance by ChordDistance (SphereGeometry.f90:67-73) on z = 0
 * coordinates, inner factor K/eps^2, then the whole field times 1/eps^2
 * (PSEDirectSum.f90:499). */
void lpm_oracle_pse_plane_laplacian(int n, const double *x, const double *y, const double *f,
                                    const double *area, const int *active, double eps, int order,
                                    int i0, int i1, double *lap) {
  const double outer = (lpm_oracle_mutation == MUT_PSE_DROP_EPS2) ? 1.0 : 1.0 / (eps * eps);
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double dx = x[j] - x[i], dy = y[j] - y[i];
      double kin = sqrt(dx * dx + dy * dy) / eps;
      double df = (lpm_oracle_mutation == MUT_PSE_SIGN) ? (f[i] - f[j]) : (f[j] - f[i]);
      acc = acc + df * lpm_oracle_pse_kernel(order, kin) / (eps * eps) * area[j];
    }
    lap[i] = acc * outer;
  }
}

/* Sphere variant used only to pin K8 against the reference's own regression
 * test: PSESphereLaplacianAtParticles (PSEDirectSum.f90:502-535) and the
 * inline particle loop of SpherePSEConvTest.f90:245-269 -- great-circle
 * distance (SphereDistance, SphereGeometry.f90:107-125) divided by eps. */
void lpm_oracle_pse_sphere_laplacian(int n, const double *x, const double *y, const double *z,
                                     const double *f, const double *area, const int *active,
                                     double eps, double radius, int i0, int i1, double *lap) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double cx = y[i] * z[j] - y[j] * z[i];
      double cy = x[j] * z[i] - x[i] * z[j];
      double cz = x[i] * y[j] - x[j] * y[i];
      double dist = atan2(sqrt(cx * cx + cy * cy + cz * cz), x[i] * x[j] + y[i] * y[j] + z[i] * z[j]) * radius;
      acc = acc + lpm_oracle_pse_kernel(8, dist / eps) / (eps * eps) * (f[j] - f[i]) * area[j] / (eps * eps);
    }
    lap[i] = acc;
  }
}

/* --------------------------------------------------------------------------
 * PSE first-derivative, interpolation and double-dot operators,
 * PSEDirectSum.f90:121-418, 537-579.
 *   delta kernel  (bivariateDeltaKernel8, :611-616)
 *     Kd(rho) = (4 - 6 rho^2 + 2 rho^4 - rho^6 / 6) exp(-rho^2) / pi
 *   derivative kernel (bivariateFirstDerivativeKernel8, :629-634)
 *     K1(rho) = (-20 + 20 rho^2 - 5 rho^4 + rho^6 / 3) exp(-rho^2) / pi
 * Distances: ChordDistance on the plane (z = 0), SphereDistance (great
 * circle, R atan2(|a x b|, a.b), SphereGeometry.f90:107-125) on the sphere.
 * ------------------------------------------------------------------------ */
double lpm_oracle_pse_delta_kernel(double rho) {
  double r2 = rho * rho;
  double c4 = (lpm_oracle_mutation == MUT_PSE_DELTA_COEF) ? 3.0 : 2.0;
  return (4.0 - 6.0 * r2 + c4 * r2 * r2 - r2 * r2 * r2 / 6.0) * exp(-r2) / M_PI;
}

double lpm_oracle_pse_deriv_kernel(double rho) {
  double r2 = rho * rho;
  double c4 = (lpm_oracle_mutation == MUT_PSE_K1_COEF) ? 4.0 : 5.0;
  return (-20.0 + 20.0 * r2 - c4 * r2 * r2 + r2 * r2 * r2 / 3.0) * exp(-r2) / M_PI;
}

static double sphere_distance(double ax, double ay, double az, double bx, double by, double bz,
                              double radius) {
  double cx = ay * bz - by * az;
  double cy = bx * az - ax * bz;
  double cz = ax * by - bx * ay;
  return atan2(sqrt(cx * cx + cy * cy + cz * cz), ax * bx + ay * by + az * bz) * radius;
}

/* x_i - x_j, or its negation under the planted sign mistake. */
static inline double sep(double a, double b) {
  return (lpm_oracle_mutation == MUT_PSE_GRAD_SIGN) ? (b - a) : (a - b);
}

/* Rows of SphereProjection(x) = I - x x^T (SphereGeometry.f90:578-590)
 * applied to g, in place; skipped under the planted mistake. */
static void sphere_project(double xi, double yi, double zi, double g[3]) {
  if (lpm_oracle_mutation == MUT_PSE_NO_PROJ) return;
  double p0 = (1.0 - xi * xi) * g[0] - xi * yi * g[1] - xi * zi * g[2];
  double p1 = -yi * xi * g[0] + (1.0 - yi * yi) * g[1] - yi * zi * g[2];
  double p2 = -zi * xi * g[0] - zi * yi * g[1] + (1.0 - zi * zi) * g[2];
  g[0] = p0;
  g[1] = p1;
  g[2] = p2;
}

/* PSEPlaneGradientAtParticles, PSEDirectSum.f90:180-217: inner factor
 * K1/eps^3, whole field times 1/eps at the end (:216). */
void lpm_oracle_pse_plane_gradient(int n, const double *x, const double *y, const double *f,
                                   const double *area, const int *active, double eps, int i0,
                                   int i1, double *gx, double *gy) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double ax = 0.0, ay = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double dx = x[i] - x[j], dy = y[i] - y[j];
      double kin = sqrt(dx * dx + dy * dy) / eps;
      double k1 = lpm_oracle_pse_deriv_kernel(kin);
      ax = ax + (f[j] + f[i]) * sep(x[i], x[j]) * k1 / (eps * eps * eps) * area[j];
      ay = ay + (f[j] + f[i]) * sep(y[i], y[j]) * k1 / (eps * eps * eps) * area[j];
    }
    gx[i] = ax / eps;
    gy[i] = ay / eps;
  }
}

/* PSESphereGradientAtParticles, PSEDirectSum.f90:220-267: K1/eps^2 of the
 * great-circle distance, projected by SphereProjection(x_i), times 1/eps^2. */
void lpm_oracle_pse_sphere_gradient(int n, const double *x, const double *y, const double *z,
                                    const double *f, const double *area, const int *active,
                                    double eps, double radius, int i0, int i1, double *gx,
                                    double *gy, double *gz) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double g[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double kin = sphere_distance(x[j], y[j], z[j], x[i], y[i], z[i], radius) / eps;
      double kout = lpm_oracle_pse_deriv_kernel(kin) / (eps * eps);
      g[0] = g[0] + (f[j] + f[i]) * sep(x[i], x[j]) * kout * area[j];
      g[1] = g[1] + (f[j] + f[i]) * sep(y[i], y[j]) * kout * area[j];
      g[2] = g[2] + (f[j] + f[i]) * sep(z[i], z[j]) * kout * area[j];
    }
    sphere_project(x[i], y[i], z[i], g);
    gx[i] = g[0] / (eps * eps);
    gy[i] = g[1] / (eps * eps);
    gz[i] = g[2] / (eps * eps);
  }
}

/* PSEPlaneSecondPartialsAtParticles, PSEDirectSum.f90:269-320: the PSE
 * gradient of each component of a gradient field; outputs dxx,
 * (dxy + dyx) / 2, dyy (xComp, yComp, zComp of the reference's Field). */
void lpm_oracle_pse_plane_second_partials(int n, const double *x, const double *y,
                                          const double *gx, const double *gy, const double *area,
                                          const int *active, double eps, int i0, int i1,
                                          double *dxx, double *dxy, double *dyy) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double axx = 0.0, axy = 0.0, ayx = 0.0, ayy = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double dx = x[i] - x[j], dy = y[i] - y[j];
      double kin = sqrt(dx * dx + dy * dy) / eps;
      double kout = lpm_oracle_pse_deriv_kernel(kin) / (eps * eps);
      axx = axx + (gx[j] + gx[i]) * sep(x[i], x[j]) / eps * kout * area[j];
      axy = axy + (gx[j] + gx[i]) * sep(y[i], y[j]) / eps * kout * area[j];
      ayx = ayx + (gy[j] + gy[i]) * sep(x[i], x[j]) / eps * kout * area[j];
      ayy = ayy + (gy[j] + gy[i]) * sep(y[i], y[j]) / eps * kout * area[j];
    }
    dxx[i] = axx / eps;
    dxy[i] = (lpm_oracle_mutation == MUT_PARTIALS_SYM ? axy : 0.5 * (axy + ayx)) / eps;
    dyy[i] = ayy / eps;
  }
}

/* PSEPlaneDoubleDotProductAtParticles, PSEDirectSum.f90:322-365:
 * (ux^2 + 2 uy vx + vy^2) / eps^2 of the PSE velocity gradient. */
void lpm_oracle_pse_plane_double_dot(int n, const double *x, const double *y, const double *u,
                                     const double *v, const double *area, const int *active,
                                     double eps, int i0, int i1, double *ddot) {
  const double two = (lpm_oracle_mutation == MUT_DDOT_CROSS) ? 1.0 : 2.0;
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double dx = x[i] - x[j], dy = y[i] - y[j];
      double kin = sqrt(dx * dx + dy * dy) / eps;
      double kout = lpm_oracle_pse_deriv_kernel(kin) / (eps * eps);
      ux = ux + (u[j] + u[i]) * sep(x[i], x[j]) / eps * kout * area[j];
      uy = uy + (u[j] + u[i]) * sep(y[i], y[j]) / eps * kout * area[j];
      vx = vx + (v[j] + v[i]) * sep(x[i], x[j]) / eps * kout * area[j];
      vy = vy + (v[j] + v[i]) * sep(y[i], y[j]) / eps * kout * area[j];
    }
    ddot[i] = (ux * ux + two * uy * vx + vy * vy) / (eps * eps);
  }
}

/* PSESphereDoubleDotProductAtParticles, PSEDirectSum.f90:367-418.  As in
 * the reference, the w row pairs w_j with v_i (zComp(j) + yComp(i),
 * :401-406), and no projection is applied. */
void lpm_oracle_pse_sphere_double_dot(int n, const double *x, const double *y, const double *z,
                                      const double *u, const double *v, const double *w,
                                      const double *area, const int *active, double eps,
                                      double radius, int i0, int i1, double *ddot) {
  const double two = (lpm_oracle_mutation == MUT_DDOT_CROSS) ? 1.0 : 2.0;
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    const double wi = (lpm_oracle_mutation == MUT_DDOT_W_ROW) ? w[i] : v[i];
    double ux = 0.0, uy = 0.0, uz = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, wx = 0.0, wy = 0.0,
           wz = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double kin = sphere_distance(x[i], y[i], z[i], x[j], y[j], z[j], radius) / eps;
      double kout = lpm_oracle_pse_deriv_kernel(kin) / (eps * eps);
      double sx = sep(x[i], x[j]) / eps, sy = sep(y[i], y[j]) / eps, sz = sep(z[i], z[j]) / eps;
      ux = ux + (u[j] + u[i]) * sx * kout * area[j];
      uy = uy + (u[j] + u[i]) * sy * kout * area[j];
      uz = uz + (u[j] + u[i]) * sz * kout * area[j];
      vx = vx + (v[j] + v[i]) * sx * kout * area[j];
      vy = vy + (v[j] + v[i]) * sy * kout * area[j];
      vz = vz + (v[j] + v[i]) * sz * kout * area[j];
      wx = wx + (w[j] + wi) * sx * kout * area[j];
      wy = wy + (w[j] + wi) * sy * kout * area[j];
      wz = wz + (w[j] + wi) * sz * kout * area[j];
    }
    ddot[i] = (ux * ux + vy * vy + wz * wz + two * (uy * vx + uz * wx + vz * wy)) / (eps * eps);
  }
}

/* PSESphereDivergenceAtParticles, PSEDirectSum.f90:537-579: per pair the
 * projected kernel gradient P(x_i) (x_i - x_j) K1 / eps^3 dotted with
 * (v_j + v_i) A_j; the field times 1/eps at the end. */
void lpm_oracle_pse_sphere_divergence(int n, const double *x, const double *y, const double *z,
                                      const double *u, const double *v, const double *w,
                                      const double *area, const int *active, double eps,
                                      double radius, int i0, int i1, double *div) {
  const double denom = eps * eps * eps;
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double kin = sphere_distance(x[j], y[j], z[j], x[i], y[i], z[i], radius) / eps;
      double k1 = lpm_oracle_pse_deriv_kernel(kin);
      double g[3] = {sep(x[i], x[j]) * k1 / denom, sep(y[i], y[j]) * k1 / denom,
                     sep(z[i], z[j]) * k1 / denom};
      sphere_project(x[i], y[i], z[i], g);
      acc = acc + (g[0] * (u[j] + u[i]) + g[1] * (v[j] + v[i]) + g[2] * (w[j] + w[i])) * area[j];
    }
    div[i] = acc / eps;
  }
}

/* PSEPlaneInterpolateScalar, PSEDirectSum.f90:128-147, at each of nt points:
 * sum over active j of f_j A_j Kd(|p - x_j| / eps) / eps^2 (no self term to
 * skip: the points are not particles). */
void lpm_oracle_pse_plane_interpolate(int n, const double *x, const double *y, const double *f,
                                      const double *area, const int *active, double eps, int nt,
                                      const double *px, const double *py, double *out) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double dx = px[k] - x[j], dy = py[k] - y[j];
      double kin = sqrt(dx * dx + dy * dy) / eps;
      acc = acc + f[j] * area[j] * lpm_oracle_pse_delta_kernel(kin) / (eps * eps);
    }
    out[k] = acc;
  }
}

/* PSESphereInterpolateScalar, PSEDirectSum.f90:149-168, at each of nt points
 * (great-circle distance). */
void lpm_oracle_pse_sphere_interpolate(int n, const double *x, const double *y, const double *z,
                                       const double *f, const double *area, const int *active,
                                       double eps, double radius, int nt, const double *px,
                                       const double *py, const double *pz, double *out) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double kin = sphere_distance(x[j], y[j], z[j], px[k], py[k], pz[k], radius) / eps;
      acc = acc + f[j] * area[j] * lpm_oracle_pse_delta_kernel(kin) / (eps * eps);
    }
    out[k] = acc;
  }
}

/* Sphere PSE Laplacian at nt points with given point values pf: the inline
 * lat-lon grid loop of SpherePSEConvTest.f90:187-207,
 *   lap(p) = sum_j K8(d(x_j, p) / eps) / eps^2 (f_j - pf) A_j / eps^2. */
void lpm_oracle_pse_sphere_laplacian_points(int n, const double *x, const double *y,
                                            const double *z, const double *f, const double *area,
                                            const int *active, double eps, double radius, int nt,
                                            const double *px, const double *py, const double *pz,
                                            const double *pf, double *out) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nt; ++k) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double kin = sphere_distance(x[j], y[j], z[j], px[k], py[k], pz[k], radius) / eps;
      acc = acc + lpm_oracle_pse_kernel(8, kin) / (eps * eps) * (f[j] - pf[k]) * area[j] / (eps * eps);
    }
    out[k] = acc;
  }
}

/* --------------------------------------------------------------------------
 * Planar shallow-water right-hand-side integrals, SWEPlaneRHSIntegrals,
 * SWEPlaneSolver.f90:457-538.  bottom[j] holds topoFn(x_j, y_j), so the
 * surface height is h_j + bottom_j.  For each target i:
 *   lapSurf: sum over active j (j = i included, its term is 0) of
 *            K8(r / eps) / eps^2 (s_j - s_i) A_j, then / eps^2;
 *   u, v and the velocity gradient (ux, uy, vx, vy) of the point vortex /
 *   point source field, sum over active j != i;
 *   doubleDot = ux^2 + 2 uy vx + vy^2.
 * ------------------------------------------------------------------------ */
void lpm_oracle_swe_plane_rhs(int n, const double *x, const double *y, const double *vort,
                              const double *divg, const double *h, const double *bottom,
                              const double *area, const int *active, double eps, int i0, int i1,
                              double *u, double *v, double *ddot, double *lap) {
  const double psgn = (lpm_oracle_mutation == MUT_SWE_POT_SIGN) ? -1.0 : 1.0;
  const double d2 = (lpm_oracle_mutation == MUT_SWE_DENOM2) ? 2.0 : 1.0;
  const double outer = (lpm_oracle_mutation == MUT_SWE_LAP_EPS) ? 1.0 : 1.0 / (eps * eps);
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double ui = 0.0, vi = 0.0, li = 0.0, ux = 0.0, uy = 0.0, vx = 0.0, vy = 0.0;
    const double si = h[i] + bottom[i];
    for (int j = 0; j < n; ++j) {
      if (!active[j]) continue;
      double sj = h[j] + bottom[j];
      double dx = x[i] - x[j], dy = y[i] - y[j];
      double sq = dx * dx + dy * dy;
      double kin = sqrt(sq) / eps;
      li = li + lpm_oracle_pse_kernel(8, kin) / (eps * eps) * (sj - si) * area[j];
      if (j == i) continue;
      double denom = 2.0 * M_PI * sq;
      double denom2 = d2 * M_PI * sq * sq;
      double rot = vort[j] * area[j] / denom;
      double pot = psgn * divg[j] * area[j] / denom;
      ui = ui - dy * rot + dx * pot;
      vi = vi + dx * rot + dy * pot;
      ux = ux + pot - (dx * (dx * divg[j] - dy * vort[j])) * area[j] / denom2;
      uy = uy - rot - (dy * (dx * divg[j] - dy * vort[j])) * area[j] / denom2;
      vx = vx + rot - (dx * (dy * divg[j] + dx * vort[j])) * area[j] / denom2;
      vy = vy + pot - (dy * (dy * divg[j] + dx * vort[j])) * area[j] / denom2;
    }
    u[i] = ui;
    v[i] = vi;
    lap[i] = li * outer;
    ddot[i] = ux * ux + 2.0 * uy * vx + vy * vy;
  }
}

/* --------------------------------------------------------------------------
 * Green's-function stream functions.
 * Sphere: SetStreamFunctionsOnMesh, SphereBVE.f90:445-485
 *   psi(x_i) = sum_{j != i} -log(R^2 - x_i.x_j) / (4 pi) * zeta_j A_j
 * Plane:  SetStreamFunctionOnMesh, PlanarIncompressible.f90:470-505
 *   psi(x_i) = sum_{j != i} log(|x_i - x_j|) / (2 pi) * omega_j A_j
 * ------------------------------------------------------------------------ */
void lpm_oracle_bve_sphere_stream(int n, const double *x, const double *y, const double *z,
                                  const double *relvort, const double *absvort, const double *area,
                                  const int *active, double radius, int i0, int i1,
                                  double *psi_rel, double *psi_abs) {
  const double fourpi = (lpm_oracle_mutation == MUT_STREAM_DROP_4PI) ? 1.0 : 4.0 * M_PI;
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double pr = 0.0, pa = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i || !active[j]) continue;
      double g = -log(radius * radius - (x[i] * x[j] + y[i] * y[j] + z[i] * z[j])) / fourpi;
      pr = pr + g * relvort[j] * area[j];
      pa = pa + g * absvort[j] * area[j];
    }
    psi_rel[i] = pr;
    psi_abs[i] = pa;
  }
}

void lpm_oracle_plane_stream(int n, const double *x, const double *y, const double *vort,
                             const double *area, const int *active, int i0, int i1, double *psi) {
#pragma omp parallel for schedule(static)
  for (int i = i0; i < i1; ++i) {
    double p = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i || !active[j]) continue;
      double g = log(sqrt((x[i] - x[j]) * (x[i] - x[j]) + (y[i] - y[j]) * (y[i] - y[j]))) / (2.0 * M_PI);
      p = p + g * vort[j] * area[j];
    }
    psi[i] = p;
  }
}

/* --------------------------------------------------------------------------
 * RK4 steps.
 * Sphere: timestepPrivate, SphereBVESolver.f90:219-353.  Stage 1 reuses the
 * stored velocity (:238-244); stages 2-4 call the Biot-Savart sum on stage
 * inputs (:249-313); combine with 1/6, 1/3, 1/3, 1/6 (:319-329); the final
 * velocity is recomputed at the new positions (:345).  d(zeta)/dt =
 * -2 Omega w / R (:243, :263, :286, :309).  Stage points are not projected
 * back onto the sphere, exactly as in the reference.
 * ------------------------------------------------------------------------ */
/* Optional projection (no reference counterpart): every stage input and the
 * new positions are mapped to R x / |x| in FP64 before the velocity sum. */
static void project3(double *x, double *y, double *z, double radius) {
  double s = radius / sqrt(*x * *x + *y * *y + *z * *z);
  *x = *x * s;
  *y = *y * s;
  *z = *z * s;
}

void lpm_oracle_bve_timestep(int n, double *x, double *y, double *z, double *relvort, double *u,
                             double *v, double *w, const double *area, const int *active,
                             double radius, double rotation_rate, double dt) {
  lpm_oracle_bve_timestep_ex(n, x, y, z, relvort, u, v, w, area, active, radius, rotation_rate, dt, 0);
}

void lpm_oracle_bve_timestep_ex(int n, double *x, double *y, double *z, double *relvort, double *u,
                                double *v, double *w, const double *area, const int *active,
                                double radius, double rotation_rate, double dt, int project) {
  size_t nb = (size_t)n * sizeof(double);
  double *buf = (double *)malloc(nb * 20);
  double *xs[4], *ys[4], *zs[4], *ks[4];
  for (int s = 0; s < 4; ++s) {
    xs[s] = buf + (size_t)n * (4 * s + 0);
    ys[s] = buf + (size_t)n * (4 * s + 1);
    zs[s] = buf + (size_t)n * (4 * s + 2);
    ks[s] = buf + (size_t)n * (4 * s + 3);
  }
  double *xin = buf + (size_t)n * 16, *yin = buf + (size_t)n * 17, *zin = buf + (size_t)n * 18,
         *kin = buf + (size_t)n * 19;
  const int zeta_from_u = (lpm_oracle_mutation == MUT_BVE_ZETA_SOURCE);

  for (int i = 0; i < n; ++i) {
    xs[0][i] = dt * u[i];
    ys[0][i] = dt * v[i];
    zs[0][i] = dt * w[i];
    ks[0][i] = -dt * 2.0 * rotation_rate * (zeta_from_u ? u[i] : w[i]) / radius;
  }
  static const double half[3] = {0.5, 0.5, 1.0};
  for (int s = 1; s < 4; ++s) {
    for (int i = 0; i < n; ++i) {
      xin[i] = x[i] + half[s - 1] * xs[s - 1][i];
      yin[i] = y[i] + half[s - 1] * ys[s - 1][i];
      zin[i] = z[i] + half[s - 1] * zs[s - 1][i];
      kin[i] = relvort[i] + half[s - 1] * ks[s - 1][i];
      if (project) project3(&xin[i], &yin[i], &zin[i], radius);
    }
    lpm_oracle_bve_sphere_velocity(n, xin, yin, zin, kin, area, active, radius, 0, n, xs[s], ys[s], zs[s]);
    for (int i = 0; i < n; ++i) {
      ks[s][i] = -dt * 2.0 * rotation_rate * (zeta_from_u ? xs[s][i] : zs[s][i]) / radius;
      xs[s][i] *= dt;
      ys[s][i] *= dt;
      zs[s][i] *= dt;
    }
  }
  /* (SphereBVESolver.f90:321-328) divisions by 6 and 3, left to right */
  double d1 = 6.0, d2 = 3.0;
  if (lpm_oracle_mutation == MUT_RK4_WEIGHTS) d1 = d2 = 4.0;
  for (int i = 0; i < n; ++i) {
    x[i] = x[i] + xs[0][i] / d1 + xs[1][i] / d2 + xs[2][i] / d2 + xs[3][i] / d1;
    y[i] = y[i] + ys[0][i] / d1 + ys[1][i] / d2 + ys[2][i] / d2 + ys[3][i] / d1;
    z[i] = z[i] + zs[0][i] / d1 + zs[1][i] / d2 + zs[2][i] / d2 + zs[3][i] / d1;
    relvort[i] = relvort[i] + ks[0][i] / d1 + ks[1][i] / d2 + ks[2][i] / d2 + ks[3][i] / d1;
    if (project) project3(&x[i], &y[i], &z[i], radius);
  }
  lpm_oracle_bve_sphere_velocity(n, x, y, z, relvort, area, active, radius, 0, n, u, v, w);
  free(buf);
}

/* Plane: timestepPrivate, PlaneIncompressibleSolver.f90:171-259; vorticity is
 * carried unchanged (:137, :196-252). */
void lpm_oracle_plane_timestep(int n, double *x, double *y, const double *vort, double *u,
                               double *v, const double *area, const int *active, double dt) {
  size_t nb = (size_t)n;
  double *buf = (double *)malloc(nb * 10 * sizeof(double));
  double *xs[4], *ys[4];
  for (int s = 0; s < 4; ++s) {
    xs[s] = buf + nb * (2 * s);
    ys[s] = buf + nb * (2 * s + 1);
  }
  double *xin = buf + nb * 8, *yin = buf + nb * 9;
  for (int i = 0; i < n; ++i) {
    xs[0][i] = dt * u[i];
    ys[0][i] = dt * v[i];
  }
  static const double half[3] = {0.5, 0.5, 1.0};
  for (int s = 1; s < 4; ++s) {
    for (int i = 0; i < n; ++i) {
      xin[i] = x[i] + half[s - 1] * xs[s - 1][i];
      yin[i] = y[i] + half[s - 1] * ys[s - 1][i];
    }
    lpm_oracle_plane_velocity(n, xin, yin, vort, area, active, 0, n, xs[s], ys[s]);
    for (int i = 0; i < n; ++i) {
      xs[s][i] *= dt;
      ys[s][i] *= dt;
    }
  }
  double d1 = 6.0, d2 = 3.0;
  if (lpm_oracle_mutation == MUT_RK4_WEIGHTS) d1 = d2 = 4.0;
  for (int i = 0; i < n; ++i) {
    x[i] = x[i] + xs[0][i] / d1 + xs[1][i] / d2 + xs[2][i] / d2 + xs[3][i] / d1;
    y[i] = y[i] + ys[0][i] / d1 + ys[1][i] / d2 + ys[2][i] / d2 + ys[3][i] / d1;
  }
  lpm_oracle_plane_velocity(n, x, y, vort, area, active, 0, n, u, v);
  free(buf);
}

/* --------------------------------------------------------------------------
 * Conserved diagnostics, FP64 serial reductions over active particles.
 * TotalKE SphereBVE.f90:410-424 / PlanarIncompressible.f90:368-383,
 * TotalEnstrophy SphereBVE.f90:429-441 / PlanarIncompressible.f90:388-401,
 * total circulation sum zeta A (BVESingleGaussianVortex.f90:350-357).
 * out = {KE, enstrophy, circulation}; ncomp = 2 (plane) or 3 (sphere).
 * ------------------------------------------------------------------------ */
void lpm_oracle_diagnostics(int n, int ncomp, const double *u, const double *v, const double *w,
                            const double *vort, const double *area, const int *active,
                            double out[3]) {
  double ke = 0.0, ens = 0.0, circ = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!active[i]) continue;
    double m = u[i] * u[i] + v[i] * v[i] + (ncomp == 3 ? w[i] * w[i] : 0.0);
    ke += m * area[i];
    ens += vort[i] * vort[i] * area[i];
    circ += vort[i] * area[i];
  }
  out[0] = 0.5 * ke;
  out[1] = 0.5 * ens;
  out[2] = circ;
}

/* LoadBalance, MPISetup.f90:132-146 (reported 1-based inclusive). */
void lpm_oracle_load_balance(int nitems, int nprocs, int *index_start, int *index_end,
                             int *message_length) {
  int chunk = nitems / nprocs;
  for (int r = 0; r < nprocs; ++r) {
    index_start[r] = r * chunk + 1;
    index_end[r] = (r + 1) * chunk;
  }
  index_end[nprocs - 1] = nitems;
  for (int r = 0; r < nprocs; ++r) message_length[r] = index_end[r] - index_start[r] + 1;
}

/* PSE newPrivate, PSEDirectSum.f90:93-113: eps = h ** pow (default pow 0.75). */
double lpm_oracle_pse_eps(double max_edge_length, double power) { return pow(max_edge_length, power); }

/* --------------------------------------------------------------------------
 * Far-field surrogate checks.  No counterpart in the reference.
 * Cell moments about the cell centroid c: for strength s_j at y_j in cell k,
 *   M0 = sum s,  M1_a = sum s d_a,  M2_ab = sum s d_a d_b  (a <= b),
 * with d = y_j - c_k; 10 moments per cell in the order
 *   [M0, Mx, My, Mz, Mxx, Mxy, Mxz, Myy, Myz, Mzz].
 * ------------------------------------------------------------------------ */
void lpm_oracle_cell_moments(int n, const double *x, const double *y, const double *z,
                             const double *strength, const int *cell, int ncell,
                             const double *centroid, double *moments) {
  memset(moments, 0, sizeof(double) * 10 * (size_t)ncell);
  for (int j = 0; j < n; ++j) {
    int k = cell[j];
    if (k < 0 || k >= ncell) continue;
    double dx = x[j] - centroid[3 * k], dy = y[j] - centroid[3 * k + 1], dz = z[j] - centroid[3 * k + 2];
    double s = strength[j];
    double *m = moments + 10 * (size_t)k;
    m[0] += s;
    m[1] += s * dx;
    m[2] += s * dy;
    m[3] += s * dz;
    m[4] += s * dx * dx;
    m[5] += s * dx * dy;
    m[6] += s * dx * dz;
    m[7] += s * dy * dy;
    m[8] += s * dy * dz;
    m[9] += s * dz * dz;
  }
}

long lpm_oracle_mlp_param_count(int nin, int width, int nhidden, int nout) {
  long c = (long)width * nin + width;
  for (int l = 1; l < nhidden; ++l) c += (long)width * width + width;
  c += (long)nout * width + nout;
  return c;
}

/* act: 0 ReLU, 1 tanh, 2 ISRU a / sqrt(1 + a^2) (derivative (1 + a^2)^(-3/2)) */
static double act_f(int act, double a) {
  if (act == 1) return tanh(a);
  if (act == 2) return a / sqrt(1.0 + a * a);
  return a > 0.0 ? a : 0.0;
}
static double act_df(int act, double a, double h) {
  if (act == 1) return 1.0 - h * h;
  if (act == 2) return 1.0 / ((1.0 + a * a) * sqrt(1.0 + a * a));
  return a > 0.0 ? 1.0 : 0.0;
}

/* Layer l maps dims[l] -> dims[l+1]; activation after every layer but the last. */
static void mlp_dims(int nin, int width, int nhidden, int nout, int *dims) {
  dims[0] = nin;
  for (int l = 1; l <= nhidden; ++l) dims[l] = width;
  dims[nhidden + 1] = nout;
}

void lpm_oracle_mlp_forward(int nin, int width, int nhidden, int nout, int act,
                            const double *params, int nbatch, const double *in, double *out) {
  int dims[64];
  mlp_dims(nin, width, nhidden, nout, dims);
  const int L = nhidden + 1;
  int maxd = width > nin ? width : nin;
  if (nout > maxd) maxd = nout;
  double *h0 = (double *)malloc(sizeof(double) * (size_t)maxd);
  double *h1 = (double *)malloc(sizeof(double) * (size_t)maxd);
  for (int b = 0; b < nbatch; ++b) {
    memcpy(h0, in + (size_t)b * nin, sizeof(double) * nin);
    const double *p = params;
    for (int l = 0; l < L; ++l) {
      int di = dims[l], dout = dims[l + 1];
      const double *W = p, *bias = p + (size_t)di * dout;
      for (int o = 0; o < dout; ++o) {
        double s = bias[o];
        for (int k = 0; k < di; ++k) s += W[(size_t)o * di + k] * h0[k];
        h1[o] = (l < L - 1) ? act_f(act, s) : s;
      }
      p += (size_t)di * dout + dout;
      if (l < L - 1) memcpy(h0, h1, sizeof(double) * dout);
    }
    memcpy(out + (size_t)b * nout, h1, sizeof(double) * nout);
  }
  free(h0);
  free(h1);
}

double lpm_oracle_mlp_loss_grad(int nin, int width, int nhidden, int nout, int act,
                                const double *params, int nbatch, const double *in,
                                const double *target, double *grad) {
  int dims[64];
  mlp_dims(nin, width, nhidden, nout, dims);
  const int L = nhidden + 1;
  long np = lpm_oracle_mlp_param_count(nin, width, nhidden, nout);
  memset(grad, 0, sizeof(double) * (size_t)np);
  int maxd = width > nin ? width : nin;
  if (nout > maxd) maxd = nout;
  /* per-layer pre-activations and activations for one sample */
  double *pre = (double *)malloc(sizeof(double) * (size_t)maxd * (L + 1));
  double *post = (double *)malloc(sizeof(double) * (size_t)maxd * (L + 1));
  double *delta = (double *)malloc(sizeof(double) * (size_t)maxd);
  double *dprev = (double *)malloc(sizeof(double) * (size_t)maxd);
  long off[64];
  off[0] = 0;
  for (int l = 0; l < L; ++l) off[l + 1] = off[l] + (long)dims[l] * dims[l + 1] + dims[l + 1];
  double loss = 0.0;
  const double scale = 1.0 / ((double)nbatch * nout);
  for (int b = 0; b < nbatch; ++b) {
    memcpy(post, in + (size_t)b * nin, sizeof(double) * nin);
    for (int l = 0; l < L; ++l) {
      int di = dims[l], dout = dims[l + 1];
      const double *W = params + off[l], *bias = W + (size_t)di * dout;
      const double *h = post + (size_t)maxd * l;
      double *z = pre + (size_t)maxd * (l + 1), *hn = post + (size_t)maxd * (l + 1);
      for (int o = 0; o < dout; ++o) {
        double s = bias[o];
        for (int k = 0; k < di; ++k) s += W[(size_t)o * di + k] * h[k];
        z[o] = s;
        hn[o] = (l < L - 1) ? act_f(act, s) : s;
      }
    }
    const double *yout = post + (size_t)maxd * L;
    for (int o = 0; o < nout; ++o) {
      double r = yout[o] - target[(size_t)b * nout + o];
      loss += r * r * scale;
      delta[o] = 2.0 * r * scale;
    }
    for (int l = L - 1; l >= 0; --l) {
      int di = dims[l], dout = dims[l + 1];
      const double *W = params + off[l];
      double *gW = grad + off[l], *gb = gW + (size_t)di * dout;
      const double *h = post + (size_t)maxd * l;
      for (int o = 0; o < dout; ++o) {
        gb[o] += delta[o];
        for (int k = 0; k < di; ++k) gW[(size_t)o * di + k] += delta[o] * h[k];
      }
      if (l > 0) {
        const double *zprev = pre + (size_t)maxd * l;
        for (int k = 0; k < di; ++k) {
          double s = 0.0;
          for (int o = 0; o < dout; ++o) s += W[(size_t)o * di + k] * delta[o];
          dprev[k] = s * act_df(act, zprev[k], h[k]);
        }
        memcpy(delta, dprev, sizeof(double) * di);
      }
    }
  }
  free(pre);
  free(post);
  free(delta);
  free(dprev);
  return loss;
}
