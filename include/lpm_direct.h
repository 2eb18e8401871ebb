/*
 * lpm_direct.h -- C ABI of the MI355X (gfx950) particle direct-sum library.
 *
 * Drop-in boundary for lpm-v2's O(T*S) direct sums.  Every entry point takes
 * plain pointers and sizes (ISO_C_BINDING compatible: double = real(c_double),
 * int = integer(c_int), int32_t = integer(c_int32_t)), returns an int status
 * (LPM_OK = 0) and never aborts; on failure lpm_last_error() gives the reason.
 *
 * Conventions shared by all entry points
 *  - arrays are 0-based SoA of length n; the target block is the half-open
 *    range [i0, i1).  A Fortran MPISetup block maps as i0 = indexStart - 1,
 *    i1 = indexEnd (MPISetup.f90:36-42, 132-146);
 *  - active[j] != 0 marks an active particle (the reference's
 *    logical(klog) isActive / activeMask; any nonzero value is true);
 *  - outputs are written for targets i0..i1-1 only, other entries untouched;
 *  - host-pointer calls are synchronous: inputs are copied to HBM, the sum
 *    runs, outputs are copied back before the call returns;
 *  - precision: LPM_PREC_FP64 runs the pair kernel in FP64 (the plain
 *    drop-in), LPM_PREC_FP32_KAHAN runs it in FP32 with compensated
 *    per-target accumulation and FP64 set-up/epilogue.
 * A context is not thread-safe; use one per host thread / MPI rank.
 */
#ifndef LPM_DIRECT_H
#define LPM_DIRECT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LPM_OK 0
#define LPM_ERR_ARG 1         /* invalid argument (null pointer, bad range, ...) */
#define LPM_ERR_HIP 2         /* HIP runtime error                                */
#define LPM_ERR_NOMEM 3       /* device allocation failed                         */
#define LPM_ERR_STATE 4       /* object in the wrong state for the call           */
#define LPM_ERR_UNSUPPORTED 5 /* option not supported                             */

/* precision modes */
#define LPM_PREC_FP64 0
#define LPM_PREC_FP32_KAHAN 1

/* pair-kernel kinds (timing statistics and tuning) */
#define LPM_K_SPHERE_BS 0    /* sphere Biot-Savart velocity     */
#define LPM_K_PLANE_BS 1     /* planar Biot-Savart velocity     */
#define LPM_K_PSE8 2         /* PSE Laplacian, 8th-order kernel */
#define LPM_K_PSE2 3         /* PSE Laplacian, 2nd-order kernel */
#define LPM_K_SPHERE_GREEN 4 /* sphere stream functions         */
#define LPM_K_PLANE_GREEN 5  /* planar stream function          */
#define LPM_K_FAR_FIELD 6    /* learned far field: near + expansion + network */
#define LPM_K_PSE_SPHERE8 7  /* sphere PSE Laplacian, 8th-order kernel */
#define LPM_K_PSE_DERIV_PLANE 8   /* PSE derivative sums, plane (gradient, partials, double dot) */
#define LPM_K_PSE_DERIV_SPHERE 9  /* PSE derivative sums, sphere (gradient, double dot, divergence) */
#define LPM_K_PSE_INTERP_PLANE 10 /* PSE interpolation at points, plane */
#define LPM_K_PSE_INTERP_SPHERE 11 /* PSE interpolation at points, sphere */
#define LPM_K_SWE_PLANE 12        /* planar shallow-water right-hand side */
#define LPM_K_COUNT 13

typedef struct lpm_ctx lpm_ctx;

/* Library version as 10000*major + 100*minor + patch. */
int lpm_version(void);

/* Context on HIP device `device` with its own stream and workspace.  On
 * failure *ctx is NULL and lpm_last_error(NULL) gives the reason. */
int lpm_create(lpm_ctx **ctx, int device);
int lpm_destroy(lpm_ctx *ctx);
/* Message of the last failed call on ctx (or of lpm_create when ctx is NULL);
 * "" when the last call succeeded. */
const char *lpm_last_error(const lpm_ctx *ctx);
int lpm_synchronize(lpm_ctx *ctx);

/* ---------------------------------------------------------------- sums -- */

/* Sphere Biot-Savart velocity.  Replaces BVESphereVelocity
 * (SphereBVESolver.f90:377-430; same arguments in the same order plus the
 * explicit block) and the public duplicate setVelocityFromVorticity
 * (SphereBVE.f90:489-531).
 *   u_i = sum_{active j != i} (x_i x x_j) (-zeta_j A_j) / (4 pi R (R^2 - x_i.x_j))
 * rotation_rate is accepted and unused, as in the reference. */
int lpm_bve_sphere_velocity(lpm_ctx *ctx, double *u, double *v, double *w, const double *x,
                            const double *y, const double *z, const double *relvort,
                            const double *area, double radius, double rotation_rate,
                            const int32_t *active, int n, int i0, int i1, int precision);

/* Allocation-free variant of lpm_bve_sphere_velocity.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code next
 * to HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as lpm_bve_sphere_velocity, except that every
 * array is DEVICE memory of the context's device (x, y, z, relvort, area,
 * active of length n; u, v, w of length n, written for targets [i0, i1)).
 * The active list, the pair records and the split partials live in a static
 * 64 MiB device array that each GPU holds for the life of the process, so
 * the call allocates nothing; n is limited to 131,072 particles
 * (LPM_NOALLOC_MAX_N; larger n fails with LPM_ERR_ARG).  The array is shared
 * by every call on one GPU (all contexts of that device) and keeps the last
 * call's contents: calls that overlap on one GPU race.  The call runs on the
 * context's stream (inputs must be complete when it is made) and returns
 * when the outputs are written.  Source splits are sized from n rather than
 * the active count, so FP32 results agree with lpm_bve_sphere_velocity to
 * rounding, not bitwise. */
#define LPM_NOALLOC_MAX_N 131072
int lpm_bve_sphere_velocity_noalloc(lpm_ctx *ctx, double *u, double *v, double *w, const double *x,
                                    const double *y, const double *z, const double *relvort,
                                    const double *area, double radius, double rotation_rate,
                                    const int32_t *active, int n, int i0, int i1, int precision);

/* Planar Biot-Savart velocity (singular point vortices).  Replaces
 * planarIncompressibleVelocity (PlaneIncompressibleSolver.f90:278-316) and
 * PlanarIncompressible.f90:426-466.
 *   u_i = -sum (y_i - y_j) s_ij,  v_i = sum (x_i - x_j) s_ij,
 *   s_ij = omega_j A_j / (2 pi |x_i - x_j|^2),  active j != i */
int lpm_plane_velocity(lpm_ctx *ctx, double *u, double *v, const double *x, const double *y,
                       const double *vort, const double *area, const int32_t *active, int n,
                       int i0, int i1, int precision);

/* PSE Laplacian on the plane.  Replaces PSEPlaneLaplacianAtParticles
 * (PSEDirectSum.f90:467-500), which takes eps from PSE%eps (PSEDirectSum.f90:56-60).
 *   lap_i = eps^-4 sum_{active j} (f_j - f_i) K(|x_i - x_j| / eps) A_j
 * order 8: K = bivariateLaplacianKernel8 (PSEDirectSum.f90:622-627);
 * order 2: K = (4/pi) exp(-rho^2). */
int lpm_pse_plane_laplacian(lpm_ctx *ctx, double *lap, const double *x, const double *y,
                            const double *f, const double *area, const int32_t *active, int n,
                            int i0, int i1, double eps, int order, int precision);

/* PSE Laplacian on the sphere of radius R.  Replaces PSESphereLaplacianAtParticles
 * (PSEDirectSum.f90:502-532):
 *   lap_i = eps^-4 sum_{active j} (f_j - f_i) K8(d_ij / eps) A_j
 * with d_ij the great-circle distance R theta_ij between the directions of x_i
 * and x_j (SphereDistance, SphereGeometry.f90:107-125), K8 as above. */
int lpm_pse_sphere_laplacian(lpm_ctx *ctx, double *lap, const double *x, const double *y,
                             const double *z, const double *f, const double *area,
                             const int32_t *active, int n, int i0, int i1, double eps,
                             double radius, int precision);

/* Sphere PSE Laplacian at nt arbitrary points p_k with given point values pf:
 *   lap_k = eps^-4 sum_{active j} (f_j - pf_k) K8(d(p_k, x_j) / eps) A_j
 * The lat-lon grid loop of the reference's spherePSETest
 * (SpherePSEConvTest.f90:187-207); no self pair is skipped. */
int lpm_pse_sphere_laplacian_at_points(lpm_ctx *ctx, double *lap, const double *px, const double *py,
                                       const double *pz, const double *pf, int nt, const double *x,
                                       const double *y, const double *z, const double *f,
                                       const double *area, const int32_t *active, int n, double eps,
                                       double radius, int precision);

/* PSE first-derivative operators.  With the derivative kernel
 *   K1(rho) = (-20 + 20 rho^2 - 5 rho^4 + rho^6 / 3) exp(-rho^2) / pi
 * (bivariateFirstDerivativeKernel8, PSEDirectSum.f90:629-634) and
 *   D[c]_i = eps^-4 sum_{active j} (c_j + c_i) (x_i - x_j) K1(d_ij / eps) A_j,
 * d_ij the chord (plane) or great-circle (sphere) distance:
 *
 * Plane gradient.  Replaces PSEPlaneGradientAtParticles (PSEDirectSum.f90:180-217):
 *   (gx, gy)_i = D[f]_i */
int lpm_pse_plane_gradient(lpm_ctx *ctx, double *gx, double *gy, const double *x, const double *y,
                           const double *f, const double *area, const int32_t *active, int n,
                           int i0, int i1, double eps, int precision);

/* Sphere gradient.  Replaces PSESphereGradientAtParticles (PSEDirectSum.f90:220-267):
 *   g_i = (I - x_i x_i^T) D[f]_i
 * (SphereProjection, SphereGeometry.f90:578-590, as written for the unit
 * sphere).  The separation x_i - x_j is taken as R (d_i - d_j) of the unit
 * directions, which is x_i - x_j for particles on the sphere. */
int lpm_pse_sphere_gradient(lpm_ctx *ctx, double *gx, double *gy, double *gz, const double *x,
                            const double *y, const double *z, const double *f, const double *area,
                            const int32_t *active, int n, int i0, int i1, double eps, double radius,
                            int precision);

/* Plane second partials of a gradient field (gx, gy).  Replaces
 * PSEPlaneSecondPartialsAtParticles (PSEDirectSum.f90:269-320):
 *   dxx = D[gx]_x,  dxy = (D[gx]_y + D[gy]_x) / 2,  dyy = D[gy]_y */
int lpm_pse_plane_second_partials(lpm_ctx *ctx, double *dxx, double *dxy, double *dyy,
                                  const double *x, const double *y, const double *gx,
                                  const double *gy, const double *area, const int32_t *active,
                                  int n, int i0, int i1, double eps, int precision);

/* Plane double dot grad u : grad u.  Replaces PSEPlaneDoubleDotProductAtParticles
 * (PSEDirectSum.f90:322-365):  ux^2 + 2 uy vx + vy^2 with (ux, uy) = D[u],
 * (vx, vy) = D[v]. */
int lpm_pse_plane_double_dot(lpm_ctx *ctx, double *ddot, const double *x, const double *y,
                             const double *u, const double *v, const double *area,
                             const int32_t *active, int n, int i0, int i1, double eps,
                             int precision);

/* Sphere double dot.  Replaces PSESphereDoubleDotProductAtParticles
 * (PSEDirectSum.f90:367-418): ux^2 + vy^2 + wz^2 + 2 (uy vx + uz wx + vz wy)
 * with the unprojected sums D[u], D[v] and, as the reference writes it
 * (zComp(j) + yComp(i), :401-406), the w row summing (w_j + v_i). */
int lpm_pse_sphere_double_dot(lpm_ctx *ctx, double *ddot, const double *x, const double *y,
                              const double *z, const double *u, const double *v, const double *w,
                              const double *area, const int32_t *active, int n, int i0, int i1,
                              double eps, double radius, int precision);

/* Sphere divergence.  Replaces PSESphereDivergenceAtParticles (PSEDirectSum.f90:537-579):
 *   div_i = sum_{a,b} P_ab(x_i) D[v_a]_b,  P = I - x_i x_i^T */
int lpm_pse_sphere_divergence(lpm_ctx *ctx, double *div, const double *x, const double *y,
                              const double *z, const double *u, const double *v, const double *w,
                              const double *area, const int32_t *active, int n, int i0, int i1,
                              double eps, double radius, int precision);

/* PSE interpolation at nt points.  Replaces PSEPlaneInterpolateScalar and
 * PSESphereInterpolateScalar (PSEDirectSum.f90:121-168), one point per call
 * there:  out_k = eps^-2 sum_{active j} f_j A_j Kd(d(p_k, x_j) / eps),
 *   Kd(rho) = (4 - 6 rho^2 + 2 rho^4 - rho^6 / 6) exp(-rho^2) / pi
 * (bivariateDeltaKernel8, PSEDirectSum.f90:611-616).  No self pair is skipped. */
int lpm_pse_plane_interpolate(lpm_ctx *ctx, double *out, const double *px, const double *py, int nt,
                              const double *x, const double *y, const double *f, const double *area,
                              const int32_t *active, int n, double eps, int precision);
int lpm_pse_sphere_interpolate(lpm_ctx *ctx, double *out, const double *px, const double *py,
                               const double *pz, int nt, const double *x, const double *y,
                               const double *z, const double *f, const double *area,
                               const int32_t *active, int n, double eps, double radius,
                               int precision);

/* Planar shallow-water right-hand-side integrals.  Replaces SWEPlaneRHSIntegrals
 * (SWEPlaneSolver.f90:457-538).  bottom[j] = topoFn(x_j, y_j) evaluated by the
 * caller (the reference passes the function; the surface height is
 * h_j + bottom_j).  For each target i, over active j:
 *   (u, v)_i  = sum_{j != i} point-vortex + point-source velocity of
 *               (vort_j A_j, div_j A_j),
 *   ddot_i    = ux^2 + 2 uy vx + vy^2 of its exact gradient,
 *   lapsurf_i = eps^-4 sum_j K8(|x_i - x_j| / eps) (s_j - s_i) A_j. */
int lpm_swe_plane_rhs(lpm_ctx *ctx, double *u, double *v, double *ddot, double *lapsurf,
                      const double *x, const double *y, const double *vort, const double *div,
                      const double *h, const double *bottom, const double *area,
                      const int32_t *active, int n, int i0, int i1, double eps, int precision);

/* Sphere stream functions.  Replaces SetStreamFunctionsOnMesh (SphereBVE.f90:445-485).
 *   psi_i = sum_{active j != i} -log(R^2 - x_i.x_j) / (4 pi) zeta_j A_j
 * for zeta = relvort (psi_rel) and zeta = absvort (psi_abs). */
int lpm_bve_sphere_stream(lpm_ctx *ctx, double *psi_rel, double *psi_abs, const double *x,
                          const double *y, const double *z, const double *relvort,
                          const double *absvort, const double *area, double radius,
                          const int32_t *active, int n, int i0, int i1, int precision);

/* Planar stream function.  Replaces SetStreamFunctionOnMesh (PlanarIncompressible.f90:470-505).
 *   psi_i = sum_{active j != i} log(|x_i - x_j|) / (2 pi) omega_j A_j */
int lpm_plane_stream(lpm_ctx *ctx, double *psi, const double *x, const double *y,
                     const double *vort, const double *area, const int32_t *active, int n,
                     int i0, int i1, int precision);

/* ------------------------------------------------------------ helpers -- */

/* PSE%eps = max_edge_length ** power (PSEDirectSum.f90:93-113, default power 0.75). */
int lpm_pse_eps(double max_edge_length, double power, double *eps);

/* MPISetup LoadBalance (MPISetup.f90:132-146): 1-based inclusive blocks. */
int lpm_load_balance(int n_items, int n_procs, int *index_start, int *index_end,
                     int *message_length);

/* FP64 conserved integrals over active particles: TotalKE and TotalEnstrophy
 * (SphereBVE.f90:410-441; PlanarIncompressible.f90:368-401) and the total
 * circulation sum zeta A.  w may be NULL (plane). */
int lpm_diagnostics(lpm_ctx *ctx, const double *u, const double *v, const double *w,
                    const double *vort, const double *area, const int32_t *active, int n,
                    double *ke, double *enstrophy, double *circulation);

/* ------------------------------------------- device-resident RK4 solvers -- */

/* Sphere BVE solver: the state (positions, relative vorticity, stored
 * velocity) lives in HBM for the solver's lifetime.  Mirrors BVESolver
 * New / Timestep / Delete (SphereBVESolver.f90:80-90, 112-353).
 * u, v, w may be NULL: the velocity is then computed from the state (the
 * SetVelocityOnMesh call drivers make before New).  absvort may be NULL when
 * no stream functions are wanted. */
typedef struct lpm_bve_solver lpm_bve_solver;

#define LPM_SOLVER_STREAM_EACH_STEP 1 /* end every step with the stream functions (SphereBVESolver.f90:352) */
#define LPM_SOLVER_USE_GRAPH 2        /* replay each RK4 step as a captured hipGraph */
/* Project every RK stage input and every new position back onto the sphere
 * of radius R in FP64 (x <- R x / |x|) before the velocity sum.  The
 * reference does not (SphereBVESolver.f90:249-329): its R^2 - x.y turns
 * negative for near pairs once the stage displacement exceeds the particle
 * spacing, so the un-projected scheme is only usable at coarse resolution. */
#define LPM_SOLVER_PROJECT_STAGES 4

int lpm_bve_solver_new(lpm_ctx *ctx, lpm_bve_solver **solver, int n, const double *x,
                       const double *y, const double *z, const double *relvort,
                       const double *absvort, const double *area, const int32_t *active,
                       const double *u, const double *v, const double *w, double radius,
                       double rotation_rate, int precision, int flags);
/* Remesh re-upload: replace the solver's particles with a new set of n
 * (the output of LagrangianRemeshBVEWithVorticityFunction, SsrfpackRemesh.f90:388-,
 * called every remeshInterval steps, BVESingleGaussianVortex.f90:218-246).
 * Radius, rotation rate, precision and flags are kept; device buffers are
 * reused and grown only when n exceeds their size; a captured step graph is
 * rebuilt at the next step.  u, v, w may be NULL (all three): the velocity
 * is then computed on the device, as SetVelocityOnMesh does after a remesh
 * (SsrfpackRemesh.f90:552).  After a failure the solver reports
 * LPM_ERR_STATE until a reload succeeds. */
int lpm_bve_solver_set_particles(lpm_bve_solver *solver, int n, const double *x, const double *y,
                                 const double *z, const double *relvort, const double *absvort,
                                 const double *area, const int32_t *active, const double *u,
                                 const double *v, const double *w);
/* One classical RK4 step (timestepPrivate, SphereBVESolver.f90:219-353). */
int lpm_bve_solver_timestep(lpm_bve_solver *solver, double dt);
int lpm_bve_solver_steps(lpm_bve_solver *solver, double dt, int nsteps);
/* Copies the state back; any pointer may be NULL. */
int lpm_bve_solver_get_state(lpm_bve_solver *solver, double *x, double *y, double *z,
                             double *relvort, double *u, double *v, double *w);
/* Recomputes the stream functions of the current state into host arrays. */
int lpm_bve_solver_stream(lpm_bve_solver *solver, double *psi_rel, double *psi_abs);
int lpm_bve_solver_diagnostics(lpm_bve_solver *solver, double *ke, double *enstrophy,
                               double *circulation);
int lpm_bve_solver_delete(lpm_bve_solver *solver);

/* Planar incompressible solver: PlaneSolver New / Timestep / Delete
 * (PlaneIncompressibleSolver.f90:66-76, 103-259).  Vorticity is carried. */
typedef struct lpm_plane_solver lpm_plane_solver;

int lpm_plane_solver_new(lpm_ctx *ctx, lpm_plane_solver **solver, int n, const double *x,
                         const double *y, const double *vort, const double *area,
                         const int32_t *active, const double *u, const double *v,
                         int precision, int flags);
/* Remesh re-upload for the planar solver (BIVARRemesh.f90 rebuilds the
 * particles on the host); same contract as lpm_bve_solver_set_particles. */
int lpm_plane_solver_set_particles(lpm_plane_solver *solver, int n, const double *x, const double *y,
                                   const double *vort, const double *area, const int32_t *active,
                                   const double *u, const double *v);
int lpm_plane_solver_timestep(lpm_plane_solver *solver, double dt);
int lpm_plane_solver_steps(lpm_plane_solver *solver, double dt, int nsteps);
int lpm_plane_solver_get_state(lpm_plane_solver *solver, double *x, double *y, double *u,
                               double *v);
int lpm_plane_solver_stream(lpm_plane_solver *solver, double *psi);
int lpm_plane_solver_diagnostics(lpm_plane_solver *solver, double *ke, double *enstrophy,
                                 double *circulation);
int lpm_plane_solver_delete(lpm_plane_solver *solver);

/* ------------------------------------------------- batched sphere problems -- */

/* Many independent sphere BVE problems of uneven size advanced together: one
 * persistent launch drains a work list of (problem, target tile, source
 * split) items per velocity evaluation, and each RK4 stage can be replayed
 * as a captured hipGraph (LPM_SOLVER_USE_GRAPH).  Problem k owns particles
 * [offsets[k], offsets[k+1]) of the concatenated arrays and has its own
 * radius[k] and rotation_rate[k]; problems never interact.  Each problem's
 * step is the BVESolver step (SphereBVESolver.f90:219-353); flags as for
 * lpm_bve_solver_new.  No counterpart in the reference (which runs one
 * problem per program). */
typedef struct lpm_bve_batch lpm_bve_batch;

int lpm_bve_batch_new(lpm_ctx *ctx, lpm_bve_batch **batch, int n_inst, const int64_t *offsets,
                      const double *x, const double *y, const double *z, const double *relvort,
                      const double *area, const int32_t *active, const double *radius,
                      const double *rotation_rate, int precision, int flags);
int lpm_bve_batch_steps(lpm_bve_batch *batch, double dt, int nsteps);
int lpm_bve_batch_get_state(lpm_bve_batch *batch, double *x, double *y, double *z,
                            double *relvort, double *u, double *v, double *w);
/* Tile shape: targets per lane (2, 4, 8), sources per compensated tile and
 * sources per work item. */
int lpm_bve_batch_set_tuning(lpm_bve_batch *batch, int targets_per_lane, int tile_sources,
                             int chunk_sources);
int lpm_bve_batch_get_tuning(lpm_bve_batch *batch, int *targets_per_lane, int *tile_sources,
                             int *chunk_sources, int *n_items);
/* Autotuner: times one velocity evaluation for each candidate tile shape and
 * keeps the fastest (any output pointer may be NULL). */
int lpm_bve_batch_autotune(lpm_bve_batch *batch, int *targets_per_lane, int *tile_sources,
                           int *chunk_sources, double *best_ms);
/* Pair interactions of one velocity evaluation (sum over problems of T*S - S). */
int lpm_bve_batch_pairs(lpm_bve_batch *batch, int64_t *pairs_per_eval);
int lpm_bve_batch_delete(lpm_bve_batch *batch);

/* -------------------------------------------- far-field surrogate network -- */

/* MLP nin -> width (x nhidden, activation tanh or ISRU) -> nout, parameters per
 * layer W[out][in] (row-major) then b[out] in FP32.  Training runs in FP32 on the GPU
 * (hand-written GEMM forward/backward and Adam); inference runs in FP16 on
 * MFMA with FP32 accumulation (nin <= 32, width 128, nhidden 3, nout <= 4).
 * No counterpart in the reference. */
typedef struct lpm_mlp lpm_mlp;
#define LPM_ACT_TANH 1 /* hidden activation tanh */
#define LPM_ACT_ISRU 2 /* hidden activation x / sqrt(1 + x^2) */

int lpm_mlp_new(lpm_ctx *ctx, lpm_mlp **mlp, int nin, int width, int nhidden, int nout,
                int activation, uint64_t seed);
int lpm_mlp_param_count(lpm_mlp *mlp, int64_t *count);
int lpm_mlp_set_params(lpm_mlp *mlp, const float *params, int64_t count);
int lpm_mlp_get_params(lpm_mlp *mlp, float *params, int64_t count);
/* in: [n][nin], out: [n][nout]; use_mfma selects the FP16 MFMA kernel
 * (inputs, weights and activations rounded to FP16, |values| <= 65504). */
int lpm_mlp_forward(lpm_mlp *mlp, int n, const float *in, float *out, int use_mfma);
/* Mean-squared error over n*nout outputs and its gradient (FP32). */
int lpm_mlp_loss_grad(lpm_mlp *mlp, int n, const float *in, const float *target, float *grad,
                      double *loss);
int lpm_mlp_train(lpm_mlp *mlp, int n, const float *in, const float *target, int epochs,
                  int batch, double lr, uint64_t seed, double *final_loss);
int lpm_mlp_delete(lpm_mlp *mlp);

/* ------------------------------------------------ learned far field (BVE) -- */

/* Sphere BVE RK4 with the far field of coarse cells replaced by an expansion
 * (+ the network's correction).  cell[i] in [0, ncell) is particle i's coarse
 * cell at set-up; it defines each cell's generator (the direction of the
 * area-weighted centroid of its active sources) and frame anchor, and every
 * evaluation bins the particles to the nearest generator, so cells stay fixed
 * regions of the sphere.  For target cell t and source cell s at
 * centroid angle a with angular radii rho_t, rho_s: a < near_radius + rho_t +
 * rho_s is summed directly (FP32 compensated), everything else by the cell's
 * expansion of order P in its moments M0..MP about the centroid (P = 2: 10
 * moments; P = 4: 35), and cells with a < band_radius + rho_t + rho_s also get
 * the network's correction of the expansion remainder (FP16 MFMA).  Stage
 * points are projected onto the sphere (as with LPM_SOLVER_PROJECT_STAGES).
 * flags: LPM_FF_ORDER(P) with P in 2..4 (0 selects 4), plus LPM_FF_SAMPLE_DRAWS(k):
 * each training sample draws k band cells of its target (k in 1..7, 0 = 1) and
 * keeps the one nearest the target, so that the near band, whose remainders
 * dominate the error, is sampled more.  No counterpart in the
 * reference.
 * The network (13 -> 128 -> 128 -> 128 -> 3, ISRU) sees a cell in the cell's
 * frame Q (rows e1, e2, e3: e3 along the centroid c, e1 towards the cell's
 * anchor, the active source farthest from c at set-up): inputs Q (x - c) / l,
 * M0 / W, Q M1 / (W l), Q M2 Q^T / (W l^2) (xx xy xz yy yz zz), with moments of
 * w = relvort * area about c, W = sum |w|, l = 2 R sin(rho / 2).  Its output o
 * gives the remainder of the cell's G / (4 pi R) as Q^T o W / (4 pi R l^2 r^(P+2)),
 * r = |x - c| / l.
 * mode: 0 plain direct sum (FP32 compensated), 1 near + expansion,
 *       2 near + expansion + network (needs training or set_params). */
typedef struct lpm_bve_ff lpm_bve_ff;
#define LPM_FF_ORDER_MASK 7
#define LPM_FF_ORDER(p) ((p) & LPM_FF_ORDER_MASK)
#define LPM_FF_SAMPLE_DRAWS_MASK 0x38
#define LPM_FF_SAMPLE_DRAWS(k) (((k) & 7) << 3)

int lpm_bve_ff_new(lpm_ctx *ctx, lpm_bve_ff **ff, int n, const double *x, const double *y,
                   const double *z, const double *relvort, const double *area,
                   const int32_t *active, const int32_t *cell, int ncell, double radius,
                   double rotation_rate, double near_radius, double band_radius, int flags);
/* Velocity of the current state by `mode` into host arrays. */
int lpm_bve_ff_velocity(lpm_bve_ff *ff, int mode, double *u, double *v, double *w);
/* Training data from the current state: n_samples (target, band cell) pairs
 * with the network's 13 inputs and, as target, the exact remainder (direct sum
 * over the cell minus its expansion, FP64) in the cell frame, times
 * 4 pi R l^2 r^(P+2) / W; appended to the set. */
int lpm_bve_ff_collect(lpm_bve_ff *ff, int n_samples, uint64_t seed, int64_t *total);
/* Train the network on the collected set (Adam with a cosine learning-rate
 * decay over the epochs, mean-squared error of the targets divided by their
 * RMS, which the first training fixes and folds into the output layer). */
int lpm_bve_ff_train(lpm_bve_ff *ff, int epochs, int batch, double lr, uint64_t seed,
                     double *final_loss);
/* Online fine-tuning in lpm_bve_ff_steps: at every judge that keeps the learned
 * mode, replace the collected set by n_samples samples of the judged state and
 * train `epochs` epochs on it.  n_samples = 0 (the default) turns it off. */
int lpm_bve_ff_set_online(lpm_bve_ff *ff, int n_samples, int epochs, int batch, double lr);
/* Fine-tune only at judges whose error is above fraction * tol (fraction in
 * [0, 1]; 0, the default, fine-tunes at every judge that keeps the learned
 * mode). */
int lpm_bve_ff_set_online_from(lpm_bve_ff *ff, double fraction);
/* Copy samples [offset, offset + count): in [count][13], target [count][3] and,
 * when ids is not NULL, ids [count][2] = (target particle, source cell). */
int lpm_bve_ff_samples(lpm_bve_ff *ff, int64_t offset, int64_t count, float *in, float *target,
                       int32_t *ids);
int lpm_bve_ff_set_params(lpm_bve_ff *ff, const float *params, int64_t count);
int lpm_bve_ff_get_params(lpm_bve_ff *ff, float *params, int64_t count);
/* nsteps RK4 steps with velocities by `mode`; every judge_every steps the plain
 * direct sum judges the stored velocity (normwise max relative error) and above
 * tol the remaining steps fall back to the plain sum.  mode_out = mode in use
 * at the end, last_error = last judged error (-1 if none), judged = count. */
int lpm_bve_ff_steps(lpm_bve_ff *ff, double dt, int nsteps, int mode, int judge_every,
                     double tol, int *mode_out, double *last_error, int *judged);
int lpm_bve_ff_get_state(lpm_bve_ff *ff, double *x, double *y, double *z, double *relvort,
                         double *u, double *v, double *w);
/* Remesh re-upload: a new particle set of n (any size) with its areas and
 * active flags (LagrangianRemeshBVEWithVorticityFunction, SsrfpackRemesh.f90:388-).
 * The cells (fixed regions of the sphere), their frames, the trained network
 * and the collected samples are kept; the cell moments follow the new
 * particles at every evaluation.  The plain velocity of the new set is
 * stored.  After a failure the object reports LPM_ERR_STATE until a reload
 * succeeds. */
int lpm_bve_ff_set_particles(lpm_bve_ff *ff, int n, const double *x, const double *y,
                             const double *z, const double *relvort, const double *area,
                             const int32_t *active);
/* Replace positions (projected onto the sphere) and relative vorticity, and
 * store the plain velocity of the new state; cells, anchors, set-up radii, the
 * network and the collected samples are kept. */
int lpm_bve_ff_set_state(lpm_bve_ff *ff, const double *x, const double *y, const double *z,
                         const double *relvort);
/* Cell of every particle (length n) as binned by the last evaluation. */
int lpm_bve_ff_particle_cells(lpm_bve_ff *ff, int32_t *cell);
/* Near and band cell counts per target cell (length ncell each). */
int lpm_bve_ff_cells(lpm_bve_ff *ff, int *near_cells, int *band_cells);
/* Coarse level of the expansion part (parent_radius > 0; 0 turns it off, the
 * default).  Cells 4p .. 4p+3 must be the four children of one coarser cell p,
 * as mesh refinement numbers them (ncell divisible by 4).  For target cell t,
 * four sibling cells that are all far (outside the band) are replaced by one
 * order-P expansion of their union about its own centroid when the angle
 * between t's and the union's directions is at least parent_radius + rho_t +
 * rho_union; near and band cells, and every group that holds one, keep their
 * own treatment.  Takes effect at the next evaluation.  Measured at
 * icosahedral L8 with 1,280 cells, near 0.1, band 0.2: parent_radius 0.6
 * evaluates 462 instead of 1,258 expansions per target. */
int lpm_bve_ff_set_parents(lpm_bve_ff *ff, double parent_radius);
/* Expansion records evaluated per target cell by the last evaluation (length
 * ncell): far and band cells, with the coarse level their parents in their
 * place where it applies. */
int lpm_bve_ff_far_counts(lpm_bve_ff *ff, int *far_records);
int lpm_bve_ff_delete(lpm_bve_ff *ff);

/* ----------------------------------------------------- timing / tuning -- */

/* When enabled, every pair-kernel launch is bracketed by HIP events on the
 * context's stream; lpm_kernel_stats returns launches and summed milliseconds
 * per kind since the last reset. */
int lpm_set_timing(lpm_ctx *ctx, int enable);
int lpm_kernel_stats(lpm_ctx *ctx, int kind, int64_t *launches, double *total_ms);
int lpm_reset_kernel_stats(lpm_ctx *ctx);

/* Launch shape of a pair kernel: targets per lane (1, 2, 4 or 8), sources
 * per compensated tile (multiple of 8) and source splits (0 = automatic:
 * chunks of about 2,048 sources, at least enough splits to fill the GPU,
 * at most 2 GiB of FP64 split partials; a forced count is not capped, and
 * its partials take splits x outputs x targets x 8 bytes of device memory).
 * Values 0 restore the per-kind
 * defaults.  The FP64 PSE derivative, interpolation and SWE kinds have no
 * 8-targets-per-lane kernel: their calls then fail with LPM_ERR_ARG.  The
 * shape changes the schedule, not the result beyond FP32 (FP64) rounding. */
int lpm_set_tuning(lpm_ctx *ctx, int kind, int precision, int targets_per_lane,
                   int tile_sources, int splits);
int lpm_get_tuning(lpm_ctx *ctx, int kind, int precision, int *targets_per_lane,
                   int *tile_sources, int *splits);

#ifdef __cplusplus
}
#endif

#endif /* LPM_DIRECT_H */
