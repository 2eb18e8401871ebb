/*
 * lpm_oracle.h -- FP64 CPU restatement of lpm-v2's particle direct sums.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Conventions: arrays are 0-based, target ranges are half-open [i0, i1)
 * (the reference's MPISetup ranges are 1-based inclusive:
 * i0 = indexStart - 1, i1 = indexEnd).  `active` holds 0/1 per particle
 * (the reference's logical(4) isActive / activeMask).
 */
#ifndef LPM_ORACLE_H
#define LPM_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Planted-mistake selector used by the oracle pin tests (0 = faithful). */
extern int lpm_oracle_mutation;
int lpm_oracle_set_mutation(int m);
int lpm_oracle_set_threads(int nthreads);

void lpm_oracle_bve_sphere_velocity(int n, const double *x, const double *y, const double *z,
                                    const double *relvort, const double *area, const int *active,
                                    double radius, int i0, int i1, double *u, double *v, double *w);

void lpm_oracle_plane_velocity(int n, const double *x, const double *y, const double *vort,
                               const double *area, const int *active, int i0, int i1, double *u,
                               double *v);

void lpm_oracle_pse_plane_laplacian(int n, const double *x, const double *y, const double *f,
                                    const double *area, const int *active, double eps, int order,
                                    int i0, int i1, double *lap);

double lpm_oracle_pse_kernel(int order, double rho);

void lpm_oracle_pse_sphere_laplacian(int n, const double *x, const double *y, const double *z,
                                     const double *f, const double *area, const int *active,
                                     double eps, double radius, int i0, int i1, double *lap);

/* PSE first-derivative, interpolation and double-dot operators
 * (PSEDirectSum.f90:121-418, 537-579) and the planar SWE right-hand side
 * (SWEPlaneSolver.f90:457-538). */
double lpm_oracle_pse_delta_kernel(double rho);
double lpm_oracle_pse_deriv_kernel(double rho);

void lpm_oracle_pse_plane_gradient(int n, const double *x, const double *y, const double *f,
                                   const double *area, const int *active, double eps, int i0,
                                   int i1, double *gx, double *gy);

void lpm_oracle_pse_sphere_gradient(int n, const double *x, const double *y, const double *z,
                                    const double *f, const double *area, const int *active,
                                    double eps, double radius, int i0, int i1, double *gx,
                                    double *gy, double *gz);

void lpm_oracle_pse_plane_second_partials(int n, const double *x, const double *y,
                                          const double *gx, const double *gy, const double *area,
                                          const int *active, double eps, int i0, int i1,
                                          double *dxx, double *dxy, double *dyy);

void lpm_oracle_pse_plane_double_dot(int n, const double *x, const double *y, const double *u,
                                     const double *v, const double *area, const int *active,
                                     double eps, int i0, int i1, double *ddot);

void lpm_oracle_pse_sphere_double_dot(int n, const double *x, const double *y, const double *z,
                                      const double *u, const double *v, const double *w,
                                      const double *area, const int *active, double eps,
                                      double radius, int i0, int i1, double *ddot);

void lpm_oracle_pse_sphere_divergence(int n, const double *x, const double *y, const double *z,
                                      const double *u, const double *v, const double *w,
                                      const double *area, const int *active, double eps,
                                      double radius, int i0, int i1, double *div);

void lpm_oracle_pse_plane_interpolate(int n, const double *x, const double *y, const double *f,
                                      const double *area, const int *active, double eps, int nt,
                                      const double *px, const double *py, double *out);

void lpm_oracle_pse_sphere_interpolate(int n, const double *x, const double *y, const double *z,
                                       const double *f, const double *area, const int *active,
                                       double eps, double radius, int nt, const double *px,
                                       const double *py, const double *pz, double *out);

void lpm_oracle_pse_sphere_laplacian_points(int n, const double *x, const double *y,
                                            const double *z, const double *f, const double *area,
                                            const int *active, double eps, double radius, int nt,
                                            const double *px, const double *py, const double *pz,
                                            const double *pf, double *out);

void lpm_oracle_swe_plane_rhs(int n, const double *x, const double *y, const double *vort,
                              const double *divg, const double *h, const double *bottom,
                              const double *area, const int *active, double eps, int i0, int i1,
                              double *u, double *v, double *ddot, double *lap);

void lpm_oracle_bve_sphere_stream(int n, const double *x, const double *y, const double *z,
                                  const double *relvort, const double *absvort, const double *area,
                                  const int *active, double radius, int i0, int i1,
                                  double *psi_rel, double *psi_abs);

void lpm_oracle_plane_stream(int n, const double *x, const double *y, const double *vort,
                             const double *area, const int *active, int i0, int i1, double *psi);

void lpm_oracle_bve_timestep(int n, double *x, double *y, double *z, double *relvort, double *u,
                             double *v, double *w, const double *area, const int *active,
                             double radius, double rotation_rate, double dt);

void lpm_oracle_bve_timestep_ex(int n, double *x, double *y, double *z, double *relvort, double *u,
                                double *v, double *w, const double *area, const int *active,
                                double radius, double rotation_rate, double dt, int project);

void lpm_oracle_plane_timestep(int n, double *x, double *y, const double *vort, double *u,
                               double *v, const double *area, const int *active, double dt);

void lpm_oracle_diagnostics(int n, int ncomp, const double *u, const double *v, const double *w,
                            const double *vort, const double *area, const int *active,
                            double out[3]);

void lpm_oracle_load_balance(int nitems, int nprocs, int *index_start, int *index_end,
                             int *message_length);

double lpm_oracle_pse_eps(double max_edge_length, double power);

/* Far-field surrogate checks (no counterpart in the reference). */
void lpm_oracle_cell_moments(int n, const double *x, const double *y, const double *z,
                             const double *strength, const int *cell, int ncell,
                             const double *centroid, double *moments);

/* MLP nin -> width (x nhidden layers) -> nout; act 0 = ReLU, 1 = tanh.
 * params: per layer W[out][in] row-major followed by b[out]. */
long lpm_oracle_mlp_param_count(int nin, int width, int nhidden, int nout);

void lpm_oracle_mlp_forward(int nin, int width, int nhidden, int nout, int act,
                            const double *params, int nbatch, const double *in, double *out);

/* Mean-squared error over nbatch*nout outputs; returns the loss, writes dL/dparams. */
double lpm_oracle_mlp_loss_grad(int nin, int width, int nhidden, int nout, int act,
                                const double *params, int nbatch, const double *in,
                                const double *target, double *grad);

#ifdef __cplusplus
}
#endif
#endif
