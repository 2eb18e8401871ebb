!> @file lpm_direct_f.f90
!> ISO_C_BINDING interfaces to liblpm_direct.so (include/lpm_direct.h).
!>
!> Fortran callers (lpm-v2's solvers) use these interfaces directly, or the
!> drop-in wrappers of module lpm_bve_solver_f.  Arrays are passed as
!> assumed-size real(c_double) / integer(c_int32_t); scalars by value.
module lpm_direct_f
  use iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: LPM_OK = 0, LPM_ERR_ARG = 1, LPM_ERR_HIP = 2, LPM_ERR_NOMEM = 3, &
                               LPM_ERR_STATE = 4, LPM_ERR_UNSUPPORTED = 5
  integer(c_int), parameter :: LPM_PREC_FP64 = 0, LPM_PREC_FP32_KAHAN = 1
  integer(c_int), parameter :: LPM_SOLVER_STREAM_EACH_STEP = 1, LPM_SOLVER_USE_GRAPH = 2, &
                               LPM_SOLVER_PROJECT_STAGES = 4

  interface
    integer(c_int) function lpm_version() bind(C, name="lpm_version")
      import :: c_int
    end function

    integer(c_int) function lpm_create(ctx, device) bind(C, name="lpm_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function lpm_destroy(ctx) bind(C, name="lpm_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function lpm_last_error_c(ctx) bind(C, name="lpm_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function lpm_bve_sphere_velocity(ctx, u, v, w, x, y, z, relvort, area, radius, &
        rotation_rate, active, n, i0, i1, precision) bind(C, name="lpm_bve_sphere_velocity")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*), w(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*)
      real(c_double), value :: radius, rotation_rate
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
    end function

    integer(c_int) function lpm_plane_velocity(ctx, u, v, x, y, vort, area, active, n, i0, i1, precision) &
        bind(C, name="lpm_plane_velocity")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*)
      real(c_double), intent(in) :: x(*), y(*), vort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
    end function

    integer(c_int) function lpm_pse_plane_laplacian(ctx, lap, x, y, f, area, active, n, i0, i1, eps, &
        order, precision) bind(C, name="lpm_pse_plane_laplacian")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: lap(*)
      real(c_double), intent(in) :: x(*), y(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, order, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_pse_sphere_laplacian(ctx, lap, x, y, z, f, area, active, n, i0, i1, eps, &
        radius, precision) bind(C, name="lpm_pse_sphere_laplacian")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: lap(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps, radius
    end function

    ! PSE operators and the planar SWE right-hand side (PSEDirectSum.f90:121-418, 537-579;
    ! SWEPlaneSolver.f90:457-538)
    integer(c_int) function lpm_pse_sphere_laplacian_at_points(ctx, lap, px, py, pz, pf, nt, x, y, z, f, &
        area, active, n, eps, radius, precision) bind(C, name="lpm_pse_sphere_laplacian_at_points")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: lap(*)
      real(c_double), intent(in) :: px(*), py(*), pz(*), pf(*), x(*), y(*), z(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: nt, n, precision
      real(c_double), value :: eps, radius
    end function

    integer(c_int) function lpm_pse_plane_gradient(ctx, gx, gy, x, y, f, area, active, n, i0, i1, eps, &
        precision) bind(C, name="lpm_pse_plane_gradient")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: gx(*), gy(*)
      real(c_double), intent(in) :: x(*), y(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_pse_sphere_gradient(ctx, gx, gy, gz, x, y, z, f, area, active, n, i0, i1, &
        eps, radius, precision) bind(C, name="lpm_pse_sphere_gradient")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: gx(*), gy(*), gz(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps, radius
    end function

    integer(c_int) function lpm_pse_plane_second_partials(ctx, dxx, dxy, dyy, x, y, gx, gy, area, active, &
        n, i0, i1, eps, precision) bind(C, name="lpm_pse_plane_second_partials")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: dxx(*), dxy(*), dyy(*)
      real(c_double), intent(in) :: x(*), y(*), gx(*), gy(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_pse_plane_double_dot(ctx, ddot, x, y, u, v, area, active, n, i0, i1, eps, &
        precision) bind(C, name="lpm_pse_plane_double_dot")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: ddot(*)
      real(c_double), intent(in) :: x(*), y(*), u(*), v(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_pse_sphere_double_dot(ctx, ddot, x, y, z, u, v, w, area, active, n, i0, i1, &
        eps, radius, precision) bind(C, name="lpm_pse_sphere_double_dot")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: ddot(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), u(*), v(*), w(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps, radius
    end function

    integer(c_int) function lpm_pse_sphere_divergence(ctx, div, x, y, z, u, v, w, area, active, n, i0, i1, &
        eps, radius, precision) bind(C, name="lpm_pse_sphere_divergence")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: div(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), u(*), v(*), w(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps, radius
    end function

    integer(c_int) function lpm_pse_plane_interpolate(ctx, out, px, py, nt, x, y, f, area, active, n, eps, &
        precision) bind(C, name="lpm_pse_plane_interpolate")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: out(*)
      real(c_double), intent(in) :: px(*), py(*), x(*), y(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: nt, n, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_pse_sphere_interpolate(ctx, out, px, py, pz, nt, x, y, z, f, area, active, &
        n, eps, radius, precision) bind(C, name="lpm_pse_sphere_interpolate")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: out(*)
      real(c_double), intent(in) :: px(*), py(*), pz(*), x(*), y(*), z(*), f(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: nt, n, precision
      real(c_double), value :: eps, radius
    end function

    integer(c_int) function lpm_swe_plane_rhs(ctx, u, v, ddot, lapsurf, x, y, vort, div, h, bottom, area, &
        active, n, i0, i1, eps, precision) bind(C, name="lpm_swe_plane_rhs")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*), ddot(*), lapsurf(*)
      real(c_double), intent(in) :: x(*), y(*), vort(*), div(*), h(*), bottom(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
      real(c_double), value :: eps
    end function

    integer(c_int) function lpm_bve_sphere_stream(ctx, psi_rel, psi_abs, x, y, z, relvort, absvort, area, &
        radius, active, n, i0, i1, precision) bind(C, name="lpm_bve_sphere_stream")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: psi_rel(*), psi_abs(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), absvort(*), area(*)
      real(c_double), value :: radius
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
    end function

    integer(c_int) function lpm_plane_stream(ctx, psi, x, y, vort, area, active, n, i0, i1, precision) &
        bind(C, name="lpm_plane_stream")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: psi(*)
      real(c_double), intent(in) :: x(*), y(*), vort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n, i0, i1, precision
    end function

    integer(c_int) function lpm_bve_solver_new(ctx, solver, n, x, y, z, relvort, absvort, area, active, &
        u, v, w, radius, rotation_rate, precision, flags) bind(C, name="lpm_bve_solver_new")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: solver
      integer(c_int), value :: n, precision, flags
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*)
      type(c_ptr), value :: absvort, u, v, w
      integer(c_int32_t), intent(in) :: active(*)
      real(c_double), value :: radius, rotation_rate
    end function

    integer(c_int) function lpm_bve_solver_timestep(solver, dt) bind(C, name="lpm_bve_solver_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), value :: dt
    end function

    ! remesh re-upload (SsrfpackRemesh.f90:388-, :552); absvort and u, v, w may be c_null_ptr
    integer(c_int) function lpm_bve_solver_set_particles(solver, n, x, y, z, relvort, absvort, area, active, &
        u, v, w) bind(C, name="lpm_bve_solver_set_particles")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: solver
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*)
      type(c_ptr), value :: absvort, u, v, w
      integer(c_int32_t), intent(in) :: active(*)
    end function

    integer(c_int) function lpm_plane_solver_set_particles(solver, n, x, y, vort, area, active, u, v) &
        bind(C, name="lpm_plane_solver_set_particles")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: solver
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(*), y(*), vort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      type(c_ptr), value :: u, v
    end function

    integer(c_int) function lpm_bve_ff_set_particles(ff, n, x, y, z, relvort, area, active) &
        bind(C, name="lpm_bve_ff_set_particles")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ff
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
    end function

    integer(c_int) function lpm_bve_solver_get_state(solver, x, y, z, relvort, u, v, w) &
        bind(C, name="lpm_bve_solver_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double) :: x(*), y(*), z(*), relvort(*), u(*), v(*), w(*)
    end function

    integer(c_int) function lpm_bve_solver_stream(solver, psi_rel, psi_abs) bind(C, name="lpm_bve_solver_stream")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double) :: psi_rel(*), psi_abs(*)
    end function

    integer(c_int) function lpm_bve_solver_delete(solver) bind(C, name="lpm_bve_solver_delete")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function

    integer(c_int) function lpm_bve_solver_steps(solver, dt, nsteps) bind(C, name="lpm_bve_solver_steps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function lpm_bve_solver_diagnostics(solver, ke, enstrophy, circulation) &
        bind(C, name="lpm_bve_solver_diagnostics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), intent(out) :: ke, enstrophy, circulation
    end function

    ! ------------------------------------------------------------- helpers
    ! PSE%eps = max_edge_length ** power (PSEDirectSum.f90:93-113)
    integer(c_int) function lpm_pse_eps(max_edge_length, power, eps) bind(C, name="lpm_pse_eps")
      import :: c_int, c_double
      real(c_double), value :: max_edge_length, power
      real(c_double), intent(out) :: eps
    end function

    ! MPISetup LoadBalance (MPISetup.f90:132-146): 1-based inclusive blocks, arrays of n_procs
    integer(c_int) function lpm_load_balance(n_items, n_procs, index_start, index_end, message_length) &
        bind(C, name="lpm_load_balance")
      import :: c_int
      integer(c_int), value :: n_items, n_procs
      integer(c_int) :: index_start(*), index_end(*), message_length(*)
    end function

    ! TotalKE / TotalEnstrophy / circulation (SphereBVE.f90:410-441); w may be c_null_ptr (plane)
    integer(c_int) function lpm_diagnostics(ctx, u, v, w, vort, area, active, n, ke, enstrophy, circulation) &
        bind(C, name="lpm_diagnostics")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx, w
      real(c_double), intent(in) :: u(*), v(*), vort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      integer(c_int), value :: n
      real(c_double), intent(out) :: ke, enstrophy, circulation
    end function

    ! ------------------------------------------------ planar solver (PlaneSolver)
    ! New / Timestep / Delete (PlaneIncompressibleSolver.f90:66-76, 103-259); u, v may be c_null_ptr
    integer(c_int) function lpm_plane_solver_new(ctx, solver, n, x, y, vort, area, active, u, v, precision, &
        flags) bind(C, name="lpm_plane_solver_new")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: solver
      integer(c_int), value :: n, precision, flags
      real(c_double), intent(in) :: x(*), y(*), vort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*)
      type(c_ptr), value :: u, v
    end function

    integer(c_int) function lpm_plane_solver_timestep(solver, dt) bind(C, name="lpm_plane_solver_timestep")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), value :: dt
    end function

    integer(c_int) function lpm_plane_solver_steps(solver, dt, nsteps) bind(C, name="lpm_plane_solver_steps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), value :: dt
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function lpm_plane_solver_get_state(solver, x, y, u, v) bind(C, name="lpm_plane_solver_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double) :: x(*), y(*), u(*), v(*)
    end function

    integer(c_int) function lpm_plane_solver_stream(solver, psi) bind(C, name="lpm_plane_solver_stream")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double) :: psi(*)
    end function

    integer(c_int) function lpm_plane_solver_diagnostics(solver, ke, enstrophy, circulation) &
        bind(C, name="lpm_plane_solver_diagnostics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: solver
      real(c_double), intent(out) :: ke, enstrophy, circulation
    end function

    integer(c_int) function lpm_plane_solver_delete(solver) bind(C, name="lpm_plane_solver_delete")
      import :: c_int, c_ptr
      type(c_ptr), value :: solver
    end function

    ! --------------------------------- batched sphere problems (no reference counterpart)
    ! problem k owns particles offsets(k+1)+1 .. offsets(k+2) (0-based offsets, n_inst + 1 of them)
    integer(c_int) function lpm_bve_batch_new(ctx, batch, n_inst, offsets, x, y, z, relvort, area, active, &
        radius, rotation_rate, precision, flags) bind(C, name="lpm_bve_batch_new")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: batch
      integer(c_int), value :: n_inst, precision, flags
      integer(c_int64_t), intent(in) :: offsets(*)
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*), radius(*), rotation_rate(*)
      integer(c_int32_t), intent(in) :: active(*)
    end function

    integer(c_int) function lpm_bve_batch_steps(batch, dt, nsteps) bind(C, name="lpm_bve_batch_steps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), value :: dt
      integer(c_int), value :: nsteps
    end function

    integer(c_int) function lpm_bve_batch_get_state(batch, x, y, z, relvort, u, v, w) &
        bind(C, name="lpm_bve_batch_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double) :: x(*), y(*), z(*), relvort(*), u(*), v(*), w(*)
    end function

    integer(c_int) function lpm_bve_batch_autotune(batch, targets_per_lane, tile_sources, chunk_sources, &
        best_ms) bind(C, name="lpm_bve_batch_autotune")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: targets_per_lane, tile_sources, chunk_sources
      real(c_double), intent(out) :: best_ms
    end function

    integer(c_int) function lpm_bve_batch_delete(batch) bind(C, name="lpm_bve_batch_delete")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    ! ------------------------------------- learned far field (no reference counterpart)
    ! cell(i) in [0, ncell) is particle i's coarse cell at set-up; flags: expansion order (0 = 4)
    integer(c_int) function lpm_bve_ff_new(ctx, ff, n, x, y, z, relvort, area, active, cell, ncell, radius, &
        rotation_rate, near_radius, band_radius, flags) bind(C, name="lpm_bve_ff_new")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: ff
      integer(c_int), value :: n, ncell, flags
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*), area(*)
      integer(c_int32_t), intent(in) :: active(*), cell(*)
      real(c_double), value :: radius, rotation_rate, near_radius, band_radius
    end function

    ! mode: 0 plain direct sum, 1 near + expansion, 2 near + expansion + network
    integer(c_int) function lpm_bve_ff_velocity(ff, mode, u, v, w) bind(C, name="lpm_bve_ff_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      integer(c_int), value :: mode
      real(c_double) :: u(*), v(*), w(*)
    end function

    integer(c_int) function lpm_bve_ff_collect(ff, n_samples, seed, total) bind(C, name="lpm_bve_ff_collect")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ff
      integer(c_int), value :: n_samples
      integer(c_int64_t), value :: seed
      integer(c_int64_t), intent(out) :: total
    end function

    integer(c_int) function lpm_bve_ff_train(ff, epochs, batch, lr, seed, final_loss) bind(C, name="lpm_bve_ff_train")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ff
      integer(c_int), value :: epochs, batch
      real(c_double), value :: lr
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: final_loss
    end function

    integer(c_int) function lpm_bve_ff_set_online(ff, n_samples, epochs, batch, lr) &
        bind(C, name="lpm_bve_ff_set_online")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      integer(c_int), value :: n_samples, epochs, batch
      real(c_double), value :: lr
    end function

    ! fine-tune only at judged errors above fraction * tol (0: at every judge)
    integer(c_int) function lpm_bve_ff_set_online_from(ff, fraction) bind(C, name="lpm_bve_ff_set_online_from")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      real(c_double), value :: fraction
    end function

    integer(c_int) function lpm_bve_ff_steps(ff, dt, nsteps, mode, judge_every, tol, mode_out, last_error, &
        judged) bind(C, name="lpm_bve_ff_steps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      real(c_double), value :: dt, tol
      integer(c_int), value :: nsteps, mode, judge_every
      integer(c_int), intent(out) :: mode_out, judged
      real(c_double), intent(out) :: last_error
    end function

    integer(c_int) function lpm_bve_ff_get_state(ff, x, y, z, relvort, u, v, w) bind(C, name="lpm_bve_ff_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      real(c_double) :: x(*), y(*), z(*), relvort(*), u(*), v(*), w(*)
    end function

    ! replace positions and relative vorticity (the plain velocity of the new state is stored)
    integer(c_int) function lpm_bve_ff_set_state(ff, x, y, z, relvort) bind(C, name="lpm_bve_ff_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      real(c_double), intent(in) :: x(*), y(*), z(*), relvort(*)
    end function

    integer(c_int) function lpm_bve_ff_get_params(ff, params, count) bind(C, name="lpm_bve_ff_get_params")
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: ff
      real(c_float) :: params(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function lpm_bve_ff_set_params(ff, params, count) bind(C, name="lpm_bve_ff_set_params")
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: ff
      real(c_float), intent(in) :: params(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function lpm_bve_ff_delete(ff) bind(C, name="lpm_bve_ff_delete")
      import :: c_int, c_ptr
      type(c_ptr), value :: ff
    end function

    ! ------------------------------------------------------- timing / tuning
    integer(c_int) function lpm_set_timing(ctx, enable) bind(C, name="lpm_set_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function lpm_kernel_stats(ctx, kind, launches, total_ms) bind(C, name="lpm_kernel_stats")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
      integer(c_int64_t), intent(out) :: launches
      real(c_double), intent(out) :: total_ms
    end function

    integer(c_int) function lpm_set_tuning(ctx, kind, precision, targets_per_lane, tile_sources, splits) &
        bind(C, name="lpm_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind, precision, targets_per_lane, tile_sources, splits
    end function

    integer(c_int) function lpm_get_tuning(ctx, kind, precision, targets_per_lane, tile_sources, splits) &
        bind(C, name="lpm_get_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind, precision
      integer(c_int), intent(out) :: targets_per_lane, tile_sources, splits
    end function

    integer(c_int) function lpm_reset_kernel_stats(ctx) bind(C, name="lpm_reset_kernel_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function lpm_synchronize(ctx) bind(C, name="lpm_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! batched solver: launch shape, work items and pairs per evaluation
    integer(c_int) function lpm_bve_batch_set_tuning(batch, targets_per_lane, tile_sources, chunk_sources) &
        bind(C, name="lpm_bve_batch_set_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: targets_per_lane, tile_sources, chunk_sources
    end function

    integer(c_int) function lpm_bve_batch_get_tuning(batch, targets_per_lane, tile_sources, chunk_sources, &
        n_items) bind(C, name="lpm_bve_batch_get_tuning")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: targets_per_lane, tile_sources, chunk_sources, n_items
    end function

    integer(c_int) function lpm_bve_batch_pairs(batch, pairs_per_eval) bind(C, name="lpm_bve_batch_pairs")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: batch
      integer(c_int64_t), intent(out) :: pairs_per_eval
    end function

    ! learned far field: particle cells (length n), near/band cell counts
    ! (length ncell each) and stored training samples
    integer(c_int) function lpm_bve_ff_particle_cells(ff, cell) bind(C, name="lpm_bve_ff_particle_cells")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ff
      integer(c_int32_t) :: cell(*)
    end function

    integer(c_int) function lpm_bve_ff_cells(ff, near_cells, band_cells) bind(C, name="lpm_bve_ff_cells")
      import :: c_int, c_ptr
      type(c_ptr), value :: ff
      integer(c_int) :: near_cells(*), band_cells(*)
    end function

    ! coarse level of the expansion part (0 turns it off) and the expansion
    ! records per target cell of the last evaluation (length ncell)
    integer(c_int) function lpm_bve_ff_set_parents(ff, parent_radius) bind(C, name="lpm_bve_ff_set_parents")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ff
      real(c_double), value :: parent_radius
    end function

    integer(c_int) function lpm_bve_ff_far_counts(ff, far_records) bind(C, name="lpm_bve_ff_far_counts")
      import :: c_int, c_ptr
      type(c_ptr), value :: ff
      integer(c_int) :: far_records(*)
    end function

    integer(c_int) function lpm_bve_ff_samples(ff, offset, count, inputs, targets, ids) &
        bind(C, name="lpm_bve_ff_samples")
      import :: c_int, c_ptr, c_int64_t, c_float, c_int32_t
      type(c_ptr), value :: ff
      integer(c_int64_t), value :: offset, count
      real(c_float) :: inputs(*), targets(*)
      integer(c_int32_t) :: ids(*)
    end function

    ! stand-alone MLP (FP32 parameters; FP16 MFMA or FP32 forward)
    integer(c_int) function lpm_mlp_new(ctx, mlp, nin, width, nhidden, nout, activation, seed) &
        bind(C, name="lpm_mlp_new")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: mlp
      integer(c_int), value :: nin, width, nhidden, nout, activation
      integer(c_int64_t), value :: seed
    end function

    integer(c_int) function lpm_mlp_delete(mlp) bind(C, name="lpm_mlp_delete")
      import :: c_int, c_ptr
      type(c_ptr), value :: mlp
    end function

    integer(c_int) function lpm_mlp_param_count(mlp, count) bind(C, name="lpm_mlp_param_count")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: mlp
      integer(c_int64_t), intent(out) :: count
    end function

    integer(c_int) function lpm_mlp_set_params(mlp, params, count) bind(C, name="lpm_mlp_set_params")
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: mlp
      real(c_float) :: params(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function lpm_mlp_get_params(mlp, params, count) bind(C, name="lpm_mlp_get_params")
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: mlp
      real(c_float) :: params(*)
      integer(c_int64_t), value :: count
    end function

    integer(c_int) function lpm_mlp_forward(mlp, n, inputs, outputs, use_mfma) bind(C, name="lpm_mlp_forward")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: mlp
      integer(c_int), value :: n, use_mfma
      real(c_float) :: inputs(*), outputs(*)
    end function

    integer(c_int) function lpm_mlp_loss_grad(mlp, n, inputs, targets, grad, loss) &
        bind(C, name="lpm_mlp_loss_grad")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: mlp
      integer(c_int), value :: n
      real(c_float) :: inputs(*), targets(*), grad(*)
      real(c_double), intent(out) :: loss
    end function

    integer(c_int) function lpm_mlp_train(mlp, n, inputs, targets, epochs, batch, lr, seed, final_loss) &
        bind(C, name="lpm_mlp_train")
      import :: c_int, c_ptr, c_float, c_double, c_int64_t
      type(c_ptr), value :: mlp
      integer(c_int), value :: n, epochs, batch
      real(c_float) :: inputs(*), targets(*)
      real(c_double), value :: lr
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: final_loss
    end function
  end interface

contains

  !> Last error message of a context as a Fortran string.
  function lpm_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: s
      end function
    end interface
    p = lpm_last_error_c(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(chars(1:n), msg)
  end function

  !> logical(4) activeMask -> int32 flags (any nonzero is active).
  subroutine lpm_mask_to_int(mask, flags)
    logical(4), intent(in) :: mask(:)
    integer(c_int32_t), intent(out) :: flags(:)
    where (mask)
      flags = 1
    elsewhere
      flags = 0
    end where
  end subroutine
end module lpm_direct_f
