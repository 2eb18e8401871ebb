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
