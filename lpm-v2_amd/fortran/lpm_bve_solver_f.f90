!> @file lpm_bve_solver_f.f90
!> Drop-in Fortran side for lpm-v2's sphere BVE solver.
!>
!> * lpm_BVESphereVelocity has the argument list of the reference's private
!>   BVESphereVelocity (SphereBVESolver.f90:377-430) with the MPISetup block
!>   passed as (indexStart, indexEnd) of this rank, 1-based inclusive; its body
!>   is one C-ABI call.  The reference's MPI_BCAST loop stays with the caller.
!> * type(GPUBVESolver) with New / Timestep / Delete mirrors BVESolver
!>   (SphereBVESolver.f90:38-90): New uploads the state once, Timestep runs the
!>   whole RK4 step in HBM, GetState copies it back for output frames.
!> Errors follow the reference convention: print the message and return.
module lpm_bve_solver_f
  use iso_c_binding
  use lpm_direct_f
  implicit none
  private
  public :: lpm_BVESphereVelocity, GPUBVESolver, New, Timestep, GetState, Delete, lpm_context

  type(c_ptr), save :: lpm_context = c_null_ptr

  type GPUBVESolver
    type(c_ptr) :: handle = c_null_ptr
    integer(c_int) :: n = 0
  end type

  interface New
    module procedure newGPUSolver
  end interface
  interface Timestep
    module procedure timestepGPUSolver
  end interface
  interface GetState
    module procedure getStateGPUSolver
  end interface
  interface Delete
    module procedure deleteGPUSolver
  end interface

contains

  subroutine ensure_context(ok)
    logical, intent(out) :: ok
    integer(c_int) :: rc
    ok = .true.
    if (c_associated(lpm_context)) return
    rc = lpm_create(lpm_context, 0_c_int)
    if (rc /= LPM_OK) then
      print *, "lpm_create ERROR : ", lpm_last_error(c_null_ptr)
      ok = .false.
    endif
  end subroutine

  !> Drop-in for BVESphereVelocity (SphereBVESolver.f90:377-430).
  subroutine lpm_BVESphereVelocity(u, v, w, x, y, z, relVortIn, areaIn, sphereRadius, rotationRate, &
                                   activeMask, indexStart, indexEnd, precision)
    real(c_double), dimension(:), intent(inout) :: u, v, w
    real(c_double), dimension(:), intent(in) :: x, y, z, relVortIn, areaIn
    real(c_double), intent(in) :: sphereRadius, rotationRate
    logical(4), dimension(:), intent(in) :: activeMask
    integer, intent(in) :: indexStart, indexEnd
    integer, intent(in), optional :: precision
    integer(c_int32_t), allocatable :: flags(:)
    integer(c_int) :: rc, prec
    logical :: ok
    call ensure_context(ok)
    if (.not. ok) return
    prec = LPM_PREC_FP64
    if (present(precision)) prec = precision
    allocate(flags(size(activeMask)))
    call lpm_mask_to_int(activeMask, flags)
    rc = lpm_bve_sphere_velocity(lpm_context, u, v, w, x, y, z, relVortIn, areaIn, sphereRadius, &
                                 rotationRate, flags, int(size(x), c_int), int(indexStart - 1, c_int), &
                                 int(indexEnd, c_int), prec)
    if (rc /= LPM_OK) print *, "lpm_BVESphereVelocity ERROR : ", lpm_last_error(lpm_context)
  end subroutine

  subroutine newGPUSolver(self, x, y, z, relVort, area, activeMask, u, v, w, radius, rotationRate, precision, flags)
    type(GPUBVESolver), intent(out) :: self
    real(c_double), dimension(:), intent(in), target :: x, y, z, relVort, area, u, v, w
    logical(4), dimension(:), intent(in) :: activeMask
    real(c_double), intent(in) :: radius, rotationRate
    integer, intent(in) :: precision, flags
    integer(c_int32_t), allocatable :: fl(:)
    integer(c_int) :: rc
    logical :: ok
    call ensure_context(ok)
    if (.not. ok) return
    allocate(fl(size(activeMask)))
    call lpm_mask_to_int(activeMask, fl)
    self%n = int(size(x), c_int)
    rc = lpm_bve_solver_new(lpm_context, self%handle, self%n, x, y, z, relVort, c_null_ptr, area, fl, &
                            c_loc(u), c_loc(v), c_loc(w), radius, rotationRate, int(precision, c_int), &
                            int(flags, c_int))
    if (rc /= LPM_OK) print *, "GPUBVESolver New ERROR : ", lpm_last_error(lpm_context)
  end subroutine

  subroutine timestepGPUSolver(self, dt)
    type(GPUBVESolver), intent(inout) :: self
    real(c_double), intent(in) :: dt
    integer(c_int) :: rc
    rc = lpm_bve_solver_timestep(self%handle, dt)
    if (rc /= LPM_OK) print *, "GPUBVESolver Timestep ERROR : ", lpm_last_error(lpm_context)
  end subroutine

  subroutine getStateGPUSolver(self, x, y, z, relVort, u, v, w)
    type(GPUBVESolver), intent(in) :: self
    real(c_double), dimension(:), intent(inout) :: x, y, z, relVort, u, v, w
    integer(c_int) :: rc
    rc = lpm_bve_solver_get_state(self%handle, x, y, z, relVort, u, v, w)
    if (rc /= LPM_OK) print *, "GPUBVESolver GetState ERROR : ", lpm_last_error(lpm_context)
  end subroutine

  subroutine deleteGPUSolver(self)
    type(GPUBVESolver), intent(inout) :: self
    integer(c_int) :: rc
    if (c_associated(self%handle)) rc = lpm_bve_solver_delete(self%handle)
    self%handle = c_null_ptr
  end subroutine
end module lpm_bve_solver_f
