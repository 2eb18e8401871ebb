!> @file SphereBVESolver_lpm.f90
!> Drop-in replacement for lpm-v2's SphereBVESolver.f90 whose Timestep runs
!> on the GPU through the C ABI.
!>
!> The module keeps the reference's name (SphereBVESolverModule), type name
!> (BVESolver) and public interface (New, Delete, Timestep with the argument
!> lists of SphereBVESolver.f90:80-90, 112, 196, 219), so a build that lists
!> this file instead of SphereBVESolver.f90 compiles the reference's drivers
!> unchanged: `call New(solver, sphere)`, `call Timestep(solver, sphere, dt)`,
!> `call Delete(solver)` (BVESingleGaussianVortex.f90:201, 240-243, 246).
!>
!> New uploads the particles, relative and absolute vorticity and velocity of
!> the BVEMesh into a device-resident solver (lpm_bve_solver_new).  Timestep
!> runs one RK4 step on the device (lpm_bve_solver_timestep) and then, as the
!> reference's timestepPrivate does (SphereBVESolver.f90:333-352), writes the
!> new positions, relative vorticity and velocity back into the mesh and sets
!> both stream functions (lpm_bve_solver_stream).  The absolute vorticity is
!> left as set up, as in the reference.  Between steps the solver continues
!> from its own state; a remesh goes through Delete and New, as the drivers do.
!>
!> Precision and stage options are module variables: lpmStepPrecision
!> (LPM_PREC_FP64, the reference's arithmetic, or LPM_PREC_FP32_KAHAN) and
!> lpmStepFlags (LPM_SOLVER_* flags; 0 keeps the reference's unprojected
!> stages).  Errors follow the reference convention: print and return.
module SphereBVESolverModule
use iso_c_binding
use NumberKindsModule
use SphereBVEModule
use lpm_direct_f
implicit none
private
public :: BVESolver, New, Delete, Timestep, lpmStepPrecision, lpmStepFlags

integer, save :: lpmStepPrecision = LPM_PREC_FP64
integer, save :: lpmStepFlags = 0
type(c_ptr), save :: lpmContext = c_null_ptr

type BVESolver
	type(c_ptr) :: handle = c_null_ptr
	integer(c_int) :: n = 0
end type

interface New
	module procedure newPrivate
end interface

interface Delete
	module procedure deletePrivate
end interface

interface Timestep
	module procedure timestepPrivate
end interface

contains

subroutine newPrivate(self, sphereBVE)
	type(BVESolver), intent(out) :: self
	type(BVEMesh), intent(in), target :: sphereBVE
	integer(c_int32_t), allocatable :: flags(:)
	integer(c_int) :: rc, n

	if ( .NOT. c_associated(lpmContext) ) then
		rc = lpm_create(lpmContext, 0_c_int)
		if ( rc /= LPM_OK ) then
			print *, "BVESolver New ERROR : ", lpm_last_error(c_null_ptr)
			return
		endif
	endif
	n = int(sphereBVE%mesh%particles%N, c_int)
	allocate(flags(n))
	call lpm_mask_to_int(sphereBVE%mesh%particles%isActive(1:n), flags)
	rc = lpm_bve_solver_new(lpmContext, self%handle, n, sphereBVE%mesh%particles%x, sphereBVE%mesh%particles%y, &
		sphereBVE%mesh%particles%z, sphereBVE%relVort%scalar, c_loc(sphereBVE%absVort%scalar), &
		sphereBVE%mesh%particles%area, flags, c_loc(sphereBVE%velocity%xComp), c_loc(sphereBVE%velocity%yComp), &
		c_loc(sphereBVE%velocity%zComp), sphereBVE%radius, sphereBVE%rotationRate, int(lpmStepPrecision, c_int), &
		int(lpmStepFlags, c_int))
	if ( rc /= LPM_OK ) then
		print *, "BVESolver New ERROR : ", lpm_last_error(lpmContext)
		return
	endif
	self%n = n
end subroutine

subroutine deletePrivate(self)
	type(BVESolver), intent(inout) :: self
	integer(c_int) :: rc
	if ( c_associated(self%handle) ) rc = lpm_bve_solver_delete(self%handle)
	self%handle = c_null_ptr
	self%n = 0
end subroutine

subroutine timestepPrivate(self, sphereBVE, dt)
	type(BVESolver), intent(inout) :: self
	type(BVEMesh), intent(inout) :: sphereBVE
	real(kreal), intent(in) :: dt
	integer(c_int) :: rc, n

	n = self%n
	if ( .NOT. c_associated(self%handle) .OR. n /= sphereBVE%mesh%particles%N ) then
		print *, "BVESolver Timestep ERROR : solver not created for this mesh"
		return
	endif
	rc = lpm_bve_solver_timestep(self%handle, dt)
	if ( rc /= LPM_OK ) then
		print *, "BVESolver Timestep ERROR : ", lpm_last_error(lpmContext)
		return
	endif
	! the new state into the mesh (SphereBVESolver.f90:333-350)
	rc = lpm_bve_solver_get_state(self%handle, sphereBVE%mesh%particles%x, sphereBVE%mesh%particles%y, &
		sphereBVE%mesh%particles%z, sphereBVE%relVort%scalar, sphereBVE%velocity%xComp, &
		sphereBVE%velocity%yComp, sphereBVE%velocity%zComp)
	if ( rc /= LPM_OK ) then
		print *, "BVESolver Timestep ERROR : ", lpm_last_error(lpmContext)
		return
	endif
	sphereBVE%velocity%N = n
	! and the stream functions of the new state (SetStreamFunctionsOnMesh, :352)
	rc = lpm_bve_solver_stream(self%handle, sphereBVE%relStream%scalar, sphereBVE%absStream%scalar)
	if ( rc /= LPM_OK ) then
		print *, "BVESolver Timestep ERROR : ", lpm_last_error(lpmContext)
		return
	endif
	sphereBVE%relStream%N = n
	sphereBVE%absStream%N = n
end subroutine

end module
