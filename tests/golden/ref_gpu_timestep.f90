! ref_gpu_timestep.f90 -- the reference's own Timestep call with the
! GPU-backed SphereBVESolverModule (lpm-v2_amd/fortran/SphereBVESolver_lpm.f90)
! in place of SphereBVESolver.f90.
!
! Test infrastructure only.  Built by oracle/ref.mk against the reference's
! other modules (compiled where they lie) and the HIP library; run by
! tests/test_gpu_fortran_dropin.py on the GPU:
!     ref_gpu_timestep <dir> <n>
! It reads the reference's set-up state of a sphere mesh (raw FP64 arrays
! <dir>/in_<name>.f64: x y z area active relvort absvort u v w, written by the
! test from a golden fixture), fills a BVEMesh with it (particles and fields
! through the reference's own Particles and Field constructors, so no seed
! file is needed), and then makes the driver's calls
! (BVESingleGaussianVortex.f90:201, 240-246):
!     call New(solver, sphere); call Timestep(solver, sphere, 0.01); call Delete(solver)
! once in FP64 and once in FP32.  The mesh fields the step wrote back go to
! <dir>/<prec>_<name>.f64.
program ref_gpu_timestep
use NumberKindsModule
use ParticlesModule
use FieldModule
use SphereBVEModule
use SphereBVESolverModule
use lpm_direct_f, only : LPM_PREC_FP64, LPM_PREC_FP32_KAHAN
implicit none
include 'mpif.h'

real(kreal), parameter :: PI_ = 3.14159265358979323846_kreal
character(len=256) :: outdir, arg
integer :: ierr, n
type(BVEMesh) :: sphere
type(BVESolver) :: solver
real(kreal), allocatable :: buf(:), x0(:), y0(:), z0(:), q0(:), u0(:), v0(:), w0(:)

call MPI_INIT(ierr)
call MPI_COMM_SIZE(MPI_COMM_WORLD, numProcs, ierr)
call MPI_COMM_RANK(MPI_COMM_WORLD, procRank, ierr)
call get_command_argument(1, outdir)
call get_command_argument(2, arg)
read(arg, *) n

allocate(buf(n))
call New(sphere%mesh%particles, n, SPHERE_GEOM)
sphere%mesh%particles%N = n
call get("x", sphere%mesh%particles%x(1:n))
call get("y", sphere%mesh%particles%y(1:n))
call get("z", sphere%mesh%particles%z(1:n))
call get("area", sphere%mesh%particles%area(1:n))
call get("active", buf)
sphere%mesh%particles%isActive(1:n) = buf > 0.5_kreal
call New(sphere%relVort, 1, n, "relVort", "1/time")
call New(sphere%absVort, 1, n, "absVort", "1/time")
call New(sphere%relStream, 1, n, "relStream", "area/time")
call New(sphere%absStream, 1, n, "absStream", "area/time")
call New(sphere%velocity, 3, n, "velocity", "length/time")
sphere%relVort%N = n
sphere%absVort%N = n
sphere%velocity%N = n
call get("relvort", sphere%relVort%scalar(1:n))
call get("absvort", sphere%absVort%scalar(1:n))
call get("u", sphere%velocity%xComp(1:n))
call get("v", sphere%velocity%yComp(1:n))
call get("w", sphere%velocity%zComp(1:n))
sphere%radius = 1.0_kreal
sphere%rotationRate = 2.0_kreal * PI_
x0 = sphere%mesh%particles%x(1:n)
y0 = sphere%mesh%particles%y(1:n)
z0 = sphere%mesh%particles%z(1:n)
q0 = sphere%relVort%scalar(1:n)
u0 = sphere%velocity%xComp(1:n)
v0 = sphere%velocity%yComp(1:n)
w0 = sphere%velocity%zComp(1:n)

call stepCase("fp64", LPM_PREC_FP64)
! back to the set-up state for the FP32 step
sphere%mesh%particles%x(1:n) = x0
sphere%mesh%particles%y(1:n) = y0
sphere%mesh%particles%z(1:n) = z0
sphere%relVort%scalar(1:n) = q0
sphere%velocity%xComp(1:n) = u0
sphere%velocity%yComp(1:n) = v0
sphere%velocity%zComp(1:n) = w0
call stepCase("fp32", LPM_PREC_FP32_KAHAN)

call MPI_FINALIZE(ierr)

contains

subroutine get(name, a)
	character(len=*), intent(in) :: name
	real(kreal), intent(out) :: a(:)
	integer :: u
	open(newunit=u, file=trim(outdir)//"/in_"//trim(name)//".f64", access="stream", form="unformatted", &
		status="old", action="read")
	read(u) a
	close(u)
end subroutine

subroutine put(name, a)
	character(len=*), intent(in) :: name
	real(kreal), intent(in) :: a(:)
	integer :: u
	open(newunit=u, file=trim(outdir)//"/"//trim(name)//".f64", access="stream", form="unformatted", &
		status="replace", action="write")
	write(u) a
	close(u)
end subroutine

subroutine stepCase(tag, prec)
	character(len=*), intent(in) :: tag
	integer, intent(in) :: prec

	lpmStepPrecision = prec
	call New(solver, sphere)
	call Timestep(solver, sphere, 0.01_kreal)
	call Delete(solver)

	call put(tag//"_x", sphere%mesh%particles%x(1:n))
	call put(tag//"_y", sphere%mesh%particles%y(1:n))
	call put(tag//"_z", sphere%mesh%particles%z(1:n))
	call put(tag//"_relvort", sphere%relVort%scalar(1:n))
	call put(tag//"_u", sphere%velocity%xComp(1:n))
	call put(tag//"_v", sphere%velocity%yComp(1:n))
	call put(tag//"_w", sphere%velocity%zComp(1:n))
	call put(tag//"_psi_rel", sphere%relStream%scalar(1:n))
	call put(tag//"_psi_abs", sphere%absStream%scalar(1:n))
end subroutine

end program
