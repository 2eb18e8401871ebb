! ref_golden.f90 -- writes golden vectors from lpm-v2's own Fortran modules.
!
! Test infrastructure only.  Built by oracle/ref.mk against the reference's
! sources (compiled where they lie) and run by tests/golden/make_golden.py in
! a scratch directory under oracle/_ref/ that links the reference's seed
! files.  For each case it writes raw little-endian FP64 arrays,
! <dir>/<case>_<name>.f64, through the reference's public entry points:
!   New / SetInitialVorticityOnMesh / SetVelocityOnMesh / SetStreamFunctionsOnMesh
!     (SphereBVE.f90:145, 334, 489, 445), BVESolver New / Timestep
!     (SphereBVESolver.f90:112, 219);
!   PlanarIncompressible New / SetInitialVorticityOnMesh / SetVelocityOnMesh /
!     SetStreamFunctionOnMesh (PlanarIncompressible.f90:139, 295, 426, 470),
!     PlaneSolver New / Timestep (PlaneIncompressibleSolver.f90:103, 171);
!   PSE New / PSEPlaneLaplacianAtParticles / PSEPlaneGradientAtParticles
!     (PSEDirectSum.f90:93, 467, 180).
! Run on one MPI rank.
module golden_fns
use NumberKindsModule
implicit none
real(kreal), parameter :: PI_ = 3.14159265358979323846_kreal
contains
	! cfg0 vortex: 4 pi exp(-16 |x - x0|^2), x0 at lat 0.3, lon 0
	real(8) function gaussVort(x, y, z)
		real(8), intent(in) :: x, y, z
		real(8) :: c(3)
		c = [cos(0.3_kreal), 0.0_kreal, sin(0.3_kreal)]
		gaussVort = 4.0_kreal * PI_ * exp(-16.0_kreal * ((x - c(1))**2 + (y - c(2))**2 + (z - c(3))**2))
	end function
	! Rossby-Haurwitz wave 4 with omega = K = 1: 2 omega sin(lat) - 30 K sin(lat) cos^4(lat) cos(4 lon)
	real(8) function rh4Vort(x, y, z)
		real(8), intent(in) :: x, y, z
		real(8) :: lat, lon
		lat = atan2(z, sqrt(x * x + y * y))
		lon = atan2(y, x)
		rh4Vort = 2.0_kreal * sin(lat) - 30.0_kreal * sin(lat) * cos(lat)**4 * cos(4.0_kreal * lon)
	end function
	! planar: two unequal Gaussians
	real(8) function planeVort(x, y)
		real(8), intent(in) :: x, y
		planeVort = exp(-8.0_kreal * (x * x + y * y)) + 0.5_kreal * exp(-20.0_kreal * ((x - 0.3_kreal)**2 + y * y))
	end function
end module

program ref_golden
use NumberKindsModule
use ParticlesModule
use FieldModule
use PolyMesh2dModule
use MPISetupModule
use SphereBVEModule, setSphereVort => SetInitialVorticityOnMesh
use SphereBVESolverModule
use PlanarIncompressibleModule, setPlaneVort => SetInitialVorticityOnMesh
use PlanarIncompressibleSolverModule
use PSEDirectSumModule
use golden_fns
implicit none
include 'mpif.h'

character(len=256) :: outdir
integer :: ierr

call MPI_INIT(ierr)
call MPI_COMM_SIZE(MPI_COMM_WORLD, numProcs, ierr)
call MPI_COMM_RANK(MPI_COMM_WORLD, procRank, ierr)
call get_command_argument(1, outdir)

call sphereCase("icos2", ICOS_TRI_SPHERE_SEED, 2, 1, .TRUE.)
call sphereCase("icos3", ICOS_TRI_SPHERE_SEED, 3, 1, .FALSE.)
call sphereCase("cubed3", CUBED_SPHERE_SEED, 3, 2, .FALSE.)
call planeCase("quad3", 3)

call MPI_FINALIZE(ierr)

contains

subroutine put(name, a)
	character(len=*), intent(in) :: name
	real(kreal), intent(in) :: a(:)
	integer :: u
	open(newunit=u, file=trim(outdir)//"/"//trim(name)//".f64", access="stream", form="unformatted", &
		status="replace", action="write")
	write(u) a
	close(u)
end subroutine

subroutine sphereCase(tag, seed, nest, vortKind, doStep)
	character(len=*), intent(in) :: tag
	integer(kint), intent(in) :: seed, nest, vortKind
	logical, intent(in) :: doStep
	type(BVEMesh) :: sphere
	type(BVESolver) :: solver
	integer(kint) :: n
	real(kreal) :: act(1)

	call New(sphere, seed, nest, nest, 0, 1.0_kreal, 2.0_kreal * PI_)
	if ( vortKind == 1 ) then
		call setSphereVort(sphere, gaussVort)
	else
		call setSphereVort(sphere, rh4Vort)
	endif
	call SetVelocityOnMesh(sphere)
	call SetStreamFunctionsOnMesh(sphere)
	n = sphere%mesh%particles%N
	call put(tag//"_x", sphere%mesh%particles%x(1:n))
	call put(tag//"_y", sphere%mesh%particles%y(1:n))
	call put(tag//"_z", sphere%mesh%particles%z(1:n))
	call put(tag//"_area", sphere%mesh%particles%area(1:n))
	call put(tag//"_active", merge(1.0_kreal, 0.0_kreal, sphere%mesh%particles%isActive(1:n)))
	call put(tag//"_relvort", sphere%relVort%scalar(1:n))
	call put(tag//"_absvort", sphere%absVort%scalar(1:n))
	call put(tag//"_u", sphere%velocity%xComp(1:n))
	call put(tag//"_v", sphere%velocity%yComp(1:n))
	call put(tag//"_w", sphere%velocity%zComp(1:n))
	call put(tag//"_psi_rel", sphere%relStream%scalar(1:n))
	call put(tag//"_psi_abs", sphere%absStream%scalar(1:n))
	if ( doStep ) then
		! one RK4 step, dt = 0.01 (it ends with the stream functions, SphereBVESolver.f90:352)
		call New(solver, sphere)
		call Timestep(solver, sphere, 0.01_kreal)
		call put(tag//"_step1_x", sphere%mesh%particles%x(1:n))
		call put(tag//"_step1_y", sphere%mesh%particles%y(1:n))
		call put(tag//"_step1_z", sphere%mesh%particles%z(1:n))
		call put(tag//"_step1_relvort", sphere%relVort%scalar(1:n))
		call put(tag//"_step1_u", sphere%velocity%xComp(1:n))
		call put(tag//"_step1_v", sphere%velocity%yComp(1:n))
		call put(tag//"_step1_w", sphere%velocity%zComp(1:n))
		call put(tag//"_step1_psi_rel", sphere%relStream%scalar(1:n))
		call put(tag//"_step1_psi_abs", sphere%absStream%scalar(1:n))
		call Delete(solver)
	endif
	act(1) = real(n, kreal)
	call put(tag//"_n", act)
	call Delete(sphere)
end subroutine

subroutine planeCase(tag, nest)
	character(len=*), intent(in) :: tag
	integer(kint), intent(in) :: nest
	type(PlaneMeshIncompressible) :: plane
	type(PlaneSolver) :: solver
	type(PSE) :: pseOp
	type(Field) :: f, lap, grad
	integer(kint) :: n, i
	real(kreal) :: eps(1)

	call New(plane, nest, nest, QUAD_RECT_SEED, 0, 1.0_kreal)
	call setPlaneVort(plane, planeVort)
	call SetVelocityOnMesh(plane)
	call SetStreamFunctionOnMesh(plane)
	n = plane%mesh%particles%N
	call put(tag//"_x", plane%mesh%particles%x(1:n))
	call put(tag//"_y", plane%mesh%particles%y(1:n))
	call put(tag//"_area", plane%mesh%particles%area(1:n))
	call put(tag//"_active", merge(1.0_kreal, 0.0_kreal, plane%mesh%particles%isActive(1:n)))
	call put(tag//"_vort", plane%vorticity%scalar(1:n))
	call put(tag//"_u", plane%velocity%xComp(1:n))
	call put(tag//"_v", plane%velocity%yComp(1:n))
	call put(tag//"_psi", plane%streamFn%scalar(1:n))

	! PSE Laplacian and gradient of exp(-8 r^2), eps = MaxEdgeLength^0.75
	call New(pseOp, plane%mesh)
	call New(f, 1, n, "f", "1")
	call New(lap, 1, n, "lap", "1")
	call New(grad, 2, n, "grad", "1")
	do i = 1, n
		call InsertScalarToField(f, exp(-8.0_kreal * (plane%mesh%particles%x(i)**2 + plane%mesh%particles%y(i)**2)))
	enddo
	call PSEPlaneLaplacianAtParticles(pseOp, plane%mesh, f, lap, plane%mpiParticles)
	call PSEPlaneGradientAtParticles(pseOp, plane%mesh, f, grad, plane%mpiParticles)
	eps(1) = pseOp%eps
	call put(tag//"_pse_eps", eps)
	call put(tag//"_pse_f", f%scalar(1:n))
	call put(tag//"_pse_lap", lap%scalar(1:n))
	call put(tag//"_pse_gx", grad%xComp(1:n))
	call put(tag//"_pse_gy", grad%yComp(1:n))

	! one RK4 step, dt = 0.005
	call New(solver, plane)
	call Timestep(solver, plane, 0.005_kreal)
	call put(tag//"_step1_x", plane%mesh%particles%x(1:n))
	call put(tag//"_step1_y", plane%mesh%particles%y(1:n))
	call put(tag//"_step1_u", plane%velocity%xComp(1:n))
	call put(tag//"_step1_v", plane%velocity%yComp(1:n))
	call put(tag//"_step1_psi", plane%streamFn%scalar(1:n))
	eps(1) = real(n, kreal)
	call put(tag//"_n", eps)
	call Delete(solver)
	call Delete(f)
	call Delete(lap)
	call Delete(grad)
	call Delete(pseOp)
	call Delete(plane)
end subroutine

end program
