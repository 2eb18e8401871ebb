! ref_cpu_bench.f90 -- times lpm-v2's own Biot-Savart loop on the CPU.
!
! Test and measurement infrastructure only (the CPU baseline beside the GPU
! benchmark).  Built by oracle/ref.mk against the reference's modules
! (compiled from its sources where they lie) into oracle/_ref/, run by
! bench.py under MPICH's mpiexec, one MPI rank per core:
!     mpiexec -n <ranks> ref_cpu_bench <dir> <n> <seconds>
! It fills a BVEMesh from raw FP64 arrays <dir>/in_<name>.f64 (x y z area
! active relvort; the reference's Particles and Field constructors, no seed
! file), then calls the reference's public SetVelocityOnMesh
! (SphereBVE.f90:489-531, the loop BVESphereVelocity repeats in
! SphereBVESolver.f90:395-420) with each rank's MPISetup block cut to a
! bounded sample of targets: first 8 targets per rank to calibrate, then as
! many as fill about <seconds>.  Every rank still sums over all n particles'
! active sources, as the reference does.  Rank 0 prints one line:
!     ranks targets sources seconds pairs_per_second
! and writes the sample's velocity to <dir>/out_{u,v,w}.f64 (targets 1..K).
program ref_cpu_bench
use NumberKindsModule
use ParticlesModule
use FieldModule
use MPISetupModule
use SphereBVEModule
implicit none
include 'mpif.h'

real(kreal), parameter :: PI_ = 3.14159265358979323846_kreal
character(len=256) :: dir, arg
integer :: ierr, n, kb, nsrc, r
real(kreal) :: budget, t, buf1(1)
real(kreal), allocatable :: buf(:)
type(BVEMesh) :: sphere

call MPI_INIT(ierr)
call MPI_COMM_SIZE(MPI_COMM_WORLD, numProcs, ierr)
call MPI_COMM_RANK(MPI_COMM_WORLD, procRank, ierr)
call get_command_argument(1, dir)
call get_command_argument(2, arg)
read(arg, *) n
call get_command_argument(3, arg)
read(arg, *) budget

allocate(buf(n))
call New(sphere%mesh%particles, n, SPHERE_GEOM)
sphere%mesh%particles%N = n
call get("x", sphere%mesh%particles%x(1:n))
call get("y", sphere%mesh%particles%y(1:n))
call get("z", sphere%mesh%particles%z(1:n))
call get("area", sphere%mesh%particles%area(1:n))
call get("active", buf)
sphere%mesh%particles%isActive(1:n) = buf > 0.5_kreal
nsrc = count(sphere%mesh%particles%isActive(1:n))
call New(sphere%relVort, 1, n, "relVort", "1/time")
call New(sphere%velocity, 3, n, "velocity", "length/time")
sphere%relVort%N = n
sphere%velocity%N = n
call get("relvort", sphere%relVort%scalar(1:n))
sphere%radius = 1.0_kreal
sphere%rotationRate = 2.0_kreal * PI_
call New(sphere%mpiParticles, n, numProcs)

! calibrate on 8 targets per rank, then size the sample for the budget
kb = 8
t = timed(kb)
if ( procRank == 0 ) buf1(1) = real(max(8, min(int(kb * budget / max(t, 1.0d-6)), n / numProcs)), kreal)
call MPI_BCAST(buf1, 1, MPI_DOUBLE_PRECISION, 0, MPI_COMM_WORLD, ierr)
kb = int(buf1(1))
t = timed(kb)
if ( procRank == 0 ) then
	print '(I0,1X,I0,1X,I0,1X,ES14.6,1X,ES14.6)', numProcs, kb * numProcs, nsrc, t, &
		real(kb, kreal) * numProcs * nsrc / t
	call put("u", sphere%velocity%xComp(1:kb * numProcs))
	call put("v", sphere%velocity%yComp(1:kb * numProcs))
	call put("w", sphere%velocity%zComp(1:kb * numProcs))
endif
call MPI_FINALIZE(ierr)

contains

! rank r owns targets r kb + 1 .. (r + 1) kb; returns the slowest rank's time
real(kreal) function timed(kb)
	integer, intent(in) :: kb
	real(kreal) :: t0, dt(1), dtmax(1)
	do r = 0, numProcs - 1
		sphere%mpiParticles%indexStart(r) = r * kb + 1
		sphere%mpiParticles%indexEnd(r) = (r + 1) * kb
		sphere%mpiParticles%messageLength(r) = kb
	enddo
	call MPI_BARRIER(MPI_COMM_WORLD, ierr)
	t0 = MPI_WTIME()
	call SetVelocityOnMesh(sphere)
	dt(1) = MPI_WTIME() - t0
	call MPI_ALLREDUCE(dt, dtmax, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)
	timed = dtmax(1)
end function

subroutine get(name, a)
	character(len=*), intent(in) :: name
	real(kreal), intent(out) :: a(:)
	integer :: u
	open(newunit=u, file=trim(dir)//"/in_"//trim(name)//".f64", access="stream", form="unformatted", &
		status="old", action="read")
	read(u) a
	close(u)
end subroutine

subroutine put(name, a)
	character(len=*), intent(in) :: name
	real(kreal), intent(in) :: a(:)
	integer :: u
	open(newunit=u, file=trim(dir)//"/out_"//trim(name)//".f64", access="stream", form="unformatted", &
		status="replace", action="write")
	write(u) a
	close(u)
end subroutine

end program
