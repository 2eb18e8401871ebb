!> @file lpm_fortran_smoke.f90
!> Fortran ISO_C_BINDING smoke program: reads particles from a raw binary file
!> (int32 n, then x, y, z, relVort, area as float64[n], then int32 active[n]),
!> calls lpm_BVESphereVelocity over two MPISetup-style blocks plus one RK4
!> step of GPUBVESolver, and writes u, v, w then the stepped x, y, z, relVort
!> (float64[n] each) to the output file.
program lpm_fortran_smoke
  use iso_c_binding
  use lpm_direct_f
  use lpm_bve_solver_f
  implicit none
  character(len=512) :: inFile, outFile
  integer(c_int32_t) :: n
  integer(c_int32_t), allocatable :: act(:)
  logical(4), allocatable :: mask(:)
  real(c_double), allocatable, target :: x(:), y(:), z(:), zeta(:), area(:), u(:), v(:), w(:)
  real(c_double), allocatable :: xs(:), ys(:), zs(:), zetas(:), us(:), vs(:), ws(:)
  type(GPUBVESolver) :: solver
  integer :: half, unit_in, unit_out

  call get_command_argument(1, inFile)
  call get_command_argument(2, outFile)
  open(newunit=unit_in, file=trim(inFile), access='stream', form='unformatted', status='old')
  read(unit_in) n
  allocate(x(n), y(n), z(n), zeta(n), area(n), act(n), mask(n), u(n), v(n), w(n))
  allocate(xs(n), ys(n), zs(n), zetas(n), us(n), vs(n), ws(n))
  read(unit_in) x, y, z, zeta, area, act
  close(unit_in)
  mask = act /= 0
  u = 0.0_c_double
  v = 0.0_c_double
  w = 0.0_c_double

  ! two "ranks": blocks [1, half] and [half+1, n] (MPISetup LoadBalance for 2 procs)
  half = n / 2
  call lpm_BVESphereVelocity(u, v, w, x, y, z, zeta, area, 1.0_c_double, 0.0_c_double, mask, 1, half)
  call lpm_BVESphereVelocity(u, v, w, x, y, z, zeta, area, 1.0_c_double, 0.0_c_double, mask, half + 1, int(n))

  call New(solver, x, y, z, zeta, area, mask, u, v, w, 1.0_c_double, 2.0_c_double * acos(-1.0_c_double), &
           LPM_PREC_FP64, 0)
  call Timestep(solver, 0.01_c_double)
  call GetState(solver, xs, ys, zs, zetas, us, vs, ws)
  call Delete(solver)

  open(newunit=unit_out, file=trim(outFile), access='stream', form='unformatted', status='replace')
  write(unit_out) u, v, w, xs, ys, zs, zetas
  close(unit_out)
  print *, "lpm_fortran_smoke: n =", n, " done"
end program
