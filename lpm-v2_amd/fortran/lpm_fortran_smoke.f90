!> @file lpm_fortran_smoke.f90
!> Fortran ISO_C_BINDING smoke program: reads particles from a raw binary file
!> (int32 n, then x, y, z, relVort, area as float64[n], int32 active[n], then
!> int32 ncell and int32 cell[n], each particle's coarse cell),
!> calls lpm_BVESphereVelocity over two MPISetup-style blocks plus one RK4
!> step of GPUBVESolver, and writes u, v, w then the stepped x, y, z, relVort
!> (float64[n] each) to the output file.  It then calls the PSE and SWE entry
!> points through their bind(C) interfaces (FP64): the sphere PSE gradient
!> (eps 0.3), the planar SWE right-hand side on (x, y) (eps 0.3), the sphere
!> interpolation of relVort at the first 10 particles, and a solver reloaded
!> with the same particles and stepped once; it appends gx, gy, gz, u, v, ddot,
!> lapSurf (n each), the 10 interpolated values and the reloaded solver's
!> stepped x, y, z, relVort (n each).  Last come one FP64 step of the planar
!> solver on (x + 0.1 z, y + 0.37 z) (x, y, u, v, then KE, enstrophy and
!> circulation), one FP64 step of a batch holding the sphere problem
!> (x, y, z, relVort), the learned far-field object's expansion-mode velocity
!> (near 0.3, band 1.5 so that the 20 coarse cells of a level-2 mesh have band
!> cells to sample; u, v, w), then its sample count, training loss, mode,
!> last judged error and judge count after 20,000 samples, 2 epochs and two
!> learned-mode steps judged at every step (tol 1), and the helpers' PSE eps
!> (0.3^0.75) and 3-rank MPISetup blocks (start, end).  Last, the batch's pair
!> count per evaluation and the number of particles whose cell, as the
!> far-field object binned them at its last evaluation, differs from the input
!> cell.
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
  real(c_double), allocatable :: gx(:), gy(:), gz(:), su(:), sv(:), sdd(:), slap(:), dv(:), hh(:), bot(:)
  real(c_double) :: pint(10)
  real(c_double), allocatable :: px(:), py(:), pu(:), pv(:)
  integer(c_int32_t), allocatable :: cell(:), cell2(:)
  integer(c_int64_t) :: npairs
  integer(c_int32_t) :: ncell
  integer(c_int64_t) :: offs(2)
  integer(c_int) :: istart(3), iend(3), mlen(3)
  real(c_double) :: diag(3), eps, loss, last_error
  integer(c_int64_t) :: total
  integer(c_int) :: mode_out, judged
  type(c_ptr) :: psolver, batch, ff
  type(GPUBVESolver) :: solver
  integer :: half, unit_in, unit_out
  integer(c_int) :: rc

  call get_command_argument(1, inFile)
  call get_command_argument(2, outFile)
  open(newunit=unit_in, file=trim(inFile), access='stream', form='unformatted', status='old')
  read(unit_in) n
  allocate(x(n), y(n), z(n), zeta(n), area(n), act(n), mask(n), u(n), v(n), w(n))
  allocate(xs(n), ys(n), zs(n), zetas(n), us(n), vs(n), ws(n))
  read(unit_in) x, y, z, zeta, area, act
  allocate(cell(n))
  read(unit_in) ncell, cell
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

  ! PSE and SWE entry points (the context was created by the calls above)
  allocate(gx(n), gy(n), gz(n), su(n), sv(n), sdd(n), slap(n), dv(n), hh(n), bot(n))
  gx = 0.0_c_double; gy = 0.0_c_double; gz = 0.0_c_double
  rc = lpm_pse_sphere_gradient(lpm_context, gx, gy, gz, x, y, z, zeta, area, act, n, 0_c_int, n, &
                               0.3_c_double, 1.0_c_double, LPM_PREC_FP64)
  if (rc /= LPM_OK) print *, "lpm_pse_sphere_gradient ERROR ", rc
  dv = 0.1_c_double * zeta
  hh = 1.0_c_double + 0.1_c_double * z
  bot = 0.05_c_double * x
  rc = lpm_swe_plane_rhs(lpm_context, su, sv, sdd, slap, x, y, zeta, dv, hh, bot, area, act, n, 0_c_int, n, &
                         0.3_c_double, LPM_PREC_FP64)
  if (rc /= LPM_OK) print *, "lpm_swe_plane_rhs ERROR ", rc
  rc = lpm_pse_sphere_interpolate(lpm_context, pint, x(1:10), y(1:10), z(1:10), 10_c_int, x, y, z, zeta, area, &
                                  act, n, 0.3_c_double, 1.0_c_double, LPM_PREC_FP64)
  if (rc /= LPM_OK) print *, "lpm_pse_sphere_interpolate ERROR ", rc
  write(unit_out) gx, gy, gz, su, sv, sdd, slap, pint

  ! remesh re-upload: a solver reloaded with the same particles steps like a new one
  call New(solver, x, y, z, zeta, area, mask, u, v, w, 1.0_c_double, 2.0_c_double * acos(-1.0_c_double), &
           LPM_PREC_FP64, 0)
  rc = lpm_bve_solver_set_particles(solver%handle, n, x, y, z, zeta, c_null_ptr, area, act, &
                                    c_loc(u), c_loc(v), c_loc(w))
  if (rc /= LPM_OK) print *, "lpm_bve_solver_set_particles ERROR ", rc
  call Timestep(solver, 0.01_c_double)
  call GetState(solver, xs, ys, zs, zetas, us, vs, ws)
  call Delete(solver)
  write(unit_out) xs, ys, zs, zetas

  ! planar solver (PlaneSolver New / Timestep / Delete) on a projection of the particles
  allocate(px(n), py(n), pu(n), pv(n))
  px = x + 0.1_c_double * z
  py = y + 0.37_c_double * z
  rc = lpm_plane_solver_new(lpm_context, psolver, n, px, py, zeta, area, act, c_null_ptr, c_null_ptr, &
                            LPM_PREC_FP64, 0_c_int)
  if (rc /= LPM_OK) print *, "lpm_plane_solver_new ERROR ", rc
  rc = lpm_plane_solver_timestep(psolver, 0.01_c_double)
  if (rc /= LPM_OK) print *, "lpm_plane_solver_timestep ERROR ", rc
  rc = lpm_plane_solver_get_state(psolver, px, py, pu, pv)
  rc = lpm_plane_solver_diagnostics(psolver, diag(1), diag(2), diag(3))
  if (rc /= LPM_OK) print *, "lpm_plane_solver_diagnostics ERROR ", rc
  rc = lpm_plane_solver_delete(psolver)
  write(unit_out) px, py, pu, pv, diag

  ! the same sphere problem as a batch of one (offsets 0, n), one FP64 step
  offs = [0_c_int64_t, int(n, c_int64_t)]
  rc = lpm_bve_batch_new(lpm_context, batch, 1_c_int, offs, x, y, z, zeta, area, act, [1.0_c_double], &
                         [2.0_c_double * acos(-1.0_c_double)], LPM_PREC_FP64, 0_c_int)
  if (rc /= LPM_OK) print *, "lpm_bve_batch_new ERROR ", rc
  rc = lpm_bve_batch_steps(batch, 0.01_c_double, 1_c_int)
  if (rc /= LPM_OK) print *, "lpm_bve_batch_steps ERROR ", rc
  rc = lpm_bve_batch_get_state(batch, xs, ys, zs, zetas, us, vs, ws)
  rc = lpm_bve_batch_pairs(batch, npairs)
  if (rc /= LPM_OK) print *, "lpm_bve_batch_pairs ERROR ", rc
  rc = lpm_bve_batch_delete(batch)
  write(unit_out) xs, ys, zs, zetas

  ! learned far-field object, expansion-mode velocity (mode 1 needs no training)
  rc = lpm_bve_ff_new(lpm_context, ff, n, x, y, z, zeta, area, act, cell, ncell, 1.0_c_double, &
                      2.0_c_double * acos(-1.0_c_double), 0.3_c_double, 1.5_c_double, 4_c_int)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_new ERROR ", rc
  rc = lpm_bve_ff_velocity(ff, 1_c_int, us, vs, ws)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_velocity ERROR ", rc
  write(unit_out) us, vs, ws
  ! a short training and two judged learned-mode steps (judge at every step)
  rc = lpm_bve_ff_collect(ff, 20000_c_int, 1_c_int64_t, total)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_collect ERROR ", rc
  rc = lpm_bve_ff_train(ff, 2_c_int, 1024_c_int, 1.0d-3, 2_c_int64_t, loss)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_train ERROR ", rc
  rc = lpm_bve_ff_steps(ff, 0.01_c_double, 2_c_int, 2_c_int, 1_c_int, 1.0_c_double, mode_out, last_error, judged)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_steps ERROR ", rc
  allocate(cell2(n))
  rc = lpm_bve_ff_particle_cells(ff, cell2)
  if (rc /= LPM_OK) print *, "lpm_bve_ff_particle_cells ERROR ", rc
  rc = lpm_bve_ff_delete(ff)
  write(unit_out) real(total, c_double), loss, real(mode_out, c_double), last_error, real(judged, c_double)

  ! helpers: PSE eps and the MPISetup blocks of 3 ranks
  rc = lpm_pse_eps(0.3_c_double, 0.75_c_double, eps)
  rc = lpm_load_balance(n, 3_c_int, istart, iend, mlen)
  write(unit_out) eps, real(istart, c_double), real(iend, c_double)
  rc = lpm_synchronize(lpm_context)
  if (rc /= LPM_OK) print *, "lpm_synchronize ERROR ", rc
  write(unit_out) real(npairs, c_double), real(count(cell2 /= cell), c_double)
  close(unit_out)
  print *, "lpm_fortran_smoke: n =", n, " done"
end program
