!> Sod shock tube (test/1D/Sod79 of the reference) on the GPU, driven from Fortran.
!>
!> x in [0,1], diaphragm at 0.5, (rho,u,p) = (1,0,1) | (0.125,0,0.1), gamma = 1.4,
!> minmod + HLLC, CFL 0.5, three-stage SSP-RK, time-accurate to t = 0.2; faces 1-4 type
!> 400, faces 5-6 type 0 (test/1D/Sod79/INPUT/bc_x*.txt).
!>
!> Usage: mose_run [output_file [ncells]]
!> Writes one line per cell: x, rho, u, p, specific internal energy (the columns of
!> test/1D/Sod79/exact.txt), and prints the final time, step count and residual norms.
program mose_run
  use iso_c_binding
  use mose_hip
  use mose_explicit
  implicit none
  type(mose_domain) :: domain
  type(mose_cfg) :: cfg
  type(mose_bc_entry), allocatable :: bc(:)
  character(len=512) :: out, arg
  integer :: n, i, u
  real(c_double) :: x, gam

  out = 'sod_profile.txt'
  n = 1000
  if (command_argument_count() >= 1) call get_command_argument(1, out)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) n
  endif

  cfg = mose_cfg(gamma=1.4_c_double, Rgas=1.0_c_double / 1.4_c_double, cfl=0.5_c_double, &
                 limiter=MOSE_LIM_MINMOD, riemann=MOSE_RS_HLLC, n_rk=3, time_accurate=1)
  gam = cfg%gamma
  call Domain_Allocate(domain, cfg, reshape([n, 1, 1], [3, 1]))
  call Cartesian_Nodes(domain%blk(1), [0.0_c_double, 0.0_c_double, 0.0_c_double], &
                       [1.0_c_double, 1.0_c_double, 1.0_c_double])
  do i = -1, n + 2
    x = (real(min(max(i, 1), n), c_double) - 0.5_c_double) / real(n, c_double)
    if (x < 0.5_c_double) then
      domain%blk(1)%P(:, i, :, :) = 0
      domain%blk(1)%P(1, i, :, :) = 1.0_c_double
      domain%blk(1)%P(5, i, :, :) = 1.0_c_double
    else
      domain%blk(1)%P(:, i, :, :) = 0
      domain%blk(1)%P(1, i, :, :) = 0.125_c_double
      domain%blk(1)%P(5, i, :, :) = 0.1_c_double
    endif
  enddo
  call Face_BC_Records(1, domain%blk(1), [400, 400, 400, 400, 0, 0], bc)
  call Domain_Setup(domain, bc)

  do while (domain%time < 0.2_c_double)
    call Explicit_Step(domain)
  enddo
  call Domain_Download(domain)

  open(newunit=u, file=trim(out), status='replace', action='write')
  do i = 1, n
    associate (P => domain%blk(1)%P(:, i, 1, 1))
      write(u, '(5ES25.16E3)') 0.5_c_double * (domain%blk(1)%node(1, i - 1, 0, 0) + &
                                                domain%blk(1)%node(1, i, 0, 0)), &
                               P(1), P(2), P(5), P(5) / ((gam - 1.0_c_double) * P(1))
    end associate
  enddo
  close(u)
  write(*, '(A,ES24.16,A,I0)') ' time = ', domain%time, '  steps = ', domain%iter
  write(*, '(A,5ES12.4)') ' residual = ', domain%res
  call Domain_Destroy(domain)
end program mose_run
