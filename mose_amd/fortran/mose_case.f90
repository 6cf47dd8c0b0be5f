!> Runs one of the reference's case directories on the GPU from Fortran: the numerics of
!> input.ini ([MOSE-Numerics], [MOSE-Parameters]), the initial blocks of INPUT/ic.tec (Tecplot
!> ASCII, ORION layout: x y z rho1..rhoN u v w p, BLOCK packing, nodal coordinates and
!> cell-centred flow variables, IO_Solution.f90:97-131) and the boundary records of
!> INPUT/bc.txt (IO_BC.f90:136-414), then MOSE's explicit loop until iter-threshold or
!> time-threshold, and writes the final blocks in the same Tecplot layout.
!>
!> Usage: mose_case input.ini ic.tec bc.txt field.tec [gamma [Rgas]]
!> gamma defaults to 1.4.  Rgas (the gas constant of the case's thermo) must be given when
!> bc.txt holds inflow records (401, 405), whose imposed temperature sets a density through it;
!> otherwise only p / rho enters the path and Rgas defaults to p / rho of the first cell.
program mose_case
  use iso_c_binding
  use mose_hip
  use mose_explicit
  implicit none
  type(mose_domain) :: domain
  type(mose_cfg) :: cfg
  type(mose_bc_entry), allocatable :: bc(:)
  character(len=1024) :: fini, fic, fbc, fout, arg
  integer, allocatable :: dims(:,:)
  integer :: nb, b, iter_max, nsp, chunk
  real(c_double) :: time_max, gam, rgas

  if (command_argument_count() < 4) then
    write(*, '(A)') 'usage: mose_case input.ini ic.tec bc.txt field.tec [gamma [Rgas]]'
    stop 1
  endif
  call get_command_argument(1, fini)
  call get_command_argument(2, fic)
  call get_command_argument(3, fbc)
  call get_command_argument(4, fout)
  gam = 1.4_c_double
  rgas = -1
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg); read(arg, *) gam
  endif
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg); read(arg, *) rgas
  endif

  cfg = mose_cfg(gamma=gam, Rgas=1.0_c_double, cfl=0.5_c_double, limiter=MOSE_LIM_MINMOD, &
                 riemann=MOSE_RS_HLLC, n_rk=3, time_accurate=1)
  call read_ini(trim(fini), cfg, iter_max, time_max)
  call scan_tecplot(trim(fic), nb, dims, nsp)
  call Domain_Allocate(domain, cfg, dims)
  call load_tecplot(trim(fic), domain, nsp)
  call read_bc(trim(fbc), bc)
  if (rgas <= 0) then
    if (any(bc%type == 401 .or. bc%type == 405)) &
      call fail('inflow records (401/405) need the gas constant: pass gamma and Rgas')
    rgas = domain%blk(1)%P(5, 1, 1, 1) / domain%blk(1)%P(1, 1, 1, 1)
  endif
  domain%cfg%Rgas = rgas
  call Domain_Setup(domain, bc)

  if (time_max > 0) then
    do while (domain%time < time_max .and. domain%iter < iter_max)
      call Explicit_Step(domain)
    enddo
  else
    do while (domain%iter < iter_max)
      chunk = min(500, iter_max - domain%iter)
      call Explicit_Advance(domain, chunk)
    enddo
  endif
  call Domain_Download(domain)
  call write_tecplot(trim(fout), domain)
  write(*, '(A,ES24.16,A,I0)') ' time = ', domain%time, '  iterations = ', domain%iter
  write(*, '(A,5ES12.4)') ' residual = ', domain%res
  call Domain_Destroy(domain)

contains

  pure function lower(s) result(t)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: q
    t = s
    do q = 1, len(s)
      if (s(q:q) >= 'A' .and. s(q:q) <= 'Z') t(q:q) = achar(iachar(s(q:q)) + 32)
    enddo
  end function lower

  !> [MOSE-Numerics] and [MOSE-Parameters] keys of input.ini (Register_Numerics.f90) that this
  !> path runs; an unsupported scheme name stops with a message.
  subroutine read_ini(path, cfg, iter_max, time_max)
    character(len=*), intent(in)    :: path
    type(mose_cfg), intent(inout)   :: cfg
    integer, intent(out)            :: iter_max
    real(c_double), intent(out)     :: time_max
    character(len=512) :: line, sec, key, val
    integer :: u, ios, eq, sc
    iter_max = huge(1)
    time_max = -1
    sec = ''
    open(newunit=u, file=path, status='old', action='read')
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      sc = index(line, ';')
      if (sc > 0) line = line(1:sc-1)
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '[') then
        sec = lower(line(2:index(line, ']')-1))
        cycle
      endif
      eq = index(line, '=')
      if (eq == 0) cycle
      key = lower(adjustl(line(1:eq-1)))
      val = adjustl(line(eq+1:))
      if (trim(sec) == 'mose-parameters') then
        if (trim(key) == 'iter-threshold') read(val, *) iter_max
        if (trim(key) == 'time-threshold') read(val, *) time_max
      else if (trim(sec) == 'mose-numerics') then
        select case (trim(key))
        case ('cfl');                read(val, *) cfg%cfl
        case ('cfl-rise-threshold'); read(val, *) cfg%rampa_cfl_iter
        case ('irs-beta');           read(val, *) cfg%irs_beta
        case ('time-accurate');      cfg%time_accurate = merge(1, 0, lower(val(1:1)) == 't')
        case ('time-scheme')
          select case (lower(trim(val)))
          case ('rk1', 'euler'); cfg%n_rk = 1
          case ('rk2');          cfg%n_rk = 2
          case ('rk3');          cfg%n_rk = 3
          case default; call fail('time-scheme ' // trim(val))
          end select
        case ('integration-variables')
          cfg%prim_vars = merge(1, 0, lower(trim(val)) == 'prim')
        case ('space-reconstruction')
          select case (lower(trim(val)))
          case ('muscl');       cfg%muscl_sd = 0
          case ('muscl-sd');    cfg%muscl_sd = 1
          case ('first-order'); cfg%limiter = MOSE_LIM_FIRST_ORDER
          case default; call fail('space-reconstruction ' // trim(val))
          end select
        case ('flux-limiter')
          if (cfg%limiter /= MOSE_LIM_FIRST_ORDER) then
            select case (lower(trim(val)))
            case ('minmod');    cfg%limiter = MOSE_LIM_MINMOD
            case ('vanleer');   cfg%limiter = MOSE_LIM_VANLEER
            case ('vanalbada'); cfg%limiter = MOSE_LIM_VANALBADA
            case ('mc');        cfg%limiter = MOSE_LIM_MC
            case ('superbee');  cfg%limiter = MOSE_LIM_SUPERBEE
            case default; call fail('flux-limiter ' // trim(val))
            end select
          endif
        case ('riemann-solver')
          select case (trim(val))
          case ('HLLC');   cfg%riemann = MOSE_RS_HLLC
          case ('AUSM+');  cfg%riemann = MOSE_RS_AUSMP
          case ('HLLE');   cfg%riemann = MOSE_RS_HLLE
          case ('SLAU');   cfg%riemann = MOSE_RS_SLAU
          case ('SLAU2');  cfg%riemann = MOSE_RS_SLAU2
          case ('HLLE++'); cfg%riemann = MOSE_RS_HLLEPP
          case ('HLLC+');  cfg%riemann = MOSE_RS_HLLCP
          case default; call fail('riemann-solver ' // trim(val))
          end select
        end select
      endif
    enddo
    close(u)
  end subroutine read_ini

  subroutine fail(what)
    character(len=*), intent(in) :: what
    write(*, '(A)') '[ERROR] not run by this path: ' // what
    stop 2
  end subroutine fail

  !> Integer after "key=" in a zone header (I=, J=, K=).
  integer function header_int(h, key)
    character(len=*), intent(in) :: h, key
    integer :: p, q
    p = index(h, key // '=')
    if (p == 0) p = index(h, key // ' =')
    if (p == 0) call fail('zone header without ' // key)
    p = p + index(h(p:), '=')
    q = p
    do while (q <= len(h))
      if (h(q:q) /= ' ') exit
      q = q + 1
    enddo
    p = q
    do while (q <= len(h))
      if (h(q:q) < '0' .or. h(q:q) > '9') exit
      q = q + 1
    enddo
    read(h(p:q-1), *) header_int
  end function header_int

  logical function is_data(line)
    character(len=*), intent(in) :: line
    character(len=1) :: c
    c = adjustl(line)
    is_data = (c >= '0' .and. c <= '9') .or. c == '-' .or. c == '+' .or. c == '.'
  end function is_data

  !> First pass: number of zones, their cell counts and the number of species densities.
  subroutine scan_tecplot(path, nb, dims, nsp)
    character(len=*), intent(in)       :: path
    integer, intent(out)               :: nb, nsp
    integer, allocatable, intent(out)  :: dims(:,:)
    character(len=4096) :: line, h
    integer :: u, ios, q, tmp(3, 4096)
    nb = 0
    nsp = 0
    open(newunit=u, file=path, status='old', action='read')
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      if (index(lower(adjustl(line)), 'variables') == 1) then
        do q = 1, 64
          if (index(lower(line), '"rho' // trim(itoa(q)) // '"') > 0) nsp = q
        enddo
      else if (index(lower(adjustl(line)), 'zone') == 1) then
        if (nb >= size(tmp, 2)) call fail('ic file with more zones than scan_tecplot holds (4096)')
        h = line
        do
          read(u, '(A)', iostat=ios) line
          if (ios /= 0 .or. is_data(line)) exit
          h = trim(h) // ' ' // trim(line)
        enddo
        nb = nb + 1
        tmp(1, nb) = header_int(h, 'I') - 1
        tmp(2, nb) = header_int(h, 'J') - 1
        tmp(3, nb) = header_int(h, 'K') - 1
      endif
    enddo
    close(u)
    if (nb == 0 .or. nsp == 0) call fail('ic file without zones or rho1')
    allocate(dims(3, nb))
    dims = tmp(:, 1:nb)
  end subroutine scan_tecplot

  function itoa(i) result(s)
    integer, intent(in) :: i
    character(len=12) :: s
    write(s, '(I0)') i
  end function itoa

  !> Second pass: nodes and primitive state (rho = sum of the species densities) per zone.
  subroutine load_tecplot(path, domain, nsp)
    character(len=*), intent(in)     :: path
    type(mose_domain), intent(inout) :: domain
    integer, intent(in)              :: nsp
    character(len=4096) :: line
    integer :: u, ios, b, c, s, nn, ncell
    real(c_double), allocatable :: a(:), rho(:)
    open(newunit=u, file=path, status='old', action='read')
    b = 0
    do while (b < domain%nb)
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) call fail('ic file ended early')
      if (index(lower(adjustl(line)), 'zone') /= 1) cycle
      do
        read(u, '(A)') line
        if (is_data(line)) exit
      enddo
      backspace(u)
      b = b + 1
      associate (k => domain%blk(b))
        nn = (k%ni + 1) * (k%nj + 1) * (k%nk + 1)
        ncell = k%ni * k%nj * k%nk
        allocate(a(nn))
        do c = 1, 3
          read(u, *) a
          k%node(c, :, :, :) = reshape(a, [k%ni + 1, k%nj + 1, k%nk + 1])
        enddo
        deallocate(a)
        allocate(a(ncell), rho(ncell))
        rho = 0
        do s = 1, nsp
          read(u, *) a
          rho = rho + a
        enddo
        k%P(1, 1:k%ni, 1:k%nj, 1:k%nk) = reshape(rho, [k%ni, k%nj, k%nk])
        do c = 2, 5
          read(u, *) a
          k%P(c, 1:k%ni, 1:k%nj, 1:k%nk) = reshape(a, [k%ni, k%nj, k%nk])
        enddo
        deallocate(a, rho)
      end associate
    enddo
    close(u)
  end subroutine load_tecplot

  !> An angle of bc.txt: a number or the token 'normal' (parse_dir_tok, IO_BC.f90:443-456).
  real(c_double) function dir_tok(t)
    character(len=*), intent(in) :: t
    if (index(lower(t), 'normal') > 0) then
      dir_tok = huge(1.0_c_double)
    else
      read(t, *) dir_tok
    endif
  end function dir_tok

  !> bc.txt records in file order (Read_BCfile, IO_BC.f90:136-414) for the types this path runs.
  subroutine read_bc(path, bc)
    character(len=*), intent(in)                    :: path
    type(mose_bc_entry), allocatable, intent(out)   :: bc(:)
    type(mose_bc_entry), allocatable :: tmp(:)
    character(len=4096) :: line
    character(len=64) :: tok(16)
    integer :: u, ios, n, nl, q
    open(newunit=u, file=path, status='old', action='read')
    nl = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      nl = nl + 1
    enddo
    rewind(u)
    allocate(tmp(nl))
    n = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      n = n + 1
      read(line, *) tmp(n)%b, tmp(n)%i, tmp(n)%j, tmp(n)%k, tmp(n)%f, tmp(n)%type
      select case (tmp(n)%type)
      case (0, 300, 400)
      case (101, 201)
        read(u, *) tmp(n)%bs, tmp(n)%is, tmp(n)%js, tmp(n)%ks, tmp(n)%fs, &
                   tmp(n)%d11, tmp(n)%d12, tmp(n)%d21, tmp(n)%d22
      case (401, 405, 406)
        read(u, '(A)') line
        do q = 1, len_trim(line)
          if (line(q:q) == ',') line(q:q) = ' '
        enddo
        call split(line, tok)
        if (tmp(n)%type == 405) then
          read(tok(1), *) tmp(n)%mach
          read(tok(2), *) tmp(n)%T
          read(tok(3), *) tmp(n)%p
          tmp(n)%alpha = dir_tok(tok(4))
          tmp(n)%beta = dir_tok(tok(5))
          read(tok(6), *) tmp(n)%rel_fac
        else if (tmp(n)%type == 401) then
          read(tok(1), *) tmp(n)%T
          read(tok(2), *) tmp(n)%p
          tmp(n)%alpha = dir_tok(tok(3))
          tmp(n)%beta = dir_tok(tok(4))
          read(tok(5), *) tmp(n)%rel_fac
        else
          read(tok(1), *) tmp(n)%p
          read(tok(2), *) tmp(n)%rel_fac
        endif
      case default
        call fail('bc type ' // trim(itoa(tmp(n)%type)))
      end select
    enddo
    close(u)
    allocate(bc(n))
    bc = tmp(1:n)
  end subroutine read_bc

  !> Blank-separated tokens of a line (at most size(tok)).
  subroutine split(line, tok)
    character(len=*), intent(in)  :: line
    character(len=*), intent(out) :: tok(:)
    integer :: q, n, p0
    tok = ''
    n = 0
    q = 1
    do while (q <= len_trim(line) .and. n < size(tok))
      if (line(q:q) == ' ') then
        q = q + 1
        cycle
      endif
      p0 = q
      do while (q <= len_trim(line))
        if (line(q:q) == ' ') exit
        q = q + 1
      enddo
      n = n + 1
      tok(n) = line(p0:q-1)
    enddo
  end subroutine split

  !> The blocks in the layout load_tecplot reads (one species), 17 significant digits.
  subroutine write_tecplot(path, domain)
    character(len=*), intent(in)  :: path
    type(mose_domain), intent(in) :: domain
    integer :: u, b, c
    open(newunit=u, file=path, status='replace', action='write')
    write(u, '(A)') ' VARIABLES ="x" "y" "z"  "rho1" "u" "v" "w" "p"'
    do b = 1, domain%nb
      associate (k => domain%blk(b))
        write(u, '(A,I0,A,I0,A,I0,A,I0,A,ES24.16E3)') ' ZONE  T = B', b, '-IG, I=', k%ni + 1, ', J=', &
          k%nj + 1, ', K=', k%nk + 1, ', DATAPACKING=BLOCK, VARLOCATION=([1-3]=NODAL,[4-8]=CELLCENTERED),' // &
          ' SOLUTIONTIME=', domain%time
        do c = 1, 3
          write(u, '(ES25.16E3)') k%node(c, :, :, :)
        enddo
        do c = 1, 5
          write(u, '(ES25.16E3)') k%P(c, 1:k%ni, 1:k%nj, 1:k%nk)
        enddo
      end associate
    enddo
    close(u)
  end subroutine write_tecplot

end program mose_case
