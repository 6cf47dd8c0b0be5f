!> Host side of the explicit residual-and-update path on one GPU, in Fortran.
!>
!> A `mose_domain` holds the blocks as the reference stores them (node(3,0:ni,0:nj,0:nk),
!> P(5,-1:ni+2,-1:nj+2,-1:nk+2) primitive with two ghost layers, src/lib/driver/
!> Mod_Allocate_Data.f90:104-136) and the handle of the device solver.  `Explicit_Step`
!> advances one time step like the reference's Explicit_Step
!> (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-146): CFL time step, time += dt,
!> n_RK stages of ghost fill, fluxes, boundary fluxes and SSP-RK update, then the residual
!> norms.  Everything runs on the GPU; the state is only copied back by `Domain_Download`.
!> A non-physical state stops the program with the reference's messages
!> (Lib_Newstate.f90:337, Lib_Convective.f90:48).
module mose_explicit
  use iso_c_binding
  use mose_hip
  implicit none
  private

  type, public :: mose_block
    integer :: ni = 0, nj = 0, nk = 0
    real(c_double), allocatable :: node(:,:,:,:)   ! (3, 0:ni, 0:nj, 0:nk)
    real(c_double), allocatable :: P(:,:,:,:)      ! (5, -1:ni+2, -1:nj+2, -1:nk+2)
  end type mose_block

  type, public :: mose_domain
    type(c_ptr)    :: ctx = c_null_ptr
    type(mose_cfg) :: cfg
    integer        :: nb = 0
    type(mose_block), allocatable :: blk(:)
    real(c_double) :: time = 0, dtglobal = 0
    integer        :: iter = 0
    real(c_double) :: res(5) = 0
  end type mose_domain

  public :: Domain_Allocate, Domain_Setup, Explicit_Step, Explicit_Advance, &
            Domain_Download, Domain_Destroy, Face_BC_Records, Cartesian_Nodes

contains

  !> Allocate nb blocks of the given sizes (dims(3, nb)).
  subroutine Domain_Allocate(domain, cfg, dims)
    type(mose_domain), intent(inout) :: domain
    type(mose_cfg), intent(in)       :: cfg
    integer, intent(in)              :: dims(:,:)
    integer :: b
    domain%cfg = cfg
    domain%nb = size(dims, 2)
    allocate(domain%blk(domain%nb))
    do b = 1, domain%nb
      associate (k => domain%blk(b))
        k%ni = dims(1, b); k%nj = dims(2, b); k%nk = dims(3, b)
        allocate(k%node(3, 0:k%ni, 0:k%nj, 0:k%nk))
        allocate(k%P(5, -1:k%ni+2, -1:k%nj+2, -1:k%nk+2))
        k%node = 0; k%P = 0
      end associate
    enddo
  end subroutine Domain_Allocate

  !> Create the device solver: blocks, boundary records (the bc.txt records of
  !> src/lib/io/IO_BC.f90), layout, then upload of every block's primitive state.
  subroutine Domain_Setup(domain, bc)
    type(mose_domain), intent(inout)      :: domain
    type(mose_bc_entry), intent(in)       :: bc(:)
    integer :: b
    integer(c_int32_t) :: d(3)
    call check(domain, mose_create(domain%cfg, domain%ctx), "mose_create")
    do b = 1, domain%nb
      d = [domain%blk(b)%ni, domain%blk(b)%nj, domain%blk(b)%nk]
      call check(domain, mose_add_block(domain%ctx, int(b, c_int32_t), d, domain%blk(b)%node), &
                 "mose_add_block")
    enddo
    call check(domain, mose_set_bc(domain%ctx, bc, int(size(bc), c_int64_t)), "mose_set_bc")
    call check(domain, mose_finalize(domain%ctx), "mose_finalize")
    do b = 1, domain%nb
      call check(domain, mose_put_state(domain%ctx, int(b, c_int32_t), domain%blk(b)%P), &
                 "mose_put_state")
    enddo
    call check(domain, mose_set_time(domain%ctx, domain%time, int(domain%iter, c_int32_t)), &
               "mose_set_time")
  end subroutine Domain_Setup

  !> One explicit time step on the GPU (Mod_Explicit.f90:10-146).
  subroutine Explicit_Step(domain)
    type(mose_domain), intent(inout) :: domain
    real(c_double) :: dt, t, norms(5)
    call check(domain, mose_step(domain%ctx, dt, t, norms), "Explicit_Step")
    domain%dtglobal = dt
    domain%time = t
    domain%iter = domain%iter + 1
    domain%res = norms
  end subroutine Explicit_Step

  !> nsteps explicit steps without host synchronisation between them.
  subroutine Explicit_Advance(domain, nsteps)
    type(mose_domain), intent(inout) :: domain
    integer, intent(in)              :: nsteps
    real(c_double) :: t, norms(5)
    call check(domain, mose_advance(domain%ctx, int(nsteps, c_int32_t), t, norms), "Explicit_Advance")
    domain%time = t
    domain%iter = domain%iter + nsteps
    domain%res = norms
  end subroutine Explicit_Advance

  !> Copy every block's primitive state (ghost layers included) back to the host.
  subroutine Domain_Download(domain)
    type(mose_domain), intent(inout) :: domain
    integer :: b
    do b = 1, domain%nb
      call check(domain, mose_get_state(domain%ctx, int(b, c_int32_t), domain%blk(b)%P), &
                 "mose_get_state")
    enddo
  end subroutine Domain_Download

  subroutine Domain_Destroy(domain)
    type(mose_domain), intent(inout) :: domain
    if (c_associated(domain%ctx)) call mose_destroy(domain%ctx)
    domain%ctx = c_null_ptr
  end subroutine Domain_Destroy

  !> Uniform Cartesian nodes of a block spanning [lo, hi]: node q = q * step + lo along each
  !> axis, step = (hi - lo) / n, the last node exactly hi.
  subroutine Cartesian_Nodes(blk, lo, hi)
    type(mose_block), intent(inout) :: blk
    real(c_double), intent(in)      :: lo(3), hi(3)
    integer :: i, j, k
    do k = 0, blk%nk
      do j = 0, blk%nj
        do i = 0, blk%ni
          blk%node(1, i, j, k) = axis_node(i, blk%ni, lo(1), hi(1))
          blk%node(2, i, j, k) = axis_node(j, blk%nj, lo(2), hi(2))
          blk%node(3, i, j, k) = axis_node(k, blk%nk, lo(3), hi(3))
        enddo
      enddo
    enddo
  contains
    pure real(c_double) function axis_node(q, n, a, b)
      integer, intent(in)        :: q, n
      real(c_double), intent(in) :: a, b
      if (q == n) then
        axis_node = b
      else
        axis_node = real(q, c_double) * ((b - a) / real(n, c_double)) + a
      endif
    end function axis_node
  end subroutine Cartesian_Nodes

  !> Boundary records of one block whose faces carry non-connection types ftype(1:6)
  !> (faces 1/2 i-min/max, 3/4 j, 5/6 k), one record per boundary cell, first tangential
  !> index fastest, as the reference's bc.txt lists them (IO_BC.f90:62-71 counts them).
  subroutine Face_BC_Records(b, blk, ftype, recs)
    integer, intent(in)                            :: b
    type(mose_block), intent(in)                   :: blk
    integer, intent(in)                            :: ftype(6)
    type(mose_bc_entry), allocatable, intent(out)  :: recs(:)
    integer :: f, d, n(3), t1, t2, a, c, m, idx(3)
    n = [blk%ni, blk%nj, blk%nk]
    allocate(recs(2 * (n(2) * n(3) + n(1) * n(3) + n(1) * n(2))))
    m = 0
    do f = 1, 6
      d = (f - 1) / 2 + 1
      t1 = mod(d, 3) + 1
      t2 = mod(d + 1, 3) + 1
      if (t1 > t2) then
        a = t1; t1 = t2; t2 = a
      endif
      do c = 1, n(t2)
        do a = 1, n(t1)
          m = m + 1
          idx(d) = merge(1, n(d), mod(f, 2) == 1)
          idx(t1) = a
          idx(t2) = c
          recs(m) = mose_bc_entry(b=b, i=idx(1), j=idx(2), k=idx(3), f=f, type=ftype(f))
        enddo
      enddo
    enddo
  end subroutine Face_BC_Records

  !> Stop with the reference's message when the device reports a failure.
  subroutine check(domain, rc, where)
    type(mose_domain), intent(in) :: domain
    integer(c_int), intent(in)    :: rc
    character(*), intent(in)      :: where
    type(mose_err) :: e
    integer(c_int) :: ierr
    if (rc == MOSE_OK) return
    e = mose_err()
    if (c_associated(domain%ctx)) ierr = mose_last_error(domain%ctx, e)
    select case (rc)
    case (MOSE_E_INTERFACE)
      write(*,*) '[ERROR] Non-physical state detected at the interface.'
      write(*,'(A,I0,A,3(I0,1X))') '        block ', e%block, ', cell ', e%i, e%j, e%k
      error stop 1
    case (MOSE_E_STATE)
      write(*,'(A,I0,A,3(I0,1X))') ' block ', e%block, ', cell ', e%i, e%j, e%k
      error stop "NaN or p<0 detected"
    case default
      write(*,'(A,A,A,I0)') ' [ERROR] ', where, ' failed with code ', rc
      error stop 1
    end select
  end subroutine check

end module mose_explicit
