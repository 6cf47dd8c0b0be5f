!> ISO_C_BINDING interface to libmose_hip.so (include/mose_hip.h).
!>
!> The derived types mirror the C structs field by field; the procedures are the C entry
!> points.  Block numbers and cell indices are 1-based; primitive state arrays have the
!> layout of the reference's blk%P(nprim, 1-gc:ni+gc, 1-gc:nj+gc, 1-gc:nk+gc)
!> (src/lib/driver/Mod_Allocate_Data.f90:104-136) and node arrays that of blk%node
!> (3, 0:ni, 0:nj, 0:nk), so a MOSE block's arrays are passed without copies.
module mose_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: MOSE_LIM_MINMOD = 0, MOSE_LIM_VANLEER = 1, &
                                           MOSE_LIM_VANALBADA = 2, MOSE_LIM_FIRST_ORDER = 3, &
                                           MOSE_LIM_MC = 4, MOSE_LIM_SUPERBEE = 5
  integer(c_int32_t), parameter, public :: MOSE_RS_HLLC = 0, MOSE_RS_AUSMP = 1, MOSE_RS_HLLE = 2, &
                                           MOSE_RS_SLAU = 3, MOSE_RS_SLAU2 = 4, MOSE_RS_HLLEPP = 5, &
                                           MOSE_RS_HLLCP = 6
  integer(c_int32_t), parameter, public :: MOSE_PREC_FP64 = 0, MOSE_PREC_FP32_FLUX = 1
  integer(c_int32_t), parameter, public :: MOSE_OK = 0, MOSE_E_ARG = 1, MOSE_E_HIP = 2, &
                                           MOSE_E_STATE = 3, MOSE_E_INTERFACE = 4, &
                                           MOSE_E_METRIC = 5, MOSE_E_BC = 6, MOSE_E_NOGPU = 7, &
                                           MOSE_E_SIZE = 8

  type, bind(C), public :: mose_cfg
    real(c_double)     :: gamma = 1.4_c_double
    real(c_double)     :: Rgas = 1.0_c_double / 1.4_c_double
    real(c_double)     :: cfl = 0.5_c_double
    integer(c_int32_t) :: limiter = 0
    integer(c_int32_t) :: riemann = 0
    integer(c_int32_t) :: n_rk = 3
    integer(c_int32_t) :: time_accurate = 1
    integer(c_int32_t) :: rampa_cfl_iter = 0
    integer(c_int32_t) :: precision = 0
    integer(c_int32_t) :: muscl_sd = 0
    integer(c_int32_t) :: prim_vars = 0
    integer(c_int32_t) :: reserved(2) = 0
    real(c_double)     :: irs_beta = 0.0_c_double
  end type mose_cfg

  type, bind(C), public :: mose_bc_entry
    integer(c_int32_t) :: b = 0, i = 0, j = 0, k = 0, f = 0, type = 0
    integer(c_int32_t) :: bs = 0, is = 0, js = 0, ks = 0, fs = 0
    integer(c_int32_t) :: d11 = 0, d12 = 0, d21 = 0, d22 = 0
    integer(c_int32_t) :: pad_ = 0
    real(c_double)     :: mach = 0, T = 0, p = 0, alpha = 0, beta = 0, rel_fac = 0
  end type mose_bc_entry

  type, bind(C), public :: mose_err
    integer(c_int32_t) :: code = 0, block = 0, i = 0, j = 0, k = 0
  end type mose_err

  public :: mose_create, mose_add_block, mose_set_viscous, mose_set_bc, mose_finalize, mose_put_state, &
            mose_get_state, mose_get_state_noalloc, mose_compute_dt, mose_step, mose_advance, mose_residual, &
            mose_put_residual_bar, mose_residual_vjp, mose_residual_vjp_geometry, mose_get_viscous_bar, &
            mose_get_state_bar, mose_get_node_bar, mose_get_inflow_bar, mose_time_residual_vjp, &
            mose_put_objective_bar, mose_adjoint_solve, mose_get_residual_bar, mose_unsteady_adjoint_ke, &
            mose_time_kernels, mose_autotune, mose_get_tiling, &
            mose_get_time, mose_set_time, mose_num_cells, mose_last_error, mose_destroy

  interface
    integer(c_int) function mose_create(cfg, ctx) bind(C, name="mose_create")
      import :: c_int, c_ptr, mose_cfg
      type(mose_cfg), intent(in) :: cfg
      type(c_ptr), intent(out)   :: ctx
    end function
    integer(c_int) function mose_add_block(ctx, b, dims, nodes) bind(C, name="mose_add_block")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value              :: ctx
      integer(c_int32_t), value       :: b
      integer(c_int32_t), intent(in)  :: dims(3)
      real(c_double), intent(in)      :: nodes(*)
    end function
    integer(c_int) function mose_set_viscous(ctx, b, params) bind(C, name="mose_set_viscous")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value         :: ctx
      integer(c_int32_t), value  :: b
      real(c_double), intent(in) :: params(5)
    end function
    integer(c_int) function mose_set_bc(ctx, entries, n) bind(C, name="mose_set_bc")
      import :: c_int, c_ptr, c_int64_t, mose_bc_entry
      type(c_ptr), value                 :: ctx
      type(mose_bc_entry), intent(in)    :: entries(*)
      integer(c_int64_t), value          :: n
    end function
    integer(c_int) function mose_finalize(ctx) bind(C, name="mose_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mose_put_state(ctx, b, P) bind(C, name="mose_put_state")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value         :: ctx
      integer(c_int32_t), value  :: b
      real(c_double), intent(in) :: P(*)
    end function
    integer(c_int) function mose_get_state(ctx, b, P) bind(C, name="mose_get_state")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value            :: ctx
      integer(c_int32_t), value     :: b
      real(c_double), intent(inout) :: P(*)
    end function
    integer(c_int) function mose_get_state_noalloc(ctx, b, P) bind(C, name="mose_get_state_noalloc")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value            :: ctx
      integer(c_int32_t), value     :: b
      real(c_double), intent(inout) :: P(*)
    end function
    integer(c_int) function mose_compute_dt(ctx, iter, dtglobal) bind(C, name="mose_compute_dt")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: iter
      real(c_double), intent(out) :: dtglobal
    end function
    integer(c_int) function mose_step(ctx, dt, time, norms) bind(C, name="mose_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: ctx
      real(c_double), intent(out) :: dt, time, norms(5)
    end function
    integer(c_int) function mose_advance(ctx, nsteps, time, norms) bind(C, name="mose_advance")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: nsteps
      real(c_double), intent(out) :: time, norms(5)
    end function
    integer(c_int) function mose_residual(ctx, b, R) bind(C, name="mose_residual")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value            :: ctx
      integer(c_int32_t), value     :: b
      real(c_double), intent(inout) :: R(*)
    end function
    ! discrete adjoint of mose_residual (state; with geometry also nodes and 405 records)
    integer(c_int) function mose_put_residual_bar(ctx, b, Rbar) bind(C, name="mose_put_residual_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value         :: ctx
      integer(c_int32_t), value  :: b
      real(c_double), intent(in) :: Rbar(*)
    end function
    integer(c_int) function mose_residual_vjp(ctx) bind(C, name="mose_residual_vjp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mose_residual_vjp_geometry(ctx) bind(C, name="mose_residual_vjp_geometry")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mose_get_viscous_bar(ctx, b, mu_ref_bar) bind(C, name="mose_get_viscous_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: b
      real(c_double), intent(out) :: mu_ref_bar
    end function
    integer(c_int) function mose_get_state_bar(ctx, b, Pbar) bind(C, name="mose_get_state_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: b
      real(c_double), intent(out) :: Pbar(*)
    end function
    integer(c_int) function mose_get_node_bar(ctx, b, Xbar) bind(C, name="mose_get_node_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: b
      real(c_double), intent(out) :: Xbar(*)
    end function
    integer(c_int) function mose_put_objective_bar(ctx, b, gP) bind(C, name="mose_put_objective_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value         :: ctx
      integer(c_int32_t), value  :: b
      real(c_double), intent(in) :: gP(*)
    end function
    integer(c_int) function mose_adjoint_solve(ctx, max_iter, rtol, check_every, iters, res) &
        bind(C, name="mose_adjoint_solve")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value              :: ctx
      integer(c_int32_t), value       :: max_iter, check_every
      real(c_double), value           :: rtol
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out)     :: res(2)
    end function
    integer(c_int) function mose_get_residual_bar(ctx, b, psi) bind(C, name="mose_get_residual_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value            :: ctx
      integer(c_int32_t), value     :: b
      real(c_double), intent(inout) :: psi(*)
    end function
    integer(c_int) function mose_unsteady_adjoint_ke(ctx, nsteps, dt, stride, ke, mu_ref_bar) &
        bind(C, name="mose_unsteady_adjoint_ke")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value            :: ctx
      integer(c_int32_t), value     :: nsteps, stride
      real(c_double), value         :: dt
      real(c_double), intent(out)   :: ke(*)
      real(c_double), intent(out)   :: mu_ref_bar(*)
    end function
    integer(c_int) function mose_get_inflow_bar(ctx, b, f, dinflow) bind(C, name="mose_get_inflow_bar")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: b, f
      real(c_double), intent(out) :: dinflow(5)
    end function
    integer(c_int) function mose_time_residual_vjp(ctx, nrep, ms) bind(C, name="mose_time_residual_vjp")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: nrep
      real(c_double), intent(out) :: ms(3)
    end function
    ! profiling and tile autotuning
    integer(c_int) function mose_time_kernels(ctx, nrep, ms) bind(C, name="mose_time_kernels")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value          :: ctx
      integer(c_int32_t), value   :: nrep
      real(c_double), intent(out) :: ms(4)
    end function
    integer(c_int) function mose_autotune(ctx, nrep, choice, best_ms) bind(C, name="mose_autotune")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value              :: ctx
      integer(c_int32_t), value       :: nrep
      integer(c_int32_t), intent(out) :: choice(3)
      real(c_double), intent(out)     :: best_ms
    end function
    integer(c_int) function mose_get_tiling(ctx, choice) bind(C, name="mose_get_tiling")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value              :: ctx
      integer(c_int32_t), intent(out) :: choice(3)
    end function
    integer(c_int) function mose_get_time(ctx, time, iter) bind(C, name="mose_get_time")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value              :: ctx
      real(c_double), intent(out)     :: time
      integer(c_int32_t), intent(out) :: iter
    end function
    integer(c_int) function mose_set_time(ctx, time, iter) bind(C, name="mose_set_time")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value        :: ctx
      real(c_double), value     :: time
      integer(c_int32_t), value :: iter
    end function
    integer(c_int64_t) function mose_num_cells(ctx) bind(C, name="mose_num_cells")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mose_last_error(ctx, err) bind(C, name="mose_last_error")
      import :: c_int, c_ptr, mose_err
      type(c_ptr), value          :: ctx
      type(mose_err), intent(out) :: err
    end function
    subroutine mose_destroy(ctx) bind(C, name="mose_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface

end module mose_hip
