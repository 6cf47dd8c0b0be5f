/*
 * mose_hip.h — C ABI of the MI355X (gfx950) explicit residual-and-update path.
 *
 * The reference (open-hydra/MOSE, Fortran) runs one explicit time step as
 *   Compute_dt -> per RK stage { Fill_Ghost_Cell -> Fluxes -> BC_Fluxes -> RK_Newstate }
 *   -> Compute_Residual
 * (src/lib/numerics/time/explicit/Mod_Explicit.f90:10-146).  Its operator API for this
 * path is a set of per-stage module calls on `type(MOSE_domain_type)` plus the
 * `rlimiter` / `Riemann` / `RK_Newstate` procedure pointers bound at setup
 * (src/lib/config/Assign_Setup.f90:41-56).  A per-face scalar callback cannot cross
 * into a GPU efficiently, so this ABI replaces the whole stage loop: the Fortran host
 * keeps MOSE's loop structure and calls these entry points through ISO_C_BINDING
 * (mose_amd/fortran/mose_hip_iface.f90).
 *
 * Conventions
 *  - Every entry point returns 0 on success and a nonzero MOSE_E_* code otherwise; the
 *    block/cell of the failure is available from mose_last_error().
 *  - Block numbers and cell indices are 1-based, as in the reference (b, i, j, k).
 *  - Primitive state arrays use the reference layout of blk%P:
 *    P(nprim, 1-gc:ni+gc, 1-gc:nj+gc, 1-gc:nk+gc), Fortran order, nprim=5, gc=2
 *    (src/lib/driver/Mod_Allocate_Data.f90:104-136; src/lib/base/Global_m.f90:17).
 *  - Node arrays use the layout of blk%node: (3, 0:ni, 0:nj, 0:nk), Fortran order.
 *  - The library owns its device buffers; host pointers are only read or written for
 *    the duration of a call.  A context is not thread-safe: call it from one thread.
 *  - The gas is single-species and calorically perfect (gamma, R).
 */
#ifndef MOSE_HIP_H
#define MOSE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Limiters: rlimiter_MINMOD/VANLEER/VANALBADA/IORD/MC/SB, src/lib/numerics/space/Lib_Limiters.f90:24-105 */
enum { MOSE_LIM_MINMOD = 0, MOSE_LIM_VANLEER = 1, MOSE_LIM_VANALBADA = 2, MOSE_LIM_FIRST_ORDER = 3,
       MOSE_LIM_MC = 4, MOSE_LIM_SUPERBEE = 5 };
/* Riemann solvers: riemann_HLLC (Lib_Riemann_HLL.f90:147-258), riemann_AUSMp
 * (Lib_Riemann_AUSM.f90:150-238), riemann_HLLE (Lib_Riemann_HLL.f90:261-320), riemann_SLAU and
 * riemann_SLAU2 (Lib_Riemann_SLAU.f90:19-140) */
enum { MOSE_RS_HLLC = 0, MOSE_RS_AUSMP = 1, MOSE_RS_HLLE = 2, MOSE_RS_SLAU = 3, MOSE_RS_SLAU2 = 4,
       MOSE_RS_HLLEPP = 5, MOSE_RS_HLLCP = 6 };
/* MOSE_RS_HLLEPP: riemann_HLLEpp (Tramel's HLLE++, Lib_Riemann_HLL.f90:321-440); MOSE_RS_HLLCP:
 * riemann_HLLCSD ('HLLC+', :486-534).  Both take the shock-detector weight SD_Tramel
 * (Lib_Shock_Detector.f90:9-38) of the face's low cell (of the block's own boundary cell on a
 * connection face), computed from the pressure of the cell and its six face neighbours. */
/* Arithmetic of reconstruction + Riemann flux; state, accumulation and reductions are always FP64. */
enum { MOSE_PREC_FP64 = 0, MOSE_PREC_FP32_FLUX = 1 };

/* Error codes (the reference `stop`s at the same places). */
enum {
  MOSE_OK = 0,
  MOSE_E_ARG = 1,          /* bad argument / unsupported configuration                        */
  MOSE_E_HIP = 2,          /* HIP runtime error                                                */
  MOSE_E_STATE = 3,        /* NaN or p<0 after an update: Lib_Newstate.f90:334-338             */
  MOSE_E_INTERFACE = 4,    /* non-physical reconstructed face state: Lib_Convective.f90:45-50  */
  MOSE_E_METRIC = 5,       /* det=0 or non-positive volume: Lib_Metrics.f90:286-288,353-356    */
  MOSE_E_BC = 6,           /* boundary list inconsistent with the blocks: IO_BC.f90:123-127    */
  MOSE_E_NOGPU = 7,        /* no HIP device                                                    */
  MOSE_E_SIZE = 8          /* input beyond a fixed maximum (mose_get_state_noalloc)            */
};

typedef struct {
  double  gamma;            /* ratio of specific heats                                         */
  double  Rgas;             /* gas constant; cp = gamma R/(gamma-1)                            */
  double  cfl;              /* [MOSE-Numerics] cfl                                             */
  int32_t limiter;          /* [MOSE-Numerics] flux-limiter       (MOSE_LIM_*)                 */
  int32_t riemann;          /* [MOSE-Numerics] riemann-solver     (MOSE_RS_*)                  */
  int32_t n_rk;             /* [MOSE-Numerics] time-scheme: 1 euler, 2 RK2, 3 RK3              */
  int32_t time_accurate;    /* [MOSE-Numerics] time-accurate: 1 global dt, 0 local dt          */
  int32_t rampa_cfl_iter;   /* CFL ramp length in iterations (0: off), Mod_dt.f90:45           */
  int32_t precision;        /* MOSE_PREC_*                                                     */
  int32_t muscl_sd;         /* [MOSE-Numerics] space-reconstruction = MUSCL-SD: limited slopes  */
                            /* scaled by the shock-detector weight (Lib_Reconstruction.f90:26-27) */
  int32_t prim_vars;        /* [MOSE-Numerics] integration-variables = prim: the RK stage on the   */
                            /* primitive variables (Newstate_Primitive, Lib_Newstate.f90:168-317)  */
  int32_t reserved[2];
  double  irs_beta;         /* [MOSE-Numerics] irs-beta: implicit residual smoothing of every   */
                            /* stage's residual when > 0 (Lib_Newstate.f90:41-75, Lib_IRS.f90)    */
} mose_cfg;

/* One boundary-face-cell record of bc.txt (src/lib/io/IO_BC.f90:136-414).
 * type 101/201 use (bs,is,js,ks,fs,d11,d12,d21,d22); type 405 uses (mach,T,p,alpha,beta,rel_fac)
 * with alpha/beta in radians, or an angle >= DBL_MAX/2 for bc.txt's 'normal' token (the
 * reference's huge() sentinel, parse_dir_tok IO_BC.f90:443-456): the angle then follows the
 * boundary face normal (Compute_Inflow_Direction, Lib_BC_Fluxes_Inflow.f90:479-496).
 * Type 401 (subsonic inlet, BC_Inlet_StagnCond, Lib_BC_Fluxes_Inflow.f90:19-75) uses T = T0,
 * p = p0, alpha, beta (angles as for 405) and rel_fac (the relaxation of p0 towards the cell
 * pressure); type 406 (outlet, BC_Outflow, Lib_BC_Fluxes_Outflow.f90:12-59) uses p = the back
 * pressure (0: extrapolation).  All records of one block face share one type, and for
 * 405/401/406 one set of values.  401/406 are refused on viscous blocks (MOSE_E_BC).
 * Supported types: 0, 101, 201, 300, 400, 401, 405, 406. */
typedef struct {
  int32_t b, i, j, k, f, type;
  int32_t bs, is, js, ks, fs;
  int32_t d11, d12, d21, d22;
  int32_t pad_;
  double  mach, T, p, alpha, beta, rel_fac;
} mose_bc_entry;

typedef struct { int32_t code, block, i, j, k; } mose_err;

typedef struct mose_ctx mose_ctx;

/* Setup: replaces MOSE_setup's Setup_Data_Structure/Setup_BC/Setup_Metrics
 * (src/lib/driver/Wrap_Setup.f90:60-82). */
int  mose_create(const mose_cfg* cfg, mose_ctx** out);
/* Allocate_Block (Mod_Allocate_Data.f90:104-136) for block b of nblocks; nodes as blk%node. */
int  mose_add_block(mose_ctx* ctx, int32_t b, const int32_t dims[3], const double* nodes);
/* Read_BCfile (IO_BC.f90:136-414) — the entries in file order. */
int  mose_set_bc(mose_ctx* ctx, const mose_bc_entry* entries, int64_t n);
/* Laminar viscous terms (the reference's model > 0 without turbulence: Diffusive_Flux,
 * Lib_Diffusive.f90:10-174; BC_Connection_Visc, Lib_BC_Fluxes_Connection.f90:115-228; the VNN
 * time-step limit, Mod_dt.f90:88-103) for block b (0: every block added so far), after
 * mose_add_block and before mose_finalize.  params = {mu_ref, T_ref, S, Pr, vnn}: Sutherland
 * viscosity mu(T) = mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S), conductivity mu cp / Pr.  The
 * reference takes mu from FLINT transport tables; Sutherland's law stands in for them.
 * Viscous blocks must be rectilinear with face types 0 / 101 / 201 (aligned connections). */
int  mose_set_viscous(mose_ctx* ctx, int32_t b, const double params[5]);
/* Setup_Metrics (Mod_Metrics.f90:8-107) + device allocation; call once after blocks and BCs. */
int  mose_finalize(mose_ctx* ctx);

/* blk%P in / out (primitive, with ghost layers; ghosts are ignored on input). */
int  mose_put_state(mose_ctx* ctx, int32_t b, const double* P);
int  mose_get_state(mose_ctx* ctx, int32_t b, double* P);
/* mose_get_state without any device allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight coupling loops,
 * harnesses that forbid hipMalloc).  The conversion stages through a fixed 32 MiB device array
 * per GPU instead of the buffer mose_get_state allocates on first use, so blocks of more than
 * 838,860 cells counting the two ghost layers on every side ((ni+4)(nj+4)(nk+4)) are refused
 * with MOSE_E_SIZE.  The array is shared by every context on a GPU: calls that overlap on one
 * GPU race.  Call from the thread and device that ran mose_finalize (else MOSE_E_ARG). */
int  mose_get_state_noalloc(mose_ctx* ctx, int32_t b, double* P);

/* Compute_dt (Mod_dt.f90:10-107) preceded by dtglobal=1e5 (Mod_Explicit.f90:44), using
 * iteration number `iter` for the CFL ramp; writes the global minimum dt. */
int  mose_compute_dt(mose_ctx* ctx, int32_t iter, double* dtglobal);

/* Explicit_Step (Mod_Explicit.f90:10-113): iter+=1, dt (+time+=dt when time-accurate),
 * Copy_State, n_rk x {Fill_Ghost_Cell, Fluxes, BC_Fluxes, RK_Newstate}, Compute_Residual.
 * dt_out / time_out / norms (5 values: |d rho|, |du|, |dv|, |dw|, |dp| L2) may be NULL. */
int  mose_step(mose_ctx* ctx, double* dt_out, double* time_out, double* norms);

/* nsteps calls of mose_step without host synchronisation in between (the step is a
 * captured hipGraph).  norms (if not NULL) receives the last step's norms. */
int  mose_advance(mose_ctx* ctx, int32_t nsteps, double* time_out, double* norms);

/* Ghost fill (every entry point that evaluates fluxes starts with it): the records run in
 * the reference's bc order semantics -- across a block one or two cells thick a record reads
 * the opposite face's ghost cells as an earlier record of the same fill wrote them, or as the
 * previous fill (or mose_put_state) left them -- so such blocks give deterministic results.
 *
 * One Fill_Ghost_Cell + Fluxes + BC_Fluxes on the current state, no update: the residual
 * R(nprim, 1-gc:ni+gc, ...) of block b in the blk%R layout (interior cells filled, ghosts 0).
 * This is the value the reference holds in blk%R right before RK_Newstate. */
int  mose_residual(mose_ctx* ctx, int32_t b, double* R);

/* Discrete adjoint of mose_residual with respect to the primitive state (no counterpart in
 * the reference, which has no adjoint; it is the backward of the residual for gradient
 * checks and adjoint-based sensitivities).  Usage: mose_put_residual_bar for every block
 * (Rbar in the blk%R layout; ghost entries ignored), mose_residual_vjp, then
 * mose_get_state_bar for every block: Pbar = (dR/dP)^T Rbar in the blk%P layout (ghost
 * entries 0), R being the residual of the current state including its ghost fill.
 * Limitations: ghost cells filled from other ghost cells (extrapolation across a block
 * thinner than three cells) are treated as constants; the schemes that use the shock
 * detector (MOSE_RS_HLLEPP, MOSE_RS_HLLCP, muscl_sd) and the 401 inlet flux have no adjoint
 * (MOSE_E_ARG); the 406 outlet flux has one (state and face normal, not p_exit). */
int  mose_put_residual_bar(mose_ctx* ctx, int32_t b, const double* Rbar);
int  mose_residual_vjp(mose_ctx* ctx);
int  mose_get_state_bar(mose_ctx* ctx, int32_t b, double* Pbar);
/* Geometry adjoint: mose_residual_vjp plus the adjoints of the residual with respect to the
 * block nodes and the imposed values of supersonic inflow faces (shape and boundary-value
 * sensitivities, the Xbar of the residual's backward).  Needs general metrics (some block
 * not rectilinear), else MOSE_E_ARG.  After it, mose_get_state_bar as above and
 *   mose_get_node_bar(b, Xbar): Xbar = (dR/dX)^T Rbar in the node layout (3,0:ni,0:nj,0:nk),
 *     through every metric the residual reads (face normals and areas, cell lengths incl.
 *     the ghost-cell lengths of each bc type, symmetry mirror normals);
 *   mose_get_inflow_bar(b, f, out): adjoints with respect to the 405 record of face f (1..6)
 *     of block b: out = (mach, T, p, alpha, beta).
 * No counterpart in the reference (no adjoint). */
int  mose_residual_vjp_geometry(mose_ctx* ctx);
/* After either vjp on a layout with viscous blocks: the adjoint of Rbar . R with respect to
 * the Sutherland reference viscosity mu_ref of block b (d/dRe = -mu_ref^2 d/dmu_ref for
 * mu_ref = 1/Re); 0 for an inviscid block.  The viscous terms of the residual (interior
 * faces and BC_Connection_Visc faces) are part of both vjps.  No counterpart in the
 * reference (no adjoint). */
int  mose_get_viscous_bar(mose_ctx* ctx, int32_t b, double* mu_ref_bar);
int  mose_get_node_bar(mose_ctx* ctx, int32_t b, double* Xbar);
int  mose_get_inflow_bar(mose_ctx* ctx, int32_t b, int32_t f, double out[5]);
/* Steady adjoint solve (no counterpart in the reference, which has no adjoint; it serves
 * shape and boundary-value gradients of a steady objective J(U*, X) at a converged state U*
 * of the local-time-step iteration, Mod_Explicit.f90:45-85):
 *   mose_put_objective_bar(b, gP): dJ/dP of block b in the blk%P layout (ghost entries
 *     ignored);
 *   mose_adjoint_solve: solves (dR/dP)^T psi = dJ/dP on the device by the transpose of the
 *     primal stages (same SSP-RK coefficients, the local dt of the current state, no CFL
 *     ramp): lambda <- (1-c) lambda^n + c (lambda + (dP/dU)^T (dJ/dP - (dR/dP)^T D lambda)),
 *     D = dt/vol per cell, psi = D lambda; it converges whenever the primal iteration does.
 *     It stops after max_iter iterations or when |Q| <= rtol |Q_0| (checked every
 *     check_every iterations; Q the conserved-variable adjoint residual); iters and
 *     res = {|Q_0|, |Q|} report it;
 *   mose_get_residual_bar(b, psi): psi in the blk%R layout (ghost entries 0).
 * Afterwards psi is the residual adjoint in place: mose_residual_vjp_geometry (without
 * mose_put_residual_bar) gives psi^T dR/dX and psi^T dR/d(inflow), and
 * dJ/dX = dJ/dX|_U - psi^T dR/dX. */
int  mose_put_objective_bar(mose_ctx* ctx, int32_t b, const double* gP);
int  mose_adjoint_solve(mose_ctx* ctx, int32_t max_iter, double rtol, int32_t check_every, int32_t* iters,
                        double res[2]);
int  mose_get_residual_bar(mose_ctx* ctx, int32_t b, double* psi);
/* Adjoint of a fixed-dt march for the kinetic energy (no counterpart in the reference, which
 * has no adjoint; it serves Reynolds-number sensitivities of unsteady runs): nsteps explicit
 * steps of the reference's stages (Lib_Newstate.f90:134-160, Lib_RK.f90:8-27) with dt held
 * fixed from the current state, then ke[b] = sum over block b's cells of rho |u|^2 vol / 2 of
 * the final state.  With mu_ref_bar non-NULL the reverse sweep follows on the device and
 * mu_ref_bar[b] = dJ/dmu_ref of block b for J = sum_b ke[b] (d/dRe = -mu_ref^2 d/dmu_ref for
 * mu_ref = 1/Re), and mose_get_state_bar(b) returns dJ/dP of the initial state.  Memory: the
 * state at the start of every `stride`-th step is kept on the device and each segment's
 * steps are recomputed from it, ceil(nsteps/stride) + stride states of 5 doubles per pool
 * cell (stride <= 0: sqrt(nsteps)); MOSE_E_HIP if they do not fit.  The context's state,
 * clock and time step are unchanged afterwards.  Time-accurate contexts without residual
 * smoothing or the primitive update only (else MOSE_E_ARG); the reverse refuses what
 * mose_residual_vjp refuses. */
int  mose_unsteady_adjoint_ke(mose_ctx* ctx, int32_t nsteps, double dt, int32_t stride, double* ke,
                              double* mu_ref_bar);
/* Average device ms of one residual evaluation (ms[0]), one adjoint evaluation (ms[1]) and,
 * with general metrics, one geometry adjoint evaluation (ms[2], device passes only; 0 on
 * rectilinear layouts) over nrep back-to-back runs on the current state and residual
 * adjoint (profiling); afterwards the adjoint results are those of the last run. */
int  mose_time_residual_vjp(mose_ctx* ctx, int32_t nrep, double ms[3]);

/* Simulation clock and iteration counter (domain%time, domain%iter). */
int  mose_get_time(mose_ctx* ctx, double* time, int32_t* iter);
int  mose_set_time(mose_ctx* ctx, double time, int32_t iter);

/* Average device time in ms of one launch of each kernel of a step, measured with HIP events
 * on the context's stream over nrep back-to-back launches on the current state:
 * ms[0] CFL time step, ms[1] ghost fill, ms[2] intermediate RK stage, ms[3] last RK stage
 * (with residual-norm partials).  The state U^n is not modified.  Profiling entry point with
 * no counterpart in the reference (which has no per-kernel timers). */
int  mose_time_kernels(mose_ctx* ctx, int32_t nrep, double ms[4]);

/* Autotuner: times the three stage launches of one step (on the current state, U^n not
 * modified) for every compiled tile shape, k-chunk of 8/16/32/64 planes and both tile orders
 * (linear, XCD-aware), keeps the fastest and re-captures the step with it.  choice (may be
 * NULL) receives {TX*1000+TY, k-chunk, xcd order}, best_ms the stage time of the winner
 * summed over the three launches.  No counterpart in the reference (its loops have fixed
 * OpenMP schedules). */
int  mose_autotune(mose_ctx* ctx, int32_t nrep, int32_t choice[3], double* best_ms);
/* choice[2] bits: 1 = XCD-aware tile order (tiles of one contiguous range per XCD queue),
 * 2 = one workgroup per tile instead of the persistent launch (as many workgroups as the
 * device holds, taking tiles from the queues). */

/* The tiling in use, in mose_autotune's encoding {TX*1000+TY, k-chunk, xcd order}.  No
 * counterpart in the reference. */
int  mose_get_tiling(const mose_ctx* ctx, int32_t choice[3]);

/* Total interior cells over all blocks. */
int64_t mose_num_cells(const mose_ctx* ctx);

int  mose_last_error(const mose_ctx* ctx, mose_err* err);
const char* mose_strerror(int32_t code);
void mose_destroy(mose_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* MOSE_HIP_H */
