/*
 * mose_oracle.h — CPU oracle of the explicit residual-and-update path (TEST INFRASTRUCTURE).
 *
 * A plain-C, FP64 behavioural restatement of the reference's hot path, following the
 * reference files cited in mose_oracle.c.  Only tests, the smoke check and the bench's CPU
 * baseline leg may load it, and only as the checker: the product library (libmose_hip.so)
 * never links, loads or calls it.
 */
#ifndef MOSE_ORACLE_H
#define MOSE_ORACLE_H

#include <stdint.h>
#include "../include/mose_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ob_domain ob_domain;

ob_domain* ob_create(const mose_cfg* cfg, int32_t nblocks);
int  ob_add_block(ob_domain* d, int32_t b, const int32_t dims[3], const double* nodes);
int  ob_set_bc(ob_domain* d, const mose_bc_entry* entries, int64_t n);
int  ob_setup(ob_domain* d);
/* Laminar viscous terms (model > 0) for block b (0: every block), before ob_setup:
 * params = {mu_ref, T_ref, S, Pr, vnn}: Sutherland mu(T) = mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S),
 * conductivity mu cp / Pr, VNN number of the time step. */
int  ob_set_viscous(ob_domain* d, int32_t b, const double params[5]);
int  ob_put_state(ob_domain* d, int32_t b, const double* P);
int  ob_get_state(ob_domain* d, int32_t b, double* P);
int  ob_fill_ghost(ob_domain* d);
int  ob_residual(ob_domain* d, int32_t b, double* R);
int  ob_compute_dt(ob_domain* d, int32_t iter, double* dtglobal);
int  ob_step(ob_domain* d, double* dt_out, double* time_out, double* norms);
int  ob_get_time(ob_domain* d, double* time, int32_t* iter);
int  ob_set_time(ob_domain* d, double time, int32_t iter);
int  ob_get_metrics(ob_domain* d, int32_t b, double* fn, double* fa, double* dl, double* M, double* vol);
int  ob_get_dtlocal(ob_domain* d, int32_t b, double* dt);
int  ob_last_error(ob_domain* d, mose_err* e);
int  ob_ndir(ob_domain* d);
int  ob_num_threads(void);   /* OpenMP threads the oracle runs its sweeps on */
void ob_destroy(ob_domain* d);

/* Point kernels, exposed for unit tests. */
double ob_limiter(int32_t kind, double x, double y);
void   ob_reconstruct(int32_t kind, const double q[4], double l0, double l1, double l2,
                      double lm, double lp, double* ql, double* qr);
void   ob_riemann(const mose_cfg* cfg, const double PL[5], const double PR[5],
                  const double n[3], double F[5]);
/* as ob_riemann with the shock-detector weight beta (HLLE++ / HLLC+ use it) */
void   ob_riemann_beta(const mose_cfg* cfg, const double PL[5], const double PR[5],
                       const double n[3], double beta, double F[5]);
double ob_sd_tramel(double p, double p1, double p2, double p3, double p4, double p5, double p6);
/* implicit residual smoothing of a residual field R (blk%R layout) of block b, in place */
int    ob_irs_smooth(ob_domain* d, int32_t b, double beta, double* R);
void   ob_prim2cons(const mose_cfg* cfg, const double P[5], double U[5]);
void   ob_cons2prim(const mose_cfg* cfg, const double U[5], double P[5]);

#ifdef __cplusplus
}
#endif
#endif
