// mose_hip.hip — kernels and C ABI of the MI355X explicit residual-and-update path.
//
// One explicit step (Explicit_Step, src/lib/numerics/time/explicit/Mod_Explicit.f90:10-146)
// is this launch sequence on one HIP stream, captured once into a hipGraph:
//   k_begin    iter += 1; dt = CFL time step of the current state; time += dt when
//              time-accurate; reset the dt accumulator to 1e5     (Mod_Explicit.f90:36-49)
//   n_rk x { k_ghost   ghost layers of every block face            (Lib_Ghost.f90:11-137)
//            k_bcface  fluxes of the non-connection boundary faces (Mod_BC_Fluxes.f90:10-218)
//            k_stage   all other face fluxes (Mod_Fluxes.f90:46-209) and the SSP-RK stage
//                      update (Lib_Newstate.f90:134-160) fused; the last stage also
//                      produces the residual-norm partial sums (Mod_Diagnostic.f90:13-50)
//                      and, from the state it writes, the CFL time step of the next step
//                      (Compute_dt, Mod_dt.f90:10-107: wave-shuffle min, one atomic per wave) }
//   k_normfin  fixed-order sum of the partials, sqrt                (Mod_Explicit.f90:111-112)
// k_dt computes the CFL time step directly when the state did not come from a step.
// Copy_State (PO = P) needs no pass: U^n stays in its own buffer while the stages run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/mose_hip.h"
#include "mose_kernels.hpp"

using namespace mose;

// ============================================================================ device
#include "mose_stage.hpp"
#include "mose_adjoint.hpp"
#include "mose_adjoint_kernels.hpp"
#include "mose_adjsolve.hpp"
#include "mose_unsteady.hpp"
#include "mose_tables.hpp"

struct DtArgs {
  const double* U;
  const Tile* tiles;
  const BlockDev* blocks;
  const LineEnt* lines;
  const double* geo;   // general metrics (GEO 1)
  unsigned long long* dtmin;
  double* dtloc;
  const int* iter;
  int64_t npool;
  int ntiles, ndir, rampa, local_dt;
  int tx, ty;        // patch shape of the tiles
  int iter_add;      // the ramp uses *iter + iter_add
  double cfl, gam, R;
};

// CFL time step of every interior cell (Compute_Dt, Mod_dt.f90:61-85) and its global
// minimum.  GEO 0: rectilinear metrics (versor = axis, dl = cell length); GEO 1: the unit
// rows of M and dl from the geometry arrays.
template <int GEO>
__global__ void __launch_bounds__(kMaxTileThreads) k_dt(DtArgs a) {
  const int tid = threadIdx.x, tx = tid % a.tx, ty = tid / a.tx;
  const double gm1 = a.gam - 1.0;
  const int it = *a.iter + a.iter_add;
  double mymin = 1e300;
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev B = a.blocks[T.b];
    const int i = T.i0 + tx, j = T.j0 + ty;
    if (i > B.ni || j > B.nj) continue;
    const LineEnt* Lx = a.lines + B.line[0];
    const LineEnt* Ly = a.lines + B.line[1];
    const LineEnt* Lz = a.lines + B.line[2];
    const int64_t cbase = B.off + (int64_t)(i + 1) + (int64_t)B.sy * (j + 1);
    for (int k = T.k0; k < T.k1; ++k) {
      const int64_t c = cbase + B.sz * (int64_t)(k + 1);
      double u[5], P[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = a.U[v * a.npool + c];
      cons2prim(u, gm1, P);
      const double snd = sqrt(a.gam * P[4] / P[0]);
      double dtc = 1e8;
      if constexpr (GEO == 0) {
        dtc = fmin(Lx[i + 1].dl / (fabs(P[1]) + snd) * a.cfl, dtc);
        if (a.ndir > 1) dtc = fmin(Ly[j + 1].dl / (fabs(P[2]) + snd) * a.cfl, dtc);
        if (a.ndir > 2) dtc = fmin(Lz[k + 1].dl / (fabs(P[3]) + snd) * a.cfl, dtc);
        if (B.visc) {   // VNN condition (Mod_dt.f90:88-103)
          const double mi = sutherland(P[4] / (P[0] * a.R), B.vis[0], B.vis[1], B.vis[2]);
          const double dls[3] = {Lx[i + 1].dl, Ly[j + 1].dl, Lz[k + 1].dl};
          for (int dd = 0; dd < a.ndir; ++dd) dtc = fmin(P[0] * dls[dd] * dls[dd] * B.vis[4] / mi, dtc);
        }
      } else {
        const double* G = a.geo;
        const size_t np = (size_t)a.npool;
        for (int dd = 0; dd < a.ndir; ++dd) {
          const double vel = fabs(P[1] * G[(GVS + 3 * dd) * np + c] + P[2] * G[(GVS + 3 * dd + 1) * np + c] +
                                  P[3] * G[(GVS + 3 * dd + 2) * np + c]);
          dtc = fmin(G[(GDL + dd) * np + c] / (vel + snd) * a.cfl, dtc);
        }
      }
      if (it < a.rampa) dtc = dtc * it / a.rampa;
      if (a.local_dt) a.dtloc[c] = dtc;
      mymin = fmin(mymin, dtc);
    }
  }
  for (int o = 32; o > 0; o >>= 1) mymin = fmin(mymin, __shfl_xor(mymin, o, 64));
  if ((tid & 63) == 0 && mymin < 1e300)
    atomicMin(a.dtmin, (unsigned long long)__double_as_longlong(mymin));
}

// Start of a step: iter += 1, take the CFL time step computed from the current state (by
// the previous step's last stage or by k_dt), time += dt when time-accurate, and reset the
// accumulator (dtglobal = 1e5, Mod_Explicit.f90:36-49) for this step's last stage.
__global__ void k_begin(unsigned long long* dtmin, int* iter, double* scal, int time_accurate) {
  *iter += 1;
  const double dt = __longlong_as_double((long long)*dtmin);
  scal[S_DT] = dt;
  if (time_accurate) scal[S_TIME] += dt;
  *dtmin = (unsigned long long)__double_as_longlong(1e5);
}

__global__ void k_dtreset(unsigned long long* dtmin) {
  *dtmin = (unsigned long long)__double_as_longlong(1e5);
}

// ---------------------------------------------------------------------------- IRS
// Implicit residual smoothing (Newstate_Conservative with IRS, Lib_Newstate.f90:41-75, and
// Residual_Smoothing, Lib_IRS.f90:14-191): the residual-only stage kernel writes the scaled
// residual R* = -R dt / vol, each direction d < ndir gets the one-layer residual ghosts
// (k_irs_ghost: the donor cell for 101 connections, the boundary cell for every other type)
// and two Jacobi sweeps R*_i <- (R_i + beta (R*_{i-1} + R*_{i+1})) / (1 + 2 beta) with the ghost
// values held (k_irs_jacobi: both sweeps in one pass from one buffer into the other; those of
// the last direction are formed inside k_irs_update, which takes the RK stage with the
// smoothed R*).
struct IrsArgs {
  double* R;
  double* dst;
  const Tile* tiles;
  const BlockDev* blocks;
  const LineEnt* lines;
  const double* geo;
  const double* scal;
  const double* dtloc;
  int64_t npool;
  int ntiles, tx, local_dt, d;
  double beta;
};

__global__ void __launch_bounds__(256) k_irs_ghost(const int64_t* __restrict__ pairs, int n, double* R, int64_t N) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int64_t dst = pairs[2 * q], src = pairs[2 * q + 1];
#pragma unroll
    for (int v = 0; v < 5; ++v) R[v * N + dst] = R[v * N + src];
  }
}

// Both Jacobi sweeps of direction d for every interior cell: the first sweep's values at the
// two neighbours along d are recomputed from R (a neighbour that is a ghost keeps its R), so
// one pass reads R (the residual ghosts filled) and writes the second sweep into dst.
__global__ void __launch_bounds__(kMaxTileThreads) k_irs_jacobi(IrsArgs a) {
  const int tx = threadIdx.x % a.tx, ty = threadIdx.x / a.tx;
  const int64_t N = a.npool;
  const double den = 1.0 + 2.0 * a.beta;
  const double* __restrict__ R = a.R;
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev B = a.blocks[T.b];
    const int i = T.i0 + tx, j = T.j0 + ty;
    if (i > B.ni || j > B.nj) continue;
    const int64_t st = a.d == 0 ? 1 : (a.d == 1 ? (int64_t)B.sy : B.sz);
    const int64_t cbase = B.off + (int64_t)(i + 1) + (int64_t)B.sy * (j + 1);
    const int n = a.d == 0 ? B.ni : (a.d == 1 ? B.nj : B.nk);
    for (int k = T.k0; k < T.k1; ++k) {
      const int64_t c = cbase + B.sz * (int64_t)(k + 1);
      const int q = a.d == 0 ? i : (a.d == 1 ? j : k);     // 1-based index along d
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double* r = R + v * N;
        const double r0 = r[c], rm = r[c - st], rp = r[c + st];
        const double sm = q > 1 ? (rm + a.beta * (r[c - 2 * st] + r0)) / den : rm;
        const double sp = q < n ? (rp + a.beta * (r0 + r[c + 2 * st])) / den : rp;
        a.dst[v * N + c] = (r0 + a.beta * (sm + sp)) / den;
      }
    }
  }
}

// RK stage of every interior cell with the smoothed R* (MODE 0 / 1 of k_stage without the
// fluxes): U = Un + rkc (Uin - Un + R*), the state check and density floor, and in the last
// stage the CFL time step of the new state and the residual-norm partials per tile.  R* is
// a.Rout after the smoothing of the directions before d; the two Jacobi sweeps of direction d
// are formed here, as k_irs_jacobi forms them (residual ghosts of a.Rout filled).
//
// PRIM (integration-variables = prim, Newstate_Primitive, Lib_Newstate.f90:168-247): the stage
// is taken on the primitive variables, P = Pn + rkc (Pin - Pn + R*), then stored as conserved;
// R* is the primitive-variable residual (prim_residual), formed here from the flux residual in
// a.Rout when d < 0 (no smoothing), else converted before the smoothing (k_prim_resid).
//
// Compute_Prim_Residual (Lib_Newstate.f90:252-299) for one species of a calorically perfect gas
// (h(T) - R T = cv T, cp - R = cv): flux residual r of a cell with primitive state P -> the
// primitive-variable residual times dt.
__device__ inline void prim_residual(const double* P, double* r, double dt, double vol, double Rg, double cp) {
  const double rho = P[0], T = P[4] / (rho * Rg);
  const double inv_volume = 1.0 / vol, inv_rhovolume = inv_volume / rho;
  r[0] = -r[0] * inv_volume;
  const double dwrotot = r[0];
  for (int c = 1; c <= 3; ++c) r[c] = -r[c] * inv_rhovolume - P[c] / rho * dwrotot;
  const double dwE = -r[4] * inv_volume;
  const double dwroe = dwE - rho * (P[1] * r[1] + P[2] * r[2] + P[3] * r[3]) -
                       0.5 * dwrotot * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]);
  const double eiroi = (cp * T - Rg * T) * r[0];
  const double cv = P[0] * (cp - Rg);
  const double dwTemp = (dwroe - eiroi) / cv;
  r[4] = Rg * (T * r[0] + P[0] * dwTemp);
#pragma unroll
  for (int v = 0; v < 5; ++v) r[v] = r[v] * dt;
}

// The flux residual in a.Rout of every interior cell -> its primitive-variable residual (the
// smoothing of integration-variables = prim acts on it).
template <int GEO>
__global__ void __launch_bounds__(kMaxTileThreads) k_prim_resid(StageArgs a, int tx_) {
  const int tid = threadIdx.x, tx = tid % tx_, ty = tid / tx_;
  const double gm1 = a.gam - 1.0, cp = a.gam * a.R / gm1;
  const int64_t N = a.npool;
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev* __restrict__ Bp = a.blocks + T.b;
    const int i = T.i0 + tx, j = T.j0 + ty;
    if (i > Bp->ni || j > Bp->nj) continue;
    const LineEnt* Lx = a.lines + Bp->line[0];
    const LineEnt* Ly = a.lines + Bp->line[1];
    const LineEnt* Lz = a.lines + Bp->line[2];
    const int64_t cbase = Bp->off + (int64_t)(i + 1) + (int64_t)Bp->sy * (j + 1);
    for (int k = T.k0; k < T.k1; ++k) {
      const int64_t c = cbase + Bp->sz * (int64_t)(k + 1);
      double u[5], P[5], r[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) { u[v] = a.Uin[v * N + c]; r[v] = a.Rout[v * N + c]; }
      cons2prim(u, gm1, P);
      const double dt = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
      const double vol = GEO != 1 ? Lx[i + 1].delta * Ly[j + 1].delta * Lz[k + 1].delta : a.geo[(size_t)GVOL * N + c];
      prim_residual(P, r, dt, vol, a.R, cp);
#pragma unroll
      for (int v = 0; v < 5; ++v) a.Rout[v * N + c] = r[v];
    }
  }
}

template <int GEO, bool LAST, bool PRIM>
__global__ void __launch_bounds__(kMaxTileThreads) k_irs_update(StageArgs a, int tx_, int d, double beta) {
  const int tid = threadIdx.x, tx = tid % tx_, ty = tid / tx_;
  const double gm1 = a.gam - 1.0;
  const int64_t N = a.npool;
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev* __restrict__ Bp = a.blocks + T.b;
    const int i = T.i0 + tx, j = T.j0 + ty;
    double nsum[5] = {0, 0, 0, 0, 0};
    double dtmin_t = 1e300;
    if (i <= Bp->ni && j <= Bp->nj) {
      const LineEnt* Lx = a.lines + Bp->line[0];
      const LineEnt* Ly = a.lines + Bp->line[1];
      const LineEnt* Lz = a.lines + Bp->line[2];
      const int64_t cbase = Bp->off + (int64_t)(i + 1) + (int64_t)Bp->sy * (j + 1);
      for (int k = T.k0; k < T.k1; ++k) {
        const int64_t c = cbase + Bp->sz * (int64_t)(k + 1);
        double un[5], ui[5], unew[5], P[5], rs[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) { un[v] = a.Un[v * N + c]; ui[v] = a.Uin[v * N + c]; }
        if (d < 0) {
#pragma unroll
          for (int v = 0; v < 5; ++v) rs[v] = a.Rout[v * N + c];
        } else {
          const int64_t st = d == 0 ? 1 : (d == 1 ? (int64_t)Bp->sy : Bp->sz);
          const int n = d == 0 ? Bp->ni : (d == 1 ? Bp->nj : Bp->nk);
          const int q = d == 0 ? i : (d == 1 ? j : k);
          const double den = 1.0 + 2.0 * beta;
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const double* r = a.Rout + v * N;
            const double r0 = r[c], rm = r[c - st], rp = r[c + st];
            const double sm = q > 1 ? (rm + beta * (r[c - 2 * st] + r0)) / den : rm;
            const double sp = q < n ? (rp + beta * (r0 + r[c + 2 * st])) / den : rp;
            rs[v] = (r0 + beta * (sm + sp)) / den;
          }
        }
        if constexpr (PRIM) {
          double Pi[5], Pn[5];
          cons2prim(ui, gm1, Pi);
          cons2prim(un, gm1, Pn);
          if (d < 0) {
            const double dt = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
            const double vol =
                GEO != 1 ? Lx[i + 1].delta * Ly[j + 1].delta * Lz[k + 1].delta : a.geo[(size_t)GVOL * N + c];
            prim_residual(Pi, rs, dt, vol, a.R, a.gam * a.R / gm1);
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) P[v] = Pn[v] + a.rkc * (Pi[v] - Pn[v] + rs[v]);
        } else {
#pragma unroll
          for (int v = 0; v < 5; ++v) unew[v] = un[v] + a.rkc * (ui[v] - un[v] + rs[v]);
          cons2prim(unew, gm1, P);
        }
        bool bad = !(P[4] >= 0.0);
#pragma unroll
        for (int v = 0; v < 5; ++v) bad = bad || (P[v] != P[v]);
        if (bad) set_error(a.err, MOSE_E_STATE, T.b + 1, i, j, k);
        if constexpr (PRIM) {
          if (P[0] < 1e-20) P[0] = 1e-20;
          prim2cons(P, 1.0 / gm1, unew);
        } else {
          if (P[0] < 1e-20) { P[0] = 1e-20; prim2cons(P, 1.0 / gm1, unew); }
        }
        if constexpr (LAST) {
          const double snd = sqrt(a.gam * P[4] / P[0]);
          double dtc = 1e8;
          if constexpr (GEO != 1) {
            dtc = fmin(Lx[i + 1].dl / (fabs(P[1]) + snd) * a.cfl, dtc);
            if (a.ndir > 1) dtc = fmin(Ly[j + 1].dl / (fabs(P[2]) + snd) * a.cfl, dtc);
            if (a.ndir > 2) dtc = fmin(Lz[k + 1].dl / (fabs(P[3]) + snd) * a.cfl, dtc);
            if (GEO == 2 && Bp->visc) {   // VNN condition (Mod_dt.f90:88-103)
              const double mi = sutherland(P[4] / (P[0] * a.R), Bp->vis[0], Bp->vis[1], Bp->vis[2]);
              const double dls[3] = {Lx[i + 1].dl, Ly[j + 1].dl, Lz[k + 1].dl};
              for (int dd = 0; dd < a.ndir; ++dd) dtc = fmin(P[0] * dls[dd] * dls[dd] * Bp->vis[4] / mi, dtc);
            }
          } else {
            for (int dd = 0; dd < a.ndir; ++dd) {
              const double vel = fabs(P[1] * a.geo[(size_t)(GVS + 3 * dd) * N + c] +
                                      P[2] * a.geo[(size_t)(GVS + 3 * dd + 1) * N + c] +
                                      P[3] * a.geo[(size_t)(GVS + 3 * dd + 2) * N + c]);
              dtc = fmin(a.geo[(size_t)(GDL + dd) * N + c] / (vel + snd) * a.cfl, dtc);
            }
          }
          const int itn = *a.iter + 1;
          if (itn < a.rampa) dtc = dtc * itn / a.rampa;
          if (a.local_dt) a.dtnext[c] = dtc;
          dtmin_t = fmin(dtmin_t, dtc);
          double pn[5];
          cons2prim(un, gm1, pn);
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const double dd = P[v] - pn[v];
            nsum[v] += dd * dd;
          }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) a.Uout[v * N + c] = unew[v];
      }
    }
    if constexpr (LAST) {   // fixed-order reduction over the workgroup
      __shared__ double sW[kMaxTileThreads / 64][5];
      const int lane = tid & 63, wave = tid >> 6, nw = (int)(blockDim.x / 64);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        double x = nsum[v];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sW[wave][v] = x;
      }
      __syncthreads();
      if (tid < 5) {
        double tot = 0.0;
        for (int w = 0; w < nw; ++w) tot += sW[w][tid];
        a.partials[5 * (int64_t)t + tid] = tot;
      }
      for (int o = 32; o > 0; o >>= 1) dtmin_t = fmin(dtmin_t, __shfl_xor(dtmin_t, o, 64));
      if (lane == 0 && dtmin_t < 1e300) atomicMin(a.dtmin, (unsigned long long)__double_as_longlong(dtmin_t));
      __syncthreads();
    }
  }
}

// Shock-detector weight of every interior cell (Fluxes_blk, Mod_Fluxes.f90:78-85): SD_Tramel of
// the pressures of the cell and its six face neighbours, ghost layers filled.  Ghost entries
// keep the 1 they were initialised with.
struct SdArgs {
  const double* pres;
  double* beta;
  const Tile* tiles;
  const BlockDev* blocks;
  int64_t npool;
  int ntiles, tx;
  double gam;
};

// Pressure of every pool cell (the detector's input, one cons2prim per cell); entries of
// pool cells no fill writes (ghost edges and corners) are computed but never read.
__global__ void __launch_bounds__(256) k_pres(const double* __restrict__ U, int64_t N, double gam,
                                              double* __restrict__ pres) {
  const double gm1 = gam - 1.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
    double u[5], P[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = U[v * N + c];
    cons2prim(u, gm1, P);
    pres[c] = P[4];
  }
}

__global__ void __launch_bounds__(kMaxTileThreads) k_sd(SdArgs a) {
  const int tx = threadIdx.x % a.tx, ty = threadIdx.x / a.tx;
  const double* __restrict__ p = a.pres;
  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev B = a.blocks[T.b];
    const int i = T.i0 + tx, j = T.j0 + ty;
    if (i > B.ni || j > B.nj) continue;
    const int64_t sy = B.sy, sz = B.sz;
    const int64_t cbase = B.off + (int64_t)(i + 1) + sy * (j + 1);
    for (int k = T.k0; k < T.k1; ++k) {
      const int64_t c = cbase + sz * (int64_t)(k + 1);
      a.beta[c] = sd_tramel(p[c], p[c - 1], p[c + 1], p[c - sy], p[c + sy], p[c - sz], p[c + sz]);
    }
  }
}

__global__ void __launch_bounds__(256) k_normfin(const double* partials, int ntiles, double* scal) {
  __shared__ double red[256][5];
  const int tid = threadIdx.x;
  double s[5] = {0, 0, 0, 0, 0};
  for (int t = tid; t < ntiles; t += 256)
#pragma unroll
    for (int v = 0; v < 5; ++v) s[v] += partials[5 * (int64_t)t + v];
#pragma unroll
  for (int v = 0; v < 5; ++v) red[tid][v] = s[v];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
#pragma unroll
      for (int v = 0; v < 5; ++v) red[tid][v] += red[tid + w][v];
    __syncthreads();
  }
  if (tid < 5) scal[S_NORM + tid] = sqrt(red[0][tid]);
}

// Ghost layers (Fill_Ghost_Cell, Lib_Ghost.f90:11-137) on the conserved SoA state.
// Copies (101/201 Ghost_Connection, zero-gradient) move conserved values, which equal the
// primitive copies of the reference exactly; extrapolation (Ghost_Extrapolate) and the
// symmetry mirror (Ghost_Symmetry) act on primitive variables as in the reference.
// The records of one launch are independent: a record that reads a ghost cell written by an
// earlier record runs in a later launch (phase), and one that reads a ghost cell written by a
// later record reads it from arc, the values at the end of the previous fill (narc cells).
__global__ void __launch_bounds__(256) k_ghost(const GhostOp* __restrict__ ops, int nops, double* U,
                                               int64_t N, double gam, const double* __restrict__ arc,
                                               int64_t narc) {
  const double gm1 = gam - 1.0, rgm1 = 1.0 / gm1;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nops; q += gridDim.x * blockDim.x) {
    const GhostOp op = ops[q];
    const int64_t d1 = op.b - op.s, d2 = op.b - 2 * op.s;
    auto rd = [&](int v, int64_t cell, int slot) { return slot >= 0 ? arc[v * narc + slot] : U[v * N + cell]; };
    if (op.type == G_COPY) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double a1 = rd(v, op.src1, op.a1), a2 = rd(v, op.src2, op.a2);
        U[v * N + d1] = a1;
        U[v * N + d2] = a2;
      }
    } else if (op.type == G_ZG) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double x = U[v * N + op.b];
        U[v * N + d1] = x;
        U[v * N + d2] = x;
      }
    } else if (op.type == G_EXTRAP) {
      double u[5], Pb[5], P1[5], P2[5], G1[5], G2[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = U[v * N + op.b];
      cons2prim(u, gm1, Pb);
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = rd(v, op.b + op.s, op.a1);
      cons2prim(u, gm1, P1);
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = rd(v, op.b + 2 * op.s, op.a2);
      cons2prim(u, gm1, P2);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        G1[v] = 3.0 * Pb[v] - 3.0 * P1[v] + P2[v];
        G2[v] = 3.0 * G1[v] - 3.0 * Pb[v] + P1[v];
      }
      prim2cons(G1, rgm1, u);
#pragma unroll
      for (int v = 0; v < 5; ++v) U[v * N + d1] = u[v];
      prim2cons(G2, rgm1, u);
#pragma unroll
      for (int v = 0; v < 5; ++v) U[v * N + d2] = u[v];
    } else {   // G_SYM: first layer only
      double u[5], P[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = U[v * N + op.b];
      cons2prim(u, gm1, P);
      const double un = P[1] * op.n[0] + P[2] * op.n[1] + P[3] * op.n[2];
      P[1] -= 2.0 * un * op.n[0];
      P[2] -= 2.0 * un * op.n[1];
      P[3] -= 2.0 * un * op.n[2];
      prim2cons(P, rgm1, u);
#pragma unroll
      for (int v = 0; v < 5; ++v) U[v * N + d1] = u[v];
    }
  }
}

// blk%P (AoS primitive, padded, Fortran order) <-> conserved SoA pool.
// Values of the archived ghost cells at the end of a fill (or after a state upload).
__global__ void k_garc_save(const double* __restrict__ U, int64_t N, const int64_t* __restrict__ cells, int64_t narc,
                            double* __restrict__ arc) {
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < narc; s += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int v = 0; v < 5; ++v) arc[v * narc + s] = U[v * N + cells[s]];
}

__global__ void k_put(const double* __restrict__ Paos, double* U, int64_t N, int64_t off, int64_t ncell,
                      double gam) {
  const double rgm1 = 1.0 / (gam - 1.0);
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    double P[5], u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) P[v] = Paos[5 * c + v];
    prim2cons(P, rgm1, u);
#pragma unroll
    for (int v = 0; v < 5; ++v) U[v * N + off + c] = u[v];
  }
}
__global__ void k_get(const double* __restrict__ U, int64_t N, int64_t off, int64_t ncell, double gam,
                      double* __restrict__ Paos) {
  const double gm1 = gam - 1.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    double u[5], P[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = U[v * N + off + c];
    if (u[0] != 0.0) cons2prim(u, gm1, P);
#pragma unroll
    for (int v = 0; v < 5; ++v) Paos[5 * c + v] = P[v];
  }
}
__global__ void k_get_raw(const double* __restrict__ R, int64_t N, int64_t off, int64_t ncell,
                          double* __restrict__ Raos) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int v = 0; v < 5; ++v) Raos[5 * c + v] = R[v * N + off + c];
}
// SoA pool -> blk%R-layout AoS of one block with the ghost layers set to zero.
__global__ void k_get_interior(const double* __restrict__ R, int64_t N, int64_t off, int ni, int nj, int nk,
                               double* __restrict__ Raos) {
  const int64_t sy = ni + 4, sz = sy * (nj + 4), ncell = sz * (nk + 4);
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % sy) - 1, j = (int)((c / sy) % (nj + 4)) - 1, k = (int)(c / sz) - 1;
    const bool in = i >= 1 && i <= ni && j >= 1 && j <= nj && k >= 1 && k <= nk;
#pragma unroll
    for (int v = 0; v < 5; ++v) Raos[5 * c + v] = in ? R[v * N + off + c] : 0.0;
  }
}
// blk%R-layout AoS of one block -> SoA pool (plain copy).
__global__ void k_put_raw(const double* __restrict__ Raos, double* R, int64_t N, int64_t off, int64_t ncell) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * blockDim.x)
#pragma unroll
    for (int v = 0; v < 5; ++v) R[v * N + off + c] = Raos[5 * c + v];
}
__global__ void k_fill(double* p, int64_t n, double x) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) p[c] = x;
}
__global__ void k_copy(const double* __restrict__ s, double* __restrict__ d, int64_t n) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) d[c] = s[c];
}

// Staging area of mose_get_state_noalloc, which must not allocate: a file-scope device array of
// kNoAllocDoubles doubles (32 MiB, blocks of up to 838,860 cells counting ghost layers).  Each
// device holds its own copy at its own address; host code never passes the array's name to a
// kernel or a copy (that is a host-side stand-in) but the address hipGetSymbolAddress gave for
// the device, looked up once per device at mose_finalize.  Every call on a device shares that
// device's array, so calls that overlap on one device race.
constexpr int64_t kNoAllocDoubles = (int64_t)1 << 22;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// ============================================================================== host
#include "mose_geometry.hpp"
#include "mose_metrics_adj.hpp"

namespace {

// every limiter's kernel variants (instantiated in mose_lim.hip, one unit per limiter and
// precision)
template <int LIM>
void fill_both(KTable& k, ATable& a) {
  fill_tables<LIM, 0>(k, a);
  fill_tables<LIM, 1>(k, a);
}
struct Tables {
  KTable k;
  ATable a;
  Tables() {
    fill_both<LIM_MINMOD>(k, a);
    fill_both<LIM_VANLEER>(k, a);
    fill_both<LIM_VANALBADA>(k, a);
    fill_both<LIM_FIRST>(k, a);
    fill_both<LIM_MC>(k, a);
    fill_both<LIM_SUPERBEE>(k, a);
  }
};
const Tables& tables() {
  static Tables t;
  return t;
}
const KTable& ktable() { return tables().k; }
const ATable& atable() { return tables().a; }

const double kRK[3][3] = {{1.0, 0.0, 0.0}, {1.0, 0.5, 0.0}, {1.0, 0.25, 2.0 / 3.0}};   // Lib_RK.f90:8-11

}  // namespace

struct mose_ctx {
  mose_cfg cfg{};
  std::map<int, HostBlock> hb;
  std::vector<mose_bc_entry> bc;
  bool finalized = false;
  mose_err err{};
  // layout
  Layout lay;
  int nb = 0;
  int64_t npool = 0;
  int ndir = 3;
  // device
  hipStream_t stream = nullptr;
  int dev = 0;                                // device current at mose_finalize
  double* scratch_ptr = nullptr;              // that device's copy of `scratch`
  double* U[3] = {nullptr, nullptr, nullptr};
  double* Rbuf = nullptr;
  double* stage = nullptr;
  int64_t stage_n = 0;
  BlockDev* d_blocks = nullptr;
  LineEnt* d_lines = nullptr;
  double* d_geo = nullptr;
  Tile* d_tiles = nullptr;
  GhostOp* d_ghost = nullptr;                 // ghost records in bc order (the adjoint's transpose)
  GhostOp* d_ghostp = nullptr;                // the same grouped by fill phase
  std::vector<int> gphase;                    // phase p: records [gphase[p], gphase[p+1]) of d_ghostp
  int64_t* d_garc_cell = nullptr;             // archived ghost cells (previous-fill values)
  double* d_garc = nullptr;
  int64_t narc = 0;
  double* d_rs1 = nullptr;                    // IRS second buffer (pool-sized), when irs_beta > 0
  int64_t* d_irsg = nullptr;                  // IRS residual ghosts: (ghost, source) per bc record
  int nirsg = 0;
  BcFace* d_bcfaces = nullptr;
  double* d_bcbuf = nullptr;
  ViscFace* d_vfaces = nullptr;
  double* d_vbuf = nullptr;
  double* d_facebc = nullptr;
  double* d_scal = nullptr;
  double* d_partials = nullptr;
  double* d_dtloc = nullptr;
  double* d_beta = nullptr;                  // shock-detector weights (pool-sized), when used
  double* d_pres = nullptr;                  // the detector's pressures (pool-sized)
  unsigned long long* d_dtmin = nullptr;
  int* d_err = nullptr;
  int* d_iter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph_n = nullptr;               // kStepsPerGraph steps in one graph (mose_advance)
  hipGraphExec_t gexec_n = nullptr;
  // discrete adjoint (allocated on first use)
  bool adj_ready = false;
  // steady adjoint solve (allocated on first use)
  double *d_gP = nullptr, *d_lam0 = nullptr, *d_lam1 = nullptr, *d_Dq = nullptr, *d_apart = nullptr;
  double* d_ares = nullptr;
  int64_t* d_icell = nullptr;                // interior pool cells
  int64_t nicell = 0;
  // adjoint of the fixed-dt march (allocated on first use)
  std::vector<double> hvol;                  // volume of each interior cell (interior-list order)
  double *d_volq = nullptr, *d_upart = nullptr, *d_ke = nullptr, *d_mu = nullptr;
  int64_t *d_icbeg = nullptr, *d_vfbeg = nullptr;
  int* d_vfi = nullptr;
  hipGraph_t agraph = nullptr;               // one pseudo-time iteration of the adjoint solve
  hipGraphExec_t agexec = nullptr;
  double *d_Rbar = nullptr, *d_W = nullptr, *d_BB = nullptr, *d_Pbar = nullptr, *d_Pp = nullptr;
  AdjSeg* d_segs = nullptr;
  AdjGhostRow* d_grows = nullptr;
  AdjGhostEnt* d_gents = nullptr;
  int nseg = 0, ngrows = 0, nents = 0;
  int64_t nwork = 0;
  // geometry adjoints (general metrics only; allocated on first use)
  std::vector<int64_t> gent_face;   // per ghost-fill entry: 3 * (cell holding the face) + d, or -1
  double *d_GF = nullptr, *d_GL = nullptr, *d_BG = nullptr, *d_GE = nullptr, *d_IB = nullptr;
  double *d_VW = nullptr, *d_VB = nullptr;   // viscous adjoint buffers (viscous layouts only)
  double* d_VBC = nullptr;                    // connection-face adjoints before their gather
  double* d_Tp = nullptr;                     // temperature of every pool cell (viscous adjoints)
  int64_t *d_vbt = nullptr, *d_vbrow = nullptr;
  int* d_vbent = nullptr;
  int nvbt = 0;
  Tile* d_vtiles = nullptr;                   // k_adj_visc's patches of the viscous blocks
  int nvtiles = 0;
  int64_t* d_vghost = nullptr;                // first ghost layer of the viscous blocks (k_adj_vcell)
  int64_t nvghost = 0;
  double* d_VPb = nullptr;                    // k_adj_visc's part of Pbar
  hipStream_t vstream = nullptr;              // k_adj_visc runs here, beside the convective passes
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool vjp_valid = false;                    // the last vjp left Pbar and the viscous buffers
  bool geo_valid = false;           // the last vjp wrote the geometry adjoints
  bool xbar_valid = false;
  std::vector<std::vector<double>> xbar;
  bool dt_primed = false;   // d_dtmin holds the CFL step of the current state
  int grid_stage = 0;
  int* d_tq = nullptr;                        // tile queues of the persistent stage launch
  int ncu = 0;                                // compute units of the device
  std::map<const void*, int> resident;        // resident workgroups per CU of each stage kernel
  int shape = 0;       // index into kShape of the stage kernel in use
  int tx = 16, ty = 8;
  int kchunk = 32;     // k-planes per tile
  int xcd = 1;         // bit 0: XCD-aware tile order; bit 1: one workgroup per tile (else persistent)
};

static int fail(mose_ctx* c, int code, int b = 0, int i = 0, int j = 0, int k = 0) {
  if (c && c->err.code == 0) c->err = mose_err{code, b, i, j, k};
  return code;
}
#define HIPCHK(ctx, x)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      fprintf(stderr, "[mose_hip] %s failed: %s\n", #x, hipGetErrorString(e_));           \
      return fail(ctx, MOSE_E_HIP);                                                       \
    }                                                                                     \
  } while (0)

static int launch_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 65536); }

// Ghost layers of the state U, then (schemes that use it) the shock detector on the filled state.
static void enqueue_garc_save(mose_ctx* c, const double* U) {
  if (c->narc > 0)
    hipLaunchKernelGGL(k_garc_save, dim3(launch_blocks(c->narc)), dim3(256), 0, c->stream, U, c->npool, c->d_garc_cell,
                       c->narc, c->d_garc);
}

static int enqueue_ghost(mose_ctx* c, double* U) {
  for (size_t p = 0; p + 1 < c->gphase.size(); ++p) {
    const int n = c->gphase[p + 1] - c->gphase[p];
    hipLaunchKernelGGL(k_ghost, dim3(launch_blocks(n)), dim3(256), 0, c->stream, c->d_ghostp + c->gphase[p], n, U,
                       c->npool, c->cfg.gamma, c->d_garc, c->narc);
  }
  enqueue_garc_save(c, U);
  if (c->d_beta) {
    hipLaunchKernelGGL(k_pres, dim3(launch_blocks(c->npool)), dim3(256), 0, c->stream, U, c->npool, c->cfg.gamma,
                       c->d_pres);
    SdArgs sa{c->d_pres, c->d_beta, c->d_tiles, c->d_blocks, c->npool, (int)c->lay.tiles.size(), c->tx, c->cfg.gamma};
    hipLaunchKernelGGL(k_sd, dim3(sa.ntiles), dim3(c->tx * c->ty), 0, c->stream, sa);
  }
  return 0;
}

static StageArgs stage_args(mose_ctx* c, const double* Uin, const double* Un, double* Uout, double rkc) {
  StageArgs a{};
  a.Uin = Uin; a.Un = Un; a.Uout = Uout; a.Rout = c->Rbuf;
  a.tiles = c->d_tiles; a.blocks = c->d_blocks; a.lines = c->d_lines; a.facebc = c->d_facebc;
  a.geo = c->d_geo;
  a.scal = c->d_scal; a.dtloc = c->d_dtloc; a.partials = c->d_partials; a.err = c->d_err;
  a.bcbuf = c->d_bcbuf;
  a.vbuf = c->d_vbuf;
  a.beta = c->d_beta;
  a.dtmin = c->d_dtmin; a.dtnext = c->d_dtloc; a.iter = c->d_iter;
  a.ndir = c->ndir; a.rampa = c->cfg.rampa_cfl_iter; a.cfl = c->cfg.cfl;
  a.npool = c->npool; a.ntiles = (int)c->lay.tiles.size(); a.local_dt = c->cfg.time_accurate ? 0 : 1;
  a.tq = c->d_tq; a.nq = (c->xcd & 1) ? kTileQueues : 1;
  a.rkc = rkc; a.gam = c->cfg.gamma; a.R = c->cfg.Rgas;
  return a;
}

// Persistent stage launch: as many workgroups as the device holds at once (never more than
// there are tiles); they take tiles from the queues (claim_tile) until all are done.
static int resident_per_cu(mose_ctx* c, StageFn fn) {
  auto it = c->resident.find((const void*)fn);
  if (it != c->resident.end()) return it->second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, c->tx * c->ty, 0) != hipSuccess)
    per_cu = 0;
  c->resident[(const void*)fn] = per_cu;
  return per_cu;
}

static void launch_stage(mose_ctx* c, StageFn fn, StageArgs a) {
  void* args[] = {&a};
  const int per_cu = resident_per_cu(c, fn);
  const int64_t cap = (int64_t)std::max(per_cu, 1) * std::max(c->ncu, 1);
  const int grid = (c->xcd & 2) ? c->grid_stage : (int)std::min<int64_t>(c->grid_stage, cap);
  (void)hipLaunchKernel((const void*)fn, dim3(std::max(grid, 1)), dim3(c->tx * c->ty), args, 0, c->stream);
}

// Boundary-face fluxes of the stage input (the BC_Fluxes pass), before the stage kernel.
static void enqueue_bcfaces(mose_ctx* c, const StageArgs& a) {
  const int nv = (int)c->lay.vfaces.size();
  if (nv > 0) {   // viscous flux of the connection faces
    StageArgs av = a;
    const ViscFace* vf = c->d_vfaces;
    void* vargs[] = {&av, &vf, (void*)&nv};
    const void* fn = c->cfg.precision == MOSE_PREC_FP32_FLUX ? (const void*)k_viscbc<float> : (const void*)k_viscbc<double>;
    (void)hipLaunchKernel(fn, dim3(launch_blocks(nv)), dim3(256), vargs, 0, c->stream);
  }
  const int n = (int)c->lay.bcfaces.size();
  if (n == 0) return;
  const int p = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  BcFn fn = ktable().bc[p][c->cfg.limiter][c->cfg.riemann];
  StageArgs aa = a;
  const BcFace* faces = c->d_bcfaces;
  void* args[] = {&aa, &faces, (void*)&n};
  (void)hipLaunchKernel((const void*)fn, dim3(launch_blocks(n)), dim3(256), args, 0, c->stream);
}

// stage-kernel geometry kind: 0 rectilinear, 1 general metrics, 2 rectilinear viscous
static int geo_kind(const mose_ctx* c) { return c->lay.curvi ? 1 : (c->lay.visc ? 2 : 0); }

static StageFn stage_fn(mose_ctx* c, int mode) {
  const int p = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  if (c->cfg.muscl_sd) return ktable().s[c->lay.curvi ? 1 : 0][c->cfg.limiter][c->cfg.riemann][mode];
  return ktable().f[geo_kind(c)][p][c->cfg.limiter][c->cfg.riemann][c->shape][mode];
}

static void launch_dt(mose_ctx* c, const DtArgs& da) {
  if (c->lay.curvi)
    hipLaunchKernelGGL(k_dt<1>, dim3(da.ntiles), dim3(c->tx * c->ty), 0, c->stream, da);
  else
    hipLaunchKernelGGL(k_dt<0>, dim3(da.ntiles), dim3(c->tx * c->ty), 0, c->stream, da);
}

static DtArgs dt_args(mose_ctx* c, int iter_add) {
  DtArgs da{};
  da.U = c->U[0]; da.tiles = c->d_tiles; da.blocks = c->d_blocks; da.lines = c->d_lines; da.geo = c->d_geo;
  da.dtmin = c->d_dtmin; da.dtloc = c->d_dtloc; da.iter = c->d_iter; da.npool = c->npool;
  da.ntiles = (int)c->lay.tiles.size(); da.ndir = c->ndir; da.rampa = c->cfg.rampa_cfl_iter;
  da.local_dt = c->cfg.time_accurate ? 0 : 1; da.cfl = c->cfg.cfl; da.gam = c->cfg.gamma;
  da.tx = c->tx; da.ty = c->ty; da.iter_add = iter_add; da.R = c->cfg.Rgas;
  return da;
}

// CFL time step of the current state for the next step (when the state did not come from a
// step: after mose_put_state, mose_set_time or mose_compute_dt).
static void enqueue_prime_dt(mose_ctx* c) {
  hipLaunchKernelGGL(k_dtreset, dim3(1), dim3(1), 0, c->stream, c->d_dtmin);
  DtArgs da = dt_args(c, 1);
  launch_dt(c, da);
}

static IrsArgs irs_args(mose_ctx* c) {
  IrsArgs ia{};
  ia.R = c->Rbuf; ia.tiles = c->d_tiles; ia.blocks = c->d_blocks; ia.lines = c->d_lines; ia.geo = c->d_geo;
  ia.scal = c->d_scal; ia.dtloc = c->d_dtloc; ia.npool = c->npool; ia.ntiles = (int)c->lay.tiles.size();
  ia.tx = c->tx; ia.local_dt = c->cfg.time_accurate ? 0 : 1; ia.beta = c->cfg.irs_beta;
  return ia;
}

// The stage update after the residual-only stage kernel: with IRS, the smoothing of the
// residual in Rbuf (R* = -R dt / vol, or the primitive-variable residual for
// integration-variables = prim) and the RK stage; without IRS (prim only), the RK stage from
// the flux residual.
static void enqueue_irs(mose_ctx* c, const StageArgs& a, bool last) {
  IrsArgs ia = irs_args(c);
  const dim3 grid(ia.ntiles), block(c->tx * c->ty);
  const bool prim = c->cfg.prim_vars != 0, irs = c->cfg.irs_beta > 0.0;
  const int geo = c->lay.curvi ? 1 : (c->lay.visc ? 2 : 0);
  const int tx = c->tx;
  double* cur = c->Rbuf;     // the directions alternate between Rbuf and d_rs1
  double* nxt = c->d_rs1;
  int dl = -1;
  if (irs) {
    if (prim) {
      if (geo == 1) hipLaunchKernelGGL(k_prim_resid<1>, grid, block, 0, c->stream, a, tx);
      else hipLaunchKernelGGL(k_prim_resid<0>, grid, block, 0, c->stream, a, tx);
    }
    for (int d = 0; d < c->ndir; ++d) {
      if (c->nirsg > 0)
        hipLaunchKernelGGL(k_irs_ghost, dim3(launch_blocks(c->nirsg)), dim3(256), 0, c->stream, c->d_irsg,
                           c->nirsg, cur, c->npool);
      if (d + 1 == c->ndir) break;          // the last direction is smoothed inside the update
      ia.d = d;
      ia.R = cur;
      ia.dst = nxt;
      hipLaunchKernelGGL(k_irs_jacobi, grid, block, 0, c->stream, ia);
      std::swap(cur, nxt);
    }
    dl = c->ndir - 1;
  }
  StageArgs u = a;
  u.Rout = cur;
  const double b = c->cfg.irs_beta;
#define MOSE_UPD(G, L, P) hipLaunchKernelGGL((k_irs_update<G, L, P>), grid, block, 0, c->stream, u, tx, dl, b)
  if (prim) {
    if (last) {
      if (geo == 1) MOSE_UPD(1, true, true);
      else if (geo == 2) MOSE_UPD(2, true, true);
      else MOSE_UPD(0, true, true);
    } else {
      if (geo == 1) MOSE_UPD(1, false, true);
      else MOSE_UPD(0, false, true);
    }
  } else {
    if (last) {
      if (geo == 1) MOSE_UPD(1, true, false);
      else if (geo == 2) MOSE_UPD(2, true, false);
      else MOSE_UPD(0, true, false);
    } else {
      if (geo == 1) MOSE_UPD(1, false, false);
      else MOSE_UPD(0, false, false);
    }
  }
#undef MOSE_UPD
}

// One explicit step enqueued on the context stream (no host synchronisation).  The CFL
// time step of the next step is produced by the last RK stage from the state it writes.
static int enqueue_step(mose_ctx* c) {
  const int nt = (int)c->lay.tiles.size();
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, c->stream, c->d_dtmin, c->d_iter, c->d_scal,
                     c->cfg.time_accurate);
  const int nrk = c->cfg.n_rk;
  for (int s = 1; s <= nrk; ++s) {
    double* in = (s == 1) ? c->U[0] : c->U[s - 1];
    double* out = (s == nrk) ? (nrk == 1 ? c->U[1] : c->U[0]) : c->U[s];
    enqueue_ghost(c, in);
    StageArgs a = stage_args(c, in, c->U[0], out, kRK[nrk - 1][s - 1]);
    enqueue_bcfaces(c, a);
    if (c->cfg.irs_beta > 0.0 || c->cfg.prim_vars) {
      StageArgs r = a;
      r.irs_scale = c->cfg.prim_vars ? 0 : 1;
      launch_stage(c, stage_fn(c, 2), r);        // R (prim) or R* = -R dt / vol into Rbuf
      enqueue_irs(c, a, s == nrk);
    } else {
      launch_stage(c, stage_fn(c, s == nrk ? 1 : 0), a);
    }
  }
  if (nrk == 1) {
    const int64_t n = 5 * c->npool;
    hipLaunchKernelGGL(k_copy, dim3(launch_blocks(n)), dim3(256), 0, c->stream, c->U[1], c->U[0], n);
  }
  hipLaunchKernelGGL(k_normfin, dim3(1), dim3(256), 0, c->stream, c->d_partials, nt, c->d_scal);
  return 0;
}

// A kernel launch that the runtime refused (launch configuration, resources) leaves its
// error for hipGetLastError; every enqueue sequence is followed by this check.
static int launch_check(mose_ctx* c, const char* where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fprintf(stderr, "[mose_hip] kernel launch failed in %s: %s\n", where, hipGetErrorString(e));
    return fail(c, MOSE_E_HIP);
  }
  return 0;
}

static int sync_and_check(mose_ctx* c) {
  if (launch_check(c, "enqueue")) return c->err.code;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int e[5];
  HIPCHK(c, hipMemcpy(e, c->d_err, sizeof e, hipMemcpyDeviceToHost));
  if (e[0]) {
    c->err = mose_err{e[0], e[1], e[2], e[3], e[4]};
    return e[0];
  }
  return 0;
}

// Graphs of one step and of kStepsPerGraph steps (the same kernels in the same order, so a run
// of steps gives the same bits either way; the longer graph saves host launches on small cases,
// whose steps are a few microseconds of work each).
constexpr int kStepsPerGraph = 8;
static int capture_steps(mose_ctx* c, int nsteps, hipGraph_t* graph, hipGraphExec_t* exec) {
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int s = 0; s < nsteps; ++s) enqueue_step(c);
  const hipError_t le = hipGetLastError();
  HIPCHK(c, hipStreamEndCapture(c->stream, graph));
  if (le != hipSuccess) {
    fprintf(stderr, "[mose_hip] kernel launch failed during step capture: %s\n", hipGetErrorString(le));
    (void)hipGraphDestroy(*graph);
    *graph = nullptr;
    return fail(c, MOSE_E_HIP);
  }
  HIPCHK(c, hipGraphInstantiate(exec, *graph, nullptr, nullptr, 0));
  return 0;
}

static int ensure_graph(mose_ctx* c) {
  if (c->gexec) return 0;
  if (capture_steps(c, 1, &c->graph, &c->gexec)) return c->err.code;
  return capture_steps(c, kStepsPerGraph, &c->graph_n, &c->gexec_n);
}

template <typename V>
static int upload(mose_ctx* c, V** dptr, const std::vector<V>& h) {
  const size_t n = std::max<size_t>(h.size(), 1);
  HIPCHK(c, hipMalloc((void**)dptr, n * sizeof(V)));
  if (!h.empty()) HIPCHK(c, hipMemcpy(*dptr, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
  return 0;
}

// Address of the current device's `scratch` (looked up once per device).
static int setup_scratch(mose_ctx* c) {
  int dev = 0;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(c, MOSE_E_ARG);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(c, MOSE_E_HIP);
    g_scratch_ptr[dev] = (double*)p;
  }
  c->dev = dev;
  c->scratch_ptr = g_scratch_ptr[dev];
  return 0;
}

static int ensure_stage(mose_ctx* c, int64_t n) {
  if (c->stage_n >= n) return 0;
  if (c->stage) HIPCHK(c, hipFree(c->stage));
  HIPCHK(c, hipMalloc((void**)&c->stage, n * sizeof(double)));
  c->stage_n = n;
  return 0;
}

static bool shape_compiled(const mose_ctx* c, int shape) {
  const int p = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  if (shape < 0 || shape >= NSHAPE) return false;
  if (c->cfg.muscl_sd) return shape == 0;
  return ktable().f[geo_kind(c)][p][c->cfg.limiter][c->cfg.riemann][shape][0] != nullptr;
}

// Every kernel launched with one thread per patch column takes tx * ty threads: refuse a
// shape that exceeds the launch bound a kernel was compiled with (such a launch fails).
static int tile_launch_fits(mose_ctx* c) {
  const int nt = c->tx * c->ty;
  if (nt > kMaxTileThreads || nt % 64 != 0) return fail(c, MOSE_E_ARG);
  const void* fns[] = {(const void*)stage_fn(c, 0), (const void*)stage_fn(c, 1), (const void*)stage_fn(c, 2),
                       (const void*)k_dt<0>, (const void*)k_dt<1>, (const void*)k_sd, (const void*)k_irs_jacobi,
                       (const void*)k_prim_resid<0>, (const void*)k_prim_resid<1>,
                       (const void*)k_irs_update<0, false, false>, (const void*)k_irs_update<0, true, false>,
                       (const void*)k_irs_update<1, true, true>, (const void*)k_irs_update<2, true, true>};
  for (const void* f : fns) {
    if (!f) continue;
    hipFuncAttributes at{};
    if (hipFuncGetAttributes(&at, f) != hipSuccess) continue;     // no device: nothing to check against
    if (at.maxThreadsPerBlock < nt) return fail(c, MOSE_E_ARG);
  }
  return 0;
}

// Rebuild the work list for another tile shape / k chunk / order (buffers sized to the
// new tile count); the captured step is discarded and re-captured on next use.
static int retile(mose_ctx* c, int shape, int kchunk, int xcd) {
  const int prec = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->shape = shape;
  c->tx = kShape[prec][shape][0];
  c->ty = kShape[prec][shape][1];
  if (tile_launch_fits(c)) return c->err.code;
  c->kchunk = kchunk;
  c->xcd = xcd;
  const size_t old = c->lay.tiles.size();
  build_tiles(c->lay, c->tx, c->ty, kchunk, xcd & 1);
  const size_t nt = c->lay.tiles.size();
  if (c->d_tiles) HIPCHK(c, hipFree(c->d_tiles));
  c->d_tiles = nullptr;
  if (upload(c, &c->d_tiles, c->lay.tiles)) return c->err.code;
  if (!c->d_partials || nt > old) {
    if (c->d_partials) HIPCHK(c, hipFree(c->d_partials));
    HIPCHK(c, hipMalloc((void**)&c->d_partials, 5 * std::max<size_t>(nt, 1) * sizeof(double)));
  }
  c->grid_stage = (int)nt;
  for (int m = 0; m < 3; ++m)     // occupancy queried here, never while a graph is captured
    if (StageFn fn = stage_fn(c, m)) (void)resident_per_cu(c, fn);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->gexec_n) (void)hipGraphExecDestroy(c->gexec_n);
  if (c->graph_n) (void)hipGraphDestroy(c->graph_n);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->gexec_n = nullptr;
  c->graph_n = nullptr;
  if (c->agexec) (void)hipGraphExecDestroy(c->agexec);
  if (c->agraph) (void)hipGraphDestroy(c->agraph);
  c->agexec = nullptr;
  c->agraph = nullptr;
  c->dt_primed = false;
  return 0;
}

// Discrete adjoint set-up: face work segments, the transposed ghost-fill map, buffers.
static int ensure_adjoint(mose_ctx* c) {
  // the adjoint of the shock-detector schemes (HLLE++, HLLC+, MUSCL-SD) and of the subsonic
  // inlet faces (401) is not built
  if (c->d_beta) return fail(c, MOSE_E_ARG);
  for (const BlockDev& B : c->lay.blocks)
    for (int f = 0; f < 6; ++f)
      if (B.bct[f] == F_INLET) return fail(c, MOSE_E_ARG);
  if (c->adj_ready) return 0;
  std::vector<AdjSeg> segs;
  int64_t start = 0;
  for (int b = 0; b < c->nb; ++b) {
    const BlockDev& B = c->lay.blocks[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    for (int d = 0; d < 3; ++d) {
      AdjSeg sg{};
      sg.start = start; sg.b = b; sg.d = d;
      sg.n0 = n[0] + (d == 0);
      sg.n1 = n[1] + (d == 1);
      sg.n2 = n[2] + (d == 2);
      segs.push_back(sg);
      start += (int64_t)sg.n0 * sg.n1 * sg.n2;
    }
  }
  c->nwork = start;
  c->nseg = (int)segs.size();
  // ghost-fill map per source cell: ghost = sum_w w P(src) (copy, zero gradient,
  // extrapolation 3a-3b+c and its second layer 6a-8b+3c) or the velocity mirror
  // (each entry is tagged with the face whose normal a mirror entry uses, for the geometry
  // adjoint: 3 * (cell holding the face as its low face) + direction)
  std::map<int64_t, std::vector<std::pair<AdjGhostEnt, int64_t>>> rows;
  auto add = [&](int64_t src, int64_t ghost, double w, int kind, const double* n, int64_t tag) {
    AdjGhostEnt e{};
    e.ghost = ghost; e.w = w; e.kind = kind;
    if (n) { e.n[0] = n[0]; e.n[1] = n[1]; e.n[2] = n[2]; }
    rows[src].push_back({e, tag});
  };
  for (size_t q = 0; q < c->lay.ghost.size(); ++q) {
    const GhostOp& op = c->lay.ghost[q];
    const int64_t g1 = op.b - op.s, g2 = op.b - 2 * op.s;
    switch (op.type) {
      case G_COPY: add(op.src1, g1, 1.0, 0, nullptr, -1); add(op.src2, g2, 1.0, 0, nullptr, -1); break;
      case G_ZG: add(op.b, g1, 1.0, 0, nullptr, -1); add(op.b, g2, 1.0, 0, nullptr, -1); break;
      case G_EXTRAP:
        add(op.b, g1, 3.0, 0, nullptr, -1); add(op.b + op.s, g1, -3.0, 0, nullptr, -1);
        add(op.b + 2 * op.s, g1, 1.0, 0, nullptr, -1);
        add(op.b, g2, 6.0, 0, nullptr, -1); add(op.b + op.s, g2, -8.0, 0, nullptr, -1);
        add(op.b + 2 * op.s, g2, 3.0, 0, nullptr, -1);
        break;
      default: {   // symmetry: one layer, mirrored with the normal of the boundary face
        const mose_bc_entry& e = c->bc[q];
        const int64_t fc = (e.f % 2) ? op.b : op.b - op.s;
        add(op.b, g1, 0.0, 1, op.n, c->lay.curvi ? 3 * fc + (e.f - 1) / 2 : -1);
        break;
      }
    }
  }
  std::vector<AdjGhostRow> grows;
  std::vector<AdjGhostEnt> gents;
  c->gent_face.clear();
  for (auto& kv : rows) {
    AdjGhostRow r{};
    r.src = kv.first;
    r.begin = (int)gents.size();
    for (auto& et : kv.second) {
      gents.push_back(et.first);
      c->gent_face.push_back(et.second);
    }
    r.end = (int)gents.size();
    grows.push_back(r);
  }
  c->ngrows = (int)grows.size();
  c->nents = (int)gents.size();
  if (upload(c, &c->d_segs, segs)) return c->err.code;
  if (upload(c, &c->d_grows, grows)) return c->err.code;
  if (upload(c, &c->d_gents, gents)) return c->err.code;
  const int64_t ns = std::max<int64_t>(c->lay.nbcslots, 1);
  HIPCHK(c, hipMalloc((void**)&c->d_Rbar, 5 * c->npool * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_Rbar, 0, 5 * c->npool * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_W, (size_t)30 * c->npool * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_BB, (size_t)10 * ns * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_Pbar, 5 * c->npool * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_Pp, 5 * c->npool * sizeof(double)));
  if (c->lay.visc) {
    HIPCHK(c, hipMalloc((void**)&c->d_VW, (size_t)3 * 18 * c->npool * sizeof(double)));
    const size_t nvf = c->lay.vfaces.size();
    HIPCHK(c, hipMalloc((void**)&c->d_VB, std::max<size_t>(nvf, 1) * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&c->d_VBC, std::max<size_t>(50 * nvf, 1) * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&c->d_Tp, c->npool * sizeof(double)));
    // receiving cells of the connection faces (the ten stencil cells of each, k_adj_vbc's
    // order), grouped per cell with the entries in (face, role) order
    std::vector<std::pair<int64_t, int>> te;
    te.reserve(10 * nvf);
    for (size_t q = 0; q < nvf; ++q) {
      const ViscFace& e = c->lay.vfaces[q];
      const int64_t cell[10] = {e.loc, e.gh, e.tn[0], e.tn[1], e.tn[2], e.tn[3], e.dn[0], e.dn[1], e.dn[2], e.dn[3]};
      for (int m = 0; m < 10; ++m) te.push_back({cell[m], (int)(10 * q + m)});
    }
    std::sort(te.begin(), te.end());
    std::vector<int64_t> tgt, row;
    std::vector<int> ent;
    for (size_t x = 0; x < te.size(); ++x) {
      if (x == 0 || te[x].first != te[x - 1].first) {
        tgt.push_back(te[x].first);
        row.push_back((int64_t)x);
      }
      ent.push_back(te[x].second);
    }
    row.push_back((int64_t)te.size());
    c->nvbt = (int)tgt.size();
    if (upload(c, &c->d_vbt, tgt) || upload(c, &c->d_vbrow, row) || upload(c, &c->d_vbent, ent)) return c->err.code;
    // k_adj_visc: kVX x kVY patches of every viscous block, 32-plane chunks, i fastest
    std::vector<Tile> vt;
    for (int b = 0; b < c->nb; ++b) {
      const BlockDev& B = c->lay.blocks[b];
      if (!B.visc) continue;
      for (int k0 = 1; k0 <= B.nk; k0 += 32)
        for (int j0 = 1; j0 <= B.nj; j0 += kVY)
          for (int i0 = 1; i0 <= B.ni; i0 += kVX) vt.push_back(Tile{b, i0, j0, k0, std::min(k0 + 32, B.nk + 1), 0, 0, 0});
    }
    c->nvtiles = (int)vt.size();
    if (upload(c, &c->d_vtiles, vt)) return c->err.code;
    std::vector<int64_t> gl;
    for (int b = 0; b < c->nb; ++b) {
      const BlockDev& B = c->lay.blocks[b];
      if (!B.visc) continue;
      const int n3[3] = {B.ni, B.nj, B.nk};
      const int64_t st[3] = {1, (int64_t)B.sy, B.sz};
      for (int d = 0; d < 3; ++d) {          // the two layers across axis d, the other indices interior
        const int a1 = (d + 1) % 3, a2 = (d + 2) % 3;
        for (int side = 0; side < 2; ++side)
          for (int u = 1; u <= n3[a2]; ++u)
            for (int t = 1; t <= n3[a1]; ++t)
              gl.push_back(B.off + st[d] * (side ? n3[d] + 2 : 1) + st[a1] * (t + 1) + st[a2] * (u + 1));
      }
    }
    std::sort(gl.begin(), gl.end());
    c->nvghost = (int64_t)gl.size();
    if (upload(c, &c->d_vghost, gl)) return c->err.code;
    HIPCHK(c, hipMalloc((void**)&c->d_VPb, 5 * c->npool * sizeof(double)));
    HIPCHK(c, hipStreamCreateWithFlags(&c->vstream, hipStreamNonBlocking));     // default priority
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  c->adj_ready = true;
  return 0;
}

static AdjArgs adj_args(mose_ctx* c, const double* U = nullptr) {
  AdjArgs a{};
  a.U = U ? U : c->U[0]; a.Pp = c->d_Pp; a.Rbar = c->d_Rbar; a.W = c->d_W; a.BB = c->d_BB; a.Pbar = c->d_Pbar;
  a.blocks = c->d_blocks; a.lines = c->d_lines; a.geo = c->d_geo; a.facebc = c->d_facebc;
  a.segs = c->d_segs; a.bcfaces = c->d_bcfaces; a.grows = c->d_grows; a.gents = c->d_gents;
  a.npool = c->npool; a.nwork = c->nwork; a.nslots = std::max<int64_t>(c->lay.nbcslots, 1);
  a.nseg = c->nseg; a.nbcf = (int)c->lay.bcfaces.size(); a.ngrows = c->ngrows; a.nblocks = c->nb;
  a.nents = std::max(c->nents, 1);
  a.VW = c->d_VW; a.VB = c->d_VB; a.vfaces = c->d_vfaces; a.nvf = (int)c->lay.vfaces.size();
  a.VBC = c->d_VBC; a.Tp = c->d_Tp; a.vbt = c->d_vbt; a.vbrow = c->d_vbrow; a.vbent = c->d_vbent; a.nvbt = c->nvbt;
  a.VPb = c->d_VPb;
  a.gam = c->cfg.gamma; a.R = c->cfg.Rgas;
  return a;
}

// Buffers of the geometry adjoints (general metrics only).
static int ensure_adjoint_geometry(mose_ctx* c) {
  if (c->d_GF) return 0;
  const int64_t ns = std::max<int64_t>(c->lay.nbcslots, 1);
  HIPCHK(c, hipMalloc((void**)&c->d_GF, (size_t)12 * c->npool * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_GF, 0, (size_t)12 * c->npool * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_GL, (size_t)6 * c->npool * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_BG, (size_t)2 * ns * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_BG, 0, (size_t)2 * ns * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_GE, (size_t)3 * std::max(c->nents, 1) * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_GE, 0, (size_t)3 * std::max(c->nents, 1) * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_IB, (size_t)6 * ns * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_IB, 0, (size_t)6 * ns * sizeof(double)));
  return 0;
}

// Pbar = (dR/dP)^T Rbar on the current state (the ghost layers are filled first); with
// geometry, also the adjoints of the geometric quantities (face normals and areas, lengths,
// mirror normals, inflow values).
static void enqueue_vjp(mose_ctx* c, bool geometry = false, double* U = nullptr) {
  if (!U) U = c->U[0];
  enqueue_ghost(c, U);
  hipLaunchKernelGGL(k_adj_prim, dim3(launch_blocks(c->npool)), dim3(256), 0, c->stream, U, c->d_Pp, c->npool,
                     c->cfg.gamma, c->d_Tp, c->cfg.Rgas);
  AdjArgs a = adj_args(c, U);
  if (geometry) {
    a.GF = c->d_GF; a.GL = c->d_GL; a.BG = c->d_BG; a.GE = c->d_GE; a.IB = c->d_IB;
  }
  const int lim = c->cfg.limiter, rs = c->cfg.riemann, geo = c->lay.curvi ? 1 : 0;
  void* args[] = {&a};
  // viscous interior faces on the second stream, concurrent with the convective face and
  // boundary passes (k_adj_visc keeps one workgroup per CU for its LDS; the memory-bound face
  // pass fills the rest); k_adj_cell waits for it and adds its part.  While the stream is being
  // captured into a graph (the steady adjoint iteration) the kernel runs in order on it instead.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(c->stream, &cap);
  const bool visc = c->lay.visc && c->nvtiles > 0;
  const bool fork = visc && cap == hipStreamCaptureStatusNone;
  if (fork) {
    (void)hipEventRecord(c->ev_fork, c->stream);
    (void)hipStreamWaitEvent(c->vstream, c->ev_fork, 0);
    hipLaunchKernelGGL(k_adj_visc, dim3(c->nvtiles), dim3(kVT), 0, c->vstream, a, (const Tile*)c->d_vtiles);
    (void)hipEventRecord(c->ev_join, c->vstream);
  } else if (visc) {
    hipLaunchKernelGGL(k_adj_visc, dim3(c->nvtiles), dim3(kVT), 0, c->stream, a, (const Tile*)c->d_vtiles);
  } else {
    a.VPb = nullptr;
  }
  (void)hipLaunchKernel((const void*)atable().face[lim][rs][geo], dim3(launch_blocks(c->nwork)), dim3(256), args, 0,
                        c->stream);
  if (a.nbcf > 0)
    (void)hipLaunchKernel((const void*)atable().bc[lim][rs], dim3(launch_blocks(a.nbcf)), dim3(256), args, 0, c->stream);
  if (fork) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
  (void)hipLaunchKernel((const void*)atable().cell[lim][geo], dim3(launch_blocks(c->npool)), dim3(256), args, 0,
                        c->stream);
  if (c->lay.visc) {   // viscous terms: ghost cells of the interior faces, then connection faces
    if (c->nvghost > 0)
      hipLaunchKernelGGL(k_adj_vcell, dim3(launch_blocks(c->nvghost)), dim3(256), 0, c->stream, a,
                         (const int64_t*)c->d_vghost, c->nvghost);
    if (a.nvf > 0) {
      hipLaunchKernelGGL(k_adj_vbc, dim3(launch_blocks(a.nvf)), dim3(256), 0, c->stream, a);
      hipLaunchKernelGGL(k_adj_vbc_gather, dim3(launch_blocks(a.nvbt)), dim3(256), 0, c->stream, a);
    }
  }
  if (c->ngrows > 0)
    hipLaunchKernelGGL(k_adj_ghostT, dim3(launch_blocks(c->ngrows)), dim3(256), 0, c->stream, a);
}

// One residual evaluation (what mose_residual computes) without the copy-out.
static void enqueue_residual(mose_ctx* c) {
  const int64_t n5 = 5 * c->npool;
  hipLaunchKernelGGL(k_fill, dim3(launch_blocks(n5)), dim3(256), 0, c->stream, c->Rbuf, n5, 0.0);
  enqueue_ghost(c, c->U[0]);
  StageArgs a = stage_args(c, c->U[0], c->U[0], c->U[1], 0.0);
  enqueue_bcfaces(c, a);
  launch_stage(c, stage_fn(c, 2), a);
}

// Fill phases and previous-fill archive of the ghost records.  The reference fills them one
// record at a time in bc order (Fill_Ghost_Cell, Lib_Ghost.f90:11-137), so across a block one or
// two cells thick an extrapolation (or a copy from such a donor) reads ghost cells of the
// opposite face: those an earlier record wrote in this fill, or, for a later record's, the
// value of the previous fill.  Here record q runs in a phase after every earlier record whose
// ghosts it reads, and reads a later record's ghosts from an archive saved after each fill.
static int plan_ghost_fill(mose_ctx* c) {
  const std::vector<GhostOp>& G = c->lay.ghost;
  const int n = (int)G.size();
  std::map<int64_t, int> writer;
  for (int q = 0; q < n; ++q) {
    writer[G[q].b - G[q].s] = q;
    if (G[q].type != G_SYM) writer[G[q].b - 2 * G[q].s] = q;
  }
  std::vector<GhostOp> ops(G);
  std::vector<int> phase(n, 0);
  std::vector<int64_t> cells;
  std::map<int64_t, int> slot;
  int np = n > 0 ? 1 : 0;
  for (int q = 0; q < n; ++q) {
    GhostOp& op = ops[q];
    op.a1 = op.a2 = -1;
    if (op.type != G_COPY && op.type != G_EXTRAP) continue;   // zero-gradient, symmetry: the boundary cell only
    const int64_t rc[2] = {op.type == G_COPY ? op.src1 : op.b + op.s, op.type == G_COPY ? op.src2 : op.b + 2 * op.s};
    for (int r = 0; r < 2; ++r) {
      auto it = writer.find(rc[r]);
      if (it == writer.end() || it->second == q) continue;
      if (it->second < q) {
        phase[q] = std::max(phase[q], phase[it->second] + 1);
      } else {
        auto sit = slot.find(rc[r]);
        int sl = (int)cells.size();
        if (sit == slot.end()) {
          slot[rc[r]] = sl;
          cells.push_back(rc[r]);
        } else {
          sl = sit->second;
        }
        (r == 0 ? op.a1 : op.a2) = sl;
      }
    }
    np = std::max(np, phase[q] + 1);
  }
  std::vector<GhostOp> sorted;
  sorted.reserve(n);
  c->gphase.assign(1, 0);
  for (int p = 0; p < np; ++p) {
    for (int q = 0; q < n; ++q)
      if (phase[q] == p) sorted.push_back(ops[q]);
    c->gphase.push_back((int)sorted.size());
  }
  if (upload(c, &c->d_ghostp, sorted)) return c->err.code;
  c->narc = (int64_t)cells.size();
  if (c->narc > 0) {
    if (upload(c, &c->d_garc_cell, cells)) return c->err.code;
    HIPCHK(c, hipMalloc((void**)&c->d_garc, 5 * c->narc * sizeof(double)));
    HIPCHK(c, hipMemset(c->d_garc, 0, 5 * c->narc * sizeof(double)));
  }
  return 0;
}

// ============================================================================= C ABI
extern "C" {

int mose_create(const mose_cfg* cfg, mose_ctx** out) {
  if (!cfg || !out) return MOSE_E_ARG;
  *out = nullptr;
  if (cfg->limiter < 0 || cfg->limiter > 5 || cfg->riemann < 0 || cfg->riemann > 6 || cfg->n_rk < 1 ||
      cfg->n_rk > 3 || !(cfg->gamma > 1.0) || !(cfg->Rgas > 0.0) || cfg->precision < 0 || cfg->precision > 1)
    return MOSE_E_ARG;
  if (!(cfg->irs_beta >= 0.0) || !(cfg->irs_beta < 1e300)) return MOSE_E_ARG;
  if (cfg->prim_vars != 0 && cfg->prim_vars != 1) return MOSE_E_ARG;
  // MUSCL-SD: FP64 kernels of the limited reconstructions only
  if (cfg->muscl_sd != 0 && (cfg->muscl_sd != 1 || cfg->limiter == MOSE_LIM_FIRST_ORDER ||
                             cfg->precision != MOSE_PREC_FP64))
    return MOSE_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MOSE_E_NOGPU;
  mose_ctx* c = new mose_ctx();
  c->cfg = *cfg;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MOSE_E_HIP;
  }
  *out = c;
  return 0;
}

int mose_add_block(mose_ctx* c, int32_t b, const int32_t dims[3], const double* nodes) {
  if (!c || c->finalized || b < 1 || !dims || !nodes || dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
    return fail(c, MOSE_E_ARG, b);
  HostBlock h;
  h.ni = dims[0]; h.nj = dims[1]; h.nk = dims[2];
  const size_t nn = (size_t)(h.ni + 1) * (h.nj + 1) * (h.nk + 1) * 3;
  h.nodes.assign(nodes, nodes + nn);
  c->hb[b] = std::move(h);
  return 0;
}

int mose_set_viscous(mose_ctx* c, int32_t b, const double params[5]) {
  if (!c || c->finalized || !params || b < 0) return fail(c, MOSE_E_ARG, b);
  if (!(params[0] >= 0.0) || !(params[1] > 0.0) || !(params[2] >= 0.0) || !(params[3] > 0.0) || !(params[4] > 0.0))
    return fail(c, MOSE_E_ARG, b);
  bool any = false;
  for (auto& kv : c->hb) {
    if (b != 0 && kv.first != b) continue;
    kv.second.visc = 1;
    for (int q = 0; q < 5; ++q) kv.second.vis[q] = params[q];
    any = true;
  }
  return any ? 0 : fail(c, MOSE_E_ARG, b);
}

int mose_set_bc(mose_ctx* c, const mose_bc_entry* e, int64_t n) {
  if (!c || c->finalized || (n > 0 && !e) || n < 0) return fail(c, MOSE_E_ARG);
  c->bc.assign(e, e + n);
  return 0;
}

int mose_finalize(mose_ctx* c) {
  if (!c || c->finalized) return fail(c, MOSE_E_ARG);
  c->nb = (int)c->hb.size();
  std::vector<HostBlock> blocks;
  for (int b = 1; b <= c->nb; ++b) {
    auto it = c->hb.find(b);
    if (it == c->hb.end()) return fail(c, MOSE_E_ARG, b);
    blocks.push_back(it->second);
  }
  const int prec = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  c->shape = 0;
  c->tx = kShape[prec][0][0];
  c->ty = kShape[prec][0][1];
  GeoError ge;
  const double cp = c->cfg.gamma * c->cfg.Rgas / (c->cfg.gamma - 1.0);
  if (!build_layout(blocks, c->bc, cp, c->lay, ge)) return fail(c, ge.code, ge.b, ge.i, ge.j, ge.k);
  if (5 * c->lay.npool >= (int64_t)1 << 31) return fail(c, MOSE_E_ARG);   // 32-bit element offsets
  c->npool = c->lay.npool;
  c->ndir = c->lay.ndir;
  if (setup_scratch(c)) return c->err.code;
  // device buffers
  for (int s = 0; s < 3; ++s) {
    HIPCHK(c, hipMalloc((void**)&c->U[s], 5 * c->npool * sizeof(double)));
    HIPCHK(c, hipMemset(c->U[s], 0, 5 * c->npool * sizeof(double)));
  }
  HIPCHK(c, hipMalloc((void**)&c->Rbuf, 5 * c->npool * sizeof(double)));
  if (upload(c, &c->d_blocks, c->lay.blocks)) return c->err.code;
  if (upload(c, &c->d_lines, c->lay.lines)) return c->err.code;
  if (c->lay.curvi && upload(c, &c->d_geo, c->lay.geo)) return c->err.code;
  std::vector<double>().swap(c->lay.geo);   // host copy no longer needed
  if (plan_ghost_fill(c)) return c->err.code;
  if (upload(c, &c->d_facebc, c->lay.facebc)) return c->err.code;
  if (upload(c, &c->d_bcfaces, c->lay.bcfaces)) return c->err.code;
  if (upload(c, &c->d_vfaces, c->lay.vfaces)) return c->err.code;
  HIPCHK(c, hipMalloc((void**)&c->d_vbuf, 5 * std::max<int64_t>(c->lay.nvslots, 1) * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_vbuf, 0, 5 * std::max<int64_t>(c->lay.nvslots, 1) * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_bcbuf, 5 * std::max<int64_t>(c->lay.nbcslots, 1) * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_bcbuf, 0, 5 * std::max<int64_t>(c->lay.nbcslots, 1) * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_scal, S_COUNT * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_scal, 0, S_COUNT * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_dtmin, sizeof(unsigned long long)));
  HIPCHK(c, hipMalloc((void**)&c->d_tq, (kTileQueues + 1) * sizeof(int)));
  HIPCHK(c, hipMemset(c->d_tq, 0, (kTileQueues + 1) * sizeof(int)));
  {
    hipDeviceProp_t prop{};
    HIPCHK(c, hipGetDeviceProperties(&prop, c->dev));
    c->ncu = prop.multiProcessorCount;
  }
  HIPCHK(c, hipMalloc((void**)&c->d_err, 8 * sizeof(int)));
  HIPCHK(c, hipMemset(c->d_err, 0, 8 * sizeof(int)));
  HIPCHK(c, hipMalloc((void**)&c->d_iter, sizeof(int)));
  HIPCHK(c, hipMemset(c->d_iter, 0, sizeof(int)));
  if (!c->cfg.time_accurate) {
    HIPCHK(c, hipMalloc((void**)&c->d_dtloc, c->npool * sizeof(double)));
    HIPCHK(c, hipMemset(c->d_dtloc, 0, c->npool * sizeof(double)));
  }
  if (c->cfg.irs_beta > 0.0) {
    HIPCHK(c, hipMalloc((void**)&c->d_rs1, 5 * c->npool * sizeof(double)));
    HIPCHK(c, hipMemset(c->d_rs1, 0, 5 * c->npool * sizeof(double)));
    // residual ghosts (Ghost_Residual_Extrapolation, Lib_IRS.f90:154-189): the first ghost
    // layer of every bc record from the donor cell (101) or the boundary cell (all others)
    std::vector<int64_t> pairs;
    for (size_t q = 0; q < c->bc.size(); ++q) {
      const GhostOp& op = c->lay.ghost[q];
      pairs.push_back(op.b - op.s);
      pairs.push_back(c->bc[q].type == 101 ? op.src1 : op.b);
    }
    c->nirsg = (int)c->bc.size();
    if (c->nirsg > 0 && upload(c, &c->d_irsg, pairs)) return c->err.code;
  }
  if (c->cfg.muscl_sd || rs_takes_beta(c->cfg.riemann)) {
    if (c->cfg.muscl_sd && c->lay.visc) return fail(c, MOSE_E_ARG);   // no viscous MUSCL-SD kernels
    HIPCHK(c, hipMalloc((void**)&c->d_beta, c->npool * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&c->d_pres, c->npool * sizeof(double)));
    hipLaunchKernelGGL(k_fill, dim3(launch_blocks(c->npool)), dim3(256), 0, c->stream, c->d_beta, c->npool, 1.0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  // tiling: default shape 0, 32-plane chunks, XCD-aware order, persistent launch;
  // MOSE_TILE_SHAPE, MOSE_KCHUNK and MOSE_TILE_XCD (bit 0 order, bit 1 one workgroup per
  // tile) override (mose_autotune picks by measurement)
  int shape = 0, kc = 32, xcd = 1;
  if (const char* e = getenv("MOSE_TILE_SHAPE")) shape = shape_compiled(c, atoi(e)) ? atoi(e) : 0;
  if (const char* e = getenv("MOSE_KCHUNK")) kc = std::max(1, atoi(e));
  if (const char* e = getenv("MOSE_TILE_XCD")) xcd = atoi(e) & 3;
  if (retile(c, shape, kc, xcd)) return c->err.code;
  c->finalized = true;
  return 0;
}

int mose_put_state(mose_ctx* c, int32_t b, const double* P) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !P) return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  HIPCHK(c, hipMemcpyAsync(c->stage, P, 5 * ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_put, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->stage, c->U[0], c->npool,
                     B.off, ncell, c->cfg.gamma);
  enqueue_garc_save(c, c->U[0]);      // the uploaded ghost values are the "previous fill" of the next
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dt_primed = false;
  return 0;
}

int mose_get_state(mose_ctx* c, int32_t b, double* P) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !P) return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  hipLaunchKernelGGL(k_get, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->U[0], c->npool, B.off, ncell,
                     c->cfg.gamma, c->stage);
  HIPCHK(c, hipMemcpyAsync(P, c->stage, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mose_get_state_noalloc(mose_ctx* c, int32_t b, double* P) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !P) return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (5 * ncell > kNoAllocDoubles) return fail(c, MOSE_E_SIZE, b);
  int dev = -1;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev != c->dev || !c->scratch_ptr) return fail(c, MOSE_E_ARG, b);   // not the context's device
  hipLaunchKernelGGL(k_get, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->U[0], c->npool, B.off, ncell,
                     c->cfg.gamma, c->scratch_ptr);
  HIPCHK(c, hipMemcpyAsync(P, c->scratch_ptr, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mose_compute_dt(mose_ctx* c, int32_t iter, double* dtglobal) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  int cur = 0;
  HIPCHK(c, hipMemcpy(&cur, c->d_iter, sizeof cur, hipMemcpyDeviceToHost));
  hipLaunchKernelGGL(k_dtreset, dim3(1), dim3(1), 0, c->stream, c->d_dtmin);
  DtArgs da = dt_args(c, iter - cur);
  launch_dt(c, da);
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->d_dtmin, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->dt_primed = false;
  double dt;
  memcpy(&dt, &bits, sizeof dt);
  if (dtglobal) *dtglobal = dt;
  return 0;
}

int mose_advance(mose_ctx* c, int32_t nsteps, double* time_out, double* norms) {
  if (!c || !c->finalized || nsteps < 0) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (ensure_graph(c)) return c->err.code;
  if (!c->dt_primed && nsteps > 0) {
    enqueue_prime_dt(c);
    c->dt_primed = true;
  }
  int s = 0;
  for (; s + kStepsPerGraph <= nsteps; s += kStepsPerGraph) HIPCHK(c, hipGraphLaunch(c->gexec_n, c->stream));
  for (; s < nsteps; ++s) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
  int rc = sync_and_check(c);
  double sc[S_COUNT];
  HIPCHK(c, hipMemcpy(sc, c->d_scal, sizeof sc, hipMemcpyDeviceToHost));
  if (time_out) *time_out = sc[S_TIME];
  if (norms) memcpy(norms, sc + S_NORM, 5 * sizeof(double));
  return rc;
}

int mose_step(mose_ctx* c, double* dt_out, double* time_out, double* norms) {
  int rc = mose_advance(c, 1, time_out, norms);
  if (dt_out) {
    double dt = 0.0;
    if (c && c->d_scal) (void)hipMemcpy(&dt, c->d_scal + S_DT, sizeof dt, hipMemcpyDeviceToHost);
    *dt_out = dt;
  }
  return rc;
}

int mose_residual(mose_ctx* c, int32_t b, double* R) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !R) return fail(c, MOSE_E_ARG, b);
  enqueue_residual(c);
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  hipLaunchKernelGGL(k_get_raw, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->Rbuf, c->npool, B.off,
                     ncell, c->stage);
  HIPCHK(c, hipMemcpyAsync(R, c->stage, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return sync_and_check(c);
}

int mose_get_time(mose_ctx* c, double* time, int32_t* iter) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  double sc[S_COUNT];
  HIPCHK(c, hipMemcpy(sc, c->d_scal, sizeof sc, hipMemcpyDeviceToHost));
  int it = 0;
  HIPCHK(c, hipMemcpy(&it, c->d_iter, sizeof it, hipMemcpyDeviceToHost));
  if (time) *time = sc[S_TIME];
  if (iter) *iter = it;
  return 0;
}

int mose_set_time(mose_ctx* c, double time, int32_t iter) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  HIPCHK(c, hipMemcpy(c->d_scal + S_TIME, &time, sizeof time, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_iter, &iter, sizeof iter, hipMemcpyHostToDevice));
  c->dt_primed = false;
  return 0;
}

int mose_time_kernels(mose_ctx* c, int32_t nrep, double ms[4]) {
  if (!c || !c->finalized || nrep < 1 || !ms) return fail(c, MOSE_E_ARG);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const int nt = (int)c->lay.tiles.size();
  DtArgs da = dt_args(c, 1);
  for (int which = 0; which < 4; ++which) {
    for (int pass = 0; pass < 2; ++pass) {      // pass 0: warm-up launch, pass 1: timed
      const int reps = pass == 0 ? 1 : nrep;
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r) {
        if (which == 0) launch_dt(c, da);
        else if (which == 1) enqueue_ghost(c, c->U[0]);
        else {
          StageArgs sa = stage_args(c, c->U[0], c->U[0], c->U[1], 0.25);
          if (which == 2) enqueue_bcfaces(c, sa);
          launch_stage(c, stage_fn(c, which == 2 ? 0 : 1), sa);
        }
      }
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
      if (pass == 1) ms[which] = (double)t / nrep;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->dt_primed = false;
  return sync_and_check(c);
}

int mose_autotune(mose_ctx* c, int32_t nrep, int32_t choice[3], double* best_ms) {
  if (!c || !c->finalized || nrep < 1) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const int kcs[4] = {8, 16, 32, 64};
  double best = 1e300;
  int bs = c->shape, bk = c->kchunk, bx = c->xcd;
  for (int shape = 0; shape < NSHAPE; ++shape) {
    if (!shape_compiled(c, shape)) continue;
    for (int kc : kcs)
      for (int xcd = 0; xcd < 4; ++xcd) {
        if (retile(c, shape, kc, xcd)) return c->err.code;
        double t = 0.0;
        for (int pass = 0; pass < 2; ++pass) {   // pass 0 warms up
          const int reps = pass == 0 ? 1 : nrep;
          HIPCHK(c, hipEventRecord(e0, c->stream));
          for (int r = 0; r < reps; ++r)
            for (int m = 0; m < 3; ++m) {     // the stage sequence of one RK3 step
              StageArgs sa = stage_args(c, c->U[0], c->U[0], c->U[1], 0.25);
              launch_stage(c, stage_fn(c, m == 2 ? 1 : 0), sa);
            }
          HIPCHK(c, hipEventRecord(e1, c->stream));
          HIPCHK(c, hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
          t = (double)ms / reps;
          if (hipGetLastError() != hipSuccess) {   // a configuration the runtime refused
            t = 1e300;
            break;
          }
        }
        if (t < best) {
          best = t;
          bs = shape; bk = kc; bx = xcd;
        }
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (retile(c, bs, bk, bx)) return c->err.code;
  if (choice) {
    choice[0] = c->tx * 1000 + c->ty;
    choice[1] = bk;
    choice[2] = bx;
  }
  if (best_ms) *best_ms = best;
  return sync_and_check(c);
}

int mose_get_tiling(const mose_ctx* c, int32_t choice[3]) {
  if (!c || !c->finalized || !choice) return MOSE_E_ARG;
  choice[0] = c->tx * 1000 + c->ty;
  choice[1] = c->kchunk;
  choice[2] = c->xcd;
  return 0;
}

int mose_put_residual_bar(mose_ctx* c, int32_t b, const double* Rbar) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !Rbar) return fail(c, MOSE_E_ARG, b);
  if (ensure_adjoint(c)) return c->err.code;
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  HIPCHK(c, hipMemcpyAsync(c->stage, Rbar, 5 * ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_put_raw, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->stage, c->d_Rbar, c->npool,
                     B.off, ncell);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mose_residual_vjp(mose_ctx* c) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (ensure_adjoint(c)) return c->err.code;
  c->geo_valid = false;
  c->xbar_valid = false;
  enqueue_vjp(c);
  c->vjp_valid = true;
  return sync_and_check(c);
}

int mose_residual_vjp_geometry(mose_ctx* c) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (!c->lay.curvi) return fail(c, MOSE_E_ARG);
  if (ensure_adjoint(c)) return c->err.code;
  if (ensure_adjoint_geometry(c)) return c->err.code;
  enqueue_vjp(c, true);
  c->vjp_valid = true;
  c->xbar_valid = false;
  const int rc = sync_and_check(c);
  c->geo_valid = rc == 0;
  return rc;
}

// Node adjoints of every block from the geometry adjoints of the last
// mose_residual_vjp_geometry: device arrays to the host, then the reverse of the metrics.
static int node_adjoints(mose_ctx* c) {
  if (c->xbar_valid) return 0;
  const int64_t np = c->npool;
  std::vector<double> GF((size_t)12 * np), GL((size_t)6 * np), RL((size_t)3 * np);
  std::vector<double> GE((size_t)3 * std::max(c->nents, 1));
  HIPCHK(c, hipMemcpy(GF.data(), c->d_GF, GF.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(GL.data(), c->d_GL, GL.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(RL.data(), c->d_geo + (size_t)GRLO * np, RL.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(GE.data(), c->d_GE, GE.size() * sizeof(double), hipMemcpyDeviceToHost));
  // mirror normals of the symmetry ghost fill are copies of boundary-face normals
  for (int e = 0; e < c->nents; ++e) {
    const int64_t t = c->gent_face[e];
    if (t < 0) continue;
    const int64_t fc = t / 3;
    const int d = (int)(t % 3);
    for (int q = 0; q < 3; ++q) GF[(size_t)(GN + 4 * d + q) * np + fc] += GE[(size_t)q * c->nents + e];
  }
  std::vector<double> RLb(GL.begin(), GL.begin() + 3 * np), DLb(GL.begin() + 3 * np, GL.end());
  std::vector<HostBlock> blocks;
  for (int b = 1; b <= c->nb; ++b) blocks.push_back(c->hb.at(b));
  geo::curvilinear_vjp(blocks, c->bc, c->lay, RL, GF, RLb, std::move(DLb), c->xbar);
  c->xbar_valid = true;
  return 0;
}

int mose_get_node_bar(mose_ctx* c, int32_t b, double* Xbar) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !Xbar || !c->geo_valid) return fail(c, MOSE_E_ARG, b);
  if (node_adjoints(c)) return c->err.code;
  const std::vector<double>& x = c->xbar[b - 1];
  std::memcpy(Xbar, x.data(), x.size() * sizeof(double));
  return 0;
}

int mose_get_inflow_bar(mose_ctx* c, int32_t b, int32_t f, double out[5]) {
  if (!c || !c->finalized || b < 1 || b > c->nb || f < 1 || f > 6 || !out || !c->geo_valid)
    return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  if (B.bct[f - 1] != F_INFLOW) return fail(c, MOSE_E_ARG, b);
  const int64_t ns = std::max<int64_t>(c->lay.nbcslots, 1);
  const int d = (f - 1) / 2;
  const int n[3] = {B.ni, B.nj, B.nk};
  const int64_t cnt = (int64_t)n[(d + 1) % 3] * n[(d + 2) % 3];
  std::vector<double> ib((size_t)6 * cnt);
  for (int r = 0; r < 6; ++r)
    HIPCHK(c, hipMemcpy(ib.data() + (size_t)r * cnt, c->d_IB + (size_t)r * ns + B.bcs[f - 1], cnt * sizeof(double),
                        hipMemcpyDeviceToHost));
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < 6; ++r)
    for (int64_t q = 0; q < cnt; ++q) s[r] += ib[(size_t)r * cnt + q];
  out[0] = s[0];
  out[1] = s[1];
  out[2] = s[2];
  const double* v = c->lay.facebc.data() + 8 * B.bcd[f - 1];
  if ((int)v[7] != 0) {
    // angle adjoints accumulated by the kernel; a 'normal' angle has none (it follows the
    // face normal, whose adjoint goes to the nodes)
    out[3] = s[3];
    out[4] = s[4];
    return 0;
  }
  // direction (cos a cos b, sin a cos b, sin b) of the record's angles
  const double ca = v[3], sa = v[4];   // cos a cos b, sin a cos b
  const double cb = std::sqrt(ca * ca + sa * sa), sb = v[5];
  const double cosa = cb > 0.0 ? ca / cb : 1.0, sina = cb > 0.0 ? sa / cb : 0.0;
  out[3] = s[3] * (-sina * cb) + s[4] * (cosa * cb);                        // d/d alpha
  out[4] = s[3] * (-cosa * sb) + s[4] * (-sina * sb) + s[5] * cb;            // d/d beta
  return 0;
}

int mose_get_viscous_bar(mose_ctx* c, int32_t b, double* mu_ref_bar) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !mu_ref_bar || !c->vjp_valid) return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  *mu_ref_bar = 0.0;
  if (!B.visc) return 0;
  const int64_t np = c->npool, ncell = B.sz * (B.nk + 4);
  std::vector<double> row((size_t)ncell);
  const int n3[3] = {B.ni, B.nj, B.nk};
  double sum = 0.0;
  for (int d = 0; d < 3; ++d) {
    HIPCHK(c, hipMemcpy(row.data(), c->d_VW + ((size_t)d * 18 + 17) * np + B.off, ncell * sizeof(double),
                        hipMemcpyDeviceToHost));
    const int tb = d == 0 ? 1 : 0, tc = d == 2 ? 1 : 2;
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          const int q[3] = {i, j, k};
          if (q[d] > n3[d] - 1 || q[tb] > n3[tb] || q[tc] > n3[tc]) continue;
          sum += row[(size_t)(i + 1) + (size_t)B.sy * (j + 1) + (size_t)B.sz * (k + 1)];
        }
  }
  const size_t nvf = c->lay.vfaces.size();
  if (nvf > 0) {
    std::vector<double> vb(nvf);
    HIPCHK(c, hipMemcpy(vb.data(), c->d_VB, nvf * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < nvf; ++q)
      if (c->lay.vfaces[q].b == b - 1) sum += vb[q];
  }
  *mu_ref_bar = sum;
  return 0;
}

int mose_get_state_bar(mose_ctx* c, int32_t b, double* Pbar) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !Pbar) return fail(c, MOSE_E_ARG, b);
  if (ensure_adjoint(c)) return c->err.code;
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  hipLaunchKernelGGL(k_get_interior, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->d_Pbar, c->npool,
                     B.off, B.ni, B.nj, B.nk, c->stage);
  HIPCHK(c, hipMemcpyAsync(Pbar, c->stage, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------------------------------------------- steady adjoint
constexpr int kAdjParts = 512;   // workgroups of k_adjsolve_rk (fixed: the norm's sum order)

static int ensure_adjsolve(mose_ctx* c) {
  if (c->d_icell) return 0;
  std::vector<int64_t> cells;
  for (const BlockDev& B : c->lay.blocks)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) cells.push_back(B.off + (i + 1) + (int64_t)B.sy * (j + 1) + B.sz * (k + 1));
  c->nicell = (int64_t)cells.size();
  if (upload(c, &c->d_icell, cells)) return c->err.code;
  const size_t n5 = 5 * (size_t)c->npool;
  for (double** p : {&c->d_gP, &c->d_lam0, &c->d_lam1}) {
    HIPCHK(c, hipMalloc((void**)p, n5 * sizeof(double)));
    HIPCHK(c, hipMemset(*p, 0, n5 * sizeof(double)));
  }
  HIPCHK(c, hipMalloc((void**)&c->d_Dq, std::max<int64_t>(c->nicell, 1) * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_apart, kAdjParts * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_ares, sizeof(double)));
  return 0;
}

int mose_put_objective_bar(mose_ctx* c, int32_t b, const double* gP) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !gP) return fail(c, MOSE_E_ARG, b);
  if (ensure_adjoint(c) || ensure_adjsolve(c)) return c->err.code;
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  HIPCHK(c, hipMemcpyAsync(c->stage, gP, 5 * ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_put_raw, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->stage, c->d_gP, c->npool,
                     B.off, ncell);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// Volume of every interior cell, in the order of the interior cell list (host copy, once).
static int interior_volumes(mose_ctx* c) {
  if (!c->hvol.empty()) return 0;
  const int64_t np = c->npool;
  std::vector<double> vol;
  if (c->lay.curvi) {
    vol.resize((size_t)np);
    HIPCHK(c, hipMemcpy(vol.data(), c->d_geo + (size_t)GVOL * np, np * sizeof(double), hipMemcpyDeviceToHost));
  }
  c->hvol.reserve((size_t)c->nicell);
  for (const BlockDev& B : c->lay.blocks) {
    const LineEnt* Lx = c->lay.lines.data() + B.line[0];
    const LineEnt* Ly = c->lay.lines.data() + B.line[1];
    const LineEnt* Lz = c->lay.lines.data() + B.line[2];
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          const int64_t cc = B.off + (i + 1) + (int64_t)B.sy * (j + 1) + B.sz * (k + 1);
          c->hvol.push_back(c->lay.curvi ? vol[(size_t)cc] : Lx[i + 1].delta * Ly[j + 1].delta * Lz[k + 1].delta);
        }
  }
  return 0;
}

// D = dt / vol of every interior cell at the current state (local dt, or the global step of
// a time-accurate context), without the CFL ramp.
static int adjsolve_dt(mose_ctx* c) {
  hipLaunchKernelGGL(k_dtreset, dim3(1), dim3(1), 0, c->stream, c->d_dtmin);
  launch_dt(c, dt_args(c, c->cfg.rampa_cfl_iter + 1));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int64_t np = c->npool;
  std::vector<double> dt;
  double dtg = 0.0;
  if (c->cfg.time_accurate) {
    unsigned long long bits = 0;
    HIPCHK(c, hipMemcpy(&bits, c->d_dtmin, sizeof bits, hipMemcpyDeviceToHost));
    memcpy(&dtg, &bits, sizeof dtg);
  } else {
    dt.resize((size_t)np);
    HIPCHK(c, hipMemcpy(dt.data(), c->d_dtloc, np * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (interior_volumes(c)) return c->err.code;
  std::vector<double> D((size_t)c->nicell);
  std::vector<int64_t> cells((size_t)c->nicell);
  HIPCHK(c, hipMemcpy(cells.data(), c->d_icell, cells.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < D.size(); ++q) D[q] = (c->cfg.time_accurate ? dtg : dt[(size_t)cells[q]]) / c->hvol[q];
  HIPCHK(c, hipMemcpy(c->d_Dq, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// One pseudo-time iteration: n_rk stages, each a residual VJP of D lambda then the stage
// update (k_adjsolve_rk); the first stage also forms |Q|^2 of the iteration's start.
static void enqueue_adjsolve_iter(mose_ctx* c) {
  const int nrk = c->cfg.n_rk;
  const double cs3[3] = {1.0, 0.25, 2.0 / 3.0}, cs2[2] = {1.0, 0.5}, cs1[1] = {1.0};
  const double* cs = nrk == 3 ? cs3 : (nrk == 2 ? cs2 : cs1);
  const int grid = (int)std::min<int64_t>(kAdjParts, std::max<int64_t>((c->nicell + 255) / 256, 1));
  for (int s = 0; s < nrk; ++s) {
    enqueue_vjp(c);
    const double* lin = s == 0 ? c->d_lam0 : c->d_lam1;
    double* lout = s == nrk - 1 ? c->d_lam0 : c->d_lam1;
    hipLaunchKernelGGL(k_adjsolve_rk, dim3(grid), dim3(256), 0, c->stream, c->d_icell, c->d_Dq, c->nicell, c->npool,
                       c->U[0], c->d_gP, c->d_Pbar, c->d_lam0, lin, lout, c->d_Rbar, 1.0 - cs[s], cs[s],
                       c->cfg.gamma, s == 0 ? c->d_apart : nullptr);
    if (s == 0) hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c->stream, c->d_apart, grid, c->d_ares);
  }
}

int mose_adjoint_solve(mose_ctx* c, int32_t max_iter, double rtol, int32_t check_every, int32_t* iters,
                       double res[2]) {
  if (!c || !c->finalized || max_iter < 1 || check_every < 1 || !(rtol >= 0.0)) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (ensure_adjoint(c) || ensure_adjsolve(c)) return c->err.code;
  if (adjsolve_dt(c)) return c->err.code;
  const int64_t n5 = 5 * c->npool;
  for (double* p : {c->d_lam0, c->d_lam1, c->d_Rbar})
    hipLaunchKernelGGL(k_fill, dim3(launch_blocks(n5)), dim3(256), 0, c->stream, p, n5, 0.0);
  if (!c->agexec) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_adjsolve_iter(c);
    HIPCHK(c, hipStreamEndCapture(c->stream, &c->agraph));
    HIPCHK(c, hipGraphInstantiate(&c->agexec, c->agraph, nullptr, nullptr, 0));
  }
  c->vjp_valid = c->geo_valid = c->xbar_valid = false;
  double r0 = 0.0, r = 0.0;
  int it = 0;
  while (it < max_iter) {
    const int n = it == 0 ? 1 : std::min(check_every, max_iter - it);
    for (int q = 0; q < n; ++q) HIPCHK(c, hipGraphLaunch(c->agexec, c->stream));
    it += n;
    if (const int rc = sync_and_check(c)) return rc;
    double a = 0.0;
    HIPCHK(c, hipMemcpy(&a, c->d_ares, sizeof a, hipMemcpyDeviceToHost));
    r = std::sqrt(a);
    if (it == 1) r0 = r;
    if (!(r == r)) return fail(c, MOSE_E_STATE);
    if (r <= rtol * r0) break;
  }
  if (iters) *iters = it;
  if (res) {
    res[0] = r0;
    res[1] = r;
  }
  return 0;
}

// ---------------------------------------------------------------------------- unsteady adjoint
constexpr int kBlockParts = 256;   // workgroups per block of the per-block reductions

static int ensure_unsteady(mose_ctx* c) {
  if (ensure_adjsolve(c) || interior_volumes(c)) return c->err.code;
  if (c->d_volq) return 0;
  if (upload(c, &c->d_volq, c->hvol)) return c->err.code;
  std::vector<int64_t> beg(1, 0);
  for (const BlockDev& B : c->lay.blocks) beg.push_back(beg.back() + (int64_t)B.ni * B.nj * B.nk);
  std::vector<int> vfi;
  std::vector<int64_t> vbeg(1, 0);
  for (int b = 0; b < c->nb; ++b) {
    for (size_t q = 0; q < c->lay.vfaces.size(); ++q)
      if (c->lay.vfaces[q].b == b) vfi.push_back((int)q);
    vbeg.push_back((int64_t)vfi.size());
  }
  if (upload(c, &c->d_icbeg, beg) || upload(c, &c->d_vfi, vfi) || upload(c, &c->d_vfbeg, vbeg)) return c->err.code;
  HIPCHK(c, hipMalloc((void**)&c->d_upart, (size_t)kBlockParts * c->nb * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_ke, c->nb * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_mu, c->nb * sizeof(double)));
  return 0;
}

// One step of the fixed-dt march from Un into Unext (stages 1..upto only when upto > 0; stage
// temporaries U[1], U[2]); the stage kernels take dt from the scalar slot.
static void enqueue_fixed_step(mose_ctx* c, double* Un, double* Unext, int upto = 0) {
  const int nrk = c->cfg.n_rk;
  const int last = upto > 0 ? upto : nrk;
  for (int s = 1; s <= last; ++s) {
    double* in = s == 1 ? Un : c->U[s - 1];
    double* out = s == nrk ? Unext : c->U[s];
    enqueue_ghost(c, in);
    StageArgs a = stage_args(c, in, Un, out, kRK[nrk - 1][s - 1]);
    enqueue_bcfaces(c, a);
    launch_stage(c, stage_fn(c, 0), a);
  }
}

// The reverse of one step from its start state Un: lamS holds dJ/dU of the step's result on
// entry and dJ/dU_n on return (lamN is the accumulator, zero on entry).
static void enqueue_reverse_step(mose_ctx* c, double* Un) {
  const int nrk = c->cfg.n_rk;
  if (nrk > 1) enqueue_fixed_step(c, Un, nullptr, nrk - 1);     // stage inputs U_1 .. U_{nrk-1}
  const int g = launch_blocks(c->nicell);
  for (int s = nrk; s >= 1; --s) {
    double* in = s == 1 ? Un : c->U[s - 1];
    const double cs = kRK[nrk - 1][s - 1];
    hipLaunchKernelGGL(k_unst_rbar, dim3(g), dim3(256), 0, c->stream, c->d_icell, c->d_Dq, c->nicell, c->npool,
                       c->d_lam0, c->d_lam1, c->d_Rbar, cs);
    enqueue_vjp(c, false, in);
    if (c->lay.visc) {
      hipLaunchKernelGGL(k_mubar, dim3(kBlockParts, c->nb), dim3(256), 0, c->stream, c->d_blocks, c->d_VW, c->d_VB,
                         c->d_vfi, c->d_vfbeg, c->npool, c->d_upart);
      hipLaunchKernelGGL(k_block_sums, dim3((c->nb + 63) / 64), dim3(64), 0, c->stream, c->d_upart, kBlockParts, c->nb,
                         c->d_mu, 1);
    }
    hipLaunchKernelGGL(k_unst_back, dim3(g), dim3(256), 0, c->stream, c->d_icell, c->nicell, c->npool, in, c->d_Pbar,
                       c->d_lam0, c->d_lam1, cs, c->cfg.gamma, s == 1 ? 1 : 0);
  }
  std::swap(c->d_lam0, c->d_lam1);       // dJ/dU_n is the next (earlier) step's result adjoint
  const int64_t n5 = 5 * c->npool;
  hipLaunchKernelGGL(k_fill, dim3(launch_blocks(n5)), dim3(256), 0, c->stream, c->d_lam1, n5, 0.0);
}

int mose_unsteady_adjoint_ke(mose_ctx* c, int32_t nsteps, double dt, int32_t stride, double* ke, double* mu_bar) {
  if (!c || !c->finalized || nsteps < 1 || !(dt > 0.0) || !(dt < 1e300) || !ke) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (!c->cfg.time_accurate || c->cfg.irs_beta > 0.0 || c->cfg.prim_vars) return fail(c, MOSE_E_ARG);
  const bool reverse = mu_bar != nullptr;
  if (reverse && ensure_adjoint(c)) return c->err.code;
  if (ensure_unsteady(c)) return c->err.code;
  {   // D = dt / vol
    std::vector<double> D(c->hvol.size());
    for (size_t q = 0; q < D.size(); ++q) D[q] = dt / c->hvol[q];
    HIPCHK(c, hipMemcpy(c->d_Dq, D.data(), D.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (stride < 1) stride = (int32_t)std::ceil(std::sqrt((double)nsteps));
  stride = std::min(stride, nsteps);
  const int nck = (nsteps + stride - 1) / stride;
  const int nseg = std::max(stride, 2);
  const int nbuf = (reverse ? nck : 1) + nseg;
  const size_t words = 5 * (size_t)c->npool;
  double* pool = nullptr;
  if (hipMalloc((void**)&pool, (size_t)nbuf * words * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    fprintf(stderr, "[mose_hip] march adjoint: %d device states of %.2f GB do not fit; use a larger stride\n", nbuf,
            words * 8e-9);
    return fail(c, MOSE_E_HIP);
  }
  auto ck = [&](int k) { return pool + (size_t)k * words; };
  auto seg = [&](int j) { return pool + (size_t)((reverse ? nck : 1) + j) * words; };
  double dt_old = 0.0;
  HIPCHK(c, hipMemcpy(&dt_old, c->d_scal + S_DT, sizeof dt_old, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(c->d_scal + S_DT, &dt, sizeof dt, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_copy, dim3(launch_blocks((int64_t)words)), dim3(256), 0, c->stream, c->U[0], ck(0),
                     (int64_t)words);
  // forward: checkpoints at the start of every segment
  double* cur = ck(0);
  for (int m = 0; m < nsteps; ++m) {
    double* nxt = (reverse && (m + 1) % stride == 0 && m + 1 < nsteps) ? ck((m + 1) / stride)
                                                                       : (cur == seg(0) ? seg(1) : seg(0));
    enqueue_fixed_step(c, cur, nxt);
    cur = nxt;
  }
  hipLaunchKernelGGL(k_ke, dim3(kBlockParts, c->nb), dim3(256), 0, c->stream, c->d_icell, c->d_volq, c->d_icbeg,
                     c->npool, cur, c->d_lam0, c->d_lam1, c->d_upart, reverse ? 1 : 0);
  hipLaunchKernelGGL(k_block_sums, dim3((c->nb + 63) / 64), dim3(64), 0, c->stream, c->d_upart, kBlockParts, c->nb,
                     c->d_ke, 0);
  int rc = sync_and_check(c);
  // MOSE_PROGRESS=1: one line per reversed segment on stderr (long sweeps)
  const char* pe = getenv("MOSE_PROGRESS");
  const bool progress = pe && atoi(pe) == 1;
  if (progress) {
    fprintf(stderr, "[mose_hip] march adjoint: %d steps marched, %d segments of %d steps to reverse\n", nsteps, nck,
            stride);
    fflush(stderr);
  }
  if (rc == 0 && reverse) {
    const int64_t n5 = (int64_t)words;
    hipLaunchKernelGGL(k_fill, dim3(launch_blocks(n5)), dim3(256), 0, c->stream, c->d_Rbar, n5, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, c->stream, c->d_mu, (int64_t)c->nb, 0.0);
    for (int k = nck - 1; k >= 0 && rc == 0; --k) {
      const int len = std::min(stride, nsteps - k * stride);
      // the segment's step-start states, recomputed from its checkpoint
      for (int j = 1; j < len; ++j) enqueue_fixed_step(c, j == 1 ? ck(k) : seg(j - 1), seg(j));
      for (int j = len - 1; j >= 0; --j) enqueue_reverse_step(c, j == 0 ? ck(k) : seg(j));
      rc = sync_and_check(c);
      if (progress) {
        fprintf(stderr, "[mose_hip] march adjoint: segment %d of %d reversed\n", nck - k, nck);
        fflush(stderr);
      }
    }
    if (rc == 0) {
      hipLaunchKernelGGL(k_prim_bar, dim3(launch_blocks(c->nicell)), dim3(256), 0, c->stream, c->d_icell, c->nicell,
                         c->npool, ck(0), c->d_lam0, c->d_Pbar, c->cfg.gamma);
      HIPCHK(c, hipMemcpy(mu_bar, c->d_mu, c->nb * sizeof(double), hipMemcpyDeviceToHost));
    }
    c->vjp_valid = c->geo_valid = c->xbar_valid = false;
  }
  if (rc == 0) HIPCHK(c, hipMemcpy(ke, c->d_ke, c->nb * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipMemcpy(c->d_scal + S_DT, &dt_old, sizeof dt_old, hipMemcpyHostToDevice);
  (void)hipFree(pool);
  return rc ? rc : sync_and_check(c);
}

int mose_get_residual_bar(mose_ctx* c, int32_t b, double* psi) {
  if (!c || !c->finalized || b < 1 || b > c->nb || !psi || !c->d_Rbar) return fail(c, MOSE_E_ARG, b);
  const BlockDev& B = c->lay.blocks[b - 1];
  const int64_t ncell = B.sz * (B.nk + 4);
  if (ensure_stage(c, 5 * ncell)) return c->err.code;
  hipLaunchKernelGGL(k_get_interior, dim3(launch_blocks(ncell)), dim3(256), 0, c->stream, c->d_Rbar, c->npool,
                     B.off, B.ni, B.nj, B.nk, c->stage);
  HIPCHK(c, hipMemcpyAsync(psi, c->stage, 5 * ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int mose_time_residual_vjp(mose_ctx* c, int32_t nrep, double ms[3]) {
  if (!c || !c->finalized || nrep < 1 || !ms) return fail(c, MOSE_E_ARG);
  if (ensure_adjoint(c)) return c->err.code;
  const bool geo = c->lay.curvi;
  if (geo && ensure_adjoint_geometry(c)) return c->err.code;
  ms[2] = 0.0;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  for (int which = 0; which < (geo ? 3 : 2); ++which)
    for (int pass = 0; pass < 2; ++pass) {
      const int reps = pass == 0 ? 1 : nrep;
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r) {
        if (which == 0) enqueue_residual(c);
        else enqueue_vjp(c, which == 2);
      }
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float t = 0.f;
      HIPCHK(c, hipEventElapsedTime(&t, e0, e1));
      if (pass == 1) ms[which] = (double)t / nrep;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->geo_valid = geo;
  c->xbar_valid = false;
  c->vjp_valid = true;
  return sync_and_check(c);
}

int64_t mose_num_cells(const mose_ctx* c) {
  if (!c) return 0;
  int64_t n = 0;
  for (const auto& kv : c->hb) n += (int64_t)kv.second.ni * kv.second.nj * kv.second.nk;
  return n;
}

int mose_last_error(const mose_ctx* c, mose_err* e) {
  if (!c || !e) return MOSE_E_ARG;
  *e = c->err;
  return c->err.code;
}

const char* mose_strerror(int32_t code) {
  switch (code) {
    case MOSE_OK: return "ok";
    case MOSE_E_ARG: return "bad argument or unsupported configuration";
    case MOSE_E_HIP: return "HIP runtime error";
    case MOSE_E_STATE: return "NaN or p<0 detected";
    case MOSE_E_INTERFACE: return "Non-physical state detected at the interface";
    case MOSE_E_METRIC: return "metric tensor det=0 or non-positive volume";
    case MOSE_E_BC: return "boundary-condition list inconsistent or unsupported";
    case MOSE_E_NOGPU: return "no HIP device";
    case MOSE_E_SIZE: return "block larger than the no-allocation staging area (838,860 cells with ghosts)";
    default: return "unknown error";
  }
}

void mose_destroy(mose_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->gexec_n) (void)hipGraphExecDestroy(c->gexec_n);
  if (c->graph_n) (void)hipGraphDestroy(c->graph_n);
  if (c->agexec) (void)hipGraphExecDestroy(c->agexec);
  if (c->agraph) (void)hipGraphDestroy(c->agraph);
  void* bufs[] = {c->U[0], c->U[1], c->U[2], c->Rbuf, c->stage, c->d_blocks, c->d_lines, c->d_geo, c->d_tiles,
                  c->d_ghost, c->d_bcfaces, c->d_vfaces, c->d_vbuf, c->d_bcbuf, c->d_facebc, c->d_scal, c->d_partials, c->d_dtloc, c->d_dtmin, c->d_err,
                  c->d_iter, c->d_tq, c->d_Rbar, c->d_W, c->d_BB, c->d_Pbar, c->d_Pp, c->d_segs, c->d_grows, c->d_gents,
                  c->d_GF, c->d_GL, c->d_BG, c->d_GE, c->d_IB, c->d_VW, c->d_VB, c->d_beta, c->d_pres, c->d_ghostp,
                  c->d_garc_cell, c->d_garc, c->d_rs1, c->d_irsg, c->d_gP, c->d_lam0, c->d_lam1, c->d_Dq,
                  c->d_apart, c->d_ares, c->d_icell, c->d_VBC, c->d_Tp, c->d_vbt, c->d_vbrow, c->d_vbent, c->d_volq,
                  c->d_upart, c->d_ke, c->d_mu, c->d_icbeg, c->d_vfbeg, c->d_vfi, c->d_vtiles, c->d_vghost, c->d_VPb};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->vstream) (void)hipStreamDestroy(c->vstream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
