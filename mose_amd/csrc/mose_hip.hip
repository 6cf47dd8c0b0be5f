// mose_hip.hip — kernels and C ABI of the MI355X explicit residual-and-update path.
//
// One explicit step (Explicit_Step, src/lib/numerics/time/explicit/Mod_Explicit.f90:10-146)
// is this launch sequence on one HIP stream, captured once into a hipGraph:
//   k_begin    iter += 1, dt accumulator = 1e5                     (Mod_Explicit.f90:36,44)
//   k_dt       CFL time step, wave-shuffle min + one atomic per wave (Mod_dt.f90:10-107)
//   k_dtfin    global dt, time += dt when time-accurate             (Mod_Explicit.f90:46-49)
//   n_rk x { k_ghost  ghost layers of every block face             (Lib_Ghost.f90:11-137)
//            k_stage  all face fluxes (interior + boundary faces, Mod_Fluxes.f90:46-209 and
//                     Mod_BC_Fluxes.f90:10-218) and the SSP-RK stage update
//                     (Lib_Newstate.f90:134-160) fused; the last stage also produces the
//                     residual-norm partial sums (Mod_Diagnostic.f90:13-50) }
//   k_normfin  fixed-order sum of the partials, sqrt                (Mod_Explicit.f90:111-112)
// Copy_State (PO = P) needs no pass: U^n stays in its own buffer while the stages run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/mose_hip.h"
#include "mose_kernels.hpp"

using namespace mose;

// ============================================================================ device
__device__ __forceinline__ void set_error(int* err, int code, int b, int i, int j, int k) {
  if (atomicCAS(err, 0, code) == 0) {
    err[1] = b; err[2] = i; err[3] = j; err[4] = k;
  }
}

template <typename FT>
__device__ __forceinline__ void load_prim(const double* U, int64_t N, int64_t c, double gm1, FT* P) {
  double u[5], p[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u[v] = U[v * N + c];
  cons2prim(u, gm1, p);
#pragma unroll
  for (int v = 0; v < 5; ++v) P[v] = (FT)p[v];
}

template <typename FT>
__device__ __forceinline__ FaceSpacing<FT> spacing_of(const LineEnt& e) {
  return FaceSpacing<FT>{FT(e.rl0), FT(e.rl1), FT(e.rl2), FT(e.lm), FT(e.lp)};
}

// Flux (times area) through the face between stencil cells q1 and q2, i.e. between cells
// idx-1 and idx along direction DIR.  bct = F_CONN for interior and connection faces,
// otherwise the boundary treatment of the block face (hi: max face).
template <int LIM, int RS, int DIR, typename FT>
__device__ __forceinline__ bool face_any(int bct, bool hi, const double* bcd, const FT* q0,
                                         const FT* q1, const FT* q2, const FT* q3,
                                         const LineEnt* L, int idx, double area,
                                         const GasT<FT>& g, double* F) {
  constexpr FT nx = DIR == 0 ? FT(1) : FT(0);
  constexpr FT ny = DIR == 1 ? FT(1) : FT(0);
  constexpr FT nz = DIR == 2 ? FT(1) : FT(0);
  FT f[5];
  bool ok = true;
  if (bct == F_CONN) {
    ok = face_flux<LIM, RS>(q0, q1, q2, q3, spacing_of<FT>(L[idx + 1]), nx, ny, nz, g, f);
  } else if (bct == F_NONE) {
#pragma unroll
    for (int v = 0; v < 5; ++v) f[v] = FT(0);
  } else {
    const FT* Pb = hi ? q1 : q2;
    const FT* Pi = hi ? q0 : q3;
    if (bct == F_EXTRAP) {
      physical_flux(Pb, nx, ny, nz, g, f);
    } else {
      bool done = false;
      if (bct == F_INFLOW) done = inflow405_flux<LIM, RS>(Pb, Pi, hi, nx, ny, nz, bcd, g, f);
      if (!done) {
        const LineEnt& e = L[idx + 1];
        const FT bdl = hi ? FT(e.lm) : FT(e.lp);
        const FT idl = FT(0.5 * (hi ? L[idx - 1].dl : L[idx + 2].dl));
        symmetry_flux<LIM, RS>(Pb, Pi, bdl, idl, hi, nx, ny, nz, g, f);
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) F[v] = double(f[v]) * area;
  return ok;
}

// MODE 0: stage update; 1: last stage update + residual-norm partials; 2: residual only.
template <int LIM, int RS, typename FT, int MODE>
__global__ void __launch_bounds__(NT) k_stage(StageArgs a) {
  __shared__ FT sP[5][TY + 4][TX + 4];
  __shared__ double sFi[5][TY][TX + 1];
  __shared__ double sFj[5][TY + 1][TX];
  __shared__ double sRed[NT / 64][5];

  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int lane = tid & 63, wave = tid >> 6;
  const double gm1 = a.gam - 1.0;
  const GasT<FT> g{FT(a.gam), FT(gm1), FT(1.0 / gm1), FT(a.R), FT(a.gam * a.R / gm1)};
  const int64_t N = a.npool;
  const double* Uin = a.Uin;

  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const Tile T = a.tiles[t];
    const BlockDev B = a.blocks[T.b];
    const int i = T.i0 + tx, j = T.j0 + ty;
    const bool colmem = (i <= B.ni + 2) && (j <= B.nj + 2);
    const bool active = (i <= B.ni) && (j <= B.nj);
    const int64_t cbase = B.off + (int64_t)(i + 1) + (int64_t)B.sy * (j + 1);
    const LineEnt* Lx = a.lines + B.line[0];
    const LineEnt* Ly = a.lines + B.line[1];
    const LineEnt* Lz = a.lines + B.line[2];
    const int k0 = T.k0, k1 = T.k1;

    FT q[4][5];
    if (colmem) {
#pragma unroll
      for (int s = 0; s < 4; ++s) load_prim(Uin, N, cbase + B.sz * (int64_t)(k0 - 1 + s), gm1, q[s]);
    }
    double Fkp[5], Fkn[5], nsum[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) { Fkp[v] = 0.0; nsum[v] = 0.0; }
    const double Az = active ? Lx[i + 1].delta * Ly[j + 1].delta : 0.0;
    if (active) {     // face k0 - 1/2
      const bool bnd = (k0 == 1);
      const int bct = bnd ? B.bct[4] : F_CONN;
      const double* bcd = (bnd && B.bcd[4] >= 0) ? a.facebc + 6 * B.bcd[4] : nullptr;
      if (!face_any<LIM, RS, 2>(bct, false, bcd, q[0], q[1], q[2], q[3], Lz, k0, Az, g, Fkp))
        set_error(a.err, MOSE_E_INTERFACE, T.b + 1, i, j, k0 - 1);
    }

    for (int k = k0; k < k1; ++k) {
      if (colmem) {
#pragma unroll
        for (int v = 0; v < 5; ++v) { q[0][v] = q[1][v]; q[1][v] = q[2][v]; q[2][v] = q[3][v]; }
        load_prim(Uin, N, cbase + B.sz * (int64_t)(k + 3), gm1, q[3]);
#pragma unroll
        for (int v = 0; v < 5; ++v) sP[v][ty + 2][tx + 2] = q[1][v];
      }
      // 2-cell i/j halo of plane k
      for (int h = tid; h < 4 * (TX + TY); h += NT) {
        int hi_, hj;
        if (h < 2 * TY) { hj = h >> 1; hi_ = -2 + (h & 1); }
        else if (h < 4 * TY) { const int h2 = h - 2 * TY; hj = h2 >> 1; hi_ = TX + (h2 & 1); }
        else if (h < 4 * TY + 2 * TX) { const int h2 = h - 4 * TY; hi_ = h2 % TX; hj = -2 + h2 / TX; }
        else { const int h2 = h - 4 * TY - 2 * TX; hi_ = h2 % TX; hj = TY + h2 / TX; }
        const int gi = T.i0 + hi_, gj = T.j0 + hj;
        if (gi <= B.ni + 2 && gj <= B.nj + 2) {
          FT tmp[5];
          load_prim(Uin, N, B.off + (int64_t)(gi + 1) + (int64_t)B.sy * (gj + 1) + B.sz * (int64_t)(k + 1), gm1, tmp);
#pragma unroll
          for (int v = 0; v < 5; ++v) sP[v][hj + 2][hi_ + 2] = tmp[v];
        }
      }
      __syncthreads();

      if (active) {   // face k + 1/2
        const bool bnd = (k == B.nk);
        const int bct = bnd ? B.bct[5] : F_CONN;
        const double* bcd = (bnd && B.bcd[5] >= 0) ? a.facebc + 6 * B.bcd[5] : nullptr;
        if (!face_any<LIM, RS, 2>(bct, bnd, bcd, q[0], q[1], q[2], q[3], Lz, k + 1, Az, g, Fkn))
          set_error(a.err, MOSE_E_INTERFACE, T.b + 1, i, j, k);
      }
      // i face on the low side of (i, j)
      if (j <= B.nj && i <= B.ni + 1) {
        const int f = (i == 1) ? 0 : ((i == B.ni + 1) ? 1 : -1);
        const int bct = f >= 0 ? B.bct[f] : F_CONN;
        const double* bcd = (f >= 0 && B.bcd[f] >= 0) ? a.facebc + 6 * B.bcd[f] : nullptr;
        FT s[4][5];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int v = 0; v < 5; ++v) s[m][v] = sP[v][ty + 2][tx + m];
        double F[5];
        const double A = Ly[j + 1].delta * Lz[k + 1].delta;
        if (!face_any<LIM, RS, 0>(bct, f == 1, bcd, s[0], s[1], s[2], s[3], Lx, i, A, g, F))
          set_error(a.err, MOSE_E_INTERFACE, T.b + 1, i - 1, j, k);
#pragma unroll
        for (int v = 0; v < 5; ++v) sFi[v][ty][tx] = F[v];
      }
      if (wave == 0 && lane < TY) {   // i face on the tile's high edge, row `lane`
        const int jr = T.j0 + lane, ie = T.i0 + TX;
        if (jr <= B.nj && ie <= B.ni + 1) {
          const int f = (ie == B.ni + 1) ? 1 : -1;
          const int bct = f >= 0 ? B.bct[f] : F_CONN;
          const double* bcd = (f >= 0 && B.bcd[f] >= 0) ? a.facebc + 6 * B.bcd[f] : nullptr;
          FT s[4][5];
#pragma unroll
          for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 5; ++v) s[m][v] = sP[v][lane + 2][TX + m];
          double F[5];
          const double A = Ly[jr + 1].delta * Lz[k + 1].delta;
          if (!face_any<LIM, RS, 0>(bct, f == 1, bcd, s[0], s[1], s[2], s[3], Lx, ie, A, g, F))
            set_error(a.err, MOSE_E_INTERFACE, T.b + 1, ie - 1, jr, k);
#pragma unroll
          for (int v = 0; v < 5; ++v) sFi[v][lane][TX] = F[v];
        }
      }
      // j face on the low side of (i, j)
      if (i <= B.ni && j <= B.nj + 1) {
        const int f = (j == 1) ? 2 : ((j == B.nj + 1) ? 3 : -1);
        const int bct = f >= 0 ? B.bct[f] : F_CONN;
        const double* bcd = (f >= 0 && B.bcd[f] >= 0) ? a.facebc + 6 * B.bcd[f] : nullptr;
        FT s[4][5];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int v = 0; v < 5; ++v) s[m][v] = sP[v][ty + m][tx + 2];
        double F[5];
        const double A = Lx[i + 1].delta * Lz[k + 1].delta;
        if (!face_any<LIM, RS, 1>(bct, f == 3, bcd, s[0], s[1], s[2], s[3], Ly, j, A, g, F))
          set_error(a.err, MOSE_E_INTERFACE, T.b + 1, i, j - 1, k);
#pragma unroll
        for (int v = 0; v < 5; ++v) sFj[v][ty][tx] = F[v];
      }
      if (wave == 1 && lane < TX) {   // j face on the tile's high edge, column `lane`
        const int ic = T.i0 + lane, je = T.j0 + TY;
        if (ic <= B.ni && je <= B.nj + 1) {
          const int f = (je == B.nj + 1) ? 3 : -1;
          const int bct = f >= 0 ? B.bct[f] : F_CONN;
          const double* bcd = (f >= 0 && B.bcd[f] >= 0) ? a.facebc + 6 * B.bcd[f] : nullptr;
          FT s[4][5];
#pragma unroll
          for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 5; ++v) s[m][v] = sP[v][TY + m][lane + 2];
          double F[5];
          const double A = Lx[ic + 1].delta * Lz[k + 1].delta;
          if (!face_any<LIM, RS, 1>(bct, f == 3, bcd, s[0], s[1], s[2], s[3], Ly, je, A, g, F))
            set_error(a.err, MOSE_E_INTERFACE, T.b + 1, ic, je - 1, k);
#pragma unroll
          for (int v = 0; v < 5; ++v) sFj[v][TY][lane] = F[v];
        }
      }
      __syncthreads();

      if (active) {
        double R[5];
#pragma unroll
        for (int v = 0; v < 5; ++v)
          R[v] = (sFi[v][ty][tx + 1] - sFi[v][ty][tx]) + (sFj[v][ty + 1][tx] - sFj[v][ty][tx]) + (Fkn[v] - Fkp[v]);
        const int64_t c = cbase + B.sz * (int64_t)(k + 1);
        if constexpr (MODE == 2) {
#pragma unroll
          for (int v = 0; v < 5; ++v) a.Rout[v * N + c] = R[v];
        } else {
          double un[5], ui[5], unew[5], P[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) { un[v] = a.Un[v * N + c]; ui[v] = Uin[v * N + c]; }
          const double dt = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
          const double vol = Lx[i + 1].delta * Ly[j + 1].delta * Lz[k + 1].delta;
          const double rv = dt / vol;
#pragma unroll
          for (int v = 0; v < 5; ++v) unew[v] = un[v] + a.rkc * (ui[v] - un[v] - R[v] * rv);
          cons2prim(unew, gm1, P);
          bool bad = !(P[4] >= 0.0);
#pragma unroll
          for (int v = 0; v < 5; ++v) bad = bad || (P[v] != P[v]);
          if (bad) set_error(a.err, MOSE_E_STATE, T.b + 1, i, j, k);
          if (P[0] < 1e-20) { P[0] = 1e-20; prim2cons(P, 1.0 / gm1, unew); }
          if constexpr (MODE == 1) {
            double pn[5];
            cons2prim(un, gm1, pn);
#pragma unroll
            for (int v = 0; v < 5; ++v) { const double d = P[v] - pn[v]; nsum[v] += d * d; }
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) a.Uout[v * N + c] = unew[v];
        }
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) Fkp[v] = Fkn[v];
    }

    if constexpr (MODE == 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        double x = nsum[v];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) sRed[wave][v] = x;
      }
      __syncthreads();
      if (tid < 5) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; ++w) s += sRed[w][tid];
        a.partials[5 * (int64_t)t + tid] = s;
      }
      __syncthreads();
    }
  }
}

struct DtArgs {
  const double* U;
  const Tile* tiles;
  const BlockDev* blocks;
  const LineEnt* lines;
  unsigned long long* dtmin;
  double* dtloc;
  const int* iter;
  int64_t npool;
  int ntiles, ndir, rampa, local_dt;
  double cfl, gam;
};

// CFL time step of every interior cell (Compute_Dt, Mod_dt.f90:61-85 with rectilinear
// metrics: versor = axis, dl = cell length) and its global minimum.
__global__ void __launch_bounds__(NT) k_dt(DtArgs a) {
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const double gm1 = a.gam - 1.0;
  const int it = *a.iter;
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
      dtc = fmin(Lx[i + 1].dl / (fabs(P[1]) + snd) * a.cfl, dtc);
      if (a.ndir > 1) dtc = fmin(Ly[j + 1].dl / (fabs(P[2]) + snd) * a.cfl, dtc);
      if (a.ndir > 2) dtc = fmin(Lz[k + 1].dl / (fabs(P[3]) + snd) * a.cfl, dtc);
      if (it < a.rampa) dtc = dtc * it / a.rampa;
      if (a.local_dt) a.dtloc[c] = dtc;
      mymin = fmin(mymin, dtc);
    }
  }
  for (int o = 32; o > 0; o >>= 1) mymin = fmin(mymin, __shfl_xor(mymin, o, 64));
  if ((tid & 63) == 0 && mymin < 1e300)
    atomicMin(a.dtmin, (unsigned long long)__double_as_longlong(mymin));
}

__global__ void k_begin(unsigned long long* dtmin, int* iter) {
  *iter += 1;
  *dtmin = (unsigned long long)__double_as_longlong(1e5);
}

__global__ void k_dtfin(const unsigned long long* dtmin, double* scal, int time_accurate) {
  const double dt = __longlong_as_double((long long)*dtmin);
  scal[S_DT] = dt;
  if (time_accurate) scal[S_TIME] += dt;
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
__global__ void __launch_bounds__(256) k_ghost(const GhostOp* __restrict__ ops, int nops, double* U,
                                               int64_t N, double gam) {
  const double gm1 = gam - 1.0, rgm1 = 1.0 / gm1;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nops; q += gridDim.x * blockDim.x) {
    const GhostOp op = ops[q];
    const int64_t d1 = op.b - op.s, d2 = op.b - 2 * op.s;
    if (op.type == G_COPY) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double a1 = U[v * N + op.src1], a2 = U[v * N + op.src2];
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
      for (int v = 0; v < 5; ++v) u[v] = U[v * N + op.b + op.s];
      cons2prim(u, gm1, P1);
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v] = U[v * N + op.b + 2 * op.s];
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
__global__ void k_fill(double* p, int64_t n, double x) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) p[c] = x;
}
__global__ void k_copy(const double* __restrict__ s, double* __restrict__ d, int64_t n) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) d[c] = s[c];
}

// ============================================================================== host
#include "mose_geometry.hpp"

namespace {

typedef void (*StageFn)(StageArgs);

template <int LIM, int RS, typename FT>
void stage_table(StageFn* t) {
  t[0] = k_stage<LIM, RS, FT, 0>;
  t[1] = k_stage<LIM, RS, FT, 1>;
  t[2] = k_stage<LIM, RS, FT, 2>;
}

// kernel table [prec][lim][rs][mode]
struct KTable {
  StageFn f[2][4][2][3];
  KTable() {
    stage_table<LIM_MINMOD, RS_HLLC, double>(f[0][0][0]);
    stage_table<LIM_MINMOD, RS_AUSMP, double>(f[0][0][1]);
    stage_table<LIM_VANLEER, RS_HLLC, double>(f[0][1][0]);
    stage_table<LIM_VANLEER, RS_AUSMP, double>(f[0][1][1]);
    stage_table<LIM_VANALBADA, RS_HLLC, double>(f[0][2][0]);
    stage_table<LIM_VANALBADA, RS_AUSMP, double>(f[0][2][1]);
    stage_table<LIM_FIRST, RS_HLLC, double>(f[0][3][0]);
    stage_table<LIM_FIRST, RS_AUSMP, double>(f[0][3][1]);
    stage_table<LIM_MINMOD, RS_HLLC, float>(f[1][0][0]);
    stage_table<LIM_MINMOD, RS_AUSMP, float>(f[1][0][1]);
    stage_table<LIM_VANLEER, RS_HLLC, float>(f[1][1][0]);
    stage_table<LIM_VANLEER, RS_AUSMP, float>(f[1][1][1]);
    stage_table<LIM_VANALBADA, RS_HLLC, float>(f[1][2][0]);
    stage_table<LIM_VANALBADA, RS_AUSMP, float>(f[1][2][1]);
    stage_table<LIM_FIRST, RS_HLLC, float>(f[1][3][0]);
    stage_table<LIM_FIRST, RS_AUSMP, float>(f[1][3][1]);
  }
};
const KTable& ktable() {
  static KTable t;
  return t;
}

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
  double* U[3] = {nullptr, nullptr, nullptr};
  double* Rbuf = nullptr;
  double* stage = nullptr;
  int64_t stage_n = 0;
  BlockDev* d_blocks = nullptr;
  LineEnt* d_lines = nullptr;
  Tile* d_tiles = nullptr;
  GhostOp* d_ghost = nullptr;
  double* d_facebc = nullptr;
  double* d_scal = nullptr;
  double* d_partials = nullptr;
  double* d_dtloc = nullptr;
  unsigned long long* d_dtmin = nullptr;
  int* d_err = nullptr;
  int* d_iter = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int grid_stage = 0;
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

static int enqueue_ghost(mose_ctx* c, double* U) {
  if (c->lay.ghost.empty()) return 0;
  const int n = (int)c->lay.ghost.size();
  hipLaunchKernelGGL(k_ghost, dim3(launch_blocks(n)), dim3(256), 0, c->stream, c->d_ghost, n, U, c->npool,
                     c->cfg.gamma);
  return 0;
}

static StageArgs stage_args(mose_ctx* c, const double* Uin, const double* Un, double* Uout, double rkc) {
  StageArgs a{};
  a.Uin = Uin; a.Un = Un; a.Uout = Uout; a.Rout = c->Rbuf;
  a.tiles = c->d_tiles; a.blocks = c->d_blocks; a.lines = c->d_lines; a.facebc = c->d_facebc;
  a.scal = c->d_scal; a.dtloc = c->d_dtloc; a.partials = c->d_partials; a.err = c->d_err;
  a.npool = c->npool; a.ntiles = (int)c->lay.tiles.size(); a.local_dt = c->cfg.time_accurate ? 0 : 1;
  a.rkc = rkc; a.gam = c->cfg.gamma; a.R = c->cfg.Rgas;
  return a;
}

static void launch_stage(mose_ctx* c, StageFn fn, StageArgs a) {
  void* args[] = {&a};
  (void)hipLaunchKernel((const void*)fn, dim3(c->grid_stage), dim3(NT), args, 0, c->stream);
}

static StageFn stage_fn(mose_ctx* c, int mode) {
  const int p = c->cfg.precision == MOSE_PREC_FP32_FLUX ? 1 : 0;
  return ktable().f[p][c->cfg.limiter][c->cfg.riemann][mode];
}

// One explicit step enqueued on the context stream (no host synchronisation).
static int enqueue_step(mose_ctx* c) {
  const int nt = (int)c->lay.tiles.size();
  hipLaunchKernelGGL(k_begin, dim3(1), dim3(1), 0, c->stream, c->d_dtmin, c->d_iter);
  DtArgs da{};
  da.U = c->U[0]; da.tiles = c->d_tiles; da.blocks = c->d_blocks; da.lines = c->d_lines;
  da.dtmin = c->d_dtmin; da.dtloc = c->d_dtloc; da.iter = c->d_iter; da.npool = c->npool;
  da.ntiles = nt; da.ndir = c->ndir; da.rampa = c->cfg.rampa_cfl_iter;
  da.local_dt = c->cfg.time_accurate ? 0 : 1; da.cfl = c->cfg.cfl; da.gam = c->cfg.gamma;
  hipLaunchKernelGGL(k_dt, dim3(nt), dim3(NT), 0, c->stream, da);
  hipLaunchKernelGGL(k_dtfin, dim3(1), dim3(1), 0, c->stream, c->d_dtmin, c->d_scal, c->cfg.time_accurate);
  const int nrk = c->cfg.n_rk;
  for (int s = 1; s <= nrk; ++s) {
    double* in = (s == 1) ? c->U[0] : c->U[s - 1];
    double* out = (s == nrk) ? (nrk == 1 ? c->U[1] : c->U[0]) : c->U[s];
    enqueue_ghost(c, in);
    StageArgs a = stage_args(c, in, c->U[0], out, kRK[nrk - 1][s - 1]);
    launch_stage(c, stage_fn(c, s == nrk ? 1 : 0), a);
  }
  if (nrk == 1) {
    const int64_t n = 5 * c->npool;
    hipLaunchKernelGGL(k_copy, dim3(launch_blocks(n)), dim3(256), 0, c->stream, c->U[1], c->U[0], n);
  }
  hipLaunchKernelGGL(k_normfin, dim3(1), dim3(256), 0, c->stream, c->d_partials, nt, c->d_scal);
  return 0;
}

static int sync_and_check(mose_ctx* c) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int e[5];
  HIPCHK(c, hipMemcpy(e, c->d_err, sizeof e, hipMemcpyDeviceToHost));
  if (e[0]) {
    c->err = mose_err{e[0], e[1], e[2], e[3], e[4]};
    return e[0];
  }
  return 0;
}

static int ensure_graph(mose_ctx* c) {
  if (c->gexec) return 0;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(c);
  HIPCHK(c, hipStreamEndCapture(c->stream, &c->graph));
  HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  return 0;
}

template <typename V>
static int upload(mose_ctx* c, V** dptr, const std::vector<V>& h) {
  const size_t n = std::max<size_t>(h.size(), 1);
  HIPCHK(c, hipMalloc((void**)dptr, n * sizeof(V)));
  if (!h.empty()) HIPCHK(c, hipMemcpy(*dptr, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
  return 0;
}

static int ensure_stage(mose_ctx* c, int64_t n) {
  if (c->stage_n >= n) return 0;
  if (c->stage) HIPCHK(c, hipFree(c->stage));
  HIPCHK(c, hipMalloc((void**)&c->stage, n * sizeof(double)));
  c->stage_n = n;
  return 0;
}

// ============================================================================= C ABI
extern "C" {

int mose_create(const mose_cfg* cfg, mose_ctx** out) {
  if (!cfg || !out) return MOSE_E_ARG;
  *out = nullptr;
  if (cfg->limiter < 0 || cfg->limiter > 3 || cfg->riemann < 0 || cfg->riemann > 1 || cfg->n_rk < 1 ||
      cfg->n_rk > 3 || !(cfg->gamma > 1.0) || !(cfg->Rgas > 0.0) || cfg->precision < 0 || cfg->precision > 1)
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
  GeoError ge;
  if (!build_layout(blocks, c->bc, TX, TY, c->lay, ge)) return fail(c, ge.code, ge.b, ge.i, ge.j, ge.k);
  c->npool = c->lay.npool;
  c->ndir = c->lay.ndir;
  // device buffers
  for (int s = 0; s < 3; ++s) {
    HIPCHK(c, hipMalloc((void**)&c->U[s], 5 * c->npool * sizeof(double)));
    HIPCHK(c, hipMemset(c->U[s], 0, 5 * c->npool * sizeof(double)));
  }
  HIPCHK(c, hipMalloc((void**)&c->Rbuf, 5 * c->npool * sizeof(double)));
  if (upload(c, &c->d_blocks, c->lay.blocks)) return c->err.code;
  if (upload(c, &c->d_lines, c->lay.lines)) return c->err.code;
  if (upload(c, &c->d_tiles, c->lay.tiles)) return c->err.code;
  if (upload(c, &c->d_ghost, c->lay.ghost)) return c->err.code;
  if (upload(c, &c->d_facebc, c->lay.facebc)) return c->err.code;
  HIPCHK(c, hipMalloc((void**)&c->d_scal, S_COUNT * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_scal, 0, S_COUNT * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_partials, 5 * std::max<size_t>(c->lay.tiles.size(), 1) * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_dtmin, sizeof(unsigned long long)));
  HIPCHK(c, hipMalloc((void**)&c->d_err, 8 * sizeof(int)));
  HIPCHK(c, hipMemset(c->d_err, 0, 8 * sizeof(int)));
  HIPCHK(c, hipMalloc((void**)&c->d_iter, sizeof(int)));
  HIPCHK(c, hipMemset(c->d_iter, 0, sizeof(int)));
  if (!c->cfg.time_accurate) {
    HIPCHK(c, hipMalloc((void**)&c->d_dtloc, c->npool * sizeof(double)));
    HIPCHK(c, hipMemset(c->d_dtloc, 0, c->npool * sizeof(double)));
  }
  c->grid_stage = (int)c->lay.tiles.size();
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
  HIPCHK(c, hipStreamSynchronize(c->stream));
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

int mose_compute_dt(mose_ctx* c, int32_t iter, double* dtglobal) {
  if (!c || !c->finalized) return fail(c, MOSE_E_ARG);
  HIPCHK(c, hipMemcpy(c->d_iter, &iter, sizeof(int), hipMemcpyHostToDevice));
  const unsigned long long init = 0x40F86A0000000000ull;   // bit pattern of 1e5
  HIPCHK(c, hipMemcpy(c->d_dtmin, &init, sizeof init, hipMemcpyHostToDevice));
  DtArgs da{};
  const int nt = (int)c->lay.tiles.size();
  da.U = c->U[0]; da.tiles = c->d_tiles; da.blocks = c->d_blocks; da.lines = c->d_lines;
  da.dtmin = c->d_dtmin; da.dtloc = c->d_dtloc; da.iter = c->d_iter; da.npool = c->npool;
  da.ntiles = nt; da.ndir = c->ndir; da.rampa = c->cfg.rampa_cfl_iter;
  da.local_dt = c->cfg.time_accurate ? 0 : 1; da.cfl = c->cfg.cfl; da.gam = c->cfg.gamma;
  hipLaunchKernelGGL(k_dt, dim3(nt), dim3(NT), 0, c->stream, da);
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->d_dtmin, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double dt;
  memcpy(&dt, &bits, sizeof dt);
  if (dtglobal) *dtglobal = dt;
  return 0;
}

int mose_advance(mose_ctx* c, int32_t nsteps, double* time_out, double* norms) {
  if (!c || !c->finalized || nsteps < 0) return fail(c, MOSE_E_ARG);
  if (c->err.code) return c->err.code;
  if (ensure_graph(c)) return c->err.code;
  for (int s = 0; s < nsteps; ++s) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
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
  const int64_t n5 = 5 * c->npool;
  hipLaunchKernelGGL(k_fill, dim3(launch_blocks(n5)), dim3(256), 0, c->stream, c->Rbuf, n5, 0.0);
  enqueue_ghost(c, c->U[0]);
  StageArgs a = stage_args(c, c->U[0], c->U[0], c->U[1], 0.0);
  launch_stage(c, stage_fn(c, 2), a);
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
  return 0;
}

int mose_time_kernels(mose_ctx* c, int32_t nrep, double ms[4]) {
  if (!c || !c->finalized || nrep < 1 || !ms) return fail(c, MOSE_E_ARG);
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const int nt = (int)c->lay.tiles.size();
  DtArgs da{};
  da.U = c->U[0]; da.tiles = c->d_tiles; da.blocks = c->d_blocks; da.lines = c->d_lines;
  da.dtmin = c->d_dtmin; da.dtloc = c->d_dtloc; da.iter = c->d_iter; da.npool = c->npool;
  da.ntiles = nt; da.ndir = c->ndir; da.rampa = c->cfg.rampa_cfl_iter;
  da.local_dt = c->cfg.time_accurate ? 0 : 1; da.cfl = c->cfg.cfl; da.gam = c->cfg.gamma;
  for (int which = 0; which < 4; ++which) {
    for (int pass = 0; pass < 2; ++pass) {      // pass 0: warm-up launch, pass 1: timed
      const int reps = pass == 0 ? 1 : nrep;
      HIPCHK(c, hipEventRecord(e0, c->stream));
      for (int r = 0; r < reps; ++r) {
        if (which == 0) hipLaunchKernelGGL(k_dt, dim3(nt), dim3(NT), 0, c->stream, da);
        else if (which == 1) enqueue_ghost(c, c->U[0]);
        else launch_stage(c, stage_fn(c, which == 2 ? 0 : 1), stage_args(c, c->U[0], c->U[0], c->U[1], 0.25));
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
    default: return "unknown error";
  }
}

void mose_destroy(mose_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  void* bufs[] = {c->U[0], c->U[1], c->U[2], c->Rbuf, c->stage, c->d_blocks, c->d_lines, c->d_tiles,
                  c->d_ghost, c->d_facebc, c->d_scal, c->d_partials, c->d_dtloc, c->d_dtmin, c->d_err,
                  c->d_iter};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

}  // extern "C"
