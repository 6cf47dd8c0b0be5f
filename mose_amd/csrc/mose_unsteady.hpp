// mose_unsteady.hpp — kernels of the adjoint of a fixed-dt explicit march (included by
// mose_hip.hip).
//
// The march repeats the reference's stages with dt held fixed (Lib_Newstate.f90:134-160,
// Lib_RK.f90:8-27): U_s = (1 - c_s) U_n + c_s (U_{s-1} - D R(U_{s-1})), U_0 = U_n,
// D = dt / vol, c = (1), (1, 1/2), (1, 1/4, 2/3).  For J(U_final), with lambda_s = dJ/dU_s,
// the reverse of one stage is
//     lambda_n     += (1 - c_s) lambda_s,
//     Rbar          = -c_s D lambda_s,   Pbar = (dR/dP)^T Rbar at U_{s-1} (the residual VJP),
//     lambda_{s-1}  = c_s lambda_s + (dP/dU)^T Pbar,
// and lambda_0 joins lambda_n.  The objective here is the kinetic energy of the final state,
// J = sum over interior cells of 1/2 rho |u|^2 vol.
#pragma once

namespace mose {

// (dP/dU)^T x for P = (rho, u, v, w, p), U = (rho, rho u, rho v, rho w, rho E).
__device__ __forceinline__ void cons_bar_of(const double* u, const double* x, double gm1, double* y) {
  const double ir = 1.0 / u[0];
  const double uu = u[1] * ir, vv = u[2] * ir, ww = u[3] * ir;
  y[0] = x[0] - (uu * x[1] + vv * x[2] + ww * x[3]) * ir + gm1 * 0.5 * (uu * uu + vv * vv + ww * ww) * x[4];
  y[1] = x[1] * ir - gm1 * uu * x[4];
  y[2] = x[2] * ir - gm1 * vv * x[4];
  y[3] = x[3] * ir - gm1 * ww * x[4];
  y[4] = gm1 * x[4];
}

// Stage reverse, first half: Rbar = -c D lambda_s (interior cells), lambda_n += (1-c) lambda_s.
__global__ void __launch_bounds__(256) k_unst_rbar(const int64_t* __restrict__ cells, const double* __restrict__ D,
                                                   int64_t n, int64_t N, const double* __restrict__ lamS,
                                                   double* __restrict__ lamN, double* __restrict__ Rbar, double cs) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    const double d = D[q];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double l = lamS[v * N + c];
      Rbar[v * N + c] = -cs * d * l;
      lamN[v * N + c] += (1.0 - cs) * l;
    }
  }
}

// Stage reverse, second half: lambda_{s-1} = c lambda_s + (dP/dU)^T Pbar at U_{s-1}, in place
// in lamS; for the first stage (U_0 = U_n) it is added to lamN instead.
__global__ void __launch_bounds__(256) k_unst_back(const int64_t* __restrict__ cells, int64_t n, int64_t N,
                                                   const double* __restrict__ U, const double* __restrict__ Pbar,
                                                   double* __restrict__ lamS, double* __restrict__ lamN, double cs,
                                                   double gam, int first) {
  const double gm1 = gam - 1.0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    double u[5], x[5], y[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      u[v] = U[v * N + c];
      x[v] = Pbar[v * N + c];
    }
    cons_bar_of(u, x, gm1, y);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double l = cs * lamS[v * N + c] + y[v];
      if (first) lamN[v * N + c] += l;
      else lamS[v * N + c] = l;
    }
  }
}

// Kinetic energy of the interior cells of block blockIdx.y (cell range [beg, end) of the
// interior list) into partial[blockIdx.y * gridDim.x + blockIdx.x]; with seed, also
// lambda = dJ/dU = (-|u|^2/2, u, v, w, 0) vol and lamN = 0.
__global__ void __launch_bounds__(256) k_ke(const int64_t* __restrict__ cells, const double* __restrict__ vol,
                                            const int64_t* __restrict__ beg, int64_t N, const double* __restrict__ U,
                                            double* __restrict__ lamS, double* __restrict__ lamN,
                                            double* __restrict__ partial, int seed) {
  const int b = blockIdx.y;
  double s = 0.0;
  for (int64_t q = beg[b] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < beg[b + 1];
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    const double ir = 1.0 / U[c];
    const double uu = U[N + c] * ir, vv = U[2 * N + c] * ir, ww = U[3 * N + c] * ir;
    const double q2 = uu * uu + vv * vv + ww * ww, w = vol[q];
    s += 0.5 * U[c] * q2 * w;
    if (seed) {
      lamS[c] = -0.5 * q2 * w;
      lamS[N + c] = uu * w;
      lamS[2 * N + c] = vv * w;
      lamS[3 * N + c] = ww * w;
      lamS[4 * N + c] = 0.0;
#pragma unroll
      for (int v = 0; v < 5; ++v) lamN[v * N + c] = 0.0;
    }
  }
  __shared__ double sred[4];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
}

// mu_ref adjoint of one VJP, block blockIdx.y: the interior viscous faces' entries (row 17 of
// each direction's face adjoints, at the low cell of faces whose two cells are interior) and
// the block's connection faces (VB), into partial[blockIdx.y * gridDim.x + blockIdx.x].
__global__ void __launch_bounds__(256) k_mubar(const BlockDev* __restrict__ blocks, const double* __restrict__ VW,
                                               const double* __restrict__ VB, const int* __restrict__ vfi,
                                               const int64_t* __restrict__ vfbeg, int64_t N,
                                               double* __restrict__ partial) {
  const int b = blockIdx.y;
  const BlockDev B = blocks[b];
  const int n3[3] = {B.ni, B.nj, B.nk};
  const int64_t ncell = B.sz * (B.nk + 4);
  double s = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < ncell; r += (int64_t)gridDim.x * blockDim.x) {
    const int q[3] = {(int)(r % B.sy) - 1, (int)((r / B.sy) % (B.nj + 4)) - 1, (int)(r / B.sz) - 1};
    if (q[0] < 1 || q[1] < 1 || q[2] < 1 || q[0] > n3[0] || q[1] > n3[1] || q[2] > n3[2]) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (q[d] <= n3[d] - 1) s += VW[((size_t)d * 18 + 17) * N + B.off + r];
  }
  for (int64_t e = vfbeg[b] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < vfbeg[b + 1];
       e += (int64_t)gridDim.x * blockDim.x)
    s += VB[vfi[e]];
  __shared__ double sred[4];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[b * gridDim.x + blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
}

// acc[b] (+)= the fixed-order sum of partial[b * np .. b * np + np), one thread per block.
__global__ void k_block_sums(const double* __restrict__ partial, int np, int nb, double* acc, int add) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0.0;
  for (int x = 0; x < np; ++x) s += partial[b * np + x];
  acc[b] = add ? acc[b] + s : s;
}

// Pbar = (dU/dP)^T lambda at U for the interior cells (the initial-state adjoint in primitive
// variables): rows of dU/dP: d rho = (1,0,0,0,0); d(rho u) = (u, rho, 0, 0, 0) (v, w alike);
// d(rho E) = (|u|^2/2, rho u, rho v, rho w, 1/(gamma-1)).
__global__ void __launch_bounds__(256) k_prim_bar(const int64_t* __restrict__ cells, int64_t n, int64_t N,
                                                  const double* __restrict__ U, const double* __restrict__ lam,
                                                  double* __restrict__ Pbar, double gam) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    const double rho = U[c], ir = 1.0 / rho;
    const double uu = U[N + c] * ir, vv = U[2 * N + c] * ir, ww = U[3 * N + c] * ir;
    double l[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) l[v] = lam[v * N + c];
    Pbar[c] = l[0] + uu * l[1] + vv * l[2] + ww * l[3] + 0.5 * (uu * uu + vv * vv + ww * ww) * l[4];
    Pbar[N + c] = rho * (l[1] + uu * l[4]);
    Pbar[2 * N + c] = rho * (l[2] + vv * l[4]);
    Pbar[3 * N + c] = rho * (l[3] + ww * l[4]);
    Pbar[4 * N + c] = l[4] / (gam - 1.0);
  }
}

}  // namespace mose
