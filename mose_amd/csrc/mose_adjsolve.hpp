// mose_adjsolve.hpp — kernels of the steady adjoint solve (included by mose_hip.hip).
//
// The steady state U* of the local-time-step iteration satisfies R(U*) = 0.  For an objective
// J(U*, X) the adjoint psi solves (dR/dU)^T psi = (dJ/dU)^T, and then
// dJ/dX = dJ/dX|_U - psi^T dR/dX for any input X of the residual (nodes, inflow record).
// The solve is the transpose of the primal iteration itself: the primal stage is
// U <- U^n + c_s (U - U^n - D R(U)) with D = diag(dt_cell / vol_cell) (Lib_Newstate.f90:134-160,
// Lib_RK.f90:15-27, local dt from Mod_dt.f90:61-85); linearised at U* its stage matrix is
// I - D A (A = dR/dU).  The same stages on lambda with the transposed matrix,
//     lambda <- (1 - c_s) lambda^n + c_s (lambda + g - A^T D lambda),
// have the spectrum of the primal's (A^T D = (D A)^T), so they converge whenever the primal
// iteration converges at U*, to A^T (D lambda) = g: psi = D lambda.  A^T r is the residual VJP
// in primitive variables, (dR/dP)^T r, taken back to conserved ones per cell with (dP/dU)^T.
#pragma once

namespace mose {

// One pseudo-time stage of the adjoint iteration over the interior cells.  Per cell c:
//   x = gP - Pbar (the P-space adjoint residual, Pbar = (dR/dP)^T D lambda_in from the VJP),
//   Q = (dP/dU)^T x at U*,
//   lambda_out = a0 lambda0 + b0 (lambda_in + Q),  Rbar = D lambda_out (the next VJP's input).
// lambda_in / lambda_out may alias (each cell reads its own entries before writing them).
// partial[blockIdx.x] receives the block's sum of |Q|^2 (fixed order; grid fixed by the host).
__global__ void __launch_bounds__(256) k_adjsolve_rk(const int64_t* __restrict__ cells, const double* __restrict__ D,
                                                     int64_t n, int64_t N, const double* __restrict__ U,
                                                     const double* __restrict__ gP, const double* __restrict__ Pbar,
                                                     const double* lam0, const double* lamIn, double* lamOut,
                                                     double* __restrict__ Rbar, double a0, double b0, double gam,
                                                     double* __restrict__ partial) {
  const double gm1 = gam - 1.0;
  double q2 = 0.0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    double u[5], x[5], y[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      u[v] = U[v * N + c];
      x[v] = gP[v * N + c] - Pbar[v * N + c];
    }
    const double ir = 1.0 / u[0];
    const double uu = u[1] * ir, vv = u[2] * ir, ww = u[3] * ir;
    // rows of dP/dU: d rho = (1,0,0,0,0); d u = (-u, 1, 0, 0, 0)/rho (v, w alike);
    // d p = (gamma-1) (|u|^2/2, -u, -v, -w, 1)
    y[0] = x[0] - (uu * x[1] + vv * x[2] + ww * x[3]) * ir + gm1 * 0.5 * (uu * uu + vv * vv + ww * ww) * x[4];
    y[1] = x[1] * ir - gm1 * uu * x[4];
    y[2] = x[2] * ir - gm1 * vv * x[4];
    y[3] = x[3] * ir - gm1 * ww * x[4];
    y[4] = gm1 * x[4];
    const double d = D[q];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double l = a0 * lam0[v * N + c] + b0 * (lamIn[v * N + c] + y[v]);
      lamOut[v * N + c] = l;
      Rbar[v * N + c] = d * l;
      q2 += y[v] * y[v];
    }
  }
  if (!partial) return;
  __shared__ double sred[4];
  for (int o = 32; o > 0; o >>= 1) q2 += __shfl_down(q2, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sred[wave] = q2;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sred[0] + sred[1] + sred[2] + sred[3];
}

// Fixed-order sum of n partials into *out (one workgroup of 256 threads).
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partial, int n, double* out) {
  __shared__ double s[256];
  double x = 0.0;
  for (int q = threadIdx.x; q < n; q += 256) x += partial[q];
  s[threadIdx.x] = x;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace mose
