// mose_stage.hpp — the fused flux + SSP-RK stage kernel and the boundary-face kernel
// (gfx950), included by mose_hip.hip.
//
// Stage kernel.  A workgroup of TXc x TYc threads owns a TXc x TYc patch of (i, j) columns
// of one block and marches along k over a chunk [k0, k1].  Per k-plane:
//   * the plane's primitive state with a 2-cell i/j halo is staged in LDS; the k-direction
//     neighbours of each column live in registers, prefetched one plane ahead;
//   * every cell computes its limited slope once per direction, sigma = phi(s+, s-), and its
//     two face states P -/+ sigma dl/2 (the reference's per-face reconstruction,
//     Lib_Reconstruction.f90:24-28, evaluated once per cell: the right state of face i-1/2 and
//     the left state of face i+1/2 come from the same limiter call);
//   * each thread solves the Riemann problems of its own cell's low i and j faces (the "+"
//     state of the neighbour comes through LDS) and of the k face above it (registers only);
//     the faces on the patch's high edges are spread over the workgroup;
//   * faces on a block boundary of non-connection type take the flux computed by
//     k_bcface (the BC_Fluxes pass);
//   * each cell adds its six face fluxes in FP64 and takes the RK stage update
//     (Lib_Newstate.f90:134-160, Lib_RK.f90:15-27).
#pragma once

namespace mose {

// Limited slope of one cell along one axis and its two face states.
// rlm / rlp: reciprocal centre distances to the low / high neighbour, hdl = dl/2.
template <int LIM, typename FT>
__device__ __forceinline__ void cell_states(const FT* Pm, const FT* P0, const FT* Pp, FT rlm, FT rlp,
                                            FT hdl, FT* Sm, FT* Sp) {
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const FT sm = (P0[v] - Pm[v]) * rlm;
    const FT sp = (Pp[v] - P0[v]) * rlp;
    const FT sg = limiter<LIM>(sp, sm) * hdl;
    Sm[v] = P0[v] - sg;
    Sp[v] = P0[v] + sg;
  }
}

// The same with MUSCL-SD (Lib_Reconstruction.f90:26-27): the slope of the low face state is
// scaled by bm, that of the high face state by bp -- the shock detector's weights of the two
// faces (the weight of a face is that of its low cell, or of the block's own boundary cell on a
// block face).
template <int LIM, bool SDR, typename FT>
__device__ __forceinline__ void cell_states_b(const FT* Pm, const FT* P0, const FT* Pp, FT rlm, FT rlp, FT hdl,
                                              FT bm, FT bp, FT* Sm, FT* Sp) {
  if constexpr (!SDR) {
    cell_states<LIM>(Pm, P0, Pp, rlm, rlp, hdl, Sm, Sp);
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const FT sm = (P0[v] - Pm[v]) * rlm;
      const FT sp = (Pp[v] - P0[v]) * rlp;
      const FT sg = limiter<LIM>(sp, sm) * hdl;
      Sm[v] = P0[v] - bm * sg;
      Sp[v] = P0[v] + bp * sg;
    }
  }
}

template <typename FT>
__device__ __forceinline__ void prim_at(const double* __restrict__ U, unsigned N, unsigned c, double gm1, FT* P) {
  double u[5], p[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u[v] = U[v * N + c];
  cons2prim(u, gm1, p);
#pragma unroll
  for (int v = 0; v < 5; ++v) P[v] = (FT)p[v];
}

// Boundary faces of non-connection type (BC_Fluxes, Mod_BC_Fluxes.f90:10-218): one thread
// per boundary face cell; the flux per unit area goes to bcbuf.
template <int LIM, int RS, typename FT>
__global__ void __launch_bounds__(256) k_bcface(StageArgs a, const BcFace* __restrict__ faces, int nfaces) {
  const double gm1 = a.gam - 1.0;
  const GasT<FT> g{FT(a.gam), FT(gm1), FT(1.0 / gm1), FT(a.R), FT(a.gam * a.R / gm1)};
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nfaces; q += gridDim.x * blockDim.x) {
    const BcFace e = faces[q];
    const BlockDev* Bp = a.blocks + e.b;
    const bool hi = e.f & 1;
    const int bct = Bp->bct[e.f];
    FT F[5] = {FT(0), FT(0), FT(0), FT(0), FT(0)};
    if (bct != F_NONE) {
      FT Pb[5], Pi[5];
      prim_at(a.Uin, (unsigned)a.npool, (unsigned)e.cb, gm1, Pb);
      prim_at(a.Uin, (unsigned)a.npool, (unsigned)e.cn, gm1, Pi);
      const FT nx = FT(e.n[0]), ny = FT(e.n[1]), nz = FT(e.n[2]);
      if (bct == F_EXTRAP) {
        physical_flux(Pb, nx, ny, nz, g, F);
      } else {
        bool done = false;
        if (bct == F_INFLOW) {
          done = inflow405_flux<LIM, RS>(Pb, Pi, hi, nx, ny, nz, a.facebc + 8 * Bp->bcd[e.f], g, F);
        } else if (bct == F_INLET || bct == F_OUTFLOW) {   // FP64 boundary state
          double P64[5], F64[5];
          prim_at(a.Uin, (unsigned)a.npool, (unsigned)e.cb, gm1, P64);
          const double* rec = a.facebc + 8 * Bp->bcd[e.f];
          done = bct == F_INLET ? inlet401_flux(P64, hi, e.n, e.t, rec, a.gam, a.R, F64)
                                : outflow406_flux(P64, hi, e.n, rec, a.gam, F64);
          if (done) {
#pragma unroll
            for (int v = 0; v < 5; ++v) F[v] = FT(F64[v]);
          }
        }
        if (!done) symmetry_flux<LIM, RS>(Pb, Pi, FT(e.bdl), FT(e.idl), hi, nx, ny, nz, g, F);
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) a.bcbuf[5 * (int64_t)e.slot + v] = (double)F[v];
  }
}

// Viscous flux of the connection faces (BC_Connection_Visc, Lib_BC_Fluxes_Connection.f90:
// 115-228): gradients at the boundary cell (normal difference to the ghost, central
// tangential differences of its neighbours) and at the ghost cell (central tangential
// differences of the donor cell's neighbours), one face flux from each, averaged.  One thread
// per boundary face cell; vbuf receives the face flux per unit area in the low-to-high
// convention of the stage kernel (-(F_loc + F_ghost)/2).
template <typename FT>
__global__ void __launch_bounds__(256) k_viscbc(StageArgs a, const ViscFace* __restrict__ vf, int nvf) {
  const double gm1 = a.gam - 1.0;
  const unsigned N = (unsigned)a.npool;
  const FT Rg = FT(a.R);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nvf; q += gridDim.x * blockDim.x) {
    const ViscFace e = vf[q];
    const BlockDev* Bp = a.blocks + e.b;
    FT Pl[5], Pg[5], Vl[4], Vg[4], Vt[4][4], Vd[4][4];
    prim_at(a.Uin, N, (unsigned)e.loc, gm1, Pl);
    prim_at(a.Uin, N, (unsigned)e.gh, gm1, Pg);
    auto vv = [Rg](const FT* P, FT* V) { V[0] = P[1]; V[1] = P[2]; V[2] = P[3]; V[3] = P[4] * rcp_(P[0] * Rg); };
    vv(Pl, Vl);
    vv(Pg, Vg);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      FT P[5];
      prim_at(a.Uin, N, (unsigned)e.tn[m], gm1, P);
      vv(P, Vt[m]);
      prim_at(a.Uin, N, (unsigned)e.dn[m], gm1, P);
      vv(P, Vd[m]);
    }
    const int d = e.d, ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const FT m2 = FT(e.modfm2);
    FT Gl[4][3], Gg[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const FT gn = (Vg[v] - Vl[v]) * m2, gn2 = (Vl[v] - Vg[v]) * (-m2);
      const FT la = (Vt[1][v] - Vt[0][v]) * FT(0.5), lb = (Vt[3][v] - Vt[2][v]) * FT(0.5);
      const FT ga = (Vd[1][v] - Vd[0][v]) * FT(0.5), gb = (Vd[3][v] - Vd[2][v]) * FT(0.5);
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        Gl[v][x] = x == d ? gn : (x == ta ? la : lb);
        Gg[v][x] = x == d ? gn2 : (x == ta ? ga : gb);
      }
    }
    FT M[3], Pm[5], vis[4], Fl[5], Fg[5];
#pragma unroll
    for (int x = 0; x < 3; ++x) M[x] = FT(0.5 * (e.ml[x] + e.mg[x]));
#pragma unroll
    for (int v = 0; v < 5; ++v) Pm[v] = (Pl[v] + Pg[v]) * FT(0.5);
#pragma unroll
    for (int x = 0; x < 4; ++x) vis[x] = FT(Bp->vis[x]);
    visc_face_flux<FT>(d, Gl, M, Pm, FT(1), Rg, vis, Fl);
    visc_face_flux<FT>(d, Gg, M, Pm, FT(1), Rg, vis, Fg);
#pragma unroll
    for (int v = 0; v < 5; ++v) a.vbuf[5 * (int64_t)e.slot + v] = -0.5 * ((double)Fl[v] + (double)Fg[v]);
  }
}

// Viscous flux of an interior face (Diffusive_Flux, Lib_Diffusive.f90:10-54, with
// Tangential_Gradient :57-87) from the LDS ring of viscous variables (u, v, w, T) of planes
// k-1, k, k+1 (slots s0, s1, s2; LDS coordinates shifted by one).  The face joins the cell at
// patch position (lx, ly) of plane k and its neighbour in direction d (for d = 2 the same
// column in plane k+1).  rho / p of the two cells in rl, pl, rr, pr; M the averaged
// diagonal metric.  Returns -Flux per unit area (the low-to-high face flux).
template <typename FT, int TXc, int TYc>
__device__ __forceinline__ void visc_interior(int d, int lx, int ly, const FT (*sV)[4][TYc + 2][TXc + 2], int s0,
                                              int s1, int s2, FT rl, FT pl, FT rr, FT pr, const FT* M, FT Rg,
                                              const FT* vis, FT* Fv) {
  FT G[4][3];
  const int x = lx + 1, y = ly + 1;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    FT VL, VR, p3, p4, p5, p6, p7, p8, p9, p10;
    if (d == 0) {
      VL = sV[s1][v][y][x]; VR = sV[s1][v][y][x + 1];
      p3 = sV[s1][v][y - 1][x]; p4 = sV[s1][v][y + 1][x]; p5 = sV[s1][v][y - 1][x + 1]; p6 = sV[s1][v][y + 1][x + 1];
      p7 = sV[s0][v][y][x]; p8 = sV[s2][v][y][x]; p9 = sV[s0][v][y][x + 1]; p10 = sV[s2][v][y][x + 1];
    } else if (d == 1) {
      VL = sV[s1][v][y][x]; VR = sV[s1][v][y + 1][x];
      p3 = sV[s1][v][y][x - 1]; p4 = sV[s1][v][y][x + 1]; p5 = sV[s1][v][y + 1][x - 1]; p6 = sV[s1][v][y + 1][x + 1];
      p7 = sV[s0][v][y][x]; p8 = sV[s2][v][y][x]; p9 = sV[s0][v][y + 1][x]; p10 = sV[s2][v][y + 1][x];
    } else {
      VL = sV[s1][v][y][x]; VR = sV[s2][v][y][x];
      p3 = sV[s1][v][y][x - 1]; p4 = sV[s1][v][y][x + 1]; p5 = sV[s2][v][y][x - 1]; p6 = sV[s2][v][y][x + 1];
      p7 = sV[s1][v][y - 1][x]; p8 = sV[s1][v][y + 1][x]; p9 = sV[s2][v][y - 1][x]; p10 = sV[s2][v][y + 1][x];
    }
    const FT gn = VR - VL;
    const FT gb = (p4 - p3 + p6 - p5) * FT(0.25);
    const FT gc = (p8 - p7 + p10 - p9) * FT(0.25);
    // (a, b, c) = (0, 1, 2), (1, 0, 2), (2, 0, 1)
    G[v][0] = d == 0 ? gn : gb;
    G[v][1] = d == 1 ? gn : (d == 0 ? gb : gc);
    G[v][2] = d == 2 ? gn : gc;
  }
  FT Pm[5];
  Pm[0] = (rl + rr) * FT(0.5);
  Pm[4] = (pl + pr) * FT(0.5);
  {
    const int xr = d == 0 ? x + 1 : x, yr = d == 1 ? y + 1 : y, sr = d == 2 ? s2 : s1;
#pragma unroll
    for (int c = 0; c < 3; ++c) Pm[1 + c] = (sV[s1][c][y][x] + sV[sr][c][yr][xr]) * FT(0.5);
  }
  FT F[5];
  visc_face_flux<FT>(d, G, M, Pm, FT(1), Rg, vis, F);
#pragma unroll
  for (int v = 0; v < 5; ++v) Fv[v] = -F[v];
}

// Viscous part of the k face k+1/2 of column (i, j): the connection-face buffer on a block
// face, else the interior formula from the LDS ring (rho / p of the two cells from the
// column queue Q0 = P(k), Q1 = P(k+1)).  Added to the face flux Fk (per unit area).
template <typename FT, int TXc, int TYc>
__device__ __forceinline__ void visc_kface(const BlockDev* __restrict__ Bp, int i, int j, int k, int ni, int nk, int tx,
                                           int ty, const FT (*sV)[4][TYc + 2][TXc + 2], const FT* Q0, const FT* Q1,
                                           const LineEnt* Lx, const LineEnt* Ly, const LineEnt* Lz,
                                           const double* __restrict__ vbuf, FT Rg, const FT* vis, FT* Fk) {
  if (k + 1 == 1 || k + 1 == nk + 1) {
    const int bf = k + 1 == 1 ? 4 : 5;
    if (Bp->bct[bf] == F_CONN) {
      const unsigned sl = (unsigned)Bp->bvs[bf] + (unsigned)(i - 1) + (unsigned)ni * (unsigned)(j - 1);
#pragma unroll
      for (int v = 0; v < 5; ++v) Fk[v] += (FT)vbuf[5 * sl + v];
    }
  } else {
    FT M[3];
    const double rzL = Lz[k + 1].rdl, rzR = Lz[k + 2].rdl;
    M[0] = FT(Lx[i + 1].rdl);
    M[1] = FT(Ly[j + 1].rdl);
    M[2] = FT(0.5 * (rzL + rzR));
    FT Fv[5];
    visc_interior<FT, TXc, TYc>(2, tx, ty, sV, (k + 2) % 3, (k + 3) % 3, (k + 4) % 3, Q0[0], Q0[4], Q1[0], Q1[4], M, Rg,
                                vis, Fv);
#pragma unroll
    for (int v = 0; v < 5; ++v) Fk[v] += Fv[v];
  }
}

// Flux of the face on the low side (direction d) of cell (fi, fj, fk) of block B (pool
// offset off, strides sy / sz): the boundary-face buffer on non-connection block faces,
// else the Riemann problem of (L, R) — a non-physical state raises MOSE_E_INTERFACE at the
// cell on the low side.  GEO 1: normal from the geometry arrays, flux scaled by the area.
// AX >= 0: the direction d = AX is known at compile time (the solver drops the products with
// the normal's zero components); -1: d at run time.
template <int RS, typename FT, int GEO, int AX = -1>
__device__ __forceinline__ void stage_face(int d, int fi, int fj, int fk, const FT* L, const FT* R, FT* F,
                                           const BlockDev* __restrict__ Bp, int bnum, int ni, int nj, int nk,
                                           unsigned off, unsigned sy, unsigned sz, const double* __restrict__ bcbuf,
                                           const double* __restrict__ geo, unsigned N, int* err,
                                           const GasT<FT>& g, const double* __restrict__ beta) {
  int idx, nd, t1, t2, n1;
  if (d == 0) { idx = fi; nd = ni; t1 = fj; t2 = fk; n1 = nj; }
  else if (d == 1) { idx = fj; nd = nj; t1 = fk; t2 = fi; n1 = nk; }
  else { idx = fk; nd = nk; t1 = fi; t2 = fj; n1 = ni; }
  const int bf = idx == 1 ? 2 * d : (idx == nd + 1 ? 2 * d + 1 : -1);
  const unsigned pf = off + (unsigned)(fi + 1) + sy * (unsigned)(fj + 1) + sz * (unsigned)(fk + 1);
  if (bf >= 0 && Bp->bct[bf] != F_CONN) {
    // tangential position on the face, cyclic order (d+1, d+2)
    const unsigned s = (unsigned)Bp->bcs[bf] + (unsigned)(t1 - 1) + (unsigned)n1 * (unsigned)(t2 - 1);
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = (FT)bcbuf[5 * s + v];
  } else {
    if (!(gas_ok(L) && gas_ok(R)))
      set_error(err, MOSE_E_INTERFACE, bnum, d == 0 ? fi - 1 : fi, d == 1 ? fj - 1 : fj, d == 2 ? fk - 1 : fk);
    FT bw = FT(1);
    if constexpr (rs_takes_beta(RS)) {
      // the low cell's weight; on the block's low face, that of its own boundary cell
      const unsigned st = d == 0 ? 1u : (d == 1 ? sy : sz);
      bw = FT(beta[idx == 1 ? pf : pf - st]);
    }
    if constexpr (GEO != 1 && AX >= 0) {
      riemann<RS>(L, R, NAx<FT, AX>{}, g, F, bw);
    } else {
      FT nx = d == 0 ? FT(1) : FT(0), ny = d == 1 ? FT(1) : FT(0), nz = d == 2 ? FT(1) : FT(0);
      if constexpr (GEO == 1) {
        nx = FT(geo[(size_t)(GN + 4 * d) * N + pf]);
        ny = FT(geo[(size_t)(GN + 4 * d + 1) * N + pf]);
        nz = FT(geo[(size_t)(GN + 4 * d + 2) * N + pf]);
      }
      riemann<RS>(L, R, NGen<FT>{nx, ny, nz}, g, F, bw);
    }
  }
  if constexpr (GEO == 1) {
    const FT area = FT(geo[(size_t)(GN + 4 * d + 3) * N + pf]);
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] = F[v] * area;
  }
}

// MODE 0: stage update; 1: last stage (update + residual-norm partials); 2: residual only.
// GEO 0: rectilinear blocks (geometry lines); 1: general metrics from the geometry arrays,
// with face fluxes scaled by their areas where they are computed; 2: rectilinear blocks
// with the laminar viscous terms (Fluxes_blk viscous sweeps, Mod_Fluxes.f90:101-207): the
// viscous variables (u, v, w, T) of planes k-1, k, k+1 with a one-cell halo (corners
// included) are kept in an LDS ring, and the connection faces take the viscous flux
// k_viscbc computed.
//
// Face ownership: every thread solves the low i face and the low j face of its own cell
// and the k face above it; the TX + TY faces on the high edges of the patch are solved by
// TX + TY threads whose choice rotates with the plane so no wave always carries them.  The
// k faces need only the thread's own registers.  The own cell's face states and faces are
// formed with the direction fixed at compile time (no per-lane direction selects); the slopes
// of the x-halo cells are formed by wave 0 and those of the y-halo cells by wave 1, each with
// its direction fixed; the spacing data along x and y of a tile sit in LDS.
// Occupancy floor: FP32 flux 3 waves per SIMD (2 for 512-thread patches, whose workgroup
// brings two waves to every SIMD); FP64 2 (256 registers: without the floor some variants take
// a few AGPRs beyond 256 and fall to one wave per SIMD); the viscous kernels 2 for the FP64
// 128-thread patch (its 51 KB of LDS let three workgroups share a CU; the floor costs a few
// spilled registers and is 16 % faster than one wave per SIMD), else none (the LDS of the
// 256-thread viscous patches allows one workgroup per CU).
template <int LIM, int RS, typename FT, int MODE, int TXc, int TYc, int GEO, bool SDR>
__global__ void __launch_bounds__(TXc* TYc)
    __attribute__((amdgpu_waves_per_eu(GEO == 2 ? (sizeof(FT) == 8 && TXc * TYc <= 128 ? 2 : 1) : (sizeof(FT) == 4 ? (TXc * TYc == 256 ? 3 : (TXc * TYc == 512 ? 2 : 1)) : 2))))
    k_stage(StageArgs a) {
  constexpr int NTc = TXc * TYc;
  constexpr int NW = NTc / 64;                   // waves per workgroup
  constexpr int NH = 4 * (TXc + TYc);            // halo cells of a plane
  constexpr int NE = TXc + TYc;                  // faces on the high edges of the patch
  static_assert(NH <= NTc && NTc % 64 == 0, "one halo / edge cell per thread at most");
  __shared__ FT sP[5][TYc + 4][TXc + 4];
  __shared__ FT sSx[5][TYc][TXc + 1];    // [y][x]: "+" i-state of cell x-1 (x = 0: the halo cell)
  __shared__ FT sSy[5][TYc + 1][TXc];    // [y][x]: "+" j-state of cell y-1
  __shared__ FT sMx[5][TYc];             // "-" i-state of the halo cell x = TX
  __shared__ FT sMy[5][TXc];             // "-" j-state of the halo cell y = TY
  __shared__ FT sFi[5][TYc][TXc + 1];    // [y][x]: flux of the low i face of cell x
  __shared__ FT sFj[5][TYc + 1][TXc];    // [y][x]: flux of the low j face of cell y
  constexpr bool VIS = GEO == 2;
  constexpr int VR = VIS ? 3 : 1, VY = VIS ? TYc + 2 : 1, VX = VIS ? TXc + 2 : 1;
  __shared__ FT sV[VR][4][VY][VX];       // viscous variables, ring over planes k-1, k, k+1
  constexpr int NL = (TXc > TYc ? TXc : TYc) + 3;
  __shared__ FT sRL[2][NL], sLP[2][NL];  // per-tile spacing data along x and y (rectilinear)
  __shared__ double sDel[2][NL], sDL[2][NL];   // node spacing delta and cell length dl along x, y
  // residual-norm sums: LDS for the FP32-flux kernels (register-bound), registers for FP64
  // (LDS-bound); reduced in a fixed order either way
  constexpr bool NLDS = MODE == 1 && sizeof(FT) == 4;
  __shared__ double sN[5][NLDS ? NTc : 1];
  static_assert(NTc <= kMaxTileThreads, "tile shape above the helper kernels' launch bound");

  const int tid = threadIdx.x, tx = tid % TXc, ty = tid / TXc;
  const double gm1 = a.gam - 1.0;
  const GasT<FT> g{FT(a.gam), FT(gm1), FT(1.0 / gm1), FT(a.R), FT(a.gam * a.R / gm1)};
  const unsigned N = (unsigned)a.npool;
  const double* __restrict__ Uin = a.Uin;
  // kernel-argument fields used inside the lambdas, as values (a reference to the argument
  // struct would force a private copy of it)
  const double* __restrict__ geo = a.geo;
  const double* __restrict__ bcbuf = a.bcbuf;
  const double* __restrict__ beta = a.beta;
  int* const err = a.err;
  // geometry array g at pool cell c (general metrics)
  auto GA = [geo, N](int g, unsigned c) -> double { return geo[(size_t)g * N + c]; };

  // the halo cell this thread stages (LDS coordinates, -2..TX+1 / -2..TY+1)
  int hx = 0, hy = 0;
  if (tid < 2 * TYc) { hy = tid >> 1; hx = -2 + (tid & 1); }
  else if (tid < 4 * TYc) { const int h = tid - 2 * TYc; hy = h >> 1; hx = TXc + (h & 1); }
  else if (tid < 4 * TYc + 2 * TXc) { const int h = tid - 4 * TYc; hx = h % TXc; hy = -2 + h / TXc; }
  else if (tid < NH) { const int h = tid - 4 * TYc - 2 * TXc; hx = h % TXc; hy = TYc + h / TXc; }
  // the cells just outside the patch whose slopes are needed: the 2 TY x-halo cells by the
  // first lanes of wave 0, the 2 TX y-halo cells by the first lanes of wave 1
  static_assert(2 * TYc <= 64 && 2 * TXc <= 64 && NTc >= 128, "halo slope tasks: one wave per direction");
  const bool xtask = tid < 2 * TYc, ytask = tid >= 64 && tid < 64 + 2 * TXc;
  const bool xplus = tid < TYc, yplus = tid - 64 < TXc;
  const int xlx = xplus ? -1 : TXc, xly = xtask ? (xplus ? tid : tid - TYc) : 0;
  const int yly = yplus ? -1 : TYc, ylx = ytask ? (yplus ? tid - 64 : tid - 64 - TXc) : 0;

  __shared__ int sTile;
  if (tid == 0) sTile = claim_tile(a.tq, a.nq, a.ntiles);
  lds_barrier();
  // the tile index is the same for every lane: held in a scalar register, so the tile's block
  // fields, line-table entries and plane offsets are scalar loads and scalar arithmetic
  for (int t = __builtin_amdgcn_readfirstlane(sTile); t >= 0;) {
    const Tile T = a.tiles[t];
    const BlockDev* __restrict__ Bp = a.blocks + T.b;
    const int ni = Bp->ni, nj = Bp->nj, nk = Bp->nk, sy_ = Bp->sy;
    const unsigned off = (unsigned)Bp->off, sz = (unsigned)Bp->sz;
    const int i = T.i0 + tx, j = T.j0 + ty;
    const bool colmem = (i <= ni + 2) && (j <= nj + 2);
    const bool active = (i <= ni) && (j <= nj);
    const bool iface = (i <= ni + 1) && (j <= nj);     // own low i face exists
    const bool jface = (i <= ni) && (j <= nj + 1);     // own low j face exists
    const unsigned cbase = off + (unsigned)(i + 1) + (unsigned)sy_ * (unsigned)(j + 1);
    const LineEnt* Lx = a.lines + Bp->line[0];
    const LineEnt* Ly = a.lines + Bp->line[1];
    const LineEnt* Lz = a.lines + Bp->line[2];
    const cptr<LineEnt> Lzc = (cptr<LineEnt>)Lz;     // Lz[k] is the same for every lane
    const int k0 = T.k0, k1 = T.k1;
    const int hgi = T.i0 + hx, hgj = T.j0 + hy;
    const bool hval = tid < NH && hgi <= ni + 2 && hgj <= nj + 2;
    const unsigned hbase = off + (unsigned)(hgi + 1) + (unsigned)sy_ * (unsigned)(hgj + 1);
    const int xci = T.i0 + xlx, xcj = T.j0 + xly, yci = T.i0 + ylx, ycj = T.j0 + yly;
    const bool xval = xtask && xcj <= nj && xci <= ni + 1;
    const bool yval = ytask && yci <= ni && ycj <= nj + 1;
    // spacing data of the patch's cells and of its halo cells along x and y (constant over the
    // tile): sRL[d][q] = reciprocal centre distance of the face on the low side of the cell at
    // offset q - 1 from the patch's first cell, sLP[d][q] its half length; read in the k-loop
    if constexpr (GEO != 1) {
      for (int q = tid; q < TXc + 3; q += NTc) {
        const bool in = T.i0 + q <= ni + 3;
        sRL[0][q] = in ? FT(Lx[T.i0 + q].rl1) : FT(0);
        sLP[0][q] = in ? FT(Lx[T.i0 + q].lp) : FT(0);
        sDel[0][q] = in ? Lx[T.i0 + q].delta : 0.0;
        sDL[0][q] = in ? Lx[T.i0 + q].dl : 0.0;
      }
      for (int q = tid; q < TYc + 3; q += NTc) {
        const bool in = T.j0 + q <= nj + 3;
        sRL[1][q] = in ? FT(Ly[T.j0 + q].rl1) : FT(0);
        sLP[1][q] = in ? FT(Ly[T.j0 + q].lp) : FT(0);
        sDel[1][q] = in ? Ly[T.j0 + q].delta : 0.0;
        sDL[1][q] = in ? Ly[T.j0 + q].dl : 0.0;
      }
    }


    // column queue P(k), P(k+1), P(k+2); "+" k-state of cell k; flux of the k face below;
    // prefetched U(k+3) of the column and U of the next plane's halo cell
    FT Q0[5], Q1[5], Q2[5], Skp[5], Fkm[5];
    double Un3[5], Uh[5], nsum[5];
    double dtmin_t = 1e300;     // last stage: min CFL time step of the new state
#pragma unroll
    for (int v = 0; v < 5; ++v) { Fkm[v] = FT(0); Skp[v] = FT(0); Un3[v] = 0.0; Uh[v] = 0.0; nsum[v] = 0.0; }
    if constexpr (NLDS) {
#pragma unroll
      for (int v = 0; v < 5; ++v) sN[v][tid] = 0.0;
    }
    if (colmem) {
      FT Qm[5], Sm_[5];
      prim_at(Uin, N, cbase + sz * (unsigned)(k0 - 1), gm1, Qm);     // P(k0-2)
      prim_at(Uin, N, cbase + sz * (unsigned)(k0), gm1, Q0);         // P(k0-1)
      prim_at(Uin, N, cbase + sz * (unsigned)(k0 + 1), gm1, Q1);     // P(k0)
      prim_at(Uin, N, cbase + sz * (unsigned)(k0 + 2), gm1, Q2);     // P(k0+1)
      const unsigned cm = cbase + sz * (unsigned)(k0);            // cell k0-1
      FT bp = FT(1);
      if constexpr (SDR) bp = FT(beta[k0 >= 2 ? cm : cm + sz]);
      if constexpr (GEO != 1) {
        cell_states_b<LIM, SDR>(Qm, Q0, Q1, FT(Lzc[k0].rl1), FT(Lzc[k0 + 1].rl1), FT(Lzc[k0].lp), FT(1), bp, Sm_, Skp);   // P+(k0-1)
      } else {
        cell_states_b<LIM, SDR>(Qm, Q0, Q1, FT(GA(GRLO + 2, cm)), FT(GA(GRLO + 2, cm + sz)), FT(0.5 * GA(GDL + 2, cm)),
                                FT(1), bp, Sm_, Skp);
      }
    }
    if (hval) {
#pragma unroll
      for (int v = 0; v < 5; ++v) Uh[v] = Uin[v * N + hbase + sz * (unsigned)(k0 + 1)];
    }
    // viscous path: one-cell-halo plane kk of the viscous variables into ring slot (kk+3)%3
    const double Rgas = a.R;
    const FT Rg = FT(Rgas);
    auto stage_visc = [=](int kk) {
      if constexpr (VIS) {
        const int slot = (kk + 3) % 3;
        for (int q = tid; q < VX * VY; q += NTc) {
          const int x = q % VX - 1, y = q / VX - 1;
          const int gi = T.i0 + x, gj = T.j0 + y;
          FT V[4] = {FT(0), FT(0), FT(0), FT(0)};
          if (gi <= ni + 2 && gj <= nj + 2) {
            double u[5], p[5];
            const unsigned c = off + (unsigned)(gi + 1) + (unsigned)sy_ * (unsigned)(gj + 1) + sz * (unsigned)(kk + 1);
#pragma unroll
            for (int v = 0; v < 5; ++v) u[v] = Uin[v * N + c];
            if (u[0] != 0.0) {
              cons2prim(u, gm1, p);
              V[0] = (FT)p[1]; V[1] = (FT)p[2]; V[2] = (FT)p[3]; V[3] = (FT)(p[4] / (p[0] * Rgas));
            }
          }
#pragma unroll
          for (int v = 0; v < 4; ++v) sV[slot][v][y + 1][x + 1] = V[v];
        }
      }
    };
    FT vis[4] = {FT(0), FT(0), FT(0), FT(0)};
    if constexpr (VIS) {
      stage_visc(k0 - 1);
      stage_visc(k0);
#pragma unroll
      for (int x = 0; x < 4; ++x) vis[x] = FT(Bp->vis[x]);
      lds_barrier();
    }

    // k = k0-1 only produces the face k0-1/2; planes k0..k1-1 are updated
    for (int k = k0 - 1; k < k1; ++k) {
      const bool plane = k >= k0;
      if (plane) {       // stage plane k
        if (colmem) {
#pragma unroll
          for (int v = 0; v < 5; ++v) sP[v][ty + 2][tx + 2] = Q0[v];
        }
        if (hval) {
          double p[5];
          cons2prim(Uh, gm1, p);
#pragma unroll
          for (int v = 0; v < 5; ++v) sP[v][hy + 2][hx + 2] = (FT)p[v];
        }
      }
      // prefetch the next plane of the column and of the halo
      if (colmem && k + 1 < k1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Un3[v] = Uin[v * N + cbase + sz * (unsigned)(k + 4)];
      }
      if (hval && plane && k + 1 < k1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Uh[v] = Uin[v * N + hbase + sz * (unsigned)(k + 2)];
      }

      if (plane) {
        if constexpr (VIS) stage_visc(k + 1);
        lds_barrier();       // plane k staged

        // ---- i and j face states: the own cell in each direction, then the cells just outside
        // the patch (x halo on wave 0, y halo on wave 1); every call has a fixed direction
        FT Smx[5], Smy[5];
        auto states = [&](const int d, int lx, int ly, FT* Sm_, FT* Sp_) {
          const int q = (d == 0 ? lx : ly) + 1;     // sRL / sLP index of the cell's low face
          FT rlm = FT(0), rlp = FT(0), hdl = FT(0);
          if constexpr (GEO != 1) {
            rlm = sRL[d][q];
            rlp = sRL[d][q + 1];
            hdl = sLP[d][q];
          }
          const int sx = d == 0 ? 1 : 0, sy = d == 0 ? 0 : 1;
          FT m[5], c[5], p[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            m[v] = sP[v][ly + 2 - sy][lx + 2 - sx];
            c[v] = sP[v][ly + 2][lx + 2];
            p[v] = sP[v][ly + 2 + sy][lx + 2 + sx];
          }
          FT bm = FT(1), bp = FT(1);
          const unsigned pc = off + (unsigned)(T.i0 + lx + 1) + (unsigned)sy_ * (unsigned)(T.j0 + ly + 1) +
                              sz * (unsigned)(k + 1);
          const unsigned st = d == 0 ? 1u : (unsigned)sy_;
          if constexpr (SDR) {
            const int cc = d == 0 ? T.i0 + lx : T.j0 + ly;
            bp = FT(beta[cc >= 1 ? pc : pc + st]);
            bm = FT(beta[cc >= 2 ? pc - st : pc]);
          }
          if constexpr (GEO == 1) {
            rlm = FT(GA(GRLO + d, pc));
            rlp = FT(GA(GRLO + d, pc + st));
            hdl = FT(0.5 * GA(GDL + d, pc));
          }
          cell_states_b<LIM, SDR>(m, c, p, rlm, rlp, hdl, bm, bp, Sm_, Sp_);
        };
        if (iface) {
          FT Sp_[5];
          states(0, tx, ty, Smx, Sp_);
#pragma unroll
          for (int v = 0; v < 5; ++v) sSx[v][ty][tx + 1] = Sp_[v];
        }
        if (jface) {
          FT Sp_[5];
          states(1, tx, ty, Smy, Sp_);
#pragma unroll
          for (int v = 0; v < 5; ++v) sSy[v][ty + 1][tx] = Sp_[v];
        }
        if (xval) {      // wave 0: the x-halo cells (low: its "+" i-state, high: its "-" i-state)
          FT Sm_[5], Sp_[5];
          states(0, xlx, xly, Sm_, Sp_);
          if (xplus) {
#pragma unroll
            for (int v = 0; v < 5; ++v) sSx[v][xly][0] = Sp_[v];
          } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) sMx[v][xly] = Sm_[v];
          }
        }
        if (yval) {      // wave 1: the y-halo cells
          FT Sm_[5], Sp_[5];
          states(1, ylx, yly, Sm_, Sp_);
          if (yplus) {
#pragma unroll
            for (int v = 0; v < 5; ++v) sSy[v][0][ylx] = Sp_[v];
          } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) sMy[v][ylx] = Sm_[v];
          }
        }
        lds_barrier();

        // ---- one Riemann problem per i / j face: the own low faces (fixed directions), then
        // the faces on the patch's high edges (a rotating wave)
        // ax: std::integral_constant of the face axis (own faces) or of -1 (edge faces, d at run time)
        auto solve = [&](auto ax, const int d, int fi, int fj, const FT* Lq, const FT* Rq, FT* F) {
          stage_face<RS, FT, GEO, decltype(ax)::value>(d, fi, fj, k, Lq, Rq, F, Bp, T.b + 1, ni, nj, nk, off, (unsigned)sy_, sz, bcbuf, geo, N, err, g, beta);
          if constexpr (VIS) {
            if (Bp->visc) {
              // cell on the low side of the face, patch coordinates
              const int lxL = fi - T.i0 - (d == 0 ? 1 : 0), lyL = fj - T.j0 - (d == 1 ? 1 : 0);
              const int idx = d == 0 ? fi : fj, nd = d == 0 ? ni : nj;
              if (idx == 1 || idx == nd + 1) {
                const int bf = idx == 1 ? 2 * d : 2 * d + 1;
                if (Bp->bct[bf] == F_CONN) {
                  const int t1 = d == 0 ? fj : k, t2 = d == 0 ? k : fi, n1 = d == 0 ? nj : nk;
                  const unsigned sl = (unsigned)Bp->bvs[bf] + (unsigned)(t1 - 1) + (unsigned)n1 * (unsigned)(t2 - 1);
#pragma unroll
                  for (int v = 0; v < 5; ++v) F[v] += (FT)a.vbuf[5 * sl + v];
                }
              } else {
                const int s1 = (k + 3) % 3, s0 = (k + 2) % 3, s2 = (k + 4) % 3;
                const FT rl = sP[0][lyL + 2][lxL + 2], pl = sP[4][lyL + 2][lxL + 2];
                const FT rr = sP[0][lyL + 2 + (d == 1)][lxL + 2 + (d == 0)], pr = sP[4][lyL + 2 + (d == 1)][lxL + 2 + (d == 0)];
                // averaged diagonal metric 1 / dl of the two cells
                FT M[3];
                const double rxL = Lx[fi - (d == 0) + 1].rdl, rxR = Lx[fi + 1].rdl;
                const double ryL = Ly[fj - (d == 1) + 1].rdl, ryR = Ly[fj + 1].rdl;
                const double rz = Lzc[k + 1].rdl;
                M[0] = FT(0.5 * (rxL + rxR)); M[1] = FT(0.5 * (ryL + ryR)); M[2] = FT(rz);
                FT Fv[5];
                visc_interior<FT, TXc, TYc>(d, lxL, lyL, sV, s0, s1, s2, rl, pl, rr, pr, M, Rg, vis, Fv);
#pragma unroll
                for (int v = 0; v < 5; ++v) F[v] += Fv[v];
              }
            }
          }
        };
        if (iface) {
          FT Lq[5], F[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) Lq[v] = sSx[v][ty][tx];
          solve(std::integral_constant<int, 0>{}, 0, i, j, Lq, Smx, F);
#pragma unroll
          for (int v = 0; v < 5; ++v) sFi[v][ty][tx] = F[v];
        }
        if (jface) {
          FT Lq[5], F[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) Lq[v] = sSy[v][ty][tx];
          solve(std::integral_constant<int, 1>{}, 1, i, j, Lq, Smy, F);
#pragma unroll
          for (int v = 0; v < 5; ++v) sFj[v][ty][tx] = F[v];
        }
        const int e = (tid + 64 * (k & (NW - 1))) % NTc;     // rotating patch-edge assignment
        if (e < NE) {
          FT Lq[5], Rq[5], F[5];
          const bool ex = e < TYc;
          const int d = ex ? 0 : 1;
          const int fi = ex ? T.i0 + TXc : T.i0 + e - TYc, fj = ex ? T.j0 + e : T.j0 + TYc;
          if (ex ? (fi <= ni + 1 && fj <= nj) : (fi <= ni && fj <= nj + 1)) {
#pragma unroll
            for (int v = 0; v < 5; ++v) {
              Lq[v] = ex ? sSx[v][e][TXc] : sSy[v][TYc][e - TYc];
              Rq[v] = ex ? sMx[v][e] : sMy[v][e - TYc];
            }
            solve(std::integral_constant<int, -1>{}, d, fi, fj, Lq, Rq, F);
            if (ex) {
#pragma unroll
              for (int v = 0; v < 5; ++v) sFi[v][e][TXc] = F[v];
            } else {
#pragma unroll
              for (int v = 0; v < 5; ++v) sFj[v][TYc][e - TYc] = F[v];
            }
          }
        }
        lds_barrier();

        // operands of the update of plane k, loaded before the k face is solved so that the
        // solve covers their latency (rectilinear inviscid stages except the last, whose
        // residual-norm terms leave no registers for them, and the HLLE++ / HLLC+ solvers,
        // whose registers are taken too: early, the loads would spill)
        const unsigned c = cbase + sz * (unsigned)(k + 1);
        constexpr bool EARLY = GEO == 0 && MODE == 0 && !rs_takes_beta(RS);
        double un[5], ui[5], dt = 0.0;
        if constexpr (EARLY) {
          if (active) {
#pragma unroll
            for (int v = 0; v < 5; ++v) { un[v] = a.Un[v * N + c]; ui[v] = Uin[v * N + c]; }
            dt = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
          }
        }

        // ---- k face k+1/2 of the own column, entirely in registers
        FT Fk[5];
        if (active) {
          FT Sm_[5], Sp_[5];
          const unsigned c1 = cbase + sz * (unsigned)(k + 2);      // cell k+1
          FT bm = FT(1), bp = FT(1);
          if constexpr (SDR) {
            bp = FT(beta[c1]);
            bm = FT(beta[k + 1 >= 2 ? c1 - sz : c1]);
          }
          if constexpr (GEO != 1) {
            cell_states_b<LIM, SDR>(Q0, Q1, Q2, FT(Lzc[k + 2].rl1), FT(Lzc[k + 3].rl1), FT(Lzc[k + 2].lp), bm, bp, Sm_, Sp_);
          } else {
            cell_states_b<LIM, SDR>(Q0, Q1, Q2, FT(GA(GRLO + 2, c1)), FT(GA(GRLO + 2, c1 + sz)), FT(0.5 * GA(GDL + 2, c1)),
                                    bm, bp, Sm_, Sp_);
          }
          stage_face<RS, FT, GEO, 2>(2, i, j, k + 1, Skp, Sm_, Fk, Bp, T.b + 1, ni, nj, nk, off, (unsigned)sy_, sz, bcbuf, geo, N, err, g, beta);
          if constexpr (VIS) {
            if (Bp->visc)
              visc_kface<FT, TXc, TYc>(Bp, i, j, k, ni, nk, tx, ty, sV, Q0, Q1, Lx, Ly, Lz, a.vbuf, Rg, vis, Fk);
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) Skp[v] = Sp_[v];
        }

        // ---- residual of the own cell and the RK update of plane k
        if (active) {
          const double dx = GEO != 1 ? sDel[0][tx + 1] : 1.0, dy = GEO != 1 ? sDel[1][ty + 1] : 1.0;
          const double dz = GEO != 1 ? Lzc[k + 1].delta : 1.0;
          const double Az = dx * dy, Ax = dy * dz, Ay = dx * dz;
          double Racc[5];
#pragma unroll
          for (int v = 0; v < 5; ++v)
            Racc[v] = Az * ((double)Fk[v] - (double)Fkm[v]) +
                      Ax * ((double)sFi[v][ty][tx + 1] - (double)sFi[v][ty][tx]) +
                      Ay * ((double)sFj[v][ty + 1][tx] - (double)sFj[v][ty][tx]);
          if constexpr (MODE == 2) {
            if (a.irs_scale) {
              const double dtc = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
              const double vol = GEO != 1 ? dx * dy * dz : GA(GVOL, c);
#pragma unroll
              for (int v = 0; v < 5; ++v) a.Rout[v * N + c] = -Racc[v] / vol * dtc;
            } else {
#pragma unroll
              for (int v = 0; v < 5; ++v) a.Rout[v * N + c] = Racc[v];
            }
          } else {
            if constexpr (!EARLY) {
#pragma unroll
              for (int v = 0; v < 5; ++v) { un[v] = a.Un[v * N + c]; ui[v] = Uin[v * N + c]; }
              dt = a.local_dt ? a.dtloc[c] : a.scal[S_DT];
            }
            double unew[5], P[5];
            const double vol = GEO != 1 ? dx * dy * dz : GA(GVOL, c);
            const double rv = dt / vol;
#pragma unroll
            for (int v = 0; v < 5; ++v) unew[v] = un[v] + a.rkc * (ui[v] - un[v] - Racc[v] * rv);
            cons2prim(unew, gm1, P);
            bool bad = !(P[4] >= 0.0);
#pragma unroll
            for (int v = 0; v < 5; ++v) bad = bad || (P[v] != P[v]);
            if (bad) set_error(a.err, MOSE_E_STATE, T.b + 1, i, j, k);
            if (P[0] < 1e-20) { P[0] = 1e-20; prim2cons(P, 1.0 / gm1, unew); }
            if constexpr (MODE == 1) {
              // CFL time step of the new state for the next step (Compute_dt, Mod_dt.f90:61-85)
              const double snd = sqrt(a.gam * P[4] / P[0]);
              double dtc = 1e8;
              if constexpr (GEO != 1) {
                dtc = fmin(sDL[0][tx + 1] / (fabs(P[1]) + snd) * a.cfl, dtc);
                if (a.ndir > 1) dtc = fmin(sDL[1][ty + 1] / (fabs(P[2]) + snd) * a.cfl, dtc);
                if (a.ndir > 2) dtc = fmin(Lzc[k + 1].dl / (fabs(P[3]) + snd) * a.cfl, dtc);
                if (GEO == 2 && Bp->visc) {   // VNN condition (Mod_dt.f90:88-103)
                  const double mi = sutherland(P[4] / (P[0] * a.R), Bp->vis[0], Bp->vis[1], Bp->vis[2]);
                  const double dls[3] = {sDL[0][tx + 1], sDL[1][ty + 1], Lzc[k + 1].dl};
                  for (int dd = 0; dd < a.ndir; ++dd) dtc = fmin(P[0] * dls[dd] * dls[dd] * Bp->vis[4] / mi, dtc);
                }
              } else {
                for (int dd = 0; dd < a.ndir; ++dd) {
                  const double vel = fabs(P[1] * GA(GVS + 3 * dd, c) + P[2] * GA(GVS + 3 * dd + 1, c) +
                                          P[3] * GA(GVS + 3 * dd + 2, c));
                  dtc = fmin(GA(GDL + dd, c) / (vel + snd) * a.cfl, dtc);
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
                if constexpr (NLDS) sN[v][tid] += dd * dd;
                else nsum[v] += dd * dd;
              }
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) a.Uout[v * N + c] = unew[v];
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) Fkm[v] = Fk[v];
        }
      }
      if (!plane) {      // k = k0-1: only the face k0-1/2
        // ---- k face k+1/2 of the own column, entirely in registers
        FT Fk[5];
        if (active) {
          FT Sm_[5], Sp_[5];
          const unsigned c1 = cbase + sz * (unsigned)(k + 2);      // cell k+1
          FT bm = FT(1), bp = FT(1);
          if constexpr (SDR) {
            bp = FT(beta[c1]);
            bm = FT(beta[k + 1 >= 2 ? c1 - sz : c1]);
          }
          if constexpr (GEO != 1) {
            cell_states_b<LIM, SDR>(Q0, Q1, Q2, FT(Lzc[k + 2].rl1), FT(Lzc[k + 3].rl1), FT(Lzc[k + 2].lp), bm, bp, Sm_, Sp_);
          } else {
            cell_states_b<LIM, SDR>(Q0, Q1, Q2, FT(GA(GRLO + 2, c1)), FT(GA(GRLO + 2, c1 + sz)), FT(0.5 * GA(GDL + 2, c1)),
                                    bm, bp, Sm_, Sp_);
          }
          stage_face<RS, FT, GEO, 2>(2, i, j, k + 1, Skp, Sm_, Fk, Bp, T.b + 1, ni, nj, nk, off, (unsigned)sy_, sz, bcbuf, geo, N, err, g, beta);
          if constexpr (VIS) {
            if (Bp->visc)
              visc_kface<FT, TXc, TYc>(Bp, i, j, k, ni, nk, tx, ty, sV, Q0, Q1, Lx, Ly, Lz, a.vbuf, Rg, vis, Fk);
          }
#pragma unroll
          for (int v = 0; v < 5; ++v) Skp[v] = Sp_[v];
        }
        if (active) {
#pragma unroll
          for (int v = 0; v < 5; ++v) Fkm[v] = Fk[v];
        }
      }
      // advance the column queue
      if (colmem && k + 1 < k1) {
        double p[5];
        cons2prim(Un3, gm1, p);
#pragma unroll
        for (int v = 0; v < 5; ++v) { Q0[v] = Q1[v]; Q1[v] = Q2[v]; Q2[v] = (FT)p[v]; }
      }
    }
    lds_barrier();

    if constexpr (MODE == 1) {   // fixed-order reductions over the workgroup: deterministic
      double tot = 0.0;
      if constexpr (NLDS) {
        for (int w = NTc / 2; w > 0; w >>= 1) {
          if (tid < w) {
#pragma unroll
            for (int v = 0; v < 5; ++v) sN[v][tid] += sN[v][tid + w];
          }
          lds_barrier();
        }
        if (tid < 5) tot = sN[tid][0];
      } else {
        const int lane = tid & 63, wave = tid >> 6;
        double* sW = reinterpret_cast<double*>(&sSx[0][0][0]);     // face states are dead here
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          double x = nsum[v];
          for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
          if (lane == 0) sW[5 * wave + v] = x;
        }
        lds_barrier();
        if (tid < 5)
          for (int w = 0; w < NW; ++w) tot += sW[5 * w + tid];
      }
      if (tid < 5) a.partials[5 * (int64_t)t + tid] = tot;
      for (int o = 32; o > 0; o >>= 1) dtmin_t = fmin(dtmin_t, __shfl_xor(dtmin_t, o, 64));
      if ((tid & 63) == 0 && dtmin_t < 1e300)
        atomicMin(a.dtmin, (unsigned long long)__double_as_longlong(dtmin_t));
      lds_barrier();
    }
    // every thread has read sTile before the barriers above; claim the next tile
    if (tid == 0) sTile = claim_tile(a.tq, a.nq, a.ntiles);
    lds_barrier();
    t = __builtin_amdgcn_readfirstlane(sTile);
  }
  if (tid == 0) finish_tiles(a.tq);
}

}  // namespace mose
