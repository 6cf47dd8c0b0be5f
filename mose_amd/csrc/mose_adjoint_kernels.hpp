// mose_adjoint_kernels.hpp — the adjoint passes that do not depend on the limiter or the
// Riemann solver (state to primitives, transpose of the ghost fill, viscous terms).  Included
// by mose_hip.hip only; the limiter- and solver-templated passes are in mose_adjoint.hpp.
#pragma once

// Primitive variables of every pool cell (ghost layers filled), read by the adjoint passes.
__global__ void __launch_bounds__(256) k_adj_prim(const double* __restrict__ U, double* __restrict__ Pp, int64_t N,
                                                  double gam, double* __restrict__ Tp, double Rg) {
  const double gm1 = gam - 1.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
    double u[5], P[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = U[v * N + c];
    if (u[0] != 0.0) cons2prim(u, gm1, P);
#pragma unroll
    for (int v = 0; v < 5; ++v) Pp[v * N + c] = P[v];
    // viscous layouts: the temperature the viscous adjoints read for every stencil cell
    if (Tp) Tp[c] = P[0] != 0.0 ? P[4] / (P[0] * Rg) : 0.0;
  }
}

// Transpose of the ghost fill: each source cell adds M^T of the ghost adjoints it feeds.
__global__ void __launch_bounds__(256) k_adj_ghostT(AdjArgs a) {
  const int64_t N = a.npool;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < a.ngrows; q += gridDim.x * blockDim.x) {
    const AdjGhostRow row = a.grows[q];
    double acc[5] = {0, 0, 0, 0, 0};
    for (int e = row.begin; e < row.end; ++e) {
      const AdjGhostEnt E = a.gents[e];
      double gb[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) gb[v] = a.Pbar[v * N + E.ghost];
      double enb[3] = {0.0, 0.0, 0.0};
      if (E.kind == 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) acc[v] += E.w * gb[v];
      } else {   // velocity mirror (I - 2 n n^T), symmetric
        const double dn = gb[1] * E.n[0] + gb[2] * E.n[1] + gb[3] * E.n[2];
        acc[0] += gb[0];
        acc[1] += gb[1] - 2.0 * dn * E.n[0];
        acc[2] += gb[2] - 2.0 * dn * E.n[1];
        acc[3] += gb[3] - 2.0 * dn * E.n[2];
        acc[4] += gb[4];
        if (a.GE) {   // d/dn of gb . (u - 2 (u.n) n)
          const double u1 = a.Pp[N + row.src], u2 = a.Pp[2 * N + row.src], u3 = a.Pp[3 * N + row.src];
          const double un = u1 * E.n[0] + u2 * E.n[1] + u3 * E.n[2];
          enb[0] = -2.0 * (un * gb[1] + dn * u1);
          enb[1] = -2.0 * (un * gb[2] + dn * u2);
          enb[2] = -2.0 * (un * gb[3] + dn * u3);
        }
      }
      if (a.GE) {
#pragma unroll
        for (int r = 0; r < 3; ++r) a.GE[(size_t)r * a.nents + e] = enb[r];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) a.Pbar[v * N + row.src] += acc[v];
  }
}

// =============================================================================== viscous
// Viscous terms of rectilinear viscous blocks (Diffusive_Flux interior faces and
// BC_Connection_Visc connection faces).  The interior part is deterministic: k_adj_visc
// reverses every interior face (both cells interior) and gathers the ten stencil roles each
// interior cell plays (low / high cell, the eight tangential neighbours) in a fixed order,
// converting (u, v, w, T) adjoints to primitive ones; it stores per face the mu_ref adjoint at
// the face's low cell (row 17 of VW) and, for the faces whose stencils reach ghost cells, the
// adjoints of the index-space gradients Gb[4][3] and of the face state Pmb[5] (rows 0-16),
// which k_adj_vcell (one thread per first-layer ghost cell) gathers the same way.  Connection faces (one per
// boundary face cell) reach cells of the donor block: k_adj_vbc stores the primitive adjoints of
// each face's ten stencil cells, and k_adj_vbc_gather adds them to every receiving cell in a
// fixed order (face, then role), so the result does not depend on the schedule.

// Stencil of the interior face on the low side of cell (i, j, k) in direction d: the two
// tangential axes in increasing order.
__device__ __forceinline__ void visc_axes(int d, int& tb, int& tc) {
  tb = d == 0 ? 1 : 0;
  tc = d == 2 ? 1 : 2;
}

// Adjoint of one interior viscous face (Diffusive_Flux, Lib_Diffusive.f90:10-54) of block B in
// direction d with high cell ijk: from the primitive states PL, PR and the residual adjoints
// RL, RR of its two cells and the viscous variables (u, v, w, T) of its ten stencil cells V[m]
// (L, R, L -/+ e_b, R -/+ e_b, L -/+ e_c, R -/+ e_c), the adjoints of its index-space gradients
// Gb, of its face state Pmb and (returned) of mu_ref.
__device__ __forceinline__ double vface_core(const AdjArgs& a, const BlockDev& B, int d, const int* ijk,
                                             const double* PL, const double* PR, const double* RL, const double* RR,
                                             const double (*V)[4], double (*Gb)[3], double* Pmb) {
  int tb, tc;
  visc_axes(d, tb, tc);
  double G[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    G[v][d] = V[1][v] - V[0][v];
    G[v][tb] = (V[3][v] - V[2][v] + V[5][v] - V[4][v]) * 0.25;
    G[v][tc] = (V[7][v] - V[6][v] + V[9][v] - V[8][v]) * 0.25;
  }
  // averaged diagonal metric and face area from the per-line data
  const LineEnt* Ln[3] = {a.lines + B.line[0], a.lines + B.line[1], a.lines + B.line[2]};
  double M[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double mR = Ln[x][ijk[x] + 1].rdl;
    M[x] = x == d ? 0.5 * (Ln[x][ijk[x]].rdl + mR) : mR;
  }
  const double area = Ln[tb][ijk[tb] + 1].delta * Ln[tc][ijk[tc] + 1].delta;
  double Pm[5];
  Pm[0] = (PL[0] + PR[0]) * 0.5;
  Pm[4] = (PL[4] + PR[4]) * 0.5;
#pragma unroll
  for (int c = 0; c < 3; ++c) Pm[1 + c] = (V[0][c] + V[1][c]) * 0.5;
  // the face adds -area F_visc to the low cell and +area F_visc to the high cell
  double Fb[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) Fb[v] = -area * (RL[v] - RR[v]);
  const double vis[4] = {B.vis[0], B.vis[1], B.vis[2], B.vis[3]};
#pragma unroll
  for (int v = 0; v < 5; ++v) Pmb[v] = 0.0;
  return visc_face_flux_adj(d, G, M, Pm, a.R, vis, Fb, Gb, Pmb);
}

// Interior viscous faces, reversed and gathered in one pass.  A workgroup owns a kVX x kVY
// patch of (i, j) columns of one viscous block over the planes [k0, k1) and marches along k:
// the cells' (rho, u, v, w, p, T) and residual adjoints of three planes with a two-cell halo in
// i and j sit in an LDS ring, the next plane's values already loaded into registers while the
// current plane is worked on; for each face plane q (the in-plane faces of plane q and the k
// faces between q-1 and q) and each direction, one thread per face position forms the face's
// adjoints (vface_core) for the patch's faces and the ring of faces around it whose stencils
// reach the patch, into an LDS face buffer (zeros where no interior face is); then every
// thread adds the roles its own cell plays in those faces (low / high cell: the normal gradient
// and half the face state; the four tangential neighbours: a quarter of a tangential gradient)
// to accumulators of planes q-1, q and q+1 in registers, in a fixed order.  Plane q-1 is then
// complete: its (u, v, w, T) adjoints are turned into primitive ones and stored in VPb, which
// k_adj_cell adds to the convective part (the two run on separate streams, concurrently).  Each
// face is formed by every patch whose cells it reaches (the halo faces twice or more) but
// written out only by the patch that holds its high cell: the mu_ref adjoint (row 17) always,
// and the whole record for faces whose stencil reaches a ghost cell, which k_adj_vcell gathers
// for the ghost cells.
// 16 x 16 patches; threads: the face positions of the direction with the most (17 x 18,
// 18 x 17, 2 x 16 + 18 x 16).  (A 14 x 14 patch, whose directions fit 256 threads, measured slower.)
constexpr int kVX = 16, kVY = 16;
constexpr int kVT = 320;
static_assert((kVX + 1) * (kVY + 2) <= kVT && (kVX + 2) * (kVY + 1) <= kVT && 2 * kVX + (kVX + 2) * kVY <= kVT,
              "one face position per thread");
__global__ void __launch_bounds__(kVT) k_adj_visc(AdjArgs a, const Tile* __restrict__ tiles) {
  constexpr int CX = kVX + 4, CY = kVY + 4;    // cell region x in [-2, kVX + 1], y in [-2, kVY + 1]
  constexpr int FX = kVX + 2, FY = kVY + 2;    // face positions x in [-1, kVX], y in [-1, kVY]
  constexpr int NST = (CX * CY + kVT - 1) / kVT;   // region cells staged per thread
  __shared__ double sC[3][6][CY][CX];
  __shared__ double sR[3][5][CY][CX];
  __shared__ double sF[17][FY][FX];
  const int tid = threadIdx.x;
  const Tile T = tiles[blockIdx.x];
  const BlockDev& B = a.blocks[T.b];
  const int64_t N = a.npool;
  const int ni = B.ni, nj = B.nj, nk = B.nk;
  const int64_t sy = B.sy, sz = B.sz;
  const int n3[3] = {ni, nj, nk};
  // plane kk (cell index, -1 <= kk <= nk + 2): its region cells into registers (load), then
  // into ring slot (kk + 3) % 3 (store); cells beyond the block's padded extent hold zeros
  double st[NST][11];
  auto load = [&](int kk) {
#pragma unroll
    for (int r = 0; r < NST; ++r) {
      const int q = tid + r * kVT;
      const int x = q % CX - 2, y = q / CX - 2;
      const int gi = T.i0 + x, gj = T.j0 + y;
#pragma unroll
      for (int v = 0; v < 11; ++v) st[r][v] = 0.0;
      if (q < CX * CY && gi <= ni + 2 && gj <= nj + 2) {
        const int64_t c = B.off + (gi + 1) + sy * (gj + 1) + sz * (kk + 1);
#pragma unroll
        for (int v = 0; v < 5; ++v) st[r][v] = a.Pp[v * N + c];
        st[r][5] = a.Tp[c];
#pragma unroll
        for (int v = 0; v < 5; ++v) st[r][6 + v] = a.Rbar[v * N + c];
      }
    }
  };
  auto store = [&](int kk) {
    const int slot = (kk + 3) % 3;
#pragma unroll
    for (int r = 0; r < NST; ++r) {
      const int q = tid + r * kVT;
      if (q < CX * CY) {
        const int x = q % CX, y = q / CX;
#pragma unroll
        for (int v = 0; v < 6; ++v) sC[slot][v][y][x] = st[r][v];
#pragma unroll
        for (int v = 0; v < 5; ++v) sR[slot][v][y][x] = st[r][6 + v];
      }
    }
  };
  // the thread's own cell (patch column tx, ty) and its accumulators of planes q-1, q, q+1:
  // (u, v, w, T) adjoints, rho and p adjoints
  const bool owner = tid < kVX * kVY;
  const int tx = tid % kVX, ty = tid / kVX;
  const int oi = T.i0 + tx, oj = T.j0 + ty;
  const bool own = owner && oi <= ni && oj <= nj;
  double am[6], a0[6], ap[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) { am[v] = 0.0; a0[v] = 0.0; ap[v] = 0.0; }
  auto F = [&](int r, int x, int y) -> double { return sF[r][y + 1][x + 1]; };

  load(T.k0 - 2);
  store(T.k0 - 2);
  load(T.k0 - 1);
  store(T.k0 - 1);
  load(T.k0);
  for (int q = T.k0 - 1; q <= T.k1; ++q) {
    store(q + 1);
    lds_barrier();
    if (q + 2 <= T.k1 + 1) load(q + 2);      // in flight while this plane's faces are formed
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == 2 && q == T.k0 - 1) continue;       // k faces below plane k0 - 1 reach no owned cell
      // ---- this thread's face position of direction d
      int x, y, npos;
      if (d == 0) { npos = (kVX + 1) * (kVY + 2); x = tid % (kVX + 1); y = tid / (kVX + 1) - 1; }
      else if (d == 1) { npos = (kVX + 2) * (kVY + 1); x = tid % (kVX + 2) - 1; y = tid / (kVX + 2); }
      else {     // the ring without its corners
        npos = 2 * kVX + (kVX + 2) * kVY;
        if (tid < kVX) { x = tid; y = -1; }
        else if (tid < 2 * kVX) { x = tid - kVX; y = kVY; }
        else { x = (tid - 2 * kVX) % (kVX + 2) - 1; y = (tid - 2 * kVX) / (kVX + 2); }
      }
      if (tid < npos) {
        const int i = T.i0 + x, j = T.j0 + y;
        const int ijk[3] = {i, j, q};
        bool valid = i >= 1 && j >= 1 && q >= 1 && i <= ni && j <= nj && q <= nk && ijk[d] >= 2;
        double Gb[4][3], Pmb[5];
        if (valid) {
          int tb, tc;
          visc_axes(d, tb, tc);
          // stencil offsets (x, y, plane) of L, R, L -/+ e_b, R -/+ e_b, L -/+ e_c, R -/+ e_c
          int o[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, ec[3] = {0, 0, 0};
          o[d] = -1;
          eb[tb] = 1;
          ec[tc] = 1;
          const int off[10][3] = {{o[0], o[1], o[2]}, {0, 0, 0},
                                  {o[0] - eb[0], o[1] - eb[1], o[2] - eb[2]}, {o[0] + eb[0], o[1] + eb[1], o[2] + eb[2]},
                                  {-eb[0], -eb[1], -eb[2]}, {eb[0], eb[1], eb[2]},
                                  {o[0] - ec[0], o[1] - ec[1], o[2] - ec[2]}, {o[0] + ec[0], o[1] + ec[1], o[2] + ec[2]},
                                  {-ec[0], -ec[1], -ec[2]}, {ec[0], ec[1], ec[2]}};
          double V[10][4], PL[5], PR[5], RL[5], RR[5];
#pragma unroll
          for (int m = 0; m < 10; ++m) {
            const int sl = (q + off[m][2] + 3) % 3, yy = y + 2 + off[m][1], xx = x + 2 + off[m][0];
#pragma unroll
            for (int v = 0; v < 3; ++v) V[m][v] = sC[sl][1 + v][yy][xx];
            V[m][3] = sC[sl][5][yy][xx];
            if (m < 2) {
              double* P = m == 0 ? PL : PR;
              double* Rq = m == 0 ? RL : RR;
#pragma unroll
              for (int v = 0; v < 5; ++v) {
                P[v] = sC[sl][v][yy][xx];
                Rq[v] = sR[sl][v][yy][xx];
              }
            }
          }
          const int64_t cR = B.off + (i + 1) + sy * (j + 1) + sz * (int64_t)(q + 1);
          const int64_t cL = cR - (d == 0 ? 1 : (d == 1 ? sy : sz));
          const double murefb = vface_core(a, B, d, ijk, PL, PR, RL, RR, V, Gb, Pmb);
          if (x >= 0 && x < kVX && y >= 0 && y < kVY && q >= T.k0 && q < T.k1) {   // the face's patch writes it out
            a.VW[((size_t)d * 18 + 17) * N + cL] = murefb;
            // stencil reaches a ghost cell: the record k_adj_vcell gathers for the ghost cells
            if (ijk[tb] == 1 || ijk[tb] == n3[tb] || ijk[tc] == 1 || ijk[tc] == n3[tc]) {
              double* W = a.VW + (size_t)d * 18 * N + cL;
#pragma unroll
              for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int z = 0; z < 3; ++z) W[(size_t)(3 * v + z) * N] = Gb[v][z];
#pragma unroll
              for (int v = 0; v < 5; ++v) W[(size_t)(12 + v) * N] = Pmb[v];
            }
          }
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int z = 0; z < 3; ++z) Gb[v][z] = 0.0;
#pragma unroll
          for (int v = 0; v < 5; ++v) Pmb[v] = 0.0;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int z = 0; z < 3; ++z) sF[3 * v + z][y + 1][x + 1] = Gb[v][z];
#pragma unroll
        for (int v = 0; v < 5; ++v) sF[12 + v][y + 1][x + 1] = Pmb[v];
      }
      lds_barrier();
      // ---- the roles of the own cell (tx, ty) in the faces of direction d at plane q
      if (owner) {
        // normal role: sg -1 as the low cell, +1 as the high cell; half the face state
        auto normal = [&](double* acc, int fx, int fy, double sg) {
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[v] += sg * F(3 * v + d, fx, fy);
          acc[4] += 0.5 * F(12, fx, fy);
#pragma unroll
          for (int v = 0; v < 3; ++v) acc[v] += 0.5 * F(13 + v, fx, fy);
          acc[5] += 0.5 * F(16, fx, fy);
        };
        // tangential role: a quarter of the gradient along axis ax, sign pm
        auto tang = [&](double* acc, int fx, int fy, int ax, double pm) {
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[v] += 0.25 * pm * F(3 * v + ax, fx, fy);
        };
        if (d == 0) {          // faces at high cells (x, y); low cell (x - 1, y); tb = y, tc = k
          normal(a0, tx, ty, 1.0);
          normal(a0, tx + 1, ty, -1.0);
          tang(a0, tx, ty - 1, 1, 1.0);
          tang(a0, tx + 1, ty - 1, 1, 1.0);
          tang(a0, tx, ty + 1, 1, -1.0);
          tang(a0, tx + 1, ty + 1, 1, -1.0);
          tang(ap, tx, ty, 2, 1.0);
          tang(ap, tx + 1, ty, 2, 1.0);
          tang(am, tx, ty, 2, -1.0);
          tang(am, tx + 1, ty, 2, -1.0);
        } else if (d == 1) {   // low cell (x, y - 1); tb = x, tc = k
          normal(a0, tx, ty, 1.0);
          normal(a0, tx, ty + 1, -1.0);
          tang(a0, tx - 1, ty, 0, 1.0);
          tang(a0, tx - 1, ty + 1, 0, 1.0);
          tang(a0, tx + 1, ty, 0, -1.0);
          tang(a0, tx + 1, ty + 1, 0, -1.0);
          tang(ap, tx, ty, 2, 1.0);
          tang(ap, tx, ty + 1, 2, 1.0);
          tang(am, tx, ty, 2, -1.0);
          tang(am, tx, ty + 1, 2, -1.0);
        } else {               // the face between planes q - 1 (low cell) and q; tb = x, tc = y
          normal(a0, tx, ty, 1.0);
          normal(am, tx, ty, -1.0);
          tang(a0, tx - 1, ty, 0, 1.0);
          tang(am, tx - 1, ty, 0, 1.0);
          tang(a0, tx + 1, ty, 0, -1.0);
          tang(am, tx + 1, ty, 0, -1.0);
          tang(a0, tx, ty - 1, 1, 1.0);
          tang(am, tx, ty - 1, 1, 1.0);
          tang(a0, tx, ty + 1, 1, -1.0);
          tang(am, tx, ty + 1, 1, -1.0);
        }
      }
      lds_barrier();
    }
    // ---- plane q - 1 is complete: (u, v, w, T) adjoints to primitive ones, into VPb (which
    // k_adj_cell adds to the convective part of Pbar)
    const int p = q - 1;
    if (own && p >= T.k0 && p < T.k1) {
      const int sl = (p + 3) % 3;
      double P[5], Pb[5] = {am[4], 0.0, 0.0, 0.0, am[5]};
#pragma unroll
      for (int v = 0; v < 5; ++v) P[v] = sC[sl][v][ty + 2][tx + 2];
      if (P[0] != 0.0) visc_vars_adj(P, a.R, am, Pb);
      const int64_t c = B.off + (oi + 1) + sy * (oj + 1) + sz * (int64_t)(p + 1);
#pragma unroll
      for (int v = 0; v < 5; ++v) a.VPb[v * N + c] = Pb[v];
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) { am[v] = a0[v]; a0[v] = ap[v]; ap[v] = 0.0; }
    lds_barrier();     // the next plane is staged into slot (q + 2) % 3 = (q - 1) % 3
  }
}

// gl: the first ghost layer of the viscous blocks (cells with one index 0 or n + 1 and the others
// interior), the only ghost cells a tangential role reaches
__global__ void __launch_bounds__(256) k_adj_vcell(AdjArgs a, const int64_t* __restrict__ gl, int64_t ngl) {
  const int64_t N = a.npool;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ngl; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = gl[e];
    int lo = 0, hi = a.nblocks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.blocks[mid].off <= c) lo = mid; else hi = mid - 1;
    }
    const BlockDev& B = a.blocks[lo];
    if (!B.visc) continue;
    const int64_t r64 = c - B.off;
    if (r64 >= B.sz * (int64_t)(B.nk + 4)) continue;
    const unsigned r = (unsigned)r64, sy = (unsigned)B.sy, rj = r / sy;
    const int q3[3] = {(int)(r - rj * sy) - 1, (int)(rj % (unsigned)(B.nj + 4)) - 1, (int)(rj / (unsigned)(B.nj + 4)) - 1};
    const int n3[3] = {B.ni, B.nj, B.nk};
    // interior cells: k_adj_visc gathers them
    if (q3[0] >= 1 && q3[1] >= 1 && q3[2] >= 1 && q3[0] <= n3[0] && q3[1] <= n3[1] && q3[2] <= n3[2]) continue;
    const int64_t st[3] = {1, (int64_t)B.sy, B.sz};
    double Vb[4] = {0, 0, 0, 0}, rhob = 0.0, pb = 0.0;
    for (int d = 0; d < 3; ++d) {
      int tb, tc;
      visc_axes(d, tb, tc);
      // is the face with low cell q3 + (dd, db, dc) along (d, tb, tc) an interior viscous face
      auto valid = [&](int dd, int db, int dc) {
        const int qd = q3[d] + dd, qb = q3[tb] + db, qc = q3[tc] + dc;
        return qd >= 1 && qd <= n3[d] - 1 && qb >= 1 && qb <= n3[tb] && qc >= 1 && qc <= n3[tc];
      };
      auto W = [&](int64_t cl, int row) { return a.VW[((size_t)d * 18 + row) * N + cl]; };
      // low cell (gn = VR - VL) and high cell, with their halves of the face state
      for (int side = 0; side < 2; ++side) {
        const int dd = side == 0 ? 0 : -1;
        if (!valid(dd, 0, 0)) continue;
        const int64_t cl = c + dd * st[d];
        const double sg = side == 0 ? -1.0 : 1.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) Vb[v] += sg * W(cl, 3 * v + d);
        rhob += 0.5 * W(cl, 12);
#pragma unroll
        for (int v = 0; v < 3; ++v) Vb[v] += 0.5 * W(cl, 13 + v);
        pb += 0.5 * W(cl, 16);
      }
      // tangential neighbours: a cell at -e_t of the low (high) cell enters with -1/4, at +e_t with +1/4
      for (int t = 0; t < 2; ++t) {
        const int ax = t == 0 ? tb : tc;
        for (int pm = -1; pm <= 1; pm += 2) {     // the face's low cell is c - pm e_t (- e_d)
          for (int side = 0; side < 2; ++side) {
            const int dd = side == 0 ? 0 : -1;
            const int db = ax == tb ? -pm : 0, dc = ax == tc ? -pm : 0;
            if (!valid(dd, db, dc)) continue;
            const int64_t cl = c + dd * st[d] - pm * st[ax];
#pragma unroll
            for (int v = 0; v < 4; ++v) Vb[v] += 0.25 * pm * W(cl, 3 * v + ax);
          }
        }
      }
    }
    double P[5], Pb[5] = {rhob, 0.0, 0.0, 0.0, pb};
    prim_load(a.Pp, N, c, P);
    if (P[0] != 0.0) visc_vars_adj(P, a.R, Vb, Pb);
#pragma unroll
    for (int v = 0; v < 5; ++v) a.Pbar[v * N + c] += Pb[v];
  }
}

__global__ void __launch_bounds__(256) k_adj_vbc(AdjArgs a) {
  const int64_t N = a.npool;
  const double Rg = a.R;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < a.nvf; q += gridDim.x * blockDim.x) {
    const ViscFace e = a.vfaces[q];
    const BlockDev& B = a.blocks[e.b];
    const int d = e.d;
    int ta, tb;
    visc_axes(d, ta, tb);
    // cells: loc, ghost, loc's tangential neighbours (4), the donor's (4)
    const int64_t cell[10] = {e.loc, e.gh, e.tn[0], e.tn[1], e.tn[2], e.tn[3], e.dn[0], e.dn[1], e.dn[2], e.dn[3]};
    double P[10][5], V[10][4];
#pragma unroll
    for (int m = 0; m < 10; ++m) {
      prim_load(a.Pp, N, cell[m], P[m]);
      V[m][0] = P[m][1]; V[m][1] = P[m][2]; V[m][2] = P[m][3]; V[m][3] = a.Tp[cell[m]];
    }
    const double m2 = (double)e.modfm2;
    double Gl[4][3], Gg[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Gl[v][d] = (V[1][v] - V[0][v]) * m2;
      Gg[v][d] = (V[0][v] - V[1][v]) * (-m2);
      Gl[v][ta] = (V[3][v] - V[2][v]) * 0.5;
      Gl[v][tb] = (V[5][v] - V[4][v]) * 0.5;
      Gg[v][ta] = (V[7][v] - V[6][v]) * 0.5;
      Gg[v][tb] = (V[9][v] - V[8][v]) * 0.5;
    }
    double M[3], Pm[5];
#pragma unroll
    for (int x = 0; x < 3; ++x) M[x] = 0.5 * (e.ml[x] + e.mg[x]);
#pragma unroll
    for (int v = 0; v < 5; ++v) Pm[v] = (P[0][v] + P[1][v]) * 0.5;
    // face flux += -(Fl + Fg)/2 (per unit area); R(loc) gets modfm2 area F
    double Fb[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) Fb[v] = -0.5 * m2 * e.area * a.Rbar[v * N + e.loc];
    const double vis[4] = {B.vis[0], B.vis[1], B.vis[2], B.vis[3]};
    double Glb[4][3], Ggb[4][3], Pmb[5] = {0, 0, 0, 0, 0};
    double murefb = visc_face_flux_adj(d, Gl, M, Pm, Rg, vis, Fb, Glb, Pmb);
    murefb += visc_face_flux_adj(d, Gg, M, Pm, Rg, vis, Fb, Ggb, Pmb);
    double Vb[10][4];
#pragma unroll
    for (int m = 0; m < 10; ++m)
#pragma unroll
      for (int v = 0; v < 4; ++v) Vb[m][v] = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double gnb = m2 * (Glb[v][d] + Ggb[v][d]);
      Vb[1][v] += gnb;
      Vb[0][v] -= gnb;
      Vb[3][v] += 0.5 * Glb[v][ta]; Vb[2][v] -= 0.5 * Glb[v][ta];
      Vb[5][v] += 0.5 * Glb[v][tb]; Vb[4][v] -= 0.5 * Glb[v][tb];
      Vb[7][v] += 0.5 * Ggb[v][ta]; Vb[6][v] -= 0.5 * Ggb[v][ta];
      Vb[9][v] += 0.5 * Ggb[v][tb]; Vb[8][v] -= 0.5 * Ggb[v][tb];
    }
#pragma unroll
    for (int m = 0; m < 10; ++m) {
      double Pb[5] = {0, 0, 0, 0, 0};
      if (m < 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Pb[v] = 0.5 * Pmb[v];
      }
      visc_vars_adj(P[m], Rg, Vb[m], Pb);
#pragma unroll
      for (int v = 0; v < 5; ++v) a.VBC[(10 * (int64_t)q + m) * 5 + v] = Pb[v];
    }
    a.VB[q] = murefb;
  }
}

// Connection-face adjoints into Pbar: one thread per receiving cell, its entries in order.
__global__ void __launch_bounds__(256) k_adj_vbc_gather(AdjArgs a) {
  const int64_t N = a.npool;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nvbt; r += gridDim.x * blockDim.x) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (int64_t e = a.vbrow[r]; e < a.vbrow[r + 1]; ++e) {
      const int64_t x = a.vbent[e];
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += a.VBC[x * 5 + v];
    }
    const int64_t cell = a.vbt[r];
#pragma unroll
    for (int v = 0; v < 5; ++v) a.Pbar[v * N + cell] += acc[v];
  }
}
