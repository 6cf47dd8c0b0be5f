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
// BC_Connection_Visc connection faces).  The interior part is deterministic: k_adj_vface
// (one thread per interior face, both cells interior) reverses the face flux and stores the
// adjoints of its index-space gradients Gb[4][3], of its face state Pmb[5] and of mu_ref at
// the face's low cell; k_adj_vcell (one thread per cell, ghost layers included) gathers them
// over the ten stencil roles the cell plays (low / high cell, the eight tangential
// neighbours) and converts (u, v, w, T) adjoints to primitive ones.  Connection faces (one per
// boundary face cell) reach cells of the donor block: k_adj_vbc stores the primitive adjoints of
// each face's ten stencil cells, and k_adj_vbc_gather adds them to every receiving cell in a
// fixed order (face, then role), so the result does not depend on the schedule.

// Stencil of the interior face on the low side of cell (i, j, k) in direction d: the two
// tangential axes in increasing order.
__device__ __forceinline__ void visc_axes(int d, int& tb, int& tc) {
  tb = d == 0 ? 1 : 0;
  tc = d == 2 ? 1 : 2;
}

__global__ void __launch_bounds__(256) k_adj_vface(AdjArgs a) {
  const int64_t N = a.npool;
  const double Rg = a.R;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.nwork; q += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.segs[mid].start <= q) lo = mid; else hi = mid - 1;
    }
    const AdjSeg S = a.segs[lo];
    const BlockDev& B = a.blocks[S.b];
    if (!B.visc) continue;
    const int d = S.d;
    unsigned r = (unsigned)(q - S.start);
    int ijk[3];
    ijk[0] = (int)(r % (unsigned)S.n0) + 1;
    r /= (unsigned)S.n0;
    ijk[1] = (int)(r % (unsigned)S.n1) + 1;
    ijk[2] = (int)(r / (unsigned)S.n1) + 1;
    const int n3[3] = {B.ni, B.nj, B.nk};
    if (ijk[d] < 2 || ijk[d] > n3[d]) continue;          // block faces: k_adj_vbc
    int tb, tc;
    visc_axes(d, tb, tc);
    const int64_t st[3] = {1, (int64_t)B.sy, B.sz};
    const int64_t cR = B.off + (ijk[0] + 1) + (int64_t)B.sy * (ijk[1] + 1) + B.sz * (int64_t)(ijk[2] + 1);
    const int64_t cL = cR - st[d];
    // stencil: L, R, L -/+ e_b, R -/+ e_b, L -/+ e_c, R -/+ e_c
    const int64_t cell[10] = {cL, cR, cL - st[tb], cL + st[tb], cR - st[tb], cR + st[tb],
                              cL - st[tc], cL + st[tc], cR - st[tc], cR + st[tc]};
    double P[10][5], V[10][4];
#pragma unroll
    for (int m = 0; m < 10; ++m) {
      prim_load(a.Pp, N, cell[m], P[m]);
      V[m][0] = P[m][1]; V[m][1] = P[m][2]; V[m][2] = P[m][3]; V[m][3] = a.Tp[cell[m]];
    }
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
    Pm[0] = (P[0][0] + P[1][0]) * 0.5;
    Pm[4] = (P[0][4] + P[1][4]) * 0.5;
#pragma unroll
    for (int c = 0; c < 3; ++c) Pm[1 + c] = (V[0][c] + V[1][c]) * 0.5;
    // the face adds -area F_visc to the low cell and +area F_visc to the high cell
    double Fb[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) Fb[v] = -area * (a.Rbar[v * N + cL] - a.Rbar[v * N + cR]);
    const double vis[4] = {B.vis[0], B.vis[1], B.vis[2], B.vis[3]};
    double Gb[4][3], Pmb[5] = {0, 0, 0, 0, 0};
    const double murefb = visc_face_flux_adj(d, G, M, Pm, Rg, vis, Fb, Gb, Pmb);
    double* W = a.VW + (size_t)d * 18 * N + cL;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int x = 0; x < 3; ++x) W[(size_t)(3 * v + x) * N] = Gb[v][x];
#pragma unroll
    for (int v = 0; v < 5; ++v) W[(size_t)(12 + v) * N] = Pmb[v];
    W[(size_t)17 * N] = murefb;
  }
}

__global__ void __launch_bounds__(256) k_adj_vcell(AdjArgs a) {
  const int64_t N = a.npool;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
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
