// mose_geometry.hpp — host-side setup of the device layout (blocks, geometry lines, tiles,
// ghost operations) from the reference's inputs: block nodes (blk%node) and the bc.txt
// records.  Host code only; included by mose_hip.hip.
//
// Geometry.  The reference computes, per cell, a metric tensor, the cell lengths dl and the
// volume from the 8 cell nodes (Compute_Metric_Tensor, src/lib/numerics/space/
// Lib_Metrics.f90:250-376) and per face a unit normal and area from the cross product of
// the face diagonals (Compute_Norm_Area, :69-247); ghost cells get metrics from donor
// cells (type 101, BC_Connect_Metrics :860-1092), from a mirror (300, BC_Symmetry_Metrics
// :800-858) or from extrapolated nodes (every other type incl. 201 periodic,
// BC_Extrapolate_Metrics :535-797).  For a rectilinear block these reduce exactly to one
// line of values per direction: dl(i) from the reference's own summation order on the two
// node planes, the face area of an i-face = (y_j - y_{j-1}) (z_k - z_{k-1}) and the
// normal = the axis unit vector.  This setup verifies rectilinearity bit-for-bit and builds
// those lines.  If any block is not rectilinear the whole context uses the general metrics
// instead (mose_metrics.hpp): per-cell arrays of face normals and areas, cell lengths,
// volumes and the inverse-Jacobian rows.
#pragma once
#pragma clang fp contract(off)

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mose_metrics.hpp"

struct HostBlock {
  int ni = 0, nj = 0, nk = 0;
  std::vector<double> nodes;   // (3, 0:ni, 0:nj, 0:nk)
  int visc = 0;                // laminar viscous terms: mu_ref, T_ref, S, Pr, vnn
  double vis[5] = {0, 0, 0, 0, 0};
  const double* node(int i, int j, int k) const {
    return &nodes[3 * ((size_t)i + (size_t)(ni + 1) * ((size_t)j + (size_t)(nj + 1) * k))];
  }
};

struct Layout {
  std::vector<mose::BlockDev> blocks;
  std::vector<mose::LineEnt> lines;
  std::vector<mose::Tile> tiles;
  std::vector<mose::GhostOp> ghost;
  std::vector<mose::BcFace> bcfaces;
  std::vector<double> facebc;
  int64_t nbcslots = 0;
  int64_t npool = 0;
  int ndir = 3;
  int curvi = 0;                 // 1: general metrics in `geo` (NGEO x npool, SoA)
  std::vector<double> geo;
  int visc = 0;                  // 1: some block has viscous terms
  std::vector<mose::ViscFace> vfaces;
  int64_t nvslots = 0;
};

struct GeoError {
  int code = 0, b = 0, i = 0, j = 0, k = 0;
};

namespace geo {

static const int guide[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// |A(d,d)| of Compute_Metric_Tensor on a rectilinear cell with node planes a (low), b (high):
// A(d,:) = N(hi) - N(lo) + N(hi) - N(lo) + ... evaluated left to right, then x 0.25.
static inline double ref_dl(double a, double b) {
  double s = b - a;
  s = s + b; s = s - a;
  s = s + b; s = s - a;
  s = s + b; s = s - a;
  return std::fabs(0.25 * s);
}

static bool set_err(GeoError& e, int code, int b, int i = 0, int j = 0, int k = 0) {
  e.code = code; e.b = b; e.i = i; e.j = j; e.k = k;
  return false;
}

static int face_kind(int type) {
  switch (type) {
    case 101: case 201: return mose::F_CONN;
    case 0: return mose::F_NONE;
    case 400: return mose::F_EXTRAP;
    case 300: return mose::F_SYM;
    case 405: return mose::F_INFLOW;
    case 401: return mose::F_INLET;
    case 406: return mose::F_OUTFLOW;
    default: return -1;
  }
}

}  // namespace geo

// Work list of the stage kernels: TX x TY column patches of every block, each marching a
// chunk of kc k-planes.  Spatial order is block, k chunk, j, i.  With xcd != 0 the list is
// permuted so that workgroup w (dispatched to accelerator complex w mod 8) gets the
// (w / 8)-th tile of the x-th eighth of the spatial order, x = w mod 8: each XCD then
// sweeps one contiguous region, and the i/j halos a tile shares with its neighbours are
// read through the same L2.
static void build_tiles(Layout& L, int TXc, int TYc, int kc, int xcd) {
  using namespace mose;
  std::vector<Tile> S;
  for (int b = 0; b < (int)L.blocks.size(); ++b) {
    const BlockDev& B = L.blocks[b];
    for (int k0 = 1; k0 <= B.nk; k0 += kc)
      for (int j0 = 1; j0 <= B.nj; j0 += TYc)
        for (int i0 = 1; i0 <= B.ni; i0 += TXc) {
          Tile t{};
          t.b = b; t.i0 = i0; t.j0 = j0; t.k0 = k0; t.k1 = std::min(k0 + kc, B.nk + 1);
          S.push_back(t);
        }
  }
  if (!xcd) {
    L.tiles.swap(S);
    return;
  }
  const int64_t n = (int64_t)S.size(), q = n / 8, r = n % 8;
  L.tiles.assign(n, Tile{});
  // XCD x owns S[start(x), start(x) + q + (x < r)); position slot*8 + x holds its slot-th tile.
  // Only the last row (slot = q) is partial, for x < r, so positions 0..n-1 are all filled.
  for (int64_t w = 0; w < n; ++w) {
    const int64_t x = w % 8, slot = w / 8;
    const int64_t start = x * q + std::min<int64_t>(x, r);
    L.tiles[w] = S[start + slot];
  }
}

// General metrics of every block into L.geo (see mose_metrics.hpp).  off: pool offset of
// each block, ftype: BC type per block face.
static bool build_curvilinear(const std::vector<HostBlock>& hb, const std::vector<mose_bc_entry>& bc,
                              const std::vector<int64_t>& off, Layout& L, GeoError& err) {
  using namespace mose;
  using metrics::V3;
  const int nb = (int)hb.size();
  const int64_t np = L.npool;
  L.geo.assign((size_t)NGEO * np, 0.0);
  double* G = L.geo.data();
  auto at = [&](int g, int64_t c) -> double& { return G[(size_t)g * np + c]; };
  auto pidx = [&](int b, int i, int j, int k) -> int64_t {
    const HostBlock& B = hb[b];
    return off[b] + (int64_t)(i + 1) + (int64_t)(B.ni + 4) * (j + 1) + (int64_t)(B.ni + 4) * (B.nj + 4) * (k + 1);
  };
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const metrics::NodeView N{B.nodes.data(), B.ni, B.nj, B.nk};
    // interior cells: dl, volume, unit rows of M (Compute_Dt versors)
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          const V3 nd[8] = {N(i - 1, j - 1, k - 1), N(i - 1, j - 1, k), N(i - 1, j, k - 1), N(i - 1, j, k),
                            N(i, j - 1, k - 1),     N(i, j - 1, k),     N(i, j, k - 1),     N(i, j, k)};
          double M[9], dl[3], vol;
          if (!metrics::cell_metric(nd, M, dl, &vol)) return geo::set_err(err, MOSE_E_METRIC, b + 1, i, j, k);
          const int64_t c = pidx(b, i, j, k);
          for (int d = 0; d < 3; ++d) {
            at(GDL + d, c) = dl[d];
            const double* m = M + 3 * d;
            const double nm = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            for (int q = 0; q < 3; ++q) at(GVS + 3 * d + q, c) = m[q] / nm;
          }
          at(GVOL, c) = vol;
        }
    // faces: the low face of cell (i,j,k) in direction d is the reference's face d at
    // node plane (index along d) - 1
    const double sg[3] = {metrics::face_sign(N, 0), metrics::face_sign(N, 1), metrics::face_sign(N, 2)};
    for (int d = 0; d < 3; ++d) {
      const int i0 = d == 0 ? 0 : 1, j0 = d == 1 ? 0 : 1, k0 = d == 2 ? 0 : 1;
      for (int k = k0; k <= B.nk; ++k)
        for (int j = j0; j <= B.nj; ++j)
          for (int i = i0; i <= B.ni; ++i) {
            double n[3], A;
            metrics::face_normal(N, d, i, j, k, sg[d], n, &A);
            const int64_t c = pidx(b, i + (d == 0), j + (d == 1), k + (d == 2));
            for (int q = 0; q < 3; ++q) at(GN + 4 * d + q, c) = n[q];
            at(GN + 4 * d + 3, c) = A;
          }
    }
  }
  // ghost-cell lengths, one bc record at a time (Mod_Metrics.f90:74-105)
  for (const mose_bc_entry& e : bc) {
    const int b = e.b - 1, d = (e.f - 1) / 2;
    const int* gm = geo::guide[e.f - 1];
    const HostBlock& B = hb[b];
    if (e.type == 101) {
      const int* gs = geo::guide[e.fs - 1];
      const int ds = (e.fs - 1) / 2;
      for (int g = 1; g <= 2; ++g) {
        const int64_t dst = pidx(b, e.i - gm[0] * g, e.j - gm[1] * g, e.k - gm[2] * g);
        const int64_t src = pidx(e.bs - 1, e.is + gs[0] * (g - 1), e.js + gs[1] * (g - 1), e.ks + gs[2] * (g - 1));
        double dl[3] = {at(GDL, src), at(GDL + 1, src), at(GDL + 2, src)};
        if (d != ds) std::swap(dl[d], dl[ds]);
        for (int q = 0; q < 3; ++q) at(GDL + q, dst) = dl[q];
      }
    } else if (e.type == 300) {
      for (int g = 1; g <= 2; ++g) {
        const int64_t dst = pidx(b, e.i - gm[0] * g, e.j - gm[1] * g, e.k - gm[2] * g);
        const int64_t src = pidx(b, e.i + gm[0] * (g - 1), e.j + gm[1] * (g - 1), e.k + gm[2] * (g - 1));
        for (int q = 0; q < 3; ++q) at(GDL + q, dst) = at(GDL + q, src);
      }
    } else {
      const int n[3] = {B.ni, B.nj, B.nk};
      const bool hi = (e.f % 2) == 0;
      bool cubic = d == 0 ? true : (d == 1 ? B.nj > 2 : L.ndir == 3);
      if (n[d] < 2) cubic = false;
      const metrics::NodeView N{B.nodes.data(), B.ni, B.nj, B.nk};
      const int cell[3] = {e.i, e.j, e.k};
      const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      // node planes along d, slot 2 = the boundary face, slots 3, 4 outward
      V3 pl[2][2][5];
      for (int a1 = 0; a1 < 2; ++a1)
        for (int a2 = 0; a2 < 2; ++a2) {
          int q[3];
          q[t1] = cell[t1] - 1 + a1;
          q[t2] = cell[t2] - 1 + a2;
          const int face = hi ? cell[d] : cell[d] - 1;
          const int s = hi ? -1 : 1;      // inward step along d
          const int known = cubic ? 3 : 2;
          for (int g = 0; g < known; ++g) {
            q[d] = face + s * g;
            pl[a1][a2][2 - g] = N(q[0], q[1], q[2]);
          }
          for (int g = 3; g <= 4; ++g) {   // outward planes
            const V3 a = pl[a1][a2][g - 1], bb = pl[a1][a2][g - 2];
            if (cubic) {
              const V3 c = pl[a1][a2][g - 3];
              pl[a1][a2][g] = V3{3.0 * a.x - 3.0 * bb.x + c.x, 3.0 * a.y - 3.0 * bb.y + c.y, 3.0 * a.z - 3.0 * bb.z + c.z};
            } else {
              pl[a1][a2][g] = V3{2.0 * a.x - bb.x, 2.0 * a.y - bb.y, 2.0 * a.z - bb.z};
            }
          }
        }
      for (int g = 1; g <= 2; ++g) {
        V3 nd[8];
        for (int n8 = 0; n8 < 8; ++n8) {
          const int ab[3] = {(n8 >> 2) & 1, (n8 >> 1) & 1, n8 & 1};
          // along d: the ghost cell spans slots 1+g (inner) and 2+g (outer); its low-index
          // node plane is the outer one on a low face
          const int slot = hi ? (1 + g + ab[d]) : (2 + g - ab[d]);
          nd[n8] = pl[ab[t1]][ab[t2]][slot];
        }
        double M[9], dl[3], vol;
        int gc[3] = {e.i, e.j, e.k};
        gc[d] += hi ? g : -g;
        if (!metrics::cell_metric(nd, M, dl, &vol)) return geo::set_err(err, MOSE_E_METRIC, 0, gc[0], gc[1], gc[2]);
        const int64_t c = pidx(b, gc[0], gc[1], gc[2]);
        for (int q = 0; q < 3; ++q) at(GDL + q, c) = dl[q];
      }
    }
  }
  // reciprocal centre distances to the low neighbour: 1 / ((dl(c - e_d) + dl(c)) / 2)
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int64_t st[3] = {1, (int64_t)(B.ni + 4), (int64_t)(B.ni + 4) * (B.nj + 4)};
    for (int k = 0; k <= B.nk + 2; ++k)
      for (int j = 0; j <= B.nj + 2; ++j)
        for (int i = 0; i <= B.ni + 2; ++i) {
          const int64_t c = pidx(b, i, j, k);
          for (int d = 0; d < 3; ++d) {
            const double s = at(GDL + d, c - st[d]) + at(GDL + d, c);
            at(GRLO + d, c) = s > 0.0 ? 1.0 / (0.5 * s) : 0.0;
          }
        }
  }
  return true;
}

static bool build_layout(const std::vector<HostBlock>& hb, const std::vector<mose_bc_entry>& bc, double cp, Layout& L,
                         GeoError& err) {
  auto cp_of = [cp](const HostBlock&) { return cp; };
  using namespace mose;
  const int nb = (int)hb.size();
  if (nb < 1) return geo::set_err(err, MOSE_E_ARG, 0);
  // ndir from block 1 (Check_Mesh_Type, Lib_Metrics.f90:13-66)
  L.ndir = hb[0].nk > 1 ? 3 : ((hb[0].nj == 1) ? 1 : 2);

  // ---- bc records: completeness and one type per block face
  std::vector<int> ftype(6 * nb, -9999), fentry(6 * nb, -1);
  std::vector<int64_t> fcount(6 * nb, 0);
  int64_t expected = 0;
  for (const HostBlock& B : hb)
    expected += 2 * ((int64_t)B.nj * B.nk + (int64_t)B.ni * B.nk + (int64_t)B.ni * B.nj);
  if ((int64_t)bc.size() != expected) return geo::set_err(err, MOSE_E_BC, 0);
  for (size_t q = 0; q < bc.size(); ++q) {
    const mose_bc_entry& e = bc[q];
    if (e.b < 1 || e.b > nb || e.f < 1 || e.f > 6) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    const HostBlock& B = hb[e.b - 1];
    const int n[3] = {B.ni, B.nj, B.nk};
    const int ijk[3] = {e.i, e.j, e.k};
    const int d = (e.f - 1) / 2;
    for (int a = 0; a < 3; ++a)
      if (ijk[a] < 1 || ijk[a] > n[a]) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    if (ijk[d] != ((e.f % 2) ? 1 : n[d])) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    const int slot = 6 * (e.b - 1) + (e.f - 1);
    if (geo::face_kind(e.type) < 0) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    if (ftype[slot] == -9999) { ftype[slot] = e.type; fentry[slot] = (int)q; }
    else if (ftype[slot] != e.type) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    if (e.type == 405 || e.type == 401 || e.type == 406) {   // one record set per block face: every cell of the face carries the same values
      const mose_bc_entry& e0 = bc[fentry[slot]];
      if (e.mach != e0.mach || e.T != e0.T || e.p != e0.p || e.alpha != e0.alpha || e.beta != e0.beta ||
          e.rel_fac != e0.rel_fac)
        return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    }
    if (e.type == 101 || e.type == 201) {
      if (e.bs < 1 || e.bs > nb || e.fs < 1 || e.fs > 6) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
      const HostBlock& S = hb[e.bs - 1];
      const int ns[3] = {S.ni, S.nj, S.nk};
      const int sijk[3] = {e.is, e.js, e.ks};
      const int ds = (e.fs - 1) / 2;
      for (int a = 0; a < 3; ++a)
        if (sijk[a] < 1 || sijk[a] > ns[a]) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
      if (ns[ds] < 2) return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);
    }
    fcount[slot]++;
  }
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int64_t fc[3] = {(int64_t)B.nj * B.nk, (int64_t)B.ni * B.nk, (int64_t)B.ni * B.nj};
    for (int f = 0; f < 6; ++f)
      if (fcount[6 * b + f] != fc[f / 2]) return geo::set_err(err, MOSE_E_BC, b + 1);
  }

  // ---- pool layout
  std::vector<int64_t> off(nb);
  int64_t pool = 0;
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    off[b] = pool;
    pool += (int64_t)(B.ni + 4) * (B.nj + 4) * (B.nk + 4);
    pool = (pool + 63) / 64 * 64;
  }
  L.npool = pool;

  // ---- geometry lines: rectilinearity and interior values
  std::vector<std::vector<double>> coord(3 * nb);
  std::vector<int> lineoff(3 * nb);
  L.lines.clear();
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    for (int d = 0; d < 3; ++d) {
      coord[3 * b + d].resize(n[d] + 1);
      for (int q = 0; q <= n[d]; ++q) {
        const int ii = d == 0 ? q : 0, jj = d == 1 ? q : 0, kk = d == 2 ? q : 0;
        coord[3 * b + d][q] = B.node(ii, jj, kk)[d];
      }
      for (int q = 1; q <= n[d]; ++q)
        if (!(coord[3 * b + d][q] > coord[3 * b + d][q - 1])) L.curvi = 1;
    }
    for (int k = 0; k <= B.nk && !L.curvi; ++k)
      for (int j = 0; j <= B.nj && !L.curvi; ++j)
        for (int i = 0; i <= B.ni; ++i) {
          const double* p = B.node(i, j, k);
          if (p[0] != coord[3 * b][i] || p[1] != coord[3 * b + 1][j] || p[2] != coord[3 * b + 2][k]) {
            L.curvi = 1;
            break;
          }
        }
    for (int d = 0; d < 3; ++d) {
      lineoff[3 * b + d] = (int)L.lines.size();
      for (int q = -1; q <= n[d] + 2; ++q) {
        LineEnt e{};
        if (q >= 1 && q <= n[d]) {
          const double a = coord[3 * b + d][q - 1], c = coord[3 * b + d][q];
          e.dl = geo::ref_dl(a, c);
          e.delta = c - a;
        }
        L.lines.push_back(e);
      }
    }
  }
  auto LE = [&](int b, int d, int q) -> LineEnt& { return L.lines[lineoff[3 * b + d] + q + 1]; };

  // ---- ghost lengths per block face
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    for (int f = 0; f < 6; ++f) {
      const int d = f / 2;
      const bool hi = (f % 2) == 1;
      const int type = ftype[6 * b + f];
      const mose_bc_entry& e0 = bc[fentry[6 * b + f]];
      const std::vector<double>& c = coord[3 * b + d];
      const int nd = n[d];
      if (type == 101) {
        const int ds = (e0.fs - 1) / 2;
        const int sijk[3] = {e0.is, e0.js, e0.ks};
        for (int g = 1; g <= 2; ++g) {
          const int sidx = sijk[ds] + geo::guide[e0.fs - 1][ds] * (g - 1);
          LE(b, d, hi ? nd + g : 1 - g).dl = LE(e0.bs - 1, ds, sidx).dl;
        }
      } else if (type == 300) {
        for (int g = 1; g <= 2; ++g) {
          if (hi) LE(b, d, nd + g).dl = LE(b, d, nd + 1 - g).dl;
          else LE(b, d, 1 - g).dl = LE(b, d, g).dl;
        }
      } else {
        bool cubic = (d == 0) ? true : (d == 1 ? (B.nj > 2) : (L.ndir == 3));
        if (nd < 2) cubic = false;
        if (nd < 1) return geo::set_err(err, MOSE_E_ARG, b + 1);
        double e1 = 0.0, e2 = 0.0;
        if (!hi) {
          if (cubic) { e1 = 3.0 * c[0] - 3.0 * c[1] + c[2]; e2 = 3.0 * e1 - 3.0 * c[0] + c[1]; }
          else { e1 = 2.0 * c[0] - c[1]; e2 = 2.0 * e1 - c[0]; }
          LE(b, d, 0).dl = geo::ref_dl(e1, c[0]);
          LE(b, d, -1).dl = geo::ref_dl(e2, e1);
        } else {
          if (cubic) { e1 = 3.0 * c[nd] - 3.0 * c[nd - 1] + c[nd - 2]; e2 = 3.0 * e1 - 3.0 * c[nd] + c[nd - 1]; }
          else { e1 = 2.0 * c[nd] - c[nd - 1]; e2 = 2.0 * e1 - c[nd]; }
          LE(b, d, nd + 1).dl = geo::ref_dl(c[nd], e1);
          LE(b, d, nd + 2).dl = geo::ref_dl(e1, e2);
        }
      }
    }
  }
  // ---- spacings (entry q: cell q and the face q-1/2 between cells q-1 and q),
  //      Lib_Convective.f90:30-34: lp = dl(q)/2, rl1 = 1 / ((dl(q-1) + dl(q))/2)
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    for (int d = 0; d < 3; ++d) {
      for (int q = -1; q <= n[d] + 2; ++q) {
        LE(b, d, q).lp = 0.5 * LE(b, d, q).dl;
        LE(b, d, q).rdl = 1.0 / LE(b, d, q).dl;    // correctly rounded, as the device's 1.0 / dl
      }
      for (int q = 0; q <= n[d] + 2; ++q) {
        LineEnt& e = LE(b, d, q);
        e.rl1 = 1.0 / (0.5 * (LE(b, d, q - 1).dl + e.dl));
        e.lm = 0.5 * LE(b, d, q - 1).dl;
        if (q >= 1) e.rl0 = 1.0 / (0.5 * (LE(b, d, q - 2).dl + LE(b, d, q - 1).dl));
        if (q <= n[d] + 1) e.rl2 = 1.0 / (0.5 * (e.dl + LE(b, d, q + 1).dl));
      }
    }
  }

  // ---- general metrics when any block is not rectilinear (the lines above are then unused)
  if (L.curvi && !build_curvilinear(hb, bc, off, L, err)) return false;
  auto G = [&](int g, int64_t c) -> double { return L.geo[(size_t)g * L.npool + c]; };

  // ---- block descriptors
  L.blocks.resize(nb);
  L.facebc.clear();
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    BlockDev& D = L.blocks[b];
    D.off = off[b];
    D.ni = B.ni; D.nj = B.nj; D.nk = B.nk;
    D.sy = B.ni + 4;
    D.sz = (int64_t)(B.ni + 4) * (B.nj + 4);
    for (int d = 0; d < 3; ++d) D.line[d] = lineoff[3 * b + d];
    D.visc = B.visc;
    for (int q = 0; q < 5; ++q) D.vis[q] = B.vis[q];
    if (B.visc) {
      // the viscous path: rectilinear blocks, faces 0 / 101 / 201 (the wall and inflow/outflow
      // viscous conditions of the reference are not part of it)
      L.visc = 1;
      if (L.curvi) return geo::set_err(err, MOSE_E_ARG, b + 1);
      for (int f = 0; f < 6; ++f) {
        const int t = ftype[6 * b + f];
        if (t != 0 && t != 101 && t != 201) return geo::set_err(err, MOSE_E_BC, b + 1);
      }
      // cp / Pr in place of Pr
      D.vis[3] = cp_of(B) / B.vis[3];
    }
    for (int f = 0; f < 6; ++f) {
      D.bct[f] = geo::face_kind(ftype[6 * b + f]);
      D.bcd[f] = -1;
      D.bcs[f] = -1;
      D.bvs[f] = -1;
      if (ftype[6 * b + f] == 405) {
        const mose_bc_entry& e = bc[fentry[6 * b + f]];
        D.bcd[f] = (int)(L.facebc.size() / 8);
        // Sup velocity direction of Supersonic_Inflow (Lib_BC_Fluxes_Inflow.f90:336-338); an
        // angle at or above half the largest double is bc.txt's 'normal' token, resolved per
        // face cell from its normal (flag bits in v[7], explicit angles in v[3], v[4])
        const int flags = (e.alpha >= 0.5 * DBL_MAX ? 1 : 0) | (e.beta >= 0.5 * DBL_MAX ? 2 : 0);
        double v[8] = {e.mach, e.T, e.p, std::cos(e.alpha) * std::cos(e.beta),
                       std::sin(e.alpha) * std::cos(e.beta), std::sin(e.beta), e.rel_fac, 0.0};
        if (flags) {
          v[3] = (flags & 1) ? 0.0 : e.alpha;
          v[4] = (flags & 2) ? 0.0 : e.beta;
          v[5] = 0.0;
          v[7] = (double)flags;
        }
        L.facebc.insert(L.facebc.end(), v, v + 8);
      } else if (ftype[6 * b + f] == 401 || ftype[6 * b + f] == 406) {
        const mose_bc_entry& e = bc[fentry[6 * b + f]];
        D.bcd[f] = (int)(L.facebc.size() / 8);
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (ftype[6 * b + f] == 401) {
          // {T0, p0, alpha, beta, rel_fac, -, -, flags}: flag bits mark the 'normal' token
          // (angles resolved per face cell from its normal, Compute_Inflow_Direction)
          const int flags = (e.alpha >= 0.5 * DBL_MAX ? 1 : 0) | (e.beta >= 0.5 * DBL_MAX ? 2 : 0);
          v[0] = e.T; v[1] = e.p; v[2] = (flags & 1) ? 0.0 : e.alpha; v[3] = (flags & 2) ? 0.0 : e.beta;
          v[4] = e.rel_fac; v[7] = (double)flags;
        } else {
          v[0] = e.p; v[1] = e.rel_fac;     // {p_exit, rel_fac}
        }
        L.facebc.insert(L.facebc.end(), v, v + 8);
      }
    }
  }

  // ---- ghost operations, one per bc record (Fill_Ghost_Cell)
  auto pidx = [&](int b, int i, int j, int k) -> int64_t {
    const BlockDev& D = L.blocks[b];
    return D.off + (int64_t)(i + 1) + (int64_t)D.sy * (j + 1) + D.sz * (int64_t)(k + 1);
  };
  auto stride = [&](int b, int f) -> int64_t {
    const BlockDev& D = L.blocks[b];
    const int* g = geo::guide[f - 1];
    return (int64_t)g[0] + (int64_t)g[1] * D.sy + (int64_t)g[2] * D.sz;
  };
  L.ghost.clear();
  L.ghost.reserve(bc.size());
  for (const mose_bc_entry& e : bc) {
    GhostOp op{};
    op.b = pidx(e.b - 1, e.i, e.j, e.k);
    op.s = stride(e.b - 1, e.f);
    switch (e.type) {
      case 101: case 201:
        op.type = G_COPY;
        op.src1 = pidx(e.bs - 1, e.is, e.js, e.ks);
        op.src2 = op.src1 + stride(e.bs - 1, e.fs);
        break;
      case 300: {
        op.type = G_SYM;
        const int d = (e.f - 1) / 2;
        if (L.curvi) {   // normal of the boundary face (Ghost_Symmetry, Lib_Ghost.f90:193-225)
          const int64_t fc = (e.f % 2) ? op.b : op.b - op.s;
          for (int q = 0; q < 3; ++q) op.n[q] = G(GN + 4 * d + q, fc);
        } else {
          op.n[d] = 1.0;
        }
        break;
      }
      case 400: op.type = G_EXTRAP; break;
      default: op.type = G_ZG; break;
    }
    L.ghost.push_back(op);
  }
  // ---- boundary faces of non-connection type: one flux slot per face cell, tangential
  //      indices in cyclic order (d+1, d+2), first one fastest
  L.bcfaces.clear();
  int64_t slot = 0;
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int n[3] = {B.ni, B.nj, B.nk};
    BlockDev& D = L.blocks[b];
    for (int f = 0; f < 6; ++f) {
      if (D.bct[f] == F_CONN) continue;
      const int d = f / 2, d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      const bool hi = f & 1;
      D.bcs[f] = (int)slot;
      for (int t2 = 1; t2 <= n[d2]; ++t2)
        for (int t1 = 1; t1 <= n[d1]; ++t1) {
          int c3[3];
          c3[d] = hi ? n[d] : 1;
          c3[d1] = t1;
          c3[d2] = t2;
          int m3[3] = {c3[0], c3[1], c3[2]};
          m3[d] = hi ? n[d] - 1 : 2;
          BcFace e{};
          e.cb = pidx(b, c3[0], c3[1], c3[2]);
          e.cn = pidx(b, m3[0], m3[1], m3[2]);
          e.b = b; e.f = f; e.slot = (int)(slot + (t1 - 1) + (int64_t)n[d1] * (t2 - 1));
          if (L.curvi) {
            e.bdl = 0.5 * G(GDL + d, e.cb);
            e.idl = 0.5 * G(GDL + d, e.cn);
            const int64_t fc = hi ? e.cb + (e.cb - e.cn) : e.cb;    // cell holding the face as its low face
            for (int q = 0; q < 3; ++q) e.n[q] = G(GN + 4 * d + q, fc);
            e.area = G(GN + 4 * d + 3, fc);
          } else {
            e.bdl = 0.5 * LE(b, d, c3[d]).dl;
            e.idl = 0.5 * LE(b, d, m3[d]).dl;
            e.n[d] = 1.0;
            e.area = LE(b, d1, c3[d1]).delta * LE(b, d2, c3[d2]).delta;
          }
          {   // face diagonal (Setup_Inflow_Geometry, Lib_BC_Fluxes_Inflow.f90: t_Vec = node of the
              // face's high corner minus node of the corner opposite along the face)
            int hiN[3] = {c3[0], c3[1], c3[2]};
            hiN[d] = hi ? n[d] : 0;
            int loN[3] = {hiN[0] - 1, hiN[1] - 1, hiN[2] - 1};
            loN[d] = hiN[d];
            const double* p1 = B.node(hiN[0], hiN[1], hiN[2]);
            const double* p0 = B.node(loN[0], loN[1], loN[2]);
            const double tv[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double tm = std::sqrt(tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2]);
            for (int q = 0; q < 3; ++q) e.t[q] = tv[q] / tm;
          }
          L.bcfaces.push_back(e);
        }
      slot += (int64_t)n[d1] * n[d2];
    }
  }
  L.nbcslots = slot;

  // ---- viscous flux of connection faces (BC_Connection_Visc), one slot per face cell
  L.vfaces.clear();
  int64_t vslot = 0;
  if (L.visc) {
    std::vector<int> seen(6 * nb, 0);
    for (const mose_bc_entry& e : bc) {
      if (e.type != 101 && e.type != 201) continue;
      const int b = e.b - 1, f = e.f - 1, d = f / 2;
      if (!hb[b].visc) continue;
      if ((e.fs - 1) / 2 != d || e.d11 != 1 || e.d12 != 0 || e.d21 != 0 || e.d22 != 1)
        return geo::set_err(err, MOSE_E_BC, e.b, e.i, e.j, e.k);   // aligned, same-direction only
      BlockDev& D = L.blocks[b];
      const int n[3] = {D.ni, D.nj, D.nk};
      const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      if (!seen[6 * b + f]) {
        seen[6 * b + f] = 1;
        D.bvs[f] = (int)vslot;
        vslot += (int64_t)n[d1] * n[d2];
      }
      const int c3[3] = {e.i, e.j, e.k};
      const int* g = geo::guide[f];
      mose::ViscFace v{};
      v.b = b; v.d = d;
      v.modfm2 = 1 - 2 * (e.f % 2);
      v.slot = D.bvs[f] + (c3[d1] - 1) + n[d1] * (c3[d2] - 1);
      v.loc = pidx(b, e.i, e.j, e.k);
      v.gh = pidx(b, e.i - g[0], e.j - g[1], e.k - g[2]);
      const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;    // tangential axes in increasing order
      int q[3];
      for (int s2 = 0; s2 < 4; ++s2) {
        q[0] = e.i; q[1] = e.j; q[2] = e.k;
        q[s2 < 2 ? ta : tb] += (s2 % 2) ? 1 : -1;
        v.tn[s2] = pidx(b, q[0], q[1], q[2]);
        q[0] = e.is; q[1] = e.js; q[2] = e.ks;                 // aligned: same axes in the donor
        q[s2 < 2 ? ta : tb] += (s2 % 2) ? 1 : -1;
        v.dn[s2] = pidx(e.bs - 1, q[0], q[1], q[2]);
      }
      const int gc[3] = {e.i - g[0], e.j - g[1], e.k - g[2]};
      for (int a = 0; a < 3; ++a) {
        v.ml[a] = 1.0 / LE(b, a, c3[a]).dl;
        v.mg[a] = 1.0 / LE(b, a, gc[a]).dl;
      }
      v.area = LE(b, d1, c3[d1]).delta * LE(b, d2, c3[d2]).delta;
      L.vfaces.push_back(v);
    }
  }
  L.nvslots = vslot;
  return true;
}
