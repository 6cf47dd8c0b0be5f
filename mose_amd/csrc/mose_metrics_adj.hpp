// mose_metrics_adj.hpp — reverse of the general (curvilinear) metric computation on the
// host: adjoints of the geometric quantities the residual reads (face normals and areas,
// cell lengths, reciprocal centre distances) -> adjoints of the block nodes.  Host code
// only, included by mose_hip.hip after mose_geometry.hpp.
//
// Forward definitions it reverses (mose_metrics.hpp, build_curvilinear in mose_geometry.hpp;
// reference src/lib/numerics/space/Lib_Metrics.f90 and Mod_Metrics.f90):
//   rl(c, d) = 1 / ((dl(c - e_d, d) + dl(c, d)) / 2)     (zero when the sum is zero)
//   dl(c, d) = |A(d, :)|, A(d, :) = 1/4 (signed sum of the cell's 8 nodes along d)
//   ghost dl: type 101 copies the donor cell (lengths permuted to the receiver's axes),
//             type 300 copies the mirrored interior cell, every other type builds the ghost
//             cell from node planes extrapolated outward (3a-3b+c or 2a-b)
//   face:     d3 = 1/2 (d1 x d2) of the face diagonals, area = |d3|, n = sgn d3 / |d3|
// The orientation signs and every branch are piecewise constant in the nodes.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace metrics {

// Adjoint accumulator over one block's nodes, (3, 0:ni, 0:nj, 0:nk) in Fortran order.
struct NodeAdj {
  double* p;
  int ni, nj, nk;
  void add(int i, int j, int k, V3 v) const {
    double* q = p + 3 * ((size_t)i + (size_t)(ni + 1) * ((size_t)j + (size_t)(nj + 1) * k));
    q[0] += v.x; q[1] += v.y; q[2] += v.z;
  }
};

static inline V3 axpy(V3 a, double s, V3 b) { return V3{a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }
static inline V3 scale(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }

// Reverse of the cell lengths of cell_metric: dlb[d] -> node adjoints Nb[0..7] (same node
// order as cell_metric: q = 4 di + 2 dj + dk).
static void cell_dl_vjp(const V3 N[8], const double dlb[3], V3 Nb[8]) {
  const double* n[8];
  for (int q = 0; q < 8; ++q) n[q] = &N[q].x;
  double A[3][3];
  for (int c = 0; c < 3; ++c) {
    A[0][c] = 0.25 * (n[4][c] - n[0][c] + n[5][c] - n[1][c] + n[6][c] - n[2][c] + n[7][c] - n[3][c]);
    A[1][c] = 0.25 * (n[2][c] - n[0][c] + n[3][c] - n[1][c] + n[6][c] - n[4][c] + n[7][c] - n[5][c]);
    A[2][c] = 0.25 * (n[1][c] - n[0][c] + n[3][c] - n[2][c] + n[5][c] - n[4][c] + n[7][c] - n[6][c]);
  }
  for (int q = 0; q < 8; ++q) Nb[q] = V3{0.0, 0.0, 0.0};
  for (int h = 0; h < 3; ++h) {
    if (dlb[h] == 0.0) continue;
    const double dl = std::sqrt(A[h][0] * A[h][0] + A[h][1] * A[h][1] + A[h][2] * A[h][2]);
    if (dl == 0.0) continue;
    const V3 Ab = V3{0.25 * dlb[h] * A[h][0] / dl, 0.25 * dlb[h] * A[h][1] / dl, 0.25 * dlb[h] * A[h][2] / dl};
    for (int q = 0; q < 8; ++q) {
      const int bit = h == 0 ? (q >> 2) & 1 : (h == 1 ? (q >> 1) & 1 : q & 1);
      Nb[q] = axpy(Nb[q], bit ? 1.0 : -1.0, Ab);
    }
  }
}

// Reverse of face_normal: (nb, Ab) of the face of direction dir at (i, j, k) -> nodes.
static void face_normal_vjp(const NodeView& N, const NodeAdj& X, int dir, int i, int j, int k, double sgn,
                            const double nb[3], double Ab) {
  if (nb[0] == 0.0 && nb[1] == 0.0 && nb[2] == 0.0 && Ab == 0.0) return;
  int a1[3], b1[3], a2[3], b2[3];   // d1 = N(a1) - N(b1), d2 = N(a2) - N(b2)
  auto set = [](int* t, int x, int y, int z) { t[0] = x; t[1] = y; t[2] = z; };
  if (dir == 0) { set(a1, i, j, k); set(b1, i, j - 1, k - 1); set(a2, i, j - 1, k); set(b2, i, j, k - 1); }
  else if (dir == 1) { set(a1, i - 1, j, k); set(b1, i, j, k - 1); set(a2, i, j, k); set(b2, i - 1, j, k - 1); }
  else { set(a1, i - 1, j, k); set(b1, i, j - 1, k); set(a2, i - 1, j - 1, k); set(b2, i, j, k); }
  const V3 d1 = sub(N(a1[0], a1[1], a1[2]), N(b1[0], b1[1], b1[2]));
  const V3 d2 = sub(N(a2[0], a2[1], a2[2]), N(b2[0], b2[1], b2[2]));
  const V3 d3 = cross(d1, d2, 0.5);
  const double A = std::sqrt(d3.x * d3.x + d3.y * d3.y + d3.z * d3.z);
  if (A == 0.0) return;
  const V3 u = scale(d3, 1.0 / A);
  // n = sgn u: dn/dd3 = sgn (I - u u^T) / A; dA/dd3 = u
  const double nu = nb[0] * u.x + nb[1] * u.y + nb[2] * u.z;
  const V3 d3b = V3{Ab * u.x + sgn * (nb[0] - nu * u.x) / A, Ab * u.y + sgn * (nb[1] - nu * u.y) / A,
                    Ab * u.z + sgn * (nb[2] - nu * u.z) / A};
  // d3 = 1/2 d1 x d2: d1b = 1/2 d2 x d3b, d2b = 1/2 d3b x d1
  const V3 d1b = cross(d2, d3b, 0.5);
  const V3 d2b = cross(d3b, d1, 0.5);
  X.add(a1[0], a1[1], a1[2], d1b);
  X.add(b1[0], b1[1], b1[2], scale(d1b, -1.0));
  X.add(a2[0], a2[1], a2[2], d2b);
  X.add(b2[0], b2[1], b2[2], scale(d2b, -1.0));
}

}  // namespace metrics

namespace geo {

// Node adjoints of every block from the adjoints of the geometry arrays:
//   GF (12 x npool)  normal (3) and area of the low face of each cell, per direction;
//   RLb (3 x npool)  reciprocal centre distance to the low neighbour, per direction;
//   DLb (3 x npool)  cell length per direction (interior and ghost cells).
// RL holds the forward reciprocal distances (3 x npool).  Xb[b] is overwritten,
// (3, 0:ni, 0:nj, 0:nk) per block.
static void curvilinear_vjp(const std::vector<HostBlock>& hb, const std::vector<mose_bc_entry>& bc, const Layout& L,
                            const std::vector<double>& RL, const std::vector<double>& GF,
                            const std::vector<double>& RLb, std::vector<double> DLb,
                            std::vector<std::vector<double>>& Xb) {
  using namespace mose;
  using metrics::V3;
  const int nb = (int)hb.size();
  const int64_t np = L.npool;
  auto pidx = [&](int b, int i, int j, int k) -> int64_t {
    const HostBlock& B = hb[b];
    return L.blocks[b].off + (int64_t)(i + 1) + (int64_t)(B.ni + 4) * (j + 1) +
           (int64_t)(B.ni + 4) * (B.nj + 4) * (k + 1);
  };
  auto dlb = [&](int d, int64_t c) -> double& { return DLb[(size_t)d * np + c]; };
  Xb.assign(nb, {});
  for (int b = 0; b < nb; ++b) Xb[b].assign((size_t)3 * (hb[b].ni + 1) * (hb[b].nj + 1) * (hb[b].nk + 1), 0.0);

  // rl = 1 / (s / 2), s = dl(c - e_d) + dl(c): d rl / d s = -rl^2 / 2
  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const int64_t st[3] = {1, (int64_t)(B.ni + 4), (int64_t)(B.ni + 4) * (B.nj + 4)};
    for (int k = 0; k <= B.nk + 2; ++k)
      for (int j = 0; j <= B.nj + 2; ++j)
        for (int i = 0; i <= B.ni + 2; ++i) {
          const int64_t c = pidx(b, i, j, k);
          for (int d = 0; d < 3; ++d) {
            const double rb = RLb[(size_t)d * np + c];
            const double rl = RL[(size_t)d * np + c];
            if (rb == 0.0 || rl == 0.0) continue;
            const double sb = -0.5 * rl * rl * rb;
            dlb(d, c - st[d]) += sb;
            dlb(d, c) += sb;
          }
        }
  }

  // ghost-cell lengths, bc records in reverse order (a later record overwrote an earlier one)
  for (size_t r = bc.size(); r-- > 0;) {
    const mose_bc_entry& e = bc[r];
    const int b = e.b - 1, d = (e.f - 1) / 2;
    const int* gm = guide[e.f - 1];
    const HostBlock& B = hb[b];
    if (e.type == 101) {
      const int* gs = guide[e.fs - 1];
      const int ds = (e.fs - 1) / 2;
      for (int g = 2; g >= 1; --g) {
        const int64_t dst = pidx(b, e.i - gm[0] * g, e.j - gm[1] * g, e.k - gm[2] * g);
        const int64_t src = pidx(e.bs - 1, e.is + gs[0] * (g - 1), e.js + gs[1] * (g - 1), e.ks + gs[2] * (g - 1));
        double gb[3] = {dlb(0, dst), dlb(1, dst), dlb(2, dst)};
        if (d != ds) std::swap(gb[d], gb[ds]);
        for (int q = 0; q < 3; ++q) { dlb(q, dst) = 0.0; dlb(q, src) += gb[q]; }
      }
    } else if (e.type == 300) {
      for (int g = 2; g >= 1; --g) {
        const int64_t dst = pidx(b, e.i - gm[0] * g, e.j - gm[1] * g, e.k - gm[2] * g);
        const int64_t src = pidx(b, e.i + gm[0] * (g - 1), e.j + gm[1] * (g - 1), e.k + gm[2] * (g - 1));
        for (int q = 0; q < 3; ++q) {
          const double v = dlb(q, dst);
          dlb(q, dst) = 0.0;
          dlb(q, src) += v;
        }
      }
    } else {
      const int n[3] = {B.ni, B.nj, B.nk};
      const bool hi = (e.f % 2) == 0;
      bool cubic = d == 0 ? true : (d == 1 ? B.nj > 2 : L.ndir == 3);
      if (n[d] < 2) cubic = false;
      const metrics::NodeView N{B.nodes.data(), B.ni, B.nj, B.nk};
      const metrics::NodeAdj X{Xb[b].data(), B.ni, B.nj, B.nk};
      const int cell[3] = {e.i, e.j, e.k};
      const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      const int face = hi ? cell[d] : cell[d] - 1;
      const int s = hi ? -1 : 1;
      const int known = cubic ? 3 : 2;
      // forward planes
      V3 pl[2][2][5];
      int qn[2][2][3][3];   // node indices of the known planes
      for (int a1 = 0; a1 < 2; ++a1)
        for (int a2 = 0; a2 < 2; ++a2) {
          for (int g = 0; g < known; ++g) {
            int* q = qn[a1][a2][g];
            q[t1] = cell[t1] - 1 + a1;
            q[t2] = cell[t2] - 1 + a2;
            q[d] = face + s * g;
            pl[a1][a2][2 - g] = N(q[0], q[1], q[2]);
          }
          for (int g = 3; g <= 4; ++g) {
            const V3 a = pl[a1][a2][g - 1], bb = pl[a1][a2][g - 2];
            if (cubic) {
              const V3 c = pl[a1][a2][g - 3];
              pl[a1][a2][g] = V3{3.0 * a.x - 3.0 * bb.x + c.x, 3.0 * a.y - 3.0 * bb.y + c.y, 3.0 * a.z - 3.0 * bb.z + c.z};
            } else {
              pl[a1][a2][g] = V3{2.0 * a.x - bb.x, 2.0 * a.y - bb.y, 2.0 * a.z - bb.z};
            }
          }
        }
      V3 plb[2][2][5];
      for (auto& x : plb)
        for (auto& y : x)
          for (auto& z : y) z = V3{0.0, 0.0, 0.0};
      bool any = false;
      for (int g = 1; g <= 2; ++g) {
        int gc[3] = {e.i, e.j, e.k};
        gc[d] += hi ? g : -g;
        const int64_t c = pidx(b, gc[0], gc[1], gc[2]);
        const double gb[3] = {dlb(0, c), dlb(1, c), dlb(2, c)};
        for (int q = 0; q < 3; ++q) dlb(q, c) = 0.0;
        if (gb[0] == 0.0 && gb[1] == 0.0 && gb[2] == 0.0) continue;
        any = true;
        V3 nd[8], ndb[8];
        int sl[8], i1[8], i2[8];
        for (int n8 = 0; n8 < 8; ++n8) {
          const int ab[3] = {(n8 >> 2) & 1, (n8 >> 1) & 1, n8 & 1};
          sl[n8] = hi ? (1 + g + ab[d]) : (2 + g - ab[d]);
          i1[n8] = ab[t1];
          i2[n8] = ab[t2];
          nd[n8] = pl[i1[n8]][i2[n8]][sl[n8]];
        }
        metrics::cell_dl_vjp(nd, gb, ndb);
        for (int n8 = 0; n8 < 8; ++n8) {
          V3& t = plb[i1[n8]][i2[n8]][sl[n8]];
          t = metrics::axpy(t, 1.0, ndb[n8]);
        }
      }
      if (!any) continue;
      for (int a1 = 0; a1 < 2; ++a1)
        for (int a2 = 0; a2 < 2; ++a2) {
          V3* p = plb[a1][a2];
          for (int g = 4; g >= 3; --g) {
            if (cubic) {
              p[g - 1] = metrics::axpy(p[g - 1], 3.0, p[g]);
              p[g - 2] = metrics::axpy(p[g - 2], -3.0, p[g]);
              p[g - 3] = metrics::axpy(p[g - 3], 1.0, p[g]);
            } else {
              p[g - 1] = metrics::axpy(p[g - 1], 2.0, p[g]);
              p[g - 2] = metrics::axpy(p[g - 2], -1.0, p[g]);
            }
          }
          for (int g = 0; g < known; ++g) {
            const int* q = qn[a1][a2][g];
            X.add(q[0], q[1], q[2], p[2 - g]);
          }
        }
    }
  }

  for (int b = 0; b < nb; ++b) {
    const HostBlock& B = hb[b];
    const metrics::NodeView N{B.nodes.data(), B.ni, B.nj, B.nk};
    const metrics::NodeAdj X{Xb[b].data(), B.ni, B.nj, B.nk};
    // interior cell lengths
    for (int k = 1; k <= B.nk; ++k)
      for (int j = 1; j <= B.nj; ++j)
        for (int i = 1; i <= B.ni; ++i) {
          const int64_t c = pidx(b, i, j, k);
          const double gb[3] = {dlb(0, c), dlb(1, c), dlb(2, c)};
          if (gb[0] == 0.0 && gb[1] == 0.0 && gb[2] == 0.0) continue;
          const V3 nd[8] = {N(i - 1, j - 1, k - 1), N(i - 1, j - 1, k), N(i - 1, j, k - 1), N(i - 1, j, k),
                            N(i, j - 1, k - 1),     N(i, j - 1, k),     N(i, j, k - 1),     N(i, j, k)};
          V3 ndb[8];
          metrics::cell_dl_vjp(nd, gb, ndb);
          for (int q = 0; q < 8; ++q) X.add(i - 1 + ((q >> 2) & 1), j - 1 + ((q >> 1) & 1), k - 1 + (q & 1), ndb[q]);
        }
    // face normals and areas
    const double sg[3] = {metrics::face_sign(N, 0), metrics::face_sign(N, 1), metrics::face_sign(N, 2)};
    for (int d = 0; d < 3; ++d) {
      const int i0 = d == 0 ? 0 : 1, j0 = d == 1 ? 0 : 1, k0 = d == 2 ? 0 : 1;
      for (int k = k0; k <= B.nk; ++k)
        for (int j = j0; j <= B.nj; ++j)
          for (int i = i0; i <= B.ni; ++i) {
            const int64_t c = pidx(b, i + (d == 0), j + (d == 1), k + (d == 2));
            const double nbv[3] = {GF[(size_t)(GN + 4 * d) * np + c], GF[(size_t)(GN + 4 * d + 1) * np + c],
                                   GF[(size_t)(GN + 4 * d + 2) * np + c]};
            metrics::face_normal_vjp(N, X, d, i, j, k, sg[d], nbv, GF[(size_t)(GN + 4 * d + 3) * np + c]);
          }
    }
  }
}

}  // namespace geo
