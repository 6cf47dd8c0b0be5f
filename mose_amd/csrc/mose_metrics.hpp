// mose_metrics.hpp — general (curvilinear) cell and face metrics on the host, written into
// the device geometry arrays (SoA over the cell pool).  Host code only, included by
// mose_geometry.hpp.
//
// Definitions (src/lib/numerics/space/Lib_Metrics.f90):
//   * cell: A = 1/4 (sum of the four edge vectors along each index direction), M = A^-1 by
//     cofactors, dl(d) = |A(d,:)|, volume = five tetrahedra (Compute_Metric_Tensor :250-376);
//   * face: normal = cross product of the face diagonals x 1/2, area = its length, oriented
//     by one sign per block and direction (Compute_Norm_Area :69-247);
//   * ghost cells: type 101 takes the donor cell (BC_Connect_Metrics :860-1092; for
//     node-matching connections that is the donor cell itself), 300 mirrors the interior
//     cells (BC_Symmetry_Metrics :800-858), every other type (201 periodic included)
//     extrapolates two node planes outward — cubic 3a-3b+c, or linear 2a-b in j when nj <= 2
//     and in k when the case is not 3-D — and computes the ghost cell from them
//     (BC_Extrapolate_Metrics :535-797; selection Mod_Metrics.f90:74-105).
// Expression order follows the reference so that the values are the same doubles.
#pragma once
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>
#include <vector>

namespace metrics {

struct V3 {
  double x, y, z;
};
static inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// Compute_Metric_Tensor on the 8 nodes N[0..7] = (i-1,j-1,k-1), (i-1,j-1,k), (i-1,j,k-1),
// (i-1,j,k), (i,j-1,k-1), (i,j-1,k), (i,j,k-1), (i,j,k).  Returns false for det = 0 or a
// non-positive volume (the reference stops).
static bool cell_metric(const V3 N[8], double M[9], double dl[3], double* vol) {
  double A[3][3];
  const double* n[8];
  for (int q = 0; q < 8; ++q) n[q] = &N[q].x;
  for (int c = 0; c < 3; ++c) {
    double a = n[4][c] - n[0][c];
    a = a + n[5][c]; a = a - n[1][c];
    a = a + n[6][c]; a = a - n[2][c];
    a = a + n[7][c]; a = a - n[3][c];
    A[0][c] = 0.25 * a;
    double b = n[2][c] - n[0][c];
    b = b + n[3][c]; b = b - n[1][c];
    b = b + n[6][c]; b = b - n[4][c];
    b = b + n[7][c]; b = b - n[5][c];
    A[1][c] = 0.25 * b;
    double e = n[1][c] - n[0][c];
    e = e + n[3][c]; e = e - n[2][c];
    e = e + n[5][c]; e = e - n[4][c];
    e = e + n[7][c]; e = e - n[6][c];
    A[2][c] = 0.25 * e;
  }
  const double det = A[0][0] * A[1][1] * A[2][2] - A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] +
                     A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] - A[0][2] * A[1][1] * A[2][0];
  if (std::fabs(det) == 0.0) return false;
  const double cof[9] = {(A[1][1] * A[2][2] - A[1][2] * A[2][1]), -(A[1][0] * A[2][2] - A[1][2] * A[2][0]),
                         (A[1][0] * A[2][1] - A[1][1] * A[2][0]), -(A[0][1] * A[2][2] - A[0][2] * A[2][1]),
                         (A[0][0] * A[2][2] - A[0][2] * A[2][0]), -(A[0][0] * A[2][1] - A[0][1] * A[2][0]),
                         (A[0][1] * A[1][2] - A[0][2] * A[1][1]), -(A[0][0] * A[1][2] - A[0][2] * A[1][0]),
                         (A[0][0] * A[1][1] - A[0][1] * A[1][0])};
  for (int q = 0; q < 9; ++q) M[q] = cof[q] / det;
  for (int h = 0; h < 3; ++h) dl[h] = std::sqrt(A[h][0] * A[h][0] + A[h][1] * A[h][1] + A[h][2] * A[h][2]);
  // tetrahedra over the vertex order (1,5,3,7,2,6,4,8) of the reference
  const int ord[8] = {0, 4, 2, 6, 1, 5, 3, 7};
  double vx[8], vy[8], vz[8];
  for (int q = 0; q < 8; ++q) { vx[q] = N[ord[q]].x; vy[q] = N[ord[q]].y; vz[q] = N[ord[q]].z; }
  auto tv = [&](int i1, int i2, int i3, int i4) {
    --i1; --i2; --i3; --i4;
    return std::fabs(((vx[i2] - vx[i1]) * ((vy[i3] - vy[i1]) * (vz[i4] - vz[i1]) - (vy[i4] - vy[i1]) * (vz[i3] - vz[i1])) +
                      (vy[i2] - vy[i1]) * ((vx[i4] - vx[i1]) * (vz[i3] - vz[i1]) - (vx[i3] - vx[i1]) * (vz[i4] - vz[i1])) +
                      (vz[i2] - vz[i1]) * ((vx[i3] - vx[i1]) * (vy[i4] - vy[i1]) - (vx[i4] - vx[i1]) * (vy[i3] - vy[i1]))) /
                     6.0);
  };
  *vol = tv(1, 2, 3, 5) + tv(2, 4, 3, 8) + tv(5, 8, 6, 2) + tv(5, 7, 8, 3) + tv(5, 8, 2, 3);
  return *vol > 0.0;
}

// d1 x d2 (times s), the reference's component formulas.
static inline V3 cross(V3 a, V3 b, double s) {
  return V3{(a.y * b.z - a.z * b.y) * s, (a.z * b.x - a.x * b.z) * s, (a.x * b.y - a.y * b.x) * s};
}

// Node access of one block: (3, 0:ni, 0:nj, 0:nk), Fortran order.
struct NodeView {
  const double* p;
  int ni, nj, nk;
  V3 operator()(int i, int j, int k) const {
    const double* q = p + 3 * ((size_t)i + (size_t)(ni + 1) * ((size_t)j + (size_t)(nj + 1) * k));
    return V3{q[0], q[1], q[2]};
  }
};

static inline V3 avg4(V3 a, V3 b, V3 c, V3 d) {
  return V3{0.25 * (a.x + b.x + c.x + d.x), 0.25 * (a.y + b.y + c.y + d.y), 0.25 * (a.z + b.z + c.z + d.z)};
}

// Orientation sign of the face normals of a block in direction dir (Compute_Norm_Area :85-149).
static double face_sign(const NodeView& N, int dir) {
  V3 d1, d2, c1, c2;
  if (dir == 0) {
    d1 = sub(N(1, 1, 1), N(1, 0, 0)); d2 = sub(N(1, 0, 1), N(1, 1, 0));
    c1 = avg4(N(1, 1, 1), N(1, 0, 1), N(1, 0, 0), N(1, 1, 0));
    c2 = avg4(N(0, 1, 1), N(0, 0, 1), N(0, 0, 0), N(0, 1, 0));
  } else if (dir == 1) {
    d1 = sub(N(0, 1, 1), N(1, 1, 0)); d2 = sub(N(1, 1, 1), N(0, 1, 0));
    c1 = avg4(N(1, 1, 1), N(0, 1, 1), N(0, 1, 0), N(1, 1, 0));
    c2 = avg4(N(1, 0, 1), N(0, 0, 1), N(0, 0, 0), N(1, 0, 0));
  } else {
    d1 = sub(N(0, 1, 1), N(1, 0, 1)); d2 = sub(N(0, 0, 1), N(1, 1, 1));
    c1 = avg4(N(1, 1, 1), N(0, 1, 1), N(0, 0, 1), N(1, 0, 1));
    c2 = avg4(N(1, 1, 0), N(0, 1, 0), N(0, 0, 0), N(1, 0, 0));
  }
  const V3 s = cross(d1, d2, 1.0);
  const V3 d3 = sub(c1, c2);
  return std::copysign(1.0, d3.x * s.x + d3.y * s.y + d3.z * s.z);
}

// Normal and area of the face of direction dir at node plane q of a cell column (a, b are
// the cell indices of the two tangential directions, 1-based), Compute_Norm_Area :160-245.
static void face_normal(const NodeView& N, int dir, int i, int j, int k, double sgn, double n[3], double* area) {
  V3 d1, d2;
  if (dir == 0) { d1 = sub(N(i, j, k), N(i, j - 1, k - 1)); d2 = sub(N(i, j - 1, k), N(i, j, k - 1)); }
  else if (dir == 1) { d1 = sub(N(i - 1, j, k), N(i, j, k - 1)); d2 = sub(N(i, j, k), N(i - 1, j, k - 1)); }
  else { d1 = sub(N(i - 1, j, k), N(i, j - 1, k)); d2 = sub(N(i - 1, j - 1, k), N(i, j, k)); }
  const V3 d3 = cross(d1, d2, 0.5);
  const double A = std::sqrt(d3.x * d3.x + d3.y * d3.y + d3.z * d3.z);
  if (A == 0.0) { n[0] = n[1] = n[2] = 0.0; }
  else { n[0] = d3.x / A * sgn; n[1] = d3.y / A * sgn; n[2] = d3.z / A * sgn; }
  *area = A;
}

}  // namespace metrics
