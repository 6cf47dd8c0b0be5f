// mose_lim.hip — the kernel variants of one limiter (MOSE_LIM_ID: 0 minmod, 1 van Leer,
// 2 van Albada, 3 first order, 4 MC, 5 superbee) in one precision (MOSE_PREC_ID: 0 FP64,
// 1 FP32 flux) for every Riemann solver, tile shape, geometry kind and mode, and their table
// entries.  Compiled once per (limiter, precision) pair.
#include <hip/hip_runtime.h>

#include "../../include/mose_hip.h"
#include "mose_kernels.hpp"

using namespace mose;

#include "mose_stage.hpp"
#include "mose_adjoint.hpp"
#include "mose_tables.hpp"

#ifndef MOSE_LIM_ID
#error "MOSE_LIM_ID must name the limiter this unit compiles"
#endif
#ifndef MOSE_PREC_ID
#error "MOSE_PREC_ID must name the precision this unit compiles"
#endif

namespace {

template <int LIM, int RS, typename FT, int TXc, int TYc, int GEO = 0, bool SDR = false>
void stage_modes(StageFn* t) {
  t[0] = k_stage<LIM, RS, FT, 0, TXc, TYc, GEO, SDR>;
  t[1] = k_stage<LIM, RS, FT, 1, TXc, TYc, GEO, SDR>;
  t[2] = k_stage<LIM, RS, FT, 2, TXc, TYc, GEO, SDR>;
}

template <int LIM, int RS, int P, int GEO, int SH>
void shape_entry(KTable& k) {
  using FT = typename std::conditional<P == 0, double, float>::type;
  if constexpr (shape_built(GEO, SH, rs_takes_beta(RS)))
    stage_modes<LIM, RS, FT, kShape[P][SH][0], kShape[P][SH][1], GEO>(k.f[GEO][P][LIM][RS][SH]);
}

template <int LIM, int RS, int P, int GEO>
void geo_entries(KTable& k) {
  shape_entry<LIM, RS, P, GEO, 0>(k);
  shape_entry<LIM, RS, P, GEO, 1>(k);
  shape_entry<LIM, RS, P, GEO, 2>(k);
  shape_entry<LIM, RS, P, GEO, 3>(k);
}

template <int LIM, int RS, int P>
void solver_entries(KTable& k, ATable& a) {
  using FT = typename std::conditional<P == 0, double, float>::type;
  geo_entries<LIM, RS, P, 0>(k);
  geo_entries<LIM, RS, P, 1>(k);
  geo_entries<LIM, RS, P, 2>(k);
  k.bc[P][LIM][RS] = k_bcface<LIM, RS, FT>;
  if constexpr (P == 0) {
    if constexpr (LIM != LIM_FIRST) {     // MUSCL-SD (no slopes to scale at first order)
      stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 0, true>(k.s[0][LIM][RS]);
      stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 1, true>(k.s[1][LIM][RS]);
    }
    if constexpr (!rs_takes_beta(RS)) {   // no adjoint of the shock-detector solvers
      a.face[LIM][RS][0] = k_adj_face<LIM, RS, 0>;
      a.face[LIM][RS][1] = k_adj_face<LIM, RS, 1>;
      a.bc[LIM][RS] = k_adj_bcface<LIM, RS>;
    }
  }
}

}  // namespace

template <int LIM, int P>
void fill_tables(KTable& k, ATable& a) {
  solver_entries<LIM, RS_HLLC, P>(k, a);
  solver_entries<LIM, RS_AUSMP, P>(k, a);
  solver_entries<LIM, RS_HLLE, P>(k, a);
  solver_entries<LIM, RS_SLAU, P>(k, a);
  solver_entries<LIM, RS_SLAU2, P>(k, a);
  solver_entries<LIM, RS_HLLEPP, P>(k, a);
  solver_entries<LIM, RS_HLLCP, P>(k, a);
  if constexpr (P == 0) {
    a.cell[LIM][0] = k_adj_cell<LIM, 0>;
    a.cell[LIM][1] = k_adj_cell<LIM, 1>;
  }
}

template void fill_tables<MOSE_LIM_ID, MOSE_PREC_ID>(KTable&, ATable&);
