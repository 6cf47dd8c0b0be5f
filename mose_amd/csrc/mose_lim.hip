// mose_lim.hip — the kernel variants of one limiter (MOSE_LIM_ID: 0 minmod, 1 van Leer,
// 2 van Albada, 3 first order, 4 MC, 5 superbee) for every Riemann solver, precision, tile
// shape, geometry kind and mode, and their table entries.  Compiled once per limiter.
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

namespace {

template <int LIM, int RS, typename FT, int TXc, int TYc, int GEO = 0, bool SDR = false>
void stage_modes(StageFn* t) {
  t[0] = k_stage<LIM, RS, FT, 0, TXc, TYc, GEO, SDR>;
  t[1] = k_stage<LIM, RS, FT, 1, TXc, TYc, GEO, SDR>;
  t[2] = k_stage<LIM, RS, FT, 2, TXc, TYc, GEO, SDR>;
}

template <int LIM, int RS>
void solver_entries(KTable& k, ATable& a) {
  stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1]>(k.f[0][LIM][RS][0]);
  stage_modes<LIM, RS, float, kShape[1][0][0], kShape[1][0][1]>(k.f[1][LIM][RS][0]);
  stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 1>(k.g[0][0][LIM][RS]);
  stage_modes<LIM, RS, float, kShape[1][0][0], kShape[1][0][1], 1>(k.g[0][1][LIM][RS]);
  stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 2>(k.g[1][0][LIM][RS]);
  stage_modes<LIM, RS, float, kShape[1][0][0], kShape[1][0][1], 2>(k.g[1][1][LIM][RS]);
  if constexpr (LIM != LIM_FIRST) {     // MUSCL-SD (no slopes to scale at first order)
    stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 0, true>(k.s[0][LIM][RS]);
    stage_modes<LIM, RS, double, kShape[0][0][0], kShape[0][0][1], 1, true>(k.s[1][LIM][RS]);
  }
  if constexpr (LIM == LIM_VANLEER) {   // the tile tuner's alternative shapes
    stage_modes<LIM, RS, double, kShape[0][1][0], kShape[0][1][1]>(k.f[0][LIM][RS][1]);
    stage_modes<LIM, RS, double, kShape[0][2][0], kShape[0][2][1]>(k.f[0][LIM][RS][2]);
    stage_modes<LIM, RS, float, kShape[1][1][0], kShape[1][1][1]>(k.f[1][LIM][RS][1]);
    stage_modes<LIM, RS, float, kShape[1][2][0], kShape[1][2][1]>(k.f[1][LIM][RS][2]);
    stage_modes<LIM, RS, double, kShape[0][3][0], kShape[0][3][1]>(k.f[0][LIM][RS][3]);
    stage_modes<LIM, RS, float, kShape[1][3][0], kShape[1][3][1]>(k.f[1][LIM][RS][3]);
  }
  k.bc[0][LIM][RS] = k_bcface<LIM, RS, double>;
  k.bc[1][LIM][RS] = k_bcface<LIM, RS, float>;
  if constexpr (!rs_takes_beta(RS)) {   // no adjoint of the shock-detector solvers
    a.face[LIM][RS][0] = k_adj_face<LIM, RS, 0>;
    a.face[LIM][RS][1] = k_adj_face<LIM, RS, 1>;
    a.bc[LIM][RS] = k_adj_bcface<LIM, RS>;
  }
}

}  // namespace

template <int LIM>
void fill_tables(KTable& k, ATable& a) {
  solver_entries<LIM, RS_HLLC>(k, a);
  solver_entries<LIM, RS_AUSMP>(k, a);
  solver_entries<LIM, RS_HLLE>(k, a);
  solver_entries<LIM, RS_SLAU>(k, a);
  solver_entries<LIM, RS_SLAU2>(k, a);
  solver_entries<LIM, RS_HLLEPP>(k, a);
  solver_entries<LIM, RS_HLLCP>(k, a);
  a.cell[LIM][0] = k_adj_cell<LIM, 0>;
  a.cell[LIM][1] = k_adj_cell<LIM, 1>;
}

template void fill_tables<MOSE_LIM_ID>(KTable&, ATable&);
