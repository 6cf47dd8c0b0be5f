// mose_tables.hpp — tables of the kernel variants the host launches: stage kernels
// [prec][limiter][riemann][tile shape][mode], general-metrics and viscous stage kernels,
// boundary-face kernels and the solver-dependent adjoint passes.  The entries of each
// limiter and precision are instantiated in a translation unit of their own (mose_lim.hip,
// compiled once per pair with MOSE_LIM_ID and MOSE_PREC_ID set) so that the variants compile
// in parallel; mose_hip.hip fills the tables once by calling fill_tables<LIM, P> for every pair.
#pragma once

typedef void (*StageFn)(mose::StageArgs);
typedef void (*BcFn)(mose::StageArgs, const mose::BcFace*, int);
typedef void (*AdjFn)(AdjArgs);

constexpr int NLIM = 6;     // minmod, van Leer, van Albada, first order, MC, superbee
constexpr int NRS = 7;      // HLLC, AUSM+, HLLE, SLAU, SLAU2, HLLE++, HLLC+
// Patch shapes (TX x TY threads) of the stage kernel per precision; shape 0 is the default
// and is compiled for every variant.  The tile tuner's alternatives, compiled for every
// limiter: rectilinear blocks shape 3 (every solver) and shape 1, general metrics shape 3,
// viscous blocks shape 2 (128 threads: half the LDS of the others), the last three for the
// solvers without the shock detector.
constexpr int NSHAPE = 4;
constexpr int kShape[2][NSHAPE][2] = {{{32, 8}, {16, 16}, {16, 8}, {32, 16}}, {{32, 8}, {16, 16}, {16, 8}, {32, 16}}};
constexpr bool shapes_within_bound() {
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < NSHAPE; ++s) {
      const int n = kShape[p][s][0] * kShape[p][s][1];
      if (n > mose::kMaxTileThreads || n % 64 != 0) return false;
    }
  return true;
}
static_assert(shapes_within_bound(), "every tile shape fits the tile kernels' launch bound in whole waves");
constexpr bool shape_built(int geo, int shape, bool detector) {
  return shape == 0 || (geo == 0 && shape == 3) ||
         (!detector && ((geo == 0 && shape == 1) || (geo == 1 && shape == 3) || (geo == 2 && shape == 2)));
}

// null where a variant is not compiled
struct KTable {
  StageFn f[3][2][NLIM][NRS][NSHAPE][3] = {};   // [geo][prec][lim][rs][shape][mode]
  StageFn s[2][NLIM][NRS][3] = {};              // MUSCL-SD, FP64, tile shape 0: [geo 0/1][lim][rs][mode]
  BcFn bc[2][NLIM][NRS] = {};
};

// adjoint kernels [lim][rs][geo] (FP64)
struct ATable {
  AdjFn face[NLIM][NRS][2] = {}, bc[NLIM][NRS] = {}, cell[NLIM][2] = {};
};

// Entries of limiter LIM in precision P (0: FP64, with the FP64-only kernels: MUSCL-SD and
// the adjoint passes; 1: FP32 flux); defined and instantiated in mose_lim.hip.
template <int LIM, int P>
void fill_tables(KTable& k, ATable& a);
