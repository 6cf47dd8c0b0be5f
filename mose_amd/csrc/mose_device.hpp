// mose_device.hpp — per-face / per-cell device math of the explicit residual path (gfx950).
//
// Everything here is templated on the flux arithmetic type T (double for the FP64 path,
// float for the FP32-flux path).  State, residual accumulation and reductions stay FP64
// in the kernels that call these functions.
//
// Reference routines restated (algebraically identical, re-arranged for the GPU):
//   limiters            src/lib/numerics/space/Lib_Limiters.f90:24-105
//   reconstruction      src/lib/numerics/space/Lib_Reconstruction.f90:10-29
//   HLLC (Batten)       src/lib/numerics/fluxes/riemann/Lib_Riemann_HLL.f90:147-258,550-647
//   AUSM+               src/lib/numerics/fluxes/riemann/Lib_Riemann_AUSM.f90:150-238
//   HLLE                src/lib/numerics/fluxes/riemann/Lib_Riemann_HLL.f90:261-320
//   SLAU, SLAU2         src/lib/numerics/fluxes/riemann/Lib_Riemann_SLAU.f90:19-140
//   HLLE++, HLLC+       src/lib/numerics/fluxes/riemann/Lib_Riemann_HLL.f90:321-534
//   shock detector      src/lib/numerics/fluxes/Lib_Shock_Detector.f90:9-38
//   interface flux      src/lib/numerics/fluxes/Lib_Convective.f90:10-85
// Gas: single species, calorically perfect (the FLINT tables of the reference are absent;
// with tables linear in T the reference's Roe averages reduce to the forms used here).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mose {

enum { LIM_MINMOD = 0, LIM_VANLEER = 1, LIM_VANALBADA = 2, LIM_FIRST = 3, LIM_MC = 4, LIM_SUPERBEE = 5 };
enum { RS_HLLC = 0, RS_AUSMP = 1, RS_HLLE = 2, RS_SLAU = 3, RS_SLAU2 = 4, RS_HLLEPP = 5, RS_HLLCP = 6 };
// Riemann solvers that take the shock detector's weight beta (obj_riemann%SD, Mod_Riemann.f90:74-81)
constexpr bool rs_takes_beta(int rs) { return rs == RS_HLLEPP || rs == RS_HLLCP; }

template <typename T>
struct GasT {
  T gam, gm1, rgm1, R, cp;   // gamma, gamma-1, 1/(gamma-1), R, cp
};

template <typename T> __device__ __forceinline__ T fabs_(T x) { return x < T(0) ? -x : x; }
// FP64: IEEE correctly rounded.  FP32 (flux variant only): the hardware square root and
// reciprocal (v_sqrt_f32, v_rcp_f32, ~1 ulp), without denormal rescaling sequences.
__device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
// FP64 reciprocal: hardware estimate refined by two Newton steps (within 1 ulp of 1/x for
// normal x; the results differ from IEEE division by at most 1 ulp).
__device__ __forceinline__ double rcp_(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
// Square root of a strictly positive, normal argument (sound speeds, sqrt(rho) of a checked
// face state).  FP64: hardware reciprocal square-root estimate refined by Goldschmidt's
// iteration (g -> sqrt(x), h -> 1/(2 sqrt(x))) and two Newton corrections of g, within 1 ulp
// of the correctly rounded root; no rescaling of denormals and no 0 / inf cases, which
// saves about half the instructions of the general square root.  FP32: v_sqrt_f32.
__device__ __forceinline__ double sqrtp_(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-g, h, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float sqrtp_(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename T> __device__ __forceinline__ T min_(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T max_(T a, T b) { return a > b ? a : b; }

// ---------------------------------------------------------------------------- limiters
// phi(x, y)*x with x the face-spanning slope and y the outer slope.
//  minmod   : reference form.
//  van Leer : (r+|r|)/(1+|r|) x with r = y/x  ==  2xy/(x+y) if xy > 0 else 0.
//  van Albada: (r^2+r)/(1+r^2) x              ==  xy(x+y)/(x^2+y^2).
//  MC       : max(0, min(2r, (1+r)/2, 2)) x   ==  sgn(x) min(2|x|, 2|y|, (|x|+|y|)/2) if xy > 0.
//  superbee : max(0, min(2r, 1), min(r, 2)) x ==  sgn(x) max(min(2|y|, |x|), min(|y|, 2|x|)) if xy > 0.
// The reference regularises r with x+1e-30; the single-division forms agree with it to
// round-off for every slope pair whose magnitude exceeds ~1e-14 of the other and differ
// by less than |x| otherwise.
template <int LIM, typename T>
__device__ __forceinline__ T limiter(T x, T y) {
  if constexpr (LIM == LIM_MINMOD) {
    T r = x;
    if (fabs_(x) > fabs_(y)) r = y;
    return (x * y < T(0)) ? T(0) : r;
  } else if constexpr (LIM == LIM_VANLEER) {
    T xy = x * y;
    T s = x + y;
    return (xy > T(0)) ? (T(2) * xy) * rcp_(s) : T(0);
  } else if constexpr (LIM == LIM_VANALBADA) {
    T q = x * x + y * y;
    return (q > T(0)) ? (x * y * (x + y)) * rcp_(q) : T(0);
  } else if constexpr (LIM == LIM_MC) {
    const T ax = fabs_(x), ay = fabs_(y);
    const T m = min_(T(2) * min_(ax, ay), T(0.5) * (ax + ay));
    return (x * y > T(0)) ? (x > T(0) ? m : -m) : T(0);
  } else if constexpr (LIM == LIM_SUPERBEE) {
    const T ax = fabs_(x), ay = fabs_(y);
    const T m = max_(min_(T(2) * ay, ax), min_(ay, T(2) * ax));
    return (x * y > T(0)) ? (x > T(0) ? m : -m) : T(0);
  } else {
    return T(0);
  }
}

// Per-face spacing data: reciprocals of the three centre distances and the two
// half-lengths of the cells adjacent to the face (Lib_Convective.f90:30-34).
template <typename T>
struct FaceSpacing {
  T rl0, rl1, rl2, lm, lp;
};

// MUSCL reconstruction of one primitive variable from the 4-cell stencil around the face.
template <int LIM, typename T>
__device__ __forceinline__ void reconstruct(T qm1, T q0, T q1, T q2, const FaceSpacing<T>& s,
                                            T& ql, T& qr) {
  T s0 = (q0 - qm1) * s.rl0;
  T s1 = (q1 - q0) * s.rl1;
  T s2 = (q2 - q1) * s.rl2;
  ql = q0 + limiter<LIM>(s1, s0) * s.lm;
  qr = q1 - limiter<LIM>(s2, s1) * s.lp;
}

// ------------------------------------------------------------------------- face normals
// A general unit normal (curvilinear faces) or a coordinate axis known at compile time (the
// faces of rectilinear blocks): dot(u, v, w) = u.n, madd(a, d, p) = a + p n_d.  For an axis,
// the products with the zero components are left out (the compiler cannot drop x * 0 itself);
// the values are those of the general form with that unit normal.
template <typename T>
struct NGen {
  T nx, ny, nz;
  __device__ __forceinline__ T x() const { return nx; }
  __device__ __forceinline__ T y() const { return ny; }
  __device__ __forceinline__ T z() const { return nz; }
  __device__ __forceinline__ T dot(T u, T v, T w) const { return u * nx + v * ny + w * nz; }
  __device__ __forceinline__ T madd(T a, int d, T p) const { return a + p * (d == 0 ? nx : (d == 1 ? ny : nz)); }
};
template <typename T, int AX>
struct NAx {
  static_assert(AX >= 0 && AX < 3, "axis");
  __device__ __forceinline__ T x() const { return AX == 0 ? T(1) : T(0); }
  __device__ __forceinline__ T y() const { return AX == 1 ? T(1) : T(0); }
  __device__ __forceinline__ T z() const { return AX == 2 ? T(1) : T(0); }
  __device__ __forceinline__ T dot(T u, T v, T w) const { return AX == 0 ? u : (AX == 1 ? v : w); }
  __device__ __forceinline__ T madd(T a, int d, T p) const { return d == AX ? a + p : a; }
};

// ------------------------------------------------------------------------- Riemann flux
// Face flux per unit area from left/right primitive states (rho,u,v,w,p), unit normal n.
// Returns false when a state is non-physical (check_gas_state: rho<=0 or p<=0).
template <typename T>
__device__ __forceinline__ bool gas_ok(const T* P) {
  return (P[0] > T(0)) && (P[4] > T(0));
}

template <typename T, typename NT>
__device__ __forceinline__ void hllc(const T* L, const T* R, const NT& n, const GasT<T>& g, T* F) {
  const T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const T rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const T irl = rcp_(rl), irr = rcp_(rr);
  const T al = sqrtp_(g.gam * pl * irl), ar = sqrtp_(g.gam * pr * irr);
  const T unl = n.dot(ul, vl, wl);
  const T unr = n.dot(ur, vr, wr);
  // total enthalpy h0 = cp T + |u|^2/2 = gamma/(gamma-1) p/rho + |u|^2/2
  const T kgl = g.gam * g.rgm1;
  const T h0l = kgl * pl * irl + T(0.5) * (ul * ul + vl * vl + wl * wl);
  const T h0r = kgl * pr * irr + T(0.5) * (ur * ur + vr * vr + wr * wr);
  // Roe averages (roe_averages)
  const T srl = sqrtp_(rl), srr = sqrtp_(rr);
  const T isr = rcp_(srl + srr);
  const T u_roe = (srr * ur + srl * ul) * isr;
  const T v_roe = (srr * vr + srl * vl) * isr;
  const T w_roe = (srr * wr + srl * wl) * isr;
  const T h0_roe = (srr * h0r + srl * h0l) * isr;
  const T a_roe = sqrtp_(g.gm1 * (h0_roe - T(0.5) * (u_roe * u_roe + v_roe * v_roe + w_roe * w_roe)));
  const T un_roe = n.dot(u_roe, v_roe, w_roe);
  const T e0l = h0l - pl * irl, e0r = h0r - pr * irr;
  // Batten signal speeds, clamped to include 0
  const T S1 = min_(T(0), min_(unl - al, un_roe - a_roe));
  const T S4 = max_(T(0), max_(unr + ar, un_roe + a_roe));
  const T dl = rl * (S1 - unl), dr = rr * (S4 - unr);
  const T Sstar = (pr - pl + dl * unl - dr * unr) * rcp_(dl - dr);
  const T pstar = rl * (unl - S1) * (unl - Sstar) + pl;
  // Side selection: S1>=0 -> F_L ; S*>=0 -> F_L + S1(U*L - U_L); S4>0 -> F_R + S4(U*R - U_R); else F_R.
  const bool left = (S1 >= T(0)) || (Sstar >= T(0));
  const T r_ = left ? rl : rr, u_ = left ? ul : ur, v_ = left ? vl : vr, w_ = left ? wl : wr;
  const T p_ = left ? pl : pr, un_ = left ? unl : unr, h0_ = left ? h0l : h0r, e0_ = left ? e0l : e0r;
  T S = left ? S1 : S4;
  const bool star = left ? (S1 < T(0)) : (S4 > T(0));
  const T Fr = r_ * un_;
  T f0 = Fr, f1 = n.madd(Fr * u_, 0, p_), f2 = n.madd(Fr * v_, 1, p_), f3 = n.madd(Fr * w_, 2, p_), f4 = Fr * h0_;
  if (star) {
    const T inv = rcp_(S - Sstar);
    const T sd = (S - un_) * r_;
    const T dp = pstar - p_;
    const T U1S = sd * inv;
    const T U2S = n.madd(sd * u_, 0, dp) * inv;
    const T U3S = n.madd(sd * v_, 1, dp) * inv;
    const T U4S = n.madd(sd * w_, 2, dp) * inv;
    const T U5S = (sd * e0_ - p_ * un_ + pstar * Sstar) * inv;
    f0 += S * (U1S - r_);
    f1 += S * (U2S - r_ * u_);
    f2 += S * (U3S - r_ * v_);
    f3 += S * (U4S - r_ * w_);
    f4 += S * (U5S - r_ * e0_);
  }
  F[0] = f0; F[1] = f1; F[2] = f2; F[3] = f3; F[4] = f4;
}

template <typename T, typename NT>
__device__ __forceinline__ void ausmp(const T* L, const T* R, const NT& n, const GasT<T>& g, T* F) {
  const T beta = T(0.125), alpha = T(0.1875);
  const T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const T rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const T irl = rcp_(rl), irr = rcp_(rr);
  const T al2 = g.gam * pl * irl, ar2 = g.gam * pr * irr;
  const T unl = n.dot(ul, vl, wl);
  const T unr = n.dot(ur, vr, wr);
  const T kgl = g.gam * g.rgm1;
  const T H_L = kgl * pl * irl + T(0.5) * (ul * ul + vl * vl + wl * wl);
  const T H_R = kgl * pr * irr + T(0.5) * (ur * ur + vr * vr + wr * wr);
  const T a_half = sqrtp_(sqrtp_(al2 * ar2));          // sqrt(aL aR), Lib_Riemann_AUSM.f90:196
  const T ia = rcp_(a_half);
  const T ML = unl * ia, MR = unr * ia;
  T MplusL, pplusL, MminusR, pminusR;
  if (fabs_(ML) < T(1)) {
    const T a = (ML + T(1)) * (ML + T(1)), b = (ML * ML - T(1)) * (ML * ML - T(1));
    MplusL = T(0.25) * a + beta * b;
    pplusL = T(0.25) * a * (T(2) - ML) + alpha * ML * b;
  } else {
    MplusL = T(0.5) * (ML + fabs_(ML));
    pplusL = ML >= T(0) ? T(1) : T(0);
  }
  if (fabs_(MR) < T(1)) {
    const T a = (MR - T(1)) * (MR - T(1)), b = (MR * MR - T(1)) * (MR * MR - T(1));
    MminusR = -T(0.25) * a - beta * b;
    pminusR = T(0.25) * a * (T(2) + MR) - alpha * MR * b;
  } else {
    MminusR = T(0.5) * (MR - fabs_(MR));
    pminusR = MR >= T(0) ? T(0) : T(1);
  }
  const T Mhalf = MplusL + MminusR;
  const T phalf = pplusL * pl + pminusR * pr;
  const T mdotp = rl * a_half * max_(Mhalf, T(0));
  const T mdotm = rr * a_half * min_(Mhalf, T(0));
  F[0] = mdotp + mdotm;
  F[1] = n.madd(mdotp * ul + mdotm * ur, 0, phalf);
  F[2] = n.madd(mdotp * vl + mdotm * vr, 1, phalf);
  F[3] = n.madd(mdotp * wl + mdotm * wr, 2, phalf);
  F[4] = mdotp * H_L + mdotm * H_R;
}

// HLLE: the two-wave HLL flux with Batten's (Einfeldt's) speed estimates, riemann_HLLE
// (Lib_Riemann_HLL.f90:261-320):  F = (S4 F_L - S1 F_R + S1 S4 (U_R - U_L)) / (S4 - S1).
template <typename T, typename NT>
__device__ __forceinline__ void hlle(const T* L, const T* R, const NT& n, const GasT<T>& g, T* F) {
  const T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const T rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const T irl = rcp_(rl), irr = rcp_(rr);
  const T al = sqrtp_(g.gam * pl * irl), ar = sqrtp_(g.gam * pr * irr);
  const T unl = n.dot(ul, vl, wl);
  const T unr = n.dot(ur, vr, wr);
  const T kgl = g.gam * g.rgm1;
  const T h0l = kgl * pl * irl + T(0.5) * (ul * ul + vl * vl + wl * wl);
  const T h0r = kgl * pr * irr + T(0.5) * (ur * ur + vr * vr + wr * wr);
  const T srl = sqrtp_(rl), srr = sqrtp_(rr);
  const T isr = rcp_(srl + srr);
  const T u_roe = (srr * ur + srl * ul) * isr;
  const T v_roe = (srr * vr + srl * vl) * isr;
  const T w_roe = (srr * wr + srl * wl) * isr;
  const T h0_roe = (srr * h0r + srl * h0l) * isr;
  const T a_roe = sqrtp_(g.gm1 * (h0_roe - T(0.5) * (u_roe * u_roe + v_roe * v_roe + w_roe * w_roe)));
  const T un_roe = n.dot(u_roe, v_roe, w_roe);
  const T S1 = min_(T(0), min_(unl - al, un_roe - a_roe));
  const T S4 = max_(T(0), max_(unr + ar, un_roe + a_roe));
  const T inv = rcp_(S4 - S1);
  const T a = S4 * inv, b = -S1 * inv, c = S1 * S4 * inv;
  const T FL = rl * unl, FR = rr * unr;
  F[0] = a * FL + b * FR + c * (rr - rl);
  F[1] = a * n.madd(FL * ul, 0, pl) + b * n.madd(FR * ur, 0, pr) + c * (rr * ur - rl * ul);
  F[2] = a * n.madd(FL * vl, 1, pl) + b * n.madd(FR * vr, 1, pr) + c * (rr * vr - rl * vl);
  F[3] = a * n.madd(FL * wl, 2, pl) + b * n.madd(FR * wr, 2, pr) + c * (rr * wr - rl * wl);
  F[4] = a * (FL * h0l) + b * (FR * h0r) + c * (rr * h0r - pr - (rl * h0l - pl));
}

// SLAU (Shima & Kitamura) and SLAU2 (Kitamura & Shima), riemann_SLAU_base / riemann_SLAU /
// riemann_SLAU2 (Lib_Riemann_SLAU.f90:19-140): mean sound speed aF = (aL + aR)/2, the
// low-Mach weight chi = (1 - min(1, |u|_rms/aF))^2, a density-weighted |u_n| blended with
// each side's by f = -clip(mL, -1, 0) clip(mR, 0, 1), cubic pressure weights BetaL/R; SLAU2
// replaces the chi term of the pressure flux by |u|_rms (BetaL + BetaR - 1) aF (rhoL + rhoR)/2.
template <bool V2, typename T, typename NT>
__device__ __forceinline__ void slau(const T* L, const T* R, const NT& n, const GasT<T>& g, T* F) {
  const T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const T rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const T irl = rcp_(rl), irr = rcp_(rr);
  const T al = sqrtp_(g.gam * pl * irl), ar = sqrtp_(g.gam * pr * irr);
  const T unl = n.dot(ul, vl, wl);
  const T unr = n.dot(ur, vr, wr);
  const T sql = ul * ul + vl * vl + wl * wl, sqr = ur * ur + vr * vr + wr * wr;
  const T aF = T(0.5) * (al + ar);
  const T iaF = rcp_(aF);
  const T mL = unl * iaF, mR = unr * iaF;
  const T vrms = sqrt_(T(0.5) * (sql + sqr));          // may be 0: the general square root
  const T mt = min_(T(1), vrms * iaF);
  const T chi = (T(1) - mt) * (T(1) - mt);
  const T aul = fabs_(unl), aur = fabs_(unr);
  const T vn = (rl * aul + rr * aur) * rcp_(rl + rr);
  const T f = -max_(min_(mL, T(0)), T(-1)) * min_(max_(mR, T(0)), T(1));
  const T vnl = (T(1) - f) * vn + f * aul, vnr = (T(1) - f) * vn + f * aur;
  const T Fr = T(0.5) * (rl * (unl + vnl) + rr * (unr - vnr) - chi * iaF * (pr - pl));
  const T bl = fabs_(mL) < T(1) ? T(0.25) * (T(2) - mL) * (mL + T(1)) * (mL + T(1)) : (mL >= T(0) ? T(1) : T(0));
  const T br = fabs_(mR) < T(1) ? T(0.25) * (T(2) + mR) * (mR - T(1)) * (mR - T(1)) : (mR >= T(0) ? T(0) : T(1));
  T ph = T(0.5) * (pl + pr) + T(0.5) * (bl - br) * (pl - pr);
  if constexpr (V2) ph += vrms * (bl + br - T(1)) * aF * T(0.5) * (rl + rr);
  else ph += (T(1) - chi) * (bl + br - T(1)) * T(0.5) * (pl + pr);
  const T kgl = g.gam * g.rgm1;
  const bool left = Fr >= T(0);
  const T u_ = left ? ul : ur, v_ = left ? vl : vr, w_ = left ? wl : wr;
  const T h0 = left ? kgl * pl * irl + T(0.5) * sql : kgl * pr * irr + T(0.5) * sqr;
  F[0] = Fr;
  F[1] = n.madd(Fr * u_, 0, ph);
  F[2] = n.madd(Fr * v_, 1, ph);
  F[3] = n.madd(Fr * w_, 2, ph);
  F[4] = Fr * h0;
}

// HLLE++ (Tramel), riemann_HLLEpp (Lib_Riemann_HLL.f90:321-440): the central flux minus the
// Roe wave decomposition (contact, two shear, two acoustic waves) with eigenvalue magnitudes
// beta |lambda|_Roe (Harten-Hyman fix, eps = 0.3 (|u_n| + a)) + (1 - beta) |lambda|_HLLE,
// |lambda|_HLLE = |((S4 + S1) lambda - 2 S4 S1) / (S4 - S1)| with Batten's S1, S4.
template <typename T>
__device__ __forceinline__ T harten_hyman(T v, T e) {
  const T av = fabs_(v);
  return av <= e ? (v * v + e * e) * rcp_(T(2) * e) : av;
}

template <typename T, typename NT>
__device__ __forceinline__ void hllepp(const T* L, const T* R, const NT& n, const GasT<T>& g, T beta, T* F) {
  const T nx = n.x(), ny = n.y(), nz = n.z();
  const T rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const T rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const T irl = rcp_(rl), irr = rcp_(rr);
  const T al = sqrtp_(g.gam * pl * irl), ar = sqrtp_(g.gam * pr * irr);
  const T unl = ul * nx + vl * ny + wl * nz;
  const T unr = ur * nx + vr * ny + wr * nz;
  const T kgl = g.gam * g.rgm1;
  const T h0l = kgl * pl * irl + T(0.5) * (ul * ul + vl * vl + wl * wl);
  const T h0r = kgl * pr * irr + T(0.5) * (ur * ur + vr * vr + wr * wr);
  const T srl = sqrtp_(rl), srr = sqrtp_(rr);
  const T isr = rcp_(srl + srr);
  const T rho_roe = srl * srr;
  const T u_roe = (srr * ur + srl * ul) * isr;
  const T v_roe = (srr * vr + srl * vl) * isr;
  const T w_roe = (srr * wr + srl * wl) * isr;
  const T h0_roe = (srr * h0r + srl * h0l) * isr;
  const T q2 = u_roe * u_roe + v_roe * v_roe + w_roe * w_roe;
  const T a_roe = sqrtp_(g.gm1 * (h0_roe - T(0.5) * q2));
  const T un_roe = u_roe * nx + v_roe * ny + w_roe * nz;
  const T Sm = min_(T(0), min_(unl - al, un_roe - a_roe));
  const T Sp = max_(T(0), max_(unr + ar, un_roe + a_roe));
  const T inv = rcp_(Sp - Sm);
  const T c1 = (Sp + Sm) * inv, c0 = T(2) * Sp * Sm * inv;
  const T l1h = fabs_(c1 * un_roe - c0), l2h = fabs_(c1 * (un_roe - a_roe) - c0), l3h = fabs_(c1 * (un_roe + a_roe) - c0);
  const T eps = T(0.3) * (fabs_(un_roe) + a_roe);
  const T l1 = beta * fabs_(un_roe) + (T(1) - beta) * l1h;
  const T l2 = beta * harten_hyman(un_roe - a_roe, eps) + (T(1) - beta) * l2h;
  const T l3 = beta * harten_hyman(un_roe + a_roe, eps) + (T(1) - beta) * l3h;
  // tangential basis
  T t1x, t1y, t1z;
  if (fabs_(nx) < T(0.9)) { t1x = T(0); t1y = -nz; t1z = ny; }
  else { t1x = -nz; t1y = T(0); t1z = nx; }
  const T it = rcp_(sqrt_(t1x * t1x + t1y * t1y + t1z * t1z));
  t1x *= it; t1y *= it; t1z *= it;
  const T t2x = ny * t1z - nz * t1y, t2y = nz * t1x - nx * t1z, t2z = nx * t1y - ny * t1x;
  // jumps and wave strengths
  const T dr = rr - rl, dp = pr - pl, du = ur - ul, dv = vr - vl, dw = wr - wl, dun = unr - unl;
  const T ia2 = rcp_(a_roe * a_roe);
  const T ra = rho_roe * a_roe;
  const T am = (dp - ra * dun) * T(0.5) * ia2, ap = (dp + ra * dun) * T(0.5) * ia2;
  const T ac = dr - dp * ia2;
  const T as1 = rho_roe * (du * t1x + dv * t1y + dw * t1z), as2 = rho_roe * (du * t2x + dv * t2y + dw * t2z);
  const T FL = rl * unl, FR = rr * unr;
  const T wm = am * l2, wp = ap * l3, wc = ac * l1;
  F[0] = T(0.5) * (FL + FR) - T(0.5) * (wc + wm + wp);
  F[1] = T(0.5) * (FL * ul + pl * nx + FR * ur + pr * nx) -
         T(0.5) * (wc * u_roe + l1 * (as1 * t1x + as2 * t2x) + wm * (u_roe - a_roe * nx) + wp * (u_roe + a_roe * nx));
  F[2] = T(0.5) * (FL * vl + pl * ny + FR * vr + pr * ny) -
         T(0.5) * (wc * v_roe + l1 * (as1 * t1y + as2 * t2y) + wm * (v_roe - a_roe * ny) + wp * (v_roe + a_roe * ny));
  F[3] = T(0.5) * (FL * wl + pl * nz + FR * wr + pr * nz) -
         T(0.5) * (wc * w_roe + l1 * (as1 * t1z + as2 * t2z) + wm * (w_roe - a_roe * nz) + wp * (w_roe + a_roe * nz));
  F[4] = T(0.5) * (FL * h0l + FR * h0r) -
         T(0.5) * (wc * T(0.5) * q2 + l1 * (as1 * (u_roe * t1x + v_roe * t1y + w_roe * t1z) +
                                            as2 * (u_roe * t2x + v_roe * t2y + w_roe * t2z)) +
                   wm * (h0_roe - a_roe * un_roe) + wp * (h0_roe + a_roe * un_roe));
}

// beta: the shock detector's weight of the face, read by HLLE++ and HLLC+ only.  HLLC+ is
// riemann_HLLCSD (Lib_Riemann_HLL.f90:486-534): HLLE at beta = 0, HLLC at beta >= 1, else
// beta F_HLLC + (1 - beta) F_HLLE.
template <int RS, typename T, typename NT>
__device__ __forceinline__ void riemann(const T* L, const T* R, const NT& n, const GasT<T>& g, T* F, T beta = T(1)) {
  if constexpr (RS == RS_HLLEPP) hllepp(L, R, n, g, beta, F);
  else if constexpr (RS == RS_HLLCP) {
    if (beta == T(0)) hlle(L, R, n, g, F);
    else if (beta >= T(1)) hllc(L, R, n, g, F);
    else {
      T Fe[5];
      hlle(L, R, n, g, Fe);
      hllc(L, R, n, g, F);
#pragma unroll
      for (int v = 0; v < 5; ++v) F[v] = beta * F[v] + (T(1) - beta) * Fe[v];
    }
  }
  else if constexpr (RS == RS_AUSMP) ausmp(L, R, n, g, F);
  else if constexpr (RS == RS_HLLE) hlle(L, R, n, g, F);
  else if constexpr (RS == RS_SLAU) slau<false>(L, R, n, g, F);
  else if constexpr (RS == RS_SLAU2) slau<true>(L, R, n, g, F);
  else hllc(L, R, n, g, F);
}
template <int RS, typename T>
__device__ __forceinline__ void riemann(const T* L, const T* R, T nx, T ny, T nz, const GasT<T>& g, T* F,
                                        T beta = T(1)) {
  riemann<RS>(L, R, NGen<T>{nx, ny, nz}, g, F, beta);
}

// Interior / connection face: reconstruction + Riemann, flux per unit area (Convective_Flux).
// For a single species the upwinded species flux equals the mass flux exactly.
template <int LIM, int RS, typename T>
__device__ __forceinline__ bool face_flux(const T* qm1, const T* q0, const T* q1, const T* q2,
                                          const FaceSpacing<T>& s, T nx, T ny, T nz,
                                          const GasT<T>& g, T* F) {
  T L[5], R[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) reconstruct<LIM>(qm1[v], q0[v], q1[v], q2[v], s, L[v], R[v]);
  const bool ok = gas_ok(L) && gas_ok(R);
  riemann<RS>(L, R, nx, ny, nz, g, F);
  return ok;
}

// Physical flux of one state through a face (BC_Extrapolation, Lib_BC_Fluxes_Extrapolation.f90:13-47).
template <typename T>
__device__ __forceinline__ void physical_flux(const T* P, T nx, T ny, T nz, const GasT<T>& g, T* F) {
  const T un = P[1] * nx + P[2] * ny + P[3] * nz;
  const T m = P[0] * un;
  const T h0 = g.gam * g.rgm1 * P[4] / P[0] + T(0.5) * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]);
  F[0] = m;
  F[1] = m * P[1] + P[4] * nx;
  F[2] = m * P[2] + P[4] * ny;
  F[3] = m * P[3] + P[4] * nz;
  F[4] = m * h0;
}

// Symmetry face (BC_Symmetry_Eul, Lib_BC_Fluxes_Symmetry.f90:14-125, Un_wall = 0): limited
// face state of the boundary cell against its mirror, Riemann problem between the face
// state and its mirror, momentum part only.  `hi` is true for a max face (modfm2 = +1).
// Pb: boundary cell, Pi: interior neighbour, bdl/idl: half lengths of the two cells.
template <int LIM, int RS, typename T>
__device__ __forceinline__ void symmetry_flux(const T* Pb, const T* Pi, T bdl, T idl, bool hi,
                                              T nx, T ny, T nz, const GasT<T>& g, T* F) {
  const T un = Pb[1] * nx + Pb[2] * ny + Pb[3] * nz;
  T Gp[5] = {Pb[0], Pb[1] - T(2) * un * nx, Pb[2] - T(2) * un * ny, Pb[3] - T(2) * un * nz, Pb[4]};
  // stencil (1,2,3) = (ghost, boundary, interior) on a min face, (interior, boundary, ghost) on a max face
  const T d12 = hi ? (bdl + idl) : T(2) * bdl;
  const T d23 = hi ? T(2) * bdl : (bdl + idl);
  const T rd12 = T(1) / d12, rd23 = T(1) / d23;
  const T dist = hi ? bdl : -bdl;
  T Fc[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const T p1 = hi ? Pi[v] : Gp[v];
    const T p3 = hi ? Gp[v] : Pi[v];
    const T s32 = (p3 - Pb[v]) * rd23;
    const T s21 = (Pb[v] - p1) * rd12;
    Fc[v] = Pb[v] + limiter<LIM>(s32, s21) * dist;
  }
  const T unf = Fc[1] * nx + Fc[2] * ny + Fc[3] * nz;
  const T un1 = hi ? unf : -unf;
  T L[5] = {Fc[0], Fc[1] - unf * nx + un1 * nx, Fc[2] - unf * ny + un1 * ny, Fc[3] - unf * nz + un1 * nz, Fc[4]};
  T R[5] = {Fc[0], Fc[1] - unf * nx - un1 * nx, Fc[2] - unf * ny - un1 * ny, Fc[3] - unf * nz - un1 * nz, Fc[4]};
  T Fr[5];
  riemann<RS>(L, R, nx, ny, nz, g, Fr);
  F[0] = T(0); F[1] = Fr[1]; F[2] = Fr[2]; F[3] = Fr[3]; F[4] = T(0);
}

// Velocity direction of a 405 record: precomputed on the host, or, where bc.txt gives the
// 'normal' token for an angle (flag bits in bc[7]: 1 alpha, 2 beta; the explicit angles then
// sit in bc[3], bc[4]), from the boundary face normal (Compute_Inflow_Direction,
// Lib_BC_Fluxes_Inflow.f90:479-496: alpha = atan(n_y / (n_x + 1e-20)), beta = atan(n_z / ...)).
template <typename T>
__device__ __forceinline__ void inflow_dir(const double* bc, T nx, T ny, T nz, T* dir) {
  const int flags = (int)bc[7];
  if (flags == 0) {
    dir[0] = T(bc[3]); dir[1] = T(bc[4]); dir[2] = T(bc[5]);
    return;
  }
  const double a = (flags & 1) ? atan((double)ny / ((double)nx + 1e-20)) : bc[3];
  const double b = (flags & 2) ? atan((double)nz / ((double)nx + 1e-20)) : bc[4];
  dir[0] = T(cos(a) * cos(b));
  dir[1] = T(sin(a) * cos(b));
  dir[2] = T(sin(b));
}

// Supersonic inflow face (BC_Inlet_Supersonic_Static + Supersonic_Inflow,
// Lib_BC_Fluxes_Inflow.f90:244-382).  Returns false when the boundary-cell flow leaves the
// domain (the caller then uses the symmetry flux).  bc = {mach, T, p, cos(alpha) cos(beta),
// sin(alpha) cos(beta), sin(beta), rel_fac} (inflow direction precomputed on the host).
template <int LIM, int RS, typename T>
__device__ __forceinline__ bool inflow405_flux(const T* Pb, const T* Pi, bool hi, T nx, T ny, T nz,
                                               const double* bc, const GasT<T>& g, T* F) {
  const T un = Pb[1] * nx + Pb[2] * ny + Pb[3] * nz;
  const T sgn = hi ? T(1) : T(-1);
  if (sgn * un >= T(0)) return false;
  const T flow_a = sqrt_(g.gam * Pb[4] / Pb[0]);
  const T flow_M = sqrt_(Pb[1] * Pb[1] + Pb[2] * Pb[2] + Pb[3] * Pb[3]) / flow_a;
  const T rel = T(bc[6]);
  const T M = rel * T(bc[0]) + (T(1) - rel) * flow_M;
  const T p = rel * T(bc[2]) + (T(1) - rel) * Pb[4];
  const T rho = p / (g.R * T(bc[1]));
  const T a = sqrt_(g.gam * p / rho);
  T dir[3];
  inflow_dir(bc, nx, ny, nz, dir);
  T Sup[5] = {rho, M * a * dir[0], M * a * dir[1], M * a * dir[2], p};
  T Fc[5], L[5], R[5];
  const T half = hi ? T(0.5) : T(-0.5);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const T p1 = hi ? Pi[v] : Sup[v];
    const T p3 = hi ? Sup[v] : Pi[v];
    Fc[v] = Pb[v] + half * limiter<LIM>(p3 - Pb[v], Pb[v] - p1);
    L[v] = hi ? Fc[v] : Sup[v];
    R[v] = hi ? Sup[v] : Fc[v];
  }
  riemann<RS>(L, R, nx, ny, nz, g, F);
  return true;
}

// Subsonic inlet with stagnation conditions, BC 401 (BC_Inlet_StagnCond + Stagn_Inflow,
// Lib_BC_Fluxes_Inflow.f90): from the outgoing Riemann invariant of the boundary cell and the
// imposed T0, p0 and flow direction, the face state on the isentrope through (T0, p0); when the
// quadratic for the face normal velocity has no real root the imposed p0 is relaxed towards the
// cell pressure (rel_fac) until it has one (at most 1000 times).  Returns false when the
// boundary-cell flow leaves the domain (the caller then uses the symmetry flux).  FP64;
// bc = {T0, p0, alpha, beta, rel_fac, -, -, flags}, t the unit face diagonal.
__device__ inline bool inlet401_flux(const double* Pb, bool hi, const double* n, const double* t, const double* bc,
                                     double gam, double Rg, double* F) {
  const double sign = hi ? 1.0 : -1.0;
  const double un = Pb[1] * n[0] + Pb[2] * n[1] + Pb[3] * n[2];
  if (sign * un >= 0.0) return false;
  const double snd = sqrt(gam * Pb[4] / Pb[0]);
  const double riem = snd + sign * 0.5 * (gam - 1.0) * un;
  const int flags = (int)bc[7];
  const double alpha = (flags & 1) ? atan(n[1] / (n[0] + 1e-20)) : bc[2];
  const double beta = (flags & 2) ? atan(n[2] / (n[0] + 1e-20)) : bc[3];
  // b = n x t; velocity components along t and b per unit normal velocity, E1 / E2
  double b[3] = {n[1] * t[2] - n[2] * t[1], n[2] * t[0] - n[0] * t[2], n[0] * t[1] - n[1] * t[0]};
  const double bm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  b[0] /= bm; b[1] /= bm; b[2] /= bm;
  const double ta = tan(alpha), tb = tan(beta);
  const double XA = n[1] - n[0] * ta, XB = t[1] - t[0] * ta, XC = b[1] - b[0] * ta;
  const double XD = n[2] - n[0] * tb, XE = t[2] - t[0] * tb, XF = b[2] - b[0] * tb;
  const double check = XC * XE - XF * XB;
  const double E1 = (XF * XA - XC * XD) / check, E2 = (XD * XB - XE * XA) / check;
  const double T0 = bc[0], rel = bc[4];
  double p0 = bc[1];
  const double snd0 = sqrt(gam * Rg * T0);
  const double cD = (1.0 + E1 * E1 + E2 * E2) * (gam - 1.0) / 2.0;
  double YA = 0.0, YB = 0.0, disc = -1.0;
  for (int it = 1; disc < 0.0; ++it) {
    double cC = pow(p0 / Pb[4], (gam - 1.0) / gam);
    cC = cC * ((snd / snd0) * (snd / snd0));
    YA = ((gam - 1.0) / 2.0) * ((gam - 1.0) / 2.0) + cC * cD;
    YB = sign * riem * ((gam - 1.0) / 2.0);
    const double YC = riem * riem - cC * snd0 * snd0;
    disc = YB * YB - YA * YC;
    if (disc < 0.0 && fabs(disc) < 1e-2) disc = 0.0;
    if (disc < 0.0 && fabs(disc) > 1e-2) p0 = rel * p0 + (1.0 - rel) * Pb[4];
    if (it > 1000) { disc = 0.0; break; }
  }
  const double un3 = (YB + (un >= 0.0 ? 1.0 : -1.0) * sqrt(disc)) / YA;
  const double snd3 = sqrt(snd0 * snd0 - cD * un3 * un3);
  const double p3 = p0 * pow(snd3 / snd0, 2.0 * gam / (gam - 1.0));
  const double rho = gam * p3 / (snd3 * snd3);
  double u[3];
  for (int q = 0; q < 3; ++q) u[q] = un3 * n[q] + un3 * E1 * t[q] + un3 * E2 * b[q];
  const double h = gam / (gam - 1.0) * p3 / rho;
  const double m = rho * un3;
  F[0] = m;
  for (int q = 0; q < 3; ++q) F[1 + q] = m * u[q] + p3 * n[q];
  F[4] = m * (0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + h);
  return true;
}

// Outlet, BC 406 (BC_Outflow + Outflow, Lib_BC_Fluxes_Outflow.f90:12-130): subsonic normal Mach
// and a back pressure p_exit > 0: the face pressure is max(p_exit, p*) (p* the cell's critical
// pressure), the face state follows the isentrope from the cell and its outgoing Riemann
// invariant, tangential velocity extrapolated; otherwise the cell state.  Returns false when the
// boundary-cell flow enters the domain (symmetry flux instead).  FP64; bc = {p_exit, rel_fac}.
__device__ inline bool outflow406_flux(const double* Pb, bool hi, const double* n, const double* bc, double gam,
                                       double* F) {
  const double sign = hi ? 1.0 : -1.0;
  const double un = Pb[1] * n[0] + Pb[2] * n[1] + Pb[3] * n[2];
  if (sign * un < 0.0) return false;
  const double snd = sqrt(gam * Pb[4] / Pb[0]);
  const double riem = snd + sign * 0.5 * (gam - 1.0) * un;
  const double vt[3] = {Pb[1] - un * n[0], Pb[2] - un * n[1], Pb[3] - un * n[2]};
  const double machn = fabs(un / snd);
  double rho = Pb[0], p = Pb[4], ufn = un, u[3] = {Pb[1], Pb[2], Pb[3]};
  double pe = bc[0];
  if (machn < 1.0 && pe > 0.0) {
    const double delta = 0.5 * (gam - 1.0);
    const double den = 1.0 / pow(1.0 + delta, gam / (gam - 1.0));
    const double pstar = Pb[4] * den * pow(1.0 + delta * machn * machn, gam / (gam - 1.0));
    if (pstar >= pe) pe = pstar;
    const double sndf = snd * pow(pe / Pb[4], (gam - 1.0) / (2.0 * gam));
    ufn = (riem - sndf) / (sign * 0.5 * (gam - 1.0));
    p = pe;
    rho = gam * p / (sndf * sndf);
    for (int q = 0; q < 3; ++q) u[q] = vt[q] + ufn * n[q];
  }
  const double h = gam / (gam - 1.0) * p / rho;
  const double m = rho * ufn;
  F[0] = m;
  for (int q = 0; q < 3; ++q) F[1 + q] = u[q] * m + p * n[q];
  F[4] = m * (0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + h);
  return true;
}

// ---------------------------------------------------------------------------- viscous
// Laminar viscosity by Sutherland's law, mu = mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S).
template <typename T>
__device__ __forceinline__ T sutherland(T temp, T mu_ref, T T_ref, T S) {
  const T t = temp * rcp_(T_ref);
  return mu_ref * t * sqrt_(t) * (T_ref + S) * rcp_(temp + S);
}

// Compute_Diffusive_Flux (Lib_Diffusive.f90:90-174, laminar, one species) with
// Stress_Vector_Std (Lib_Fluid.f90:23-46) on an axis-aligned face: G[v][a] are the
// index-space gradients of (u, v, w, T) along axis a, M[a] the averaged diagonal metric
// (1/dl), Pm = (rho, u, v, w, p) the face-averaged primitive state, vis = {mu_ref, T_ref, S,
// cp/Pr}.  Returns the flux through the face (area-scaled); the face adds -Flux to the cell
// on its low side and +Flux to the cell on its high side.
template <typename T>
__device__ __forceinline__ void visc_face_flux(int d, const T G[4][3], const T* M, const T* Pm, T area, T Rg,
                                               const T* vis, T* F) {
  T Gp[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int a = 0; a < 3; ++a) Gp[v][a] = G[v][a] * M[a];
  const T temp = Pm[4] * rcp_(Pm[0] * Rg);
  const T mu = sutherland(temp, vis[0], vis[1], vis[2]);
  const T kap = mu * vis[3];
  const T lam = T(-2.0 / 3.0) * mu;
  const T ld = lam * (Gp[0][0] + Gp[1][1] + Gp[2][2]);
  // stress vector tau . n with n the axis d: column d of the (symmetric) stress tensor
  T st[3];
  if (d == 0) {
    st[0] = ld + T(2) * mu * Gp[0][0]; st[1] = mu * (Gp[0][1] + Gp[1][0]); st[2] = mu * (Gp[0][2] + Gp[2][0]);
  } else if (d == 1) {
    st[0] = mu * (Gp[0][1] + Gp[1][0]); st[1] = ld + T(2) * mu * Gp[1][1]; st[2] = mu * (Gp[1][2] + Gp[2][1]);
  } else {
    st[0] = mu * (Gp[0][2] + Gp[2][0]); st[1] = mu * (Gp[1][2] + Gp[2][1]); st[2] = ld + T(2) * mu * Gp[2][2];
  }
  const T gTn = d == 0 ? Gp[3][0] : (d == 1 ? Gp[3][1] : Gp[3][2]);
  F[0] = T(0);
  F[1] = area * st[0];
  F[2] = area * st[1];
  F[3] = area * st[2];
  F[4] = area * ((st[0] * Pm[1] + st[1] * Pm[2] + st[2] * Pm[3]) + kap * gTn);
}

// --------------------------------------------------------------------------- thermo (FP64)
// Conservative <-> primitive for the calorically perfect gas; always FP64 (the pressure
// recovery p = (gamma-1)(E - |m|^2/(2 rho)) stays FP64 near vacuum).
__device__ __forceinline__ void cons2prim(const double* U, double gm1, double* P) {
  const double ir = rcp_(U[0]);
  const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  P[0] = U[0]; P[1] = u; P[2] = v; P[3] = w;
  P[4] = gm1 * (U[4] - 0.5 * (U[1] * u + U[2] * v + U[3] * w));
}
__device__ __forceinline__ void prim2cons(const double* P, double rgm1, double* U) {
  U[0] = P[0]; U[1] = P[0] * P[1]; U[2] = P[0] * P[2]; U[3] = P[0] * P[3];
  U[4] = P[4] * rgm1 + 0.5 * P[0] * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]);
}

// SD_Tramel (Lib_Shock_Detector.f90:9-38, DELTA = 20): Jameson-type pressure sensor of a cell
// from its own pressure and its six face neighbours' (p1/p2 along i, p3/p4 along j, p5/p6
// along k); 1 - tanh(10 (sw/DELTA)^3) below sw = 1/DELTA, else 0.
__device__ __forceinline__ double sd_tramel(double p, double p1, double p2, double p3, double p4, double p5,
                                            double p6) {
  const double kp1 = fabs((p2 - 2.0 * p + p1) / (p2 + 2.0 * p + p1));
  const double kp2 = fabs((p4 - 2.0 * p + p3) / (p4 + 2.0 * p + p3));
  const double kp3 = fabs((p6 - 2.0 * p + p5) / (p6 + 2.0 * p + p5));
  const double sw = fmax(kp1, fmax(kp2, kp3));
  if (sw < 1.0 / 20.0) {
    const double phi = fmax(sw / 20.0, 0.0);
    return 1.0 - tanh(10.0 * phi * phi * phi);
  }
  return 0.0;
}

}  // namespace mose
