// mose_adjoint.hpp — reverse-mode (discrete adjoint) of the face-flux math, FP64, gfx950.
//
// Each *_adj routine recomputes its forward pass (branch decisions included: limiter
// regimes, min/max of the wave speeds, the HLLC side, AUSM+ Mach regimes — nothing is
// stored by the primal) and then propagates the output adjoint back to its inputs,
// accumulating (+=) into the input-adjoint arrays.  Forward formulas: mose_device.hpp.
#pragma once

namespace mose {

// d phi / d(x, y) of limiter<LIM>(x, y), times gb, added to xb / yb.
template <int LIM>
__device__ __forceinline__ void limiter_adj(double x, double y, double gb, double& xb, double& yb) {
  if constexpr (LIM == LIM_MINMOD) {
    // where a slope is exactly zero (uniform regions) minmod clips perturbations of either
    // sign of the other argument to first order; the adjoint takes the zero subgradient there,
    // as the van Leer and van Albada adjoints do
    if (!(x * y > 0.0)) return;
    if (fabs(x) > fabs(y)) yb += gb;
    else xb += gb;
  } else if constexpr (LIM == LIM_VANLEER) {
    const double xy = x * y;
    if (!(xy > 0.0)) return;
    const double s = x + y, is2 = 1.0 / (s * s);
    xb += gb * 2.0 * y * y * is2;     // d(2xy/(x+y))/dx = 2 y^2/(x+y)^2
    yb += gb * 2.0 * x * x * is2;
  } else if constexpr (LIM == LIM_VANALBADA) {
    const double q = x * x + y * y;
    if (!(q > 0.0)) return;
    const double n = x * y * (x + y), iq = 1.0 / q;
    // f = n / q, n = x^2 y + x y^2
    const double nx = 2.0 * x * y + y * y, ny = x * x + 2.0 * x * y;
    xb += gb * (nx * q - n * 2.0 * x) * iq * iq;
    yb += gb * (ny * q - n * 2.0 * y) * iq * iq;
  } else if constexpr (LIM == LIM_MC) {
    // sgn(x) min_(2 min_(|x|, |y|), (|x|+|y|)/2): the selected piece's slope (x, y same sign)
    if (!(x * y > 0.0)) return;
    const double ax = fabs(x), ay = fabs(y);
    if (2.0 * min_(ax, ay) < 0.5 * (ax + ay)) {
      if (ax < ay) xb += 2.0 * gb;
      else yb += 2.0 * gb;
    } else {
      xb += 0.5 * gb;
      yb += 0.5 * gb;
    }
  } else if constexpr (LIM == LIM_SUPERBEE) {
    // sgn(x) max_(min_(2|y|, |x|), min_(|y|, 2|x|))
    if (!(x * y > 0.0)) return;
    const double ax = fabs(x), ay = fabs(y);
    if (min_(2.0 * ay, ax) > min_(ay, 2.0 * ax)) {
      if (2.0 * ay < ax) yb += 2.0 * gb;
      else xb += gb;
    } else {
      if (ay < 2.0 * ax) yb += gb;
      else xb += 2.0 * gb;
    }
  }
}

// Per-cell slopes and face states (cell_states): Sm = P0 - sigma hdl, Sp = P0 + sigma hdl,
// sigma = phi((Pp - P0) rlp, (P0 - Pm) rlm).  Adjoints Smb / Spb -> Pmb, P0b, Ppb and, when
// gb3 is given, the spacing adjoints gb3 += (rlm_bar, rlp_bar, hdl_bar).
template <int LIM>
__device__ __forceinline__ void cell_states_adj(const double* Pm, const double* P0, const double* Pp, double rlm,
                                                double rlp, double hdl, const double* Smb, const double* Spb,
                                                double* Pmb, double* P0b, double* Ppb, double* gb3 = nullptr) {
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double sm = (P0[v] - Pm[v]) * rlm;
    const double sp = (Pp[v] - P0[v]) * rlp;
    const double gb = (Spb[v] - Smb[v]) * hdl;
    double spb = 0.0, smb = 0.0;
    limiter_adj<LIM>(sp, sm, gb, spb, smb);
    P0b[v] += Smb[v] + Spb[v] - spb * rlp + smb * rlm;
    Ppb[v] += spb * rlp;
    Pmb[v] -= smb * rlm;
    if (gb3) {
      gb3[0] += smb * (P0[v] - Pm[v]);
      gb3[1] += spb * (Pp[v] - P0[v]);
      gb3[2] += (Spb[v] - Smb[v]) * limiter<LIM>(sp, sm);
    }
  }
}

// Adjoints of one side's quantities of the HLL-family front end.
struct SideBar { double r, u, v, w, p, un, h0, ir; };

// Reverse of the front end HLLC and HLLE share: a = sqrt(gam p/rho), un = u.n,
// h0 = gam/(gam-1) p/rho + |u|^2/2, the Roe averages and Batten's speeds
// S1 = min(0, unl - al, un_roe - a_roe), S4 = max(0, unr + ar, un_roe + a_roe).
// l / r carry the adjoints already gathered on each side's (rho, u, v, w, p, un, h0, 1/rho);
// the result is added to Lb, Rb and nb.
__device__ __forceinline__ void roe_batten_adj(const double* L, const double* R, double nx, double ny, double nz,
                                               const GasT<double>& g, double S1b, double S4b, SideBar l, SideBar r,
                                               double* Lb, double* Rb, double* nb) {
  const double rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const double rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double al = sqrt(g.gam * pl * irl), ar = sqrt(g.gam * pr * irr);
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double kg = g.gam * g.rgm1;
  const double h0l = kg * pl * irl + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double h0r = kg * pr * irr + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double srl = sqrt(rl), srr = sqrt(rr), isr = 1.0 / (srl + srr);
  const double u_roe = (srr * ur + srl * ul) * isr, v_roe = (srr * vr + srl * vl) * isr, w_roe = (srr * wr + srl * wl) * isr;
  const double h0_roe = (srr * h0r + srl * h0l) * isr;
  const double a_roe = sqrt(g.gm1 * (h0_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe + w_roe * w_roe)));
  const double un_roe = u_roe * nx + v_roe * ny + w_roe * nz;
  const double x1 = unl - al, y1 = un_roe - a_roe, m1 = min_(x1, y1);
  const double x4 = unr + ar, y4 = un_roe + a_roe, m4 = max_(x4, y4);
  double rlb = l.r, ulb = l.u, vlb = l.v, wlb = l.w, plb = l.p, unlb = l.un, h0lb = l.h0, irlb = l.ir;
  double rrb = r.r, urb = r.u, vrb = r.v, wrb = r.w, prb = r.p, unrb = r.un, h0rb = r.h0, irrb = r.ir;
  // S1 = min_(0, min_(x1, y1)), S4 = max_(0, max_(x4, y4)): the selected argument gets the adjoint
  double alb = 0, arb = 0, un_roe_b = 0, a_roe_b = 0;
  if (!(0.0 < m1)) {
    if (x1 < y1) { unlb += S1b; alb -= S1b; }
    else { un_roe_b += S1b; a_roe_b -= S1b; }
  }
  if (!(0.0 > m4)) {
    if (x4 > y4) { unrb += S4b; arb += S4b; }
    else { un_roe_b += S4b; a_roe_b += S4b; }
  }
  // Roe averages
  double u_roe_b = un_roe_b * nx, v_roe_b = un_roe_b * ny, w_roe_b = un_roe_b * nz;
  nb[0] += un_roe_b * u_roe;
  nb[1] += un_roe_b * v_roe;
  nb[2] += un_roe_b * w_roe;
  double srrb = 0, srlb = 0, isrb = 0;
  if (a_roe > 0.0) {      // a_roe = 0: only the un_roe path (through u_roe) carries sensitivity
    const double hb = a_roe_b * 0.5 / a_roe * g.gm1;
    u_roe_b -= hb * u_roe;
    v_roe_b -= hb * v_roe;
    w_roe_b -= hb * w_roe;
    h0rb += hb * srr * isr;
    h0lb += hb * srl * isr;
    srrb += hb * h0r * isr;
    srlb += hb * h0l * isr;
    isrb += hb * (srr * h0r + srl * h0l);
  }
  srrb += (u_roe_b * ur + v_roe_b * vr + w_roe_b * wr) * isr;
  srlb += (u_roe_b * ul + v_roe_b * vl + w_roe_b * wl) * isr;
  urb += u_roe_b * srr * isr; vrb += v_roe_b * srr * isr; wrb += w_roe_b * srr * isr;
  ulb += u_roe_b * srl * isr; vlb += v_roe_b * srl * isr; wlb += w_roe_b * srl * isr;
  isrb += u_roe_b * (srr * ur + srl * ul) + v_roe_b * (srr * vr + srl * vl) + w_roe_b * (srr * wr + srl * wl);
  const double sb = -isr * isr * isrb;
  srlb += sb;
  srrb += sb;
  rlb += srlb * 0.5 / srl;
  rrb += srrb * 0.5 / srr;
  // h0 = kg p / rho + |u|^2 / 2
  plb += h0lb * kg * irl; irlb += h0lb * kg * pl;
  ulb += h0lb * ul; vlb += h0lb * vl; wlb += h0lb * wl;
  prb += h0rb * kg * irr; irrb += h0rb * kg * pr;
  urb += h0rb * ur; vrb += h0rb * vr; wrb += h0rb * wr;
  // un = u . n
  ulb += unlb * nx; vlb += unlb * ny; wlb += unlb * nz;
  urb += unrb * nx; vrb += unrb * ny; wrb += unrb * nz;
  nb[0] += unlb * ul + unrb * ur;
  nb[1] += unlb * vl + unrb * vr;
  nb[2] += unlb * wl + unrb * wr;
  // a = sqrt(gam p / rho)
  {
    const double tl = alb * 0.5 / al * g.gam, tr = arb * 0.5 / ar * g.gam;
    plb += tl * irl; irlb += tl * pl;
    prb += tr * irr; irrb += tr * pr;
  }
  rlb -= irlb * irl * irl;
  rrb -= irrb * irr * irr;
  Lb[0] += rlb; Lb[1] += ulb; Lb[2] += vlb; Lb[3] += wlb; Lb[4] += plb;
  Rb[0] += rrb; Rb[1] += urb; Rb[2] += vrb; Rb[3] += wrb; Rb[4] += prb;
}

// ------------------------------------------------------------------------------ HLLC
__device__ void hllc_adj(const double* L, const double* R, double nx, double ny, double nz, const GasT<double>& g,
                         const double* Fb, double* Lb, double* Rb, double* nb) {
  const double rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const double rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  // ---- forward
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double al = sqrt(g.gam * pl * irl), ar = sqrt(g.gam * pr * irr);
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double kg = g.gam * g.rgm1;
  const double h0l = kg * pl * irl + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double h0r = kg * pr * irr + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double srl = sqrt(rl), srr = sqrt(rr), isr = 1.0 / (srl + srr);
  const double u_roe = (srr * ur + srl * ul) * isr, v_roe = (srr * vr + srl * vl) * isr, w_roe = (srr * wr + srl * wl) * isr;
  const double h0_roe = (srr * h0r + srl * h0l) * isr;
  const double a_roe = sqrt(g.gm1 * (h0_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe + w_roe * w_roe)));
  const double un_roe = u_roe * nx + v_roe * ny + w_roe * nz;
  const double e0l = h0l - pl * irl, e0r = h0r - pr * irr;
  const double x1 = unl - al, y1 = un_roe - a_roe;
  const double m1 = min_(x1, y1);
  const double S1 = min_(0.0, m1);
  const double x4 = unr + ar, y4 = un_roe + a_roe;
  const double m4 = max_(x4, y4);
  const double S4 = max_(0.0, m4);
  const double dl = rl * (S1 - unl), dr = rr * (S4 - unr);
  const double Nn = pr - pl + dl * unl - dr * unr, Dd = dl - dr;
  const double Sstar = Nn / Dd;
  const double pstar = rl * (unl - S1) * (unl - Sstar) + pl;
  const bool left = (S1 >= 0.0) || (Sstar >= 0.0);
  const double r_ = left ? rl : rr, u_ = left ? ul : ur, v_ = left ? vl : vr, w_ = left ? wl : wr;
  const double p_ = left ? pl : pr, un_ = left ? unl : unr, h0_ = left ? h0l : h0r, e0_ = left ? e0l : e0r;
  const double S = left ? S1 : S4;
  const bool star = left ? (S1 < 0.0) : (S4 > 0.0);
  const double Fr = r_ * un_;
  // ---- reverse: side quantities
  double r_b = 0, u_b = 0, v_b = 0, w_b = 0, p_b = 0, un_b = 0, h0_b = 0, e0_b = 0, S_b = 0;
  double pstar_b = 0, Sstar_b = 0;
  if (star) {
    const double q = S - Sstar, inv = 1.0 / q;
    const double sd = (S - un_) * r_;
    const double dp = pstar - p_;
    const double n2 = sd * u_ + dp * nx, n3 = sd * v_ + dp * ny, n4 = sd * w_ + dp * nz;
    const double n5 = sd * e0_ - p_ * un_ + pstar * Sstar;
    const double U1S = sd * inv, U2S = n2 * inv, U3S = n3 * inv, U4S = n4 * inv, U5S = n5 * inv;
    // f_k += S (UkS - Uk), U = (r, r u, r v, r w, r e0)
    S_b += Fb[0] * (U1S - r_) + Fb[1] * (U2S - r_ * u_) + Fb[2] * (U3S - r_ * v_) + Fb[3] * (U4S - r_ * w_) +
           Fb[4] * (U5S - r_ * e0_);
    const double U1b = Fb[0] * S, U2b = Fb[1] * S, U3b = Fb[2] * S, U4b = Fb[3] * S, U5b = Fb[4] * S;
    r_b -= S * (Fb[0] + Fb[1] * u_ + Fb[2] * v_ + Fb[3] * w_ + Fb[4] * e0_);
    u_b -= S * Fb[1] * r_;
    v_b -= S * Fb[2] * r_;
    w_b -= S * Fb[3] * r_;
    e0_b -= S * Fb[4] * r_;
    const double sd_b = (U1b + U2b * u_ + U3b * v_ + U4b * w_ + U5b * e0_) * inv;
    const double inv_b = U1b * sd + U2b * n2 + U3b * n3 + U4b * n4 + U5b * n5;
    u_b += U2b * inv * sd;
    v_b += U3b * inv * sd;
    w_b += U4b * inv * sd;
    e0_b += U5b * inv * sd;
    const double dp_b = (U2b * nx + U3b * ny + U4b * nz) * inv;
    nb[0] += U2b * inv * dp;
    nb[1] += U3b * inv * dp;
    nb[2] += U4b * inv * dp;
    p_b -= U5b * inv * un_;
    un_b -= U5b * inv * p_;
    pstar_b += U5b * inv * Sstar;
    Sstar_b += U5b * inv * pstar;
    pstar_b += dp_b;
    p_b -= dp_b;
    S_b += sd_b * r_;
    un_b -= sd_b * r_;
    r_b += sd_b * (S - un_);
    const double q_b = -inv * inv * inv_b;
    S_b += q_b;
    Sstar_b -= q_b;
  }
  {
    const double Fr_b = Fb[0] + Fb[1] * u_ + Fb[2] * v_ + Fb[3] * w_ + Fb[4] * h0_;
    u_b += Fb[1] * Fr;
    v_b += Fb[2] * Fr;
    w_b += Fb[3] * Fr;
    h0_b += Fb[4] * Fr;
    p_b += Fb[1] * nx + Fb[2] * ny + Fb[3] * nz;
    nb[0] += Fb[1] * p_;
    nb[1] += Fb[2] * p_;
    nb[2] += Fb[3] * p_;
    r_b += Fr_b * un_;
    un_b += Fr_b * r_;
  }
  // scatter the side adjoints
  double rlb = 0, ulb = 0, vlb = 0, wlb = 0, plb = 0, unlb = 0, h0lb = 0, e0lb = 0, S1b = 0;
  double rrb = 0, urb = 0, vrb = 0, wrb = 0, prb = 0, unrb = 0, h0rb = 0, e0rb = 0, S4b = 0;
  if (left) { rlb = r_b; ulb = u_b; vlb = v_b; wlb = w_b; plb = p_b; unlb = un_b; h0lb = h0_b; e0lb = e0_b; S1b = S_b; }
  else { rrb = r_b; urb = u_b; vrb = v_b; wrb = w_b; prb = p_b; unrb = un_b; h0rb = h0_b; e0rb = e0_b; S4b = S_b; }
  // pstar = rl (unl - S1)(unl - Sstar) + pl
  {
    const double a1 = unl - S1, a2 = unl - Sstar;
    rlb += pstar_b * a1 * a2;
    unlb += pstar_b * rl * (a2 + a1);
    S1b -= pstar_b * rl * a2;
    Sstar_b -= pstar_b * rl * a1;
    plb += pstar_b;
  }
  // Sstar = Nn / Dd
  double dlb = 0, drb = 0;
  {
    const double Nb = Sstar_b / Dd, Db = -Sstar_b * Sstar / Dd;
    prb += Nb;
    plb -= Nb;
    dlb += Nb * unl;
    unlb += Nb * dl;
    drb -= Nb * unr;
    unrb -= Nb * dr;
    dlb += Db;
    drb -= Db;
  }
  rlb += dlb * (S1 - unl);
  S1b += dlb * rl;
  unlb -= dlb * rl;
  rrb += drb * (S4 - unr);
  S4b += drb * rr;
  unrb -= drb * rr;
  // e0 = h0 - p / rho
  h0lb += e0lb; plb -= e0lb * irl;
  h0rb += e0rb; prb -= e0rb * irr;
  const SideBar lb{rlb, ulb, vlb, wlb, plb, unlb, h0lb, -e0lb * pl};
  const SideBar rb{rrb, urb, vrb, wrb, prb, unrb, h0rb, -e0rb * pr};
  roe_batten_adj(L, R, nx, ny, nz, g, S1b, S4b, lb, rb, Lb, Rb, nb);
}

// HLLE (hlle in mose_device.hpp): F = a F_L + b F_R + c (U_R - U_L) with
// a = S4/(S4-S1), b = -S1/(S4-S1), c = S1 S4/(S4-S1).
__device__ void hlle_adj(const double* L, const double* R, double nx, double ny, double nz, const GasT<double>& g,
                         const double* Fb, double* Lb, double* Rb, double* nb) {
  const double rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const double rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  // ---- forward
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double al = sqrt(g.gam * pl * irl), ar = sqrt(g.gam * pr * irr);
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double kg = g.gam * g.rgm1;
  const double h0l = kg * pl * irl + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double h0r = kg * pr * irr + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double srl = sqrt(rl), srr = sqrt(rr), isr = 1.0 / (srl + srr);
  const double u_roe = (srr * ur + srl * ul) * isr, v_roe = (srr * vr + srl * vl) * isr, w_roe = (srr * wr + srl * wl) * isr;
  const double h0_roe = (srr * h0r + srl * h0l) * isr;
  const double a_roe = sqrt(g.gm1 * (h0_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe + w_roe * w_roe)));
  const double un_roe = u_roe * nx + v_roe * ny + w_roe * nz;
  const double S1 = min_(0.0, min_(unl - al, un_roe - a_roe));
  const double S4 = max_(0.0, max_(unr + ar, un_roe + a_roe));
  const double inv = 1.0 / (S4 - S1);
  const double FL = rl * unl, FR = rr * unr;
  const double fl[5] = {FL, FL * ul + pl * nx, FL * vl + pl * ny, FL * wl + pl * nz, FL * h0l};
  const double fr[5] = {FR, FR * ur + pr * nx, FR * vr + pr * ny, FR * wr + pr * nz, FR * h0r};
  const double du[5] = {rr - rl, rr * ur - rl * ul, rr * vr - rl * vl, rr * wr - rl * wl,
                        rr * h0r - pr - (rl * h0l - pl)};
  // ---- reverse: coefficients -> S1, S4
  double a_b = 0, b_b = 0, c_b = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) { a_b += Fb[k] * fl[k]; b_b += Fb[k] * fr[k]; c_b += Fb[k] * du[k]; }
  const double inv_b = a_b * S4 - b_b * S1 + c_b * S1 * S4;
  const double q_b = -inv * inv * inv_b;
  double S4b = a_b * inv + c_b * S1 * inv + q_b;
  double S1b = -b_b * inv + c_b * S4 * inv - q_b;
  // physical fluxes, weights a Fb (left) and b Fb (right); jumps, weight c Fb
  const double A = S4 * inv, B = -S1 * inv, C = S1 * S4 * inv;
  SideBar lb{0, 0, 0, 0, 0, 0, 0, 0}, rb{0, 0, 0, 0, 0, 0, 0, 0};
  {
    const double w0 = A * Fb[0], w1 = A * Fb[1], w2 = A * Fb[2], w3 = A * Fb[3], w4 = A * Fb[4];
    const double Fr_b = w0 + w1 * ul + w2 * vl + w3 * wl + w4 * h0l;
    lb.u += w1 * FL; lb.v += w2 * FL; lb.w += w3 * FL; lb.h0 += w4 * FL;
    lb.p += w1 * nx + w2 * ny + w3 * nz;
    nb[0] += w1 * pl; nb[1] += w2 * pl; nb[2] += w3 * pl;
    lb.r += Fr_b * unl; lb.un += Fr_b * rl;
  }
  {
    const double w0 = B * Fb[0], w1 = B * Fb[1], w2 = B * Fb[2], w3 = B * Fb[3], w4 = B * Fb[4];
    const double Fr_b = w0 + w1 * ur + w2 * vr + w3 * wr + w4 * h0r;
    rb.u += w1 * FR; rb.v += w2 * FR; rb.w += w3 * FR; rb.h0 += w4 * FR;
    rb.p += w1 * nx + w2 * ny + w3 * nz;
    nb[0] += w1 * pr; nb[1] += w2 * pr; nb[2] += w3 * pr;
    rb.r += Fr_b * unr; rb.un += Fr_b * rr;
  }
  {
    const double w0 = C * Fb[0], w1 = C * Fb[1], w2 = C * Fb[2], w3 = C * Fb[3], w4 = C * Fb[4];
    rb.r += w0 + w1 * ur + w2 * vr + w3 * wr + w4 * h0r;
    rb.u += w1 * rr; rb.v += w2 * rr; rb.w += w3 * rr; rb.h0 += w4 * rr; rb.p -= w4;
    lb.r -= w0 + w1 * ul + w2 * vl + w3 * wl + w4 * h0l;
    lb.u -= w1 * rl; lb.v -= w2 * rl; lb.w -= w3 * rl; lb.h0 -= w4 * rl; lb.p += w4;
  }
  roe_batten_adj(L, R, nx, ny, nz, g, S1b, S4b, lb, rb, Lb, Rb, nb);
}

// ------------------------------------------------------------------------------ SLAU / SLAU2
template <bool V2>
__device__ void slau_adj(const double* L, const double* R, double nx, double ny, double nz, const GasT<double>& g,
                         const double* Fb, double* Lb, double* Rb, double* nb) {
  const double rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const double rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  // ---- forward (slau in mose_device.hpp)
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double al = sqrt(g.gam * pl * irl), ar = sqrt(g.gam * pr * irr);
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double sql = ul * ul + vl * vl + wl * wl, sqr = ur * ur + vr * vr + wr * wr;
  const double aF = 0.5 * (al + ar), iaF = 1.0 / aF;
  const double mL = unl * iaF, mR = unr * iaF;
  const double vrms = sqrt(0.5 * (sql + sqr));
  const double mraw = vrms * iaF;
  const double mt = min_(1.0, mraw);
  const double chi = (1.0 - mt) * (1.0 - mt);
  const double aul = fabs(unl), aur = fabs(unr);
  const double rs = rl + rr, irs = 1.0 / rs;
  const double vn = (rl * aul + rr * aur) * irs;
  const double A = max_(min_(mL, 0.0), -1.0), B = min_(max_(mR, 0.0), 1.0);
  const double f = -A * B;
  const double vnl = (1.0 - f) * vn + f * aul, vnr = (1.0 - f) * vn + f * aur;
  const double Fr = 0.5 * (rl * (unl + vnl) + rr * (unr - vnr) - chi * iaF * (pr - pl));
  const bool inL = fabs(mL) < 1.0, inR = fabs(mR) < 1.0;
  const double bl = inL ? 0.25 * (2.0 - mL) * (mL + 1.0) * (mL + 1.0) : (mL >= 0.0 ? 1.0 : 0.0);
  const double br = inR ? 0.25 * (2.0 + mR) * (mR - 1.0) * (mR - 1.0) : (mR >= 0.0 ? 0.0 : 1.0);
  const double kg = g.gam * g.rgm1;
  const bool left = Fr >= 0.0;
  // ---- reverse: F = (Fr, Fr u_s + n ph, Fr h0_s)
  double rlb = 0, ulb = 0, vlb = 0, wlb = 0, plb = 0, rrb = 0, urb = 0, vrb = 0, wrb = 0, prb = 0;
  double unlb = 0, unrb = 0, sqlb = 0, sqrb = 0, aFb = 0, iaFb = 0, vrmsb = 0, chib = 0;
  double Fr_b = Fb[0];
  const double phb = Fb[1] * nx + Fb[2] * ny + Fb[3] * nz;
  double ph = 0.5 * (pl + pr) + 0.5 * (bl - br) * (pl - pr);
  if constexpr (V2) ph += vrms * (bl + br - 1.0) * aF * 0.5 * rs;
  else ph += (1.0 - chi) * (bl + br - 1.0) * 0.5 * (pl + pr);
  nb[0] += Fb[1] * ph; nb[1] += Fb[2] * ph; nb[2] += Fb[3] * ph;
  if (left) {
    const double h0 = kg * pl * irl + 0.5 * sql;
    Fr_b += Fb[1] * ul + Fb[2] * vl + Fb[3] * wl + Fb[4] * h0;
    ulb += Fb[1] * Fr; vlb += Fb[2] * Fr; wlb += Fb[3] * Fr;
    const double hb = Fb[4] * Fr;
    plb += hb * kg * irl; rlb -= hb * kg * pl * irl * irl; sqlb += 0.5 * hb;
  } else {
    const double h0 = kg * pr * irr + 0.5 * sqr;
    Fr_b += Fb[1] * ur + Fb[2] * vr + Fb[3] * wr + Fb[4] * h0;
    urb += Fb[1] * Fr; vrb += Fb[2] * Fr; wrb += Fb[3] * Fr;
    const double hb = Fb[4] * Fr;
    prb += hb * kg * irr; rrb -= hb * kg * pr * irr * irr; sqrb += 0.5 * hb;
  }
  // pressure flux
  double blb = 0, brb = 0;
  {
    const double S = pl + pr, D = pl - pr, W = bl + br - 1.0;
    plb += phb * (0.5 + 0.5 * (bl - br)); prb += phb * (0.5 - 0.5 * (bl - br));
    blb += phb * 0.5 * D; brb -= phb * 0.5 * D;
    if constexpr (V2) {
      const double c = vrms * aF * 0.5 * rs;
      blb += phb * c; brb += phb * c;
      vrmsb += phb * W * aF * 0.5 * rs;
      aFb += phb * vrms * W * 0.5 * rs;
      rlb += phb * vrms * W * aF * 0.5; rrb += phb * vrms * W * aF * 0.5;
    } else {
      const double c = (1.0 - chi) * 0.5 * S;
      blb += phb * c; brb += phb * c;
      chib -= phb * W * 0.5 * S;
      plb += phb * (1.0 - chi) * W * 0.5; prb += phb * (1.0 - chi) * W * 0.5;
    }
  }
  double mLb = 0, mRb = 0;
  if (inL) mLb += blb * 0.75 * (1.0 - mL * mL);
  if (inR) mRb += brb * 0.75 * (mR * mR - 1.0);
  // mass flux
  double vnlb, vnrb;
  {
    const double h = 0.5 * Fr_b;
    rlb += h * (unl + vnl); unlb += h * rl; vnlb = h * rl;
    rrb += h * (unr - vnr); unrb += h * rr; vnrb = -h * rr;
    chib -= h * iaF * (pr - pl); iaFb -= h * chi * (pr - pl);
    prb -= h * chi * iaF; plb += h * chi * iaF;
  }
  const double vnb = (vnlb + vnrb) * (1.0 - f);
  const double fb = vnlb * (aul - vn) + vnrb * (aur - vn);
  double aulb = vnlb * f, aurb = vnrb * f;
  // f = -A B, A = clip(mL, -1, 0), B = clip(mR, 0, 1)
  if (mL < 0.0 && mL > -1.0) mLb -= fb * B;
  if (mR > 0.0 && mR < 1.0) mRb -= fb * A;
  // vn = (rl |unl| + rr |unr|) / (rl + rr)
  rlb += vnb * (aul - vn) * irs; rrb += vnb * (aur - vn) * irs;
  aulb += vnb * rl * irs; aurb += vnb * rr * irs;
  unlb += unl >= 0.0 ? aulb : -aulb;
  unrb += unr >= 0.0 ? aurb : -aurb;
  // chi = (1 - mt)^2, mt = min_(1, vrms / aF)
  if (!(1.0 < mraw)) {
    const double mtb = -2.0 * (1.0 - mt) * chib;
    vrmsb += mtb * iaF; iaFb += mtb * vrms;
  }
  if (vrms > 0.0) {
    const double qb = vrmsb * 0.5 / vrms;
    sqlb += 0.5 * qb; sqrb += 0.5 * qb;
  }
  // Mach numbers, aF
  unlb += mLb * iaF; unrb += mRb * iaF; iaFb += mLb * unl + mRb * unr;
  aFb -= iaFb * iaF * iaF;
  {
    const double tl = 0.5 * aFb * 0.5 / al * g.gam, tr = 0.5 * aFb * 0.5 / ar * g.gam;
    plb += tl * irl; rlb -= tl * pl * irl * irl;
    prb += tr * irr; rrb -= tr * pr * irr * irr;
  }
  ulb += unlb * nx + 2.0 * ul * sqlb; vlb += unlb * ny + 2.0 * vl * sqlb; wlb += unlb * nz + 2.0 * wl * sqlb;
  urb += unrb * nx + 2.0 * ur * sqrb; vrb += unrb * ny + 2.0 * vr * sqrb; wrb += unrb * nz + 2.0 * wr * sqrb;
  nb[0] += unlb * ul + unrb * ur;
  nb[1] += unlb * vl + unrb * vr;
  nb[2] += unlb * wl + unrb * wr;
  Lb[0] += rlb; Lb[1] += ulb; Lb[2] += vlb; Lb[3] += wlb; Lb[4] += plb;
  Rb[0] += rrb; Rb[1] += urb; Rb[2] += vrb; Rb[3] += wrb; Rb[4] += prb;
}

// ------------------------------------------------------------------------------ AUSM+
__device__ void ausmp_adj(const double* L, const double* R, double nx, double ny, double nz, const GasT<double>& g,
                          const double* Fb, double* Lb, double* Rb, double* nb) {
  const double beta = 0.125, alpha = 0.1875;
  const double rl = L[0], ul = L[1], vl = L[2], wl = L[3], pl = L[4];
  const double rr = R[0], ur = R[1], vr = R[2], wr = R[3], pr = R[4];
  // ---- forward
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double al2 = g.gam * pl * irl, ar2 = g.gam * pr * irr;
  const double unl = ul * nx + vl * ny + wl * nz, unr = ur * nx + vr * ny + wr * nz;
  const double kg = g.gam * g.rgm1;
  const double H_L = kg * pl * irl + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double H_R = kg * pr * irr + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double prod = al2 * ar2, sp = sqrt(prod), a_half = sqrt(sp);
  const double ia = 1.0 / a_half;
  const double ML = unl * ia, MR = unr * ia;
  double MplusL, pplusL, dMplusL, dpplusL;   // values and d/dML
  if (fabs(ML) < 1.0) {
    const double a = (ML + 1.0) * (ML + 1.0), b = (ML * ML - 1.0) * (ML * ML - 1.0);
    const double da = 2.0 * (ML + 1.0), db = 4.0 * ML * (ML * ML - 1.0);
    MplusL = 0.25 * a + beta * b;
    pplusL = 0.25 * a * (2.0 - ML) + alpha * ML * b;
    dMplusL = 0.25 * da + beta * db;
    dpplusL = 0.25 * da * (2.0 - ML) - 0.25 * a + alpha * b + alpha * ML * db;
  } else {
    MplusL = 0.5 * (ML + fabs(ML));
    pplusL = ML >= 0.0 ? 1.0 : 0.0;
    dMplusL = ML >= 0.0 ? 1.0 : 0.0;
    dpplusL = 0.0;
  }
  double MminusR, pminusR, dMminusR, dpminusR;
  if (fabs(MR) < 1.0) {
    const double a = (MR - 1.0) * (MR - 1.0), b = (MR * MR - 1.0) * (MR * MR - 1.0);
    const double da = 2.0 * (MR - 1.0), db = 4.0 * MR * (MR * MR - 1.0);
    MminusR = -0.25 * a - beta * b;
    pminusR = 0.25 * a * (2.0 + MR) - alpha * MR * b;
    dMminusR = -0.25 * da - beta * db;
    dpminusR = 0.25 * da * (2.0 + MR) + 0.25 * a - alpha * b - alpha * MR * db;
  } else {
    MminusR = 0.5 * (MR - fabs(MR));
    pminusR = MR >= 0.0 ? 0.0 : 1.0;
    dMminusR = MR >= 0.0 ? 0.0 : 1.0;
    dpminusR = 0.0;
  }
  const double Mhalf = MplusL + MminusR;
  const double phalf = pplusL * pl + pminusR * pr;
  const double mp = max_(Mhalf, 0.0), mm = min_(Mhalf, 0.0);
  const double mdotp = rl * a_half * mp, mdotm = rr * a_half * mm;
  // ---- reverse
  const double mdotp_b = Fb[0] + Fb[1] * ul + Fb[2] * vl + Fb[3] * wl + Fb[4] * H_L;
  const double mdotm_b = Fb[0] + Fb[1] * ur + Fb[2] * vr + Fb[3] * wr + Fb[4] * H_R;
  const double phalf_b = Fb[1] * nx + Fb[2] * ny + Fb[3] * nz;
  nb[0] += Fb[1] * phalf;
  nb[1] += Fb[2] * phalf;
  nb[2] += Fb[3] * phalf;
  double ulb = Fb[1] * mdotp, vlb = Fb[2] * mdotp, wlb = Fb[3] * mdotp, HLb = Fb[4] * mdotp;
  double urb = Fb[1] * mdotm, vrb = Fb[2] * mdotm, wrb = Fb[3] * mdotm, HRb = Fb[4] * mdotm;
  double rlb = mdotp_b * a_half * mp, rrb = mdotm_b * a_half * mm;
  double a_half_b = mdotp_b * rl * mp + mdotm_b * rr * mm;
  const double Mhalf_b = (Mhalf > 0.0 ? mdotp_b * rl * a_half : 0.0) + (Mhalf < 0.0 ? mdotm_b * rr * a_half : 0.0);
  double plb = phalf_b * pplusL, prb = phalf_b * pminusR;
  const double MLb = Mhalf_b * dMplusL + phalf_b * pl * dpplusL;
  const double MRb = Mhalf_b * dMminusR + phalf_b * pr * dpminusR;
  // ML = unl / a_half, MR = unr / a_half
  const double unlb = MLb * ia, unrb = MRb * ia;
  a_half_b -= (MLb * unl + MRb * unr) * ia * ia;
  // a_half = (al2 ar2)^(1/4)
  const double prod_b = a_half_b * 0.5 / a_half * 0.5 / sp;
  const double al2b = prod_b * ar2, ar2b = prod_b * al2;
  double irlb = al2b * g.gam * pl, irrb = ar2b * g.gam * pr;
  plb += al2b * g.gam * irl;
  prb += ar2b * g.gam * irr;
  // H = kg p / rho + |u|^2/2
  plb += HLb * kg * irl; irlb += HLb * kg * pl;
  ulb += HLb * ul; vlb += HLb * vl; wlb += HLb * wl;
  prb += HRb * kg * irr; irrb += HRb * kg * pr;
  urb += HRb * ur; vrb += HRb * vr; wrb += HRb * wr;
  ulb += unlb * nx; vlb += unlb * ny; wlb += unlb * nz;
  urb += unrb * nx; vrb += unrb * ny; wrb += unrb * nz;
  nb[0] += unlb * ul + unrb * ur;
  nb[1] += unlb * vl + unrb * vr;
  nb[2] += unlb * wl + unrb * wr;
  rlb -= irlb * irl * irl;
  rrb -= irrb * irr * irr;
  Lb[0] += rlb; Lb[1] += ulb; Lb[2] += vlb; Lb[3] += wlb; Lb[4] += plb;
  Rb[0] += rrb; Rb[1] += urb; Rb[2] += vrb; Rb[3] += wrb; Rb[4] += prb;
}

// nb (may be null) receives the adjoint of the face normal.
template <int RS>
__device__ __forceinline__ void riemann_adj(const double* L, const double* R, double nx, double ny, double nz,
                                            const GasT<double>& g, const double* Fb, double* Lb, double* Rb,
                                            double* nb = nullptr) {
  double nd[3] = {0.0, 0.0, 0.0};
  if constexpr (RS == RS_AUSMP) ausmp_adj(L, R, nx, ny, nz, g, Fb, Lb, Rb, nd);
  else if constexpr (RS == RS_HLLE) hlle_adj(L, R, nx, ny, nz, g, Fb, Lb, Rb, nd);
  else if constexpr (RS == RS_SLAU) slau_adj<false>(L, R, nx, ny, nz, g, Fb, Lb, Rb, nd);
  else if constexpr (RS == RS_SLAU2) slau_adj<true>(L, R, nx, ny, nz, g, Fb, Lb, Rb, nd);
  else hllc_adj(L, R, nx, ny, nz, g, Fb, Lb, Rb, nd);
  if (nb) { nb[0] += nd[0]; nb[1] += nd[1]; nb[2] += nd[2]; }
}

// Physical flux of one state (BC_Extrapolation).
__device__ __forceinline__ void physical_flux_adj(const double* P, double nx, double ny, double nz,
                                                  const GasT<double>& g, const double* Fb, double* Pb, double* nb) {
  const double un = P[1] * nx + P[2] * ny + P[3] * nz;
  const double m = P[0] * un;
  const double ir = 1.0 / P[0];
  const double h0 = g.gam * g.rgm1 * P[4] * ir + 0.5 * (P[1] * P[1] + P[2] * P[2] + P[3] * P[3]);
  const double mb = Fb[0] + Fb[1] * P[1] + Fb[2] * P[2] + Fb[3] * P[3] + Fb[4] * h0;
  const double h0b = Fb[4] * m;
  const double kg = g.gam * g.rgm1;
  const double unb = mb * P[0];
  Pb[0] += mb * un - h0b * kg * P[4] * ir * ir;
  Pb[1] += Fb[1] * m + unb * nx + h0b * P[1];
  Pb[2] += Fb[2] * m + unb * ny + h0b * P[2];
  Pb[3] += Fb[3] * m + unb * nz + h0b * P[3];
  Pb[4] += Fb[1] * nx + Fb[2] * ny + Fb[3] * nz + h0b * kg * ir;
  nb[0] += unb * P[1] + Fb[1] * P[4];
  nb[1] += unb * P[2] + Fb[2] * P[4];
  nb[2] += unb * P[3] + Fb[3] * P[4];
}

// Symmetry face (symmetry_flux): momentum part of the mirrored Riemann problem on the
// limited face state of the boundary cell.
template <int LIM, int RS>
__device__ void symmetry_flux_adj(const double* Pb, const double* Pi, double bdl, double idl, bool hi, double nx,
                                  double ny, double nz, const GasT<double>& g, const double* Fb_in, double* Pbb,
                                  double* Pib, double* nb, double* bdlb, double* idlb) {
  const double un = Pb[1] * nx + Pb[2] * ny + Pb[3] * nz;
  const double Gp[5] = {Pb[0], Pb[1] - 2.0 * un * nx, Pb[2] - 2.0 * un * ny, Pb[3] - 2.0 * un * nz, Pb[4]};
  const double d12 = hi ? (bdl + idl) : 2.0 * bdl;
  const double d23 = hi ? 2.0 * bdl : (bdl + idl);
  const double rd12 = 1.0 / d12, rd23 = 1.0 / d23;
  const double dist = hi ? bdl : -bdl;
  double Fc[5], s32[5], s21[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double p1 = hi ? Pi[v] : Gp[v];
    const double p3 = hi ? Gp[v] : Pi[v];
    s32[v] = (p3 - Pb[v]) * rd23;
    s21[v] = (Pb[v] - p1) * rd12;
    Fc[v] = Pb[v] + limiter<LIM>(s32[v], s21[v]) * dist;
  }
  const double unf = Fc[1] * nx + Fc[2] * ny + Fc[3] * nz;
  const double un1 = hi ? unf : -unf;
  const double L[5] = {Fc[0], Fc[1] - unf * nx + un1 * nx, Fc[2] - unf * ny + un1 * ny, Fc[3] - unf * nz + un1 * nz, Fc[4]};
  const double R[5] = {Fc[0], Fc[1] - unf * nx - un1 * nx, Fc[2] - unf * ny - un1 * ny, Fc[3] - unf * nz - un1 * nz, Fc[4]};
  // reverse
  const double Fb[5] = {0.0, Fb_in[1], Fb_in[2], Fb_in[3], 0.0};
  double Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
  riemann_adj<RS>(L, R, nx, ny, nz, g, Fb, Lb, Rb, nb);
  double Fcb[5];
  Fcb[0] = Lb[0] + Rb[0];
  Fcb[4] = Lb[4] + Rb[4];
  Fcb[1] = Lb[1] + Rb[1];
  Fcb[2] = Lb[2] + Rb[2];
  Fcb[3] = Lb[3] + Rb[3];
  // L/R velocity = Fc - unf n +/- un1 n
  const double nvec[3] = {nx, ny, nz};
  double unfb = 0.0, un1b = 0.0;
  for (int c = 0; c < 3; ++c) {
    unfb -= (Lb[1 + c] + Rb[1 + c]) * nvec[c];
    un1b += (Lb[1 + c] - Rb[1 + c]) * nvec[c];
  }
  unfb += hi ? un1b : -un1b;
  Fcb[1] += unfb * nx;
  Fcb[2] += unfb * ny;
  Fcb[3] += unfb * nz;
  // normal: L/R velocity terms (un1 - unf) n, -(unf + un1) n, and unf = Fc . n
  for (int c = 0; c < 3; ++c)
    nb[c] += (un1 - unf) * Lb[1 + c] - (unf + un1) * Rb[1 + c] + unfb * Fc[1 + c];
  double Gpb[5] = {0, 0, 0, 0, 0};
  double distb = 0.0, rd12b = 0.0, rd23b = 0.0;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    double xb = 0.0, yb = 0.0;
    const double phi = limiter<LIM>(s32[v], s21[v]);
    distb += Fcb[v] * phi;
    limiter_adj<LIM>(s32[v], s21[v], Fcb[v] * dist, xb, yb);
    Pbb[v] += Fcb[v] - xb * rd23 + yb * rd12;
    const double p3b = xb * rd23, p1b = -yb * rd12;
    rd23b += xb * s32[v] / rd23;
    rd12b += yb * s21[v] / rd12;
    if (hi) { Pib[v] += p1b; Gpb[v] += p3b; }
    else { Gpb[v] += p1b; Pib[v] += p3b; }
  }
  // spacings: dist = +/-bdl, d12 / d23 = 2 bdl or bdl + idl, rd = 1/d
  {
    const double d12b = -rd12b * rd12 * rd12, d23b = -rd23b * rd23 * rd23;
    *bdlb += (hi ? 1.0 : -1.0) * distb;
    if (hi) { *bdlb += d12b + 2.0 * d23b; *idlb += d12b; }
    else { *bdlb += 2.0 * d12b + d23b; *idlb += d23b; }
  }
  // Gp = mirror(Pb)
  Pbb[0] += Gpb[0];
  Pbb[4] += Gpb[4];
  const double gdn = Gpb[1] * nx + Gpb[2] * ny + Gpb[3] * nz;
  Pbb[1] += Gpb[1] - 2.0 * gdn * nx;
  Pbb[2] += Gpb[2] - 2.0 * gdn * ny;
  Pbb[3] += Gpb[3] - 2.0 * gdn * nz;
  for (int c = 0; c < 3; ++c) nb[c] += -2.0 * (Pb[1 + c] * gdn + un * Gpb[1 + c]);
}

// Reverse of outflow406_flux (BC 406, Lib_BC_Fluxes_Outflow.f90:12-130) for the flux adjoint Fb
// (per unit area): accumulates into Pbb (boundary-cell primitives) and nb (face normal).
// The branch (subsonic with p_exit > 0, and whether p* replaces p_exit) is taken as the
// forward pass takes it; returns false, touching nothing, where the forward falls back to
// the symmetry flux.
__device__ bool outflow406_adj(const double* Pb, bool hi, const double* n, const double* bc, double gam,
                               const double* Fb, double* Pbb, double* nb) {
  const double sgn = hi ? 1.0 : -1.0;
  const double un = Pb[1] * n[0] + Pb[2] * n[1] + Pb[3] * n[2];
  if (sgn * un < 0.0) return false;
  const double dl = 0.5 * (gam - 1.0), E = gam / (gam - 1.0), X = (gam - 1.0) / (2.0 * gam);
  const double snd = sqrt(gam * Pb[4] / Pb[0]);
  const double riem = snd + sgn * dl * un;
  const double vt[3] = {Pb[1] - un * n[0], Pb[2] - un * n[1], Pb[3] - un * n[2]};
  const double machn = fabs(un / snd);
  const bool sub = machn < 1.0 && bc[0] > 0.0;
  double rho = Pb[0], p = Pb[4], ufn = un, u[3] = {Pb[1], Pb[2], Pb[3]};
  double pe = bc[0], sndf = 0.0, ratio = 0.0, base = 0.0, den = 0.0;
  bool clip = false;
  if (sub) {
    den = 1.0 / pow(1.0 + dl, E);
    base = 1.0 + dl * machn * machn;
    const double pstar = Pb[4] * den * pow(base, E);
    clip = pstar >= pe;
    if (clip) pe = pstar;
    ratio = pow(pe / Pb[4], X);
    sndf = snd * ratio;
    ufn = (riem - sndf) / (sgn * dl);
    p = pe;
    rho = gam * p / (sndf * sndf);
    for (int q = 0; q < 3; ++q) u[q] = vt[q] + ufn * n[q];
  }
  const double h = E * p / rho;
  const double m = rho * ufn;
  const double q2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  // reverse of F0 = m, F(1+q) = m u_q + p n_q, F4 = m (q2/2 + h)
  const double mb = Fb[0] + Fb[1] * u[0] + Fb[2] * u[1] + Fb[3] * u[2] + Fb[4] * (0.5 * q2 + h);
  double ub[3], pb = 0.0, rhob = 0.0, ufnb = 0.0, unb = 0.0, sndb = 0.0;
  for (int q = 0; q < 3; ++q) {
    ub[q] = Fb[1 + q] * m + Fb[4] * m * u[q];
    pb += Fb[1 + q] * n[q];
    nb[q] += Fb[1 + q] * p;
  }
  const double hb = Fb[4] * m;
  pb += hb * E / rho;
  rhob -= hb * E * p / (rho * rho);
  rhob += mb * ufn;
  ufnb += mb * rho;
  if (sub) {
    double vtb[3];
    for (int q = 0; q < 3; ++q) {
      vtb[q] = ub[q];
      ufnb += ub[q] * n[q];
      nb[q] += ub[q] * ufn;
    }
    // rho = gam pe / sndf^2, p = pe
    double peb = pb + rhob * gam / (sndf * sndf);
    double sndfb = -2.0 * rhob * gam * p / (sndf * sndf * sndf);
    // ufn = (riem - sndf) / (sgn dl)
    const double riemb = ufnb / (sgn * dl);
    sndfb -= ufnb / (sgn * dl);
    // sndf = snd (pe / p_cell)^X
    sndb += sndfb * ratio;
    peb += sndfb * sndf * X / pe;
    double pcb = -sndfb * sndf * X / Pb[4];
    double machnb = 0.0;
    if (clip) {   // pe = p* = p_cell den base^E, base = 1 + dl machn^2
      pcb += peb * den * pow(base, E);
      machnb = peb * Pb[4] * den * E * pow(base, E - 1.0) * 2.0 * dl * machn;
    }
    // machn = |un / snd|
    const double sr = (un / snd) >= 0.0 ? 1.0 : -1.0;
    unb += machnb * sr / snd;
    sndb -= machnb * sr * un / (snd * snd);
    // vt_q = P(1+q) - un n_q
    for (int q = 0; q < 3; ++q) {
      Pbb[1 + q] += vtb[q];
      unb -= vtb[q] * n[q];
      nb[q] -= vtb[q] * un;
    }
    // riem = snd + sgn dl un
    sndb += riemb;
    unb += riemb * sgn * dl;
    Pbb[4] += pcb;
  } else {
    Pbb[0] += rhob;
    Pbb[4] += pb;
    unb += ufnb;
    for (int q = 0; q < 3; ++q) Pbb[1 + q] += ub[q];
  }
  // snd = sqrt(gam p / rho)
  Pbb[4] += sndb * gam / (2.0 * snd * Pb[0]);
  Pbb[0] -= sndb * snd / (2.0 * Pb[0]);
  // un = P . n
  for (int q = 0; q < 3; ++q) {
    Pbb[1 + q] += unb * n[q];
    nb[q] += unb * Pb[1 + q];
  }
  return true;
}

// Supersonic inflow face (inflow405_flux); returns false when the symmetry flux applies.
// bcb (6) receives the adjoints of the imposed Mach number, temperature, pressure and the
// three components of the velocity direction.
template <int LIM, int RS>
__device__ bool inflow405_adj(const double* Pb, const double* Pi, bool hi, double nx, double ny, double nz,
                              const double* bc, const GasT<double>& g, const double* Fb, double* Pbb, double* Pib,
                              double* nb, double* bcb) {
  const double un = Pb[1] * nx + Pb[2] * ny + Pb[3] * nz;
  const double sgn = hi ? 1.0 : -1.0;
  if (sgn * un >= 0.0) return false;
  const double snd2 = g.gam * Pb[4] / Pb[0];
  const double flow_a = sqrt(snd2);
  const double vmag = sqrt(Pb[1] * Pb[1] + Pb[2] * Pb[2] + Pb[3] * Pb[3]);
  const double flow_M = vmag / flow_a;
  const double rel = bc[6];
  const double M = rel * bc[0] + (1.0 - rel) * flow_M;
  const double p = rel * bc[2] + (1.0 - rel) * Pb[4];
  const double RT = g.R * bc[1];
  const double rho = p / RT;
  const double a = sqrt(g.gam * p / rho);
  double dir[3];
  inflow_dir(bc, nx, ny, nz, dir);
  const double Sup[5] = {rho, M * a * dir[0], M * a * dir[1], M * a * dir[2], p};
  double Fc[5], L[5], R[5], d32[5], d21[5];
  const double half = hi ? 0.5 : -0.5;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double p1 = hi ? Pi[v] : Sup[v];
    const double p3 = hi ? Sup[v] : Pi[v];
    d32[v] = p3 - Pb[v];
    d21[v] = Pb[v] - p1;
    Fc[v] = Pb[v] + half * limiter<LIM>(d32[v], d21[v]);
    L[v] = hi ? Fc[v] : Sup[v];
    R[v] = hi ? Sup[v] : Fc[v];
  }
  double Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
  riemann_adj<RS>(L, R, nx, ny, nz, g, Fb, Lb, Rb, nb);
  double Supb[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double Fcb = hi ? Lb[v] : Rb[v];
    Supb[v] = hi ? Rb[v] : Lb[v];
    double xb = 0.0, yb = 0.0;
    limiter_adj<LIM>(d32[v], d21[v], Fcb * half, xb, yb);
    Pbb[v] += Fcb - xb + yb;
    if (hi) { Pib[v] -= yb; Supb[v] += xb; }
    else { Supb[v] -= yb; Pib[v] += xb; }
  }
  // Sup = (p / RT, M a dir, p) with a = sqrt(gam p / rho) = sqrt(gam R T)
  const double Mb = a * (Supb[1] * dir[0] + Supb[2] * dir[1] + Supb[3] * dir[2]);
  const double pb = Supb[4] + Supb[0] / RT;
  {
    const double ab = M * (Supb[1] * dir[0] + Supb[2] * dir[1] + Supb[3] * dir[2]);
    bcb[0] += rel * Mb;
    bcb[1] += -Supb[0] * rho / bc[1] + ab * 0.5 * a / bc[1];
    bcb[2] += rel * pb;
    const double db[3] = {M * a * Supb[1], M * a * Supb[2], M * a * Supb[3]};
    const int flags = (int)bc[7];
    if (flags == 0) {   // adjoints of the three direction components
      bcb[3] += db[0];
      bcb[4] += db[1];
      bcb[5] += db[2];
    } else {
      // angles: explicit ones get their adjoint in bcb[3] (alpha) / bcb[4] (beta); 'normal'
      // ones pass it to the face normal, alpha = atan(ny / (nx + 1e-20)), beta = atan(nz / ...)
      const double al = (flags & 1) ? atan(ny / (nx + 1e-20)) : bc[3];
      const double be = (flags & 2) ? atan(nz / (nx + 1e-20)) : bc[4];
      const double ca = cos(al), sa = sin(al), cb = cos(be), sb = sin(be);
      const double alb = db[0] * (-sa * cb) + db[1] * (ca * cb);
      const double beb = db[0] * (-ca * sb) + db[1] * (-sa * sb) + db[2] * cb;
      const double nx_ = nx + 1e-20;
      if (flags & 1) {
        const double q = 1.0 / (nx_ * nx_ + ny * ny);
        nb[0] += alb * (-ny * q);
        nb[1] += alb * (nx_ * q);
      } else {
        bcb[3] += alb;
      }
      if (flags & 2) {
        const double q = 1.0 / (nx_ * nx_ + nz * nz);
        nb[0] += beb * (-nz * q);
        nb[2] += beb * (nx_ * q);
      } else {
        bcb[4] += beb;
      }
    }
  }
  if (rel != 1.0) {
    // Sup depends on the boundary cell through M (flow Mach) and p
    const double fMb = (1.0 - rel) * Mb;
    Pbb[4] += (1.0 - rel) * pb;
    if (vmag > 0.0) {
      const double vb = fMb / flow_a;
      Pbb[1] += vb * Pb[1] / vmag;
      Pbb[2] += vb * Pb[2] / vmag;
      Pbb[3] += vb * Pb[3] / vmag;
    }
    // flow_a = sqrt(gam p / rho)
    const double ab = -fMb * vmag / (flow_a * flow_a);
    const double s2b = ab * 0.5 / flow_a;
    Pbb[4] += s2b * g.gam / Pb[0];
    Pbb[0] -= s2b * g.gam * Pb[4] / (Pb[0] * Pb[0]);
  }
  return true;
}

// ------------------------------------------------------------------------------ viscous
// Reverse of visc_face_flux with area 1 (the forward's tau.n and heat flux with Sutherland's
// mu(T)): Fb -> Gb, the adjoints of the index-space gradients G[v][a] (written), and Pmb,
// the adjoints of the face state (rho, u, v, w, p) (accumulated); returns the adjoint of
// mu_ref.  The stress column is written st[c] = lambda div delta_cd + mu (Gp[c][d] + Gp[d][c]),
// Gp = G M, which is the forward's per-direction expression.
__device__ inline double visc_face_flux_adj(int d, const double (*G)[3], const double* M, const double* Pm, double Rg,
                                            const double* vis, const double* Fb, double (*Gb)[3], double* Pmb) {
  double Gp[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int x = 0; x < 3; ++x) Gp[v][x] = G[v][x] * M[x];
  const double temp = Pm[4] / (Pm[0] * Rg);
  const double mu = sutherland(temp, vis[0], vis[1], vis[2]);
  const double kap = mu * vis[3];
  const double lam = -2.0 / 3.0 * mu;
  const double div = Gp[0][0] + Gp[1][1] + Gp[2][2];
  double st[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) st[c] = (c == d ? lam * div : 0.0) + mu * (Gp[c][d] + Gp[d][c]);
  // F = (0, st, st . u + kap Gp[3][d])
  double Gpb[4][3];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int x = 0; x < 3; ++x) Gpb[v][x] = 0.0;
  double stb[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    stb[c] = Fb[1 + c] + Fb[4] * Pm[1 + c];
    Pmb[1 + c] += Fb[4] * st[c];
  }
  Gpb[3][d] += Fb[4] * kap;
  double mub = Fb[4] * Gp[3][d] * vis[3];
  const double ldb = stb[d];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mub += stb[c] * (Gp[c][d] + Gp[d][c]);
    Gpb[c][d] += mu * stb[c];
    Gpb[d][c] += mu * stb[c];
  }
  const double lamb = ldb * div;
#pragma unroll
  for (int x = 0; x < 3; ++x) Gpb[x][x] += lam * ldb;
  mub += -2.0 / 3.0 * lamb;
  // mu = mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S): d ln mu / dT = 1.5/T - 1/(T + S)
  const double tempb = mub * mu * (1.5 / temp - 1.0 / (temp + vis[2]));
  Pmb[4] += tempb / (Pm[0] * Rg);
  Pmb[0] -= tempb * temp / Pm[0];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int x = 0; x < 3; ++x) Gb[v][x] = Gpb[v][x] * M[x];
  return mub * mu / vis[0];
}

// Adjoint of the viscous variables V = (u, v, w, T = p / (rho R)) of a cell added to the
// adjoint of its primitive state P.
__device__ __forceinline__ void visc_vars_adj(const double* P, double Rg, const double* Vb, double* Pb) {
  const double iRr = 1.0 / (P[0] * Rg);
  const double T = P[4] * iRr;
  Pb[0] -= Vb[3] * T / P[0];
  Pb[1] += Vb[0];
  Pb[2] += Vb[1];
  Pb[3] += Vb[2];
  Pb[4] += Vb[3] * iRr;
}

}  // namespace mose

// =============================================================================== kernels
// Discrete adjoint of the residual R(P) = Fill_Ghost_Cell + Fluxes + BC_Fluxes (the value
// mose_residual returns) with respect to the primitive state P of every interior cell:
// Pbar = (dR/dP)^T Rbar.  Four passes on the current state (ghost layers filled first):
//   k_adj_face    one thread per interior or connection face: recompute the two face states
//                 and the Riemann flux, reverse them, store the adjoints of the face states
//                 (the "+" state of the left cell, the "-" state of the right cell);
//   k_adj_bcface  one thread per non-connection boundary face: reverse of the boundary flux
//                 into adjoints of the boundary cell and its interior neighbour;
//   k_adj_cell    one thread per cell (ghost layers included): gathers the reverse of the
//                 limited slopes of itself and its two neighbours along each direction plus
//                 the boundary-face adjoints that name it (deterministic, no atomics);
//   k_adj_ghostT  one thread per cell that feeds ghost cells: adds the transpose of the
//                 ghost-fill map (copy, extrapolation, mirror) applied to the ghost adjoints.
// With the geometry outputs set (general metrics only), the same passes also write the
// adjoints of every geometric quantity the residual reads: face normals and areas (GF, by
// the face and boundary-face passes), reciprocal centre distances and cell lengths (GL, by
// the cell pass, which gathers the boundary-face half lengths from BG), the mirror normals
// of the symmetry ghost fill (GE) and the imposed values of inflow faces (IB).  The host
// turns them into node adjoints by the reverse of the metric computation.

struct AdjSeg {       // the low faces of one block in one direction
  int64_t start;      // first work item
  int b, d;
  int n0, n1, n2;     // extents along i, j, k (n + 1 along d: the faces), i fastest
  int pad;
};

struct AdjGhostRow {  // one source cell of the ghost fill and its entry range
  int64_t src;
  int begin, end;
};
struct AdjGhostEnt {
  int64_t ghost;      // pool index of the ghost cell
  double w;           // weight (kind 0) ...
  double n[3];        // ... or mirror normal (kind 1)
  int kind, pad;
};

struct AdjArgs {
  const double* __restrict__ U;         // state with filled ghosts
  const double* __restrict__ Pp;        // 5 x npool: its primitive variables
  const double* __restrict__ Rbar;      // 5 x npool
  double* __restrict__ W;               // 30 x npool: face-state adjoints (3 dirs x {Sp, Sm} x 5)
  double* __restrict__ BB;              // 10 x nslots: boundary-face adjoints (cell, neighbour)
  double* __restrict__ Pbar;            // 5 x npool
  const BlockDev* __restrict__ blocks;
  const LineEnt* __restrict__ lines;
  const double* __restrict__ geo;
  const double* __restrict__ facebc;
  const AdjSeg* __restrict__ segs;
  const BcFace* __restrict__ bcfaces;
  const AdjGhostRow* __restrict__ grows;
  const AdjGhostEnt* __restrict__ gents;
  // geometry adjoints (null: not requested)
  double* __restrict__ GF;              // 12 x npool: low-face normal (3) and area, per direction
  double* __restrict__ GL;              // 6 x npool: rl(d) and dl(d) per direction
  double* __restrict__ BG;              // 2 x nslots: boundary-face bdl, idl
  double* __restrict__ GE;              // 3 x nents: mirror normal of the ghost-fill entries
  double* __restrict__ IB;              // 6 x nslots: inflow M, T, p, direction (3)
  // viscous blocks (null / 0 without)
  double* __restrict__ VW;              // 3 x 18 x npool: interior viscous face adjoints at the low cell
  double* __restrict__ VB;              // nvf: mu_ref adjoint of each connection viscous face
  double* __restrict__ VBC;             // 50 x nvf: each connection face's primitive adjoints of
                                        // its ten stencil cells, gathered by k_adj_vbc_gather
  const double* __restrict__ Tp;        // npool: temperature p / (rho R) of every cell
  double* VPb;                          // 5 x npool: k_adj_visc's part of Pbar (interior cells of
                                        // viscous blocks), folded in by k_adj_cell; null without
  const int64_t* __restrict__ vbt;      // cells receiving connection-face adjoints
  const int64_t* __restrict__ vbrow;    // their entries vbent[vbrow[r] .. vbrow[r+1]) (face*10 + role)
  const int* __restrict__ vbent;
  const ViscFace* __restrict__ vfaces;
  int64_t npool, nwork, nslots;
  int nseg, nbcf, ngrows, nblocks, nents, nvf, nvbt;
  double gam, R;
};

// Geometry of the path in both layouts: reciprocal low-neighbour distance, half length and
// low-face normal/area of cell c (padded index q along d, tangential padded indices).
template <int GEO>
struct AdjGeo {
  const AdjArgs& a;
  const BlockDev& B;
  __device__ double rl(int d, int64_t c, int q) const {
    if constexpr (GEO == 0) return a.lines[B.line[d] + q + 1].rl1;
    else return a.geo[(size_t)(GRLO + d) * a.npool + c];
  }
  __device__ double hdl(int d, int64_t c, int q) const {
    if constexpr (GEO == 0) return a.lines[B.line[d] + q + 1].lp;
    else return 0.5 * a.geo[(size_t)(GDL + d) * a.npool + c];
  }
  __device__ void face(int d, int64_t c, int i, int j, int k, double n[3], double& area) const {
    if constexpr (GEO == 0) {
      n[0] = d == 0 ? 1.0 : 0.0; n[1] = d == 1 ? 1.0 : 0.0; n[2] = d == 2 ? 1.0 : 0.0;
      const double dx = a.lines[B.line[0] + i + 1].delta, dy = a.lines[B.line[1] + j + 1].delta,
                   dz = a.lines[B.line[2] + k + 1].delta;
      area = d == 0 ? dy * dz : (d == 1 ? dx * dz : dx * dy);
    } else {
      const size_t np = (size_t)a.npool;
      n[0] = a.geo[(GN + 4 * d) * np + c]; n[1] = a.geo[(GN + 4 * d + 1) * np + c];
      n[2] = a.geo[(GN + 4 * d + 2) * np + c]; area = a.geo[(GN + 4 * d + 3) * np + c];
    }
  }
};

__device__ __forceinline__ void prim_load(const double* __restrict__ Pp, int64_t N, int64_t c, double* P) {
#pragma unroll
  for (int v = 0; v < 5; ++v) P[v] = Pp[v * N + c];
}


template <int LIM, int RS, int GEO>
__global__ void __launch_bounds__(256) k_adj_face(AdjArgs a) {
  const double gm1 = a.gam - 1.0;
  const GasT<double> g{a.gam, gm1, 1.0 / gm1, a.R, a.gam * a.R / gm1};
  const int64_t N = a.npool;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < a.nwork; q += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = a.nseg - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.segs[mid].start <= q) lo = mid; else hi = mid - 1;
    }
    const AdjSeg S = a.segs[lo];
    const BlockDev& B = a.blocks[S.b];
    const int d = S.d;
    // i fastest for every direction, so that a wavefront reads consecutive cells (32-bit
    // arithmetic: a segment holds fewer than 2^32 faces)
    unsigned r = (unsigned)(q - S.start);
    int ijk[3];
    ijk[0] = (int)(r % (unsigned)S.n0) + 1;
    r /= (unsigned)S.n0;
    ijk[1] = (int)(r % (unsigned)S.n1) + 1;
    ijk[2] = (int)(r / (unsigned)S.n1) + 1;
    const int qd = ijk[d];                       // cell index along d whose low face this is
    const int nd = d == 0 ? B.ni : (d == 1 ? B.nj : B.nk);
    const bool blo = qd == 1, bhi = qd == nd + 1;
    const bool skip = (blo && B.bct[2 * d] != F_CONN) || (bhi && B.bct[2 * d + 1] != F_CONN);
    const int64_t s = d == 0 ? 1 : (d == 1 ? (int64_t)B.sy : B.sz);
    const int64_t c = B.off + (ijk[0] + 1) + (int64_t)B.sy * (ijk[1] + 1) + B.sz * (int64_t)(ijk[2] + 1);
    double Lb[5] = {0, 0, 0, 0, 0}, Rb[5] = {0, 0, 0, 0, 0};
    if (!skip) {
      double Pm2[5], Pm1[5], P0[5], Pp1[5];
      prim_load(a.Pp, N, c - 2 * s, Pm2);
      prim_load(a.Pp, N, c - s, Pm1);
      prim_load(a.Pp, N, c, P0);
      prim_load(a.Pp, N, c + s, Pp1);
      const AdjGeo<GEO> G{a, B};
      const double rlA = G.rl(d, c - s, qd - 1), rlB = G.rl(d, c, qd), rlC = G.rl(d, c + s, qd + 1);
      const double hA = G.hdl(d, c - s, qd - 1), hB = G.hdl(d, c, qd);
      double PL[5], PR[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double sgA = limiter<LIM>((P0[v] - Pm1[v]) * rlB, (Pm1[v] - Pm2[v]) * rlA) * hA;
        const double sgB = limiter<LIM>((Pp1[v] - P0[v]) * rlC, (P0[v] - Pm1[v]) * rlB) * hB;
        PL[v] = Pm1[v] + sgA;
        PR[v] = P0[v] - sgB;
      }
      double n[3], area;
      G.face(d, c, ijk[0], ijk[1], ijk[2], n, area);
      double Fb[5], dR[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double rl_ = blo ? 0.0 : a.Rbar[v * N + c - s];
        const double rr_ = bhi ? 0.0 : a.Rbar[v * N + c];
        dR[v] = rl_ - rr_;
        Fb[v] = area * dR[v];
      }
      if constexpr (GEO == 1) {
        if (a.GF) {
          double nb[3] = {0.0, 0.0, 0.0}, F[5];
          riemann_adj<RS>(PL, PR, n[0], n[1], n[2], g, Fb, Lb, Rb, nb);
          riemann<RS>(PL, PR, n[0], n[1], n[2], g, F);
          double Ab = 0.0;
#pragma unroll
          for (int v = 0; v < 5; ++v) Ab += F[v] * dR[v];
          const size_t np = (size_t)N;
          a.GF[(4 * d) * np + c] = nb[0];
          a.GF[(4 * d + 1) * np + c] = nb[1];
          a.GF[(4 * d + 2) * np + c] = nb[2];
          a.GF[(4 * d + 3) * np + c] = Ab;
        } else {
          riemann_adj<RS>(PL, PR, n[0], n[1], n[2], g, Fb, Lb, Rb);
        }
      } else {
        riemann_adj<RS>(PL, PR, n[0], n[1], n[2], g, Fb, Lb, Rb);
      }
    } else if (GEO == 1 && a.GF) {
      const size_t np = (size_t)N;
#pragma unroll
      for (int q = 0; q < 4; ++q) a.GF[(4 * d + q) * np + c] = 0.0;   // a non-connection face: k_adj_bcface
    }
    // every "+" / "-" slot that the cell pass reads is written here (zero when no face uses it)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      a.W[(size_t)(10 * d + v) * N + c - s] = Lb[v];       // "+" state of the left cell
      a.W[(size_t)(10 * d + 5 + v) * N + c] = Rb[v];       // "-" state of the right cell
      if (blo) a.W[(size_t)(10 * d + 5 + v) * N + c - s] = 0.0;   // "-" slot of cell 0
      if (bhi) a.W[(size_t)(10 * d + v) * N + c] = 0.0;           // "+" slot of cell n+1
    }
  }
}

template <int LIM, int RS>
__global__ void __launch_bounds__(256) k_adj_bcface(AdjArgs a) {
  const double gm1 = a.gam - 1.0;
  const GasT<double> g{a.gam, gm1, 1.0 / gm1, a.R, a.gam * a.R / gm1};
  const int64_t N = a.npool;
  const bool want_geo = a.GF != nullptr;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < a.nbcf; q += gridDim.x * blockDim.x) {
    const BcFace e = a.bcfaces[q];
    const BlockDev& B = a.blocks[e.b];
    const bool hi = e.f & 1;
    const int bct = B.bct[e.f];
    double Pbb[5] = {0, 0, 0, 0, 0}, Pib[5] = {0, 0, 0, 0, 0};
    double nb[3] = {0, 0, 0}, bcb[6] = {0, 0, 0, 0, 0, 0}, Ab = 0.0, bdlb = 0.0, idlb = 0.0;
    if (bct != F_NONE) {
      double Pb[5], Pi[5], Fb[5], F[5];
      prim_load(a.Pp, N, e.cb, Pb);
      prim_load(a.Pp, N, e.cn, Pi);
      const double sg = hi ? 1.0 : -1.0;
      // R(cb) += (hi ? +1 : -1) area F
#pragma unroll
      for (int v = 0; v < 5; ++v) Fb[v] = sg * e.area * a.Rbar[v * N + e.cb];
      if (bct == F_EXTRAP) {
        physical_flux_adj(Pb, e.n[0], e.n[1], e.n[2], g, Fb, Pbb, nb);
        if (want_geo) physical_flux(Pb, e.n[0], e.n[1], e.n[2], g, F);
      } else {
        const double* bc = a.facebc + 8 * B.bcd[e.f];
        bool done = false;
        if (bct == F_INFLOW)
          done = inflow405_adj<LIM, RS>(Pb, Pi, hi, e.n[0], e.n[1], e.n[2], bc, g, Fb, Pbb, Pib, nb, bcb);
        else if (bct == F_OUTFLOW)
          done = outflow406_adj(Pb, hi, e.n, bc, a.gam, Fb, Pbb, nb);
        if (!done)
          symmetry_flux_adj<LIM, RS>(Pb, Pi, e.bdl, e.idl, hi, e.n[0], e.n[1], e.n[2], g, Fb, Pbb, Pib, nb, &bdlb,
                                     &idlb);
        if (want_geo) {
          bool fdone = false;
          if (bct == F_INFLOW) fdone = inflow405_flux<LIM, RS>(Pb, Pi, hi, e.n[0], e.n[1], e.n[2], bc, g, F);
          else if (bct == F_OUTFLOW) fdone = outflow406_flux(Pb, hi, e.n, bc, a.gam, F);
          if (!fdone) symmetry_flux<LIM, RS>(Pb, Pi, e.bdl, e.idl, hi, e.n[0], e.n[1], e.n[2], g, F);
        }
      }
      if (want_geo) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ab += sg * F[v] * a.Rbar[v * N + e.cb];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      a.BB[(size_t)v * a.nslots + e.slot] = Pbb[v];
      a.BB[(size_t)(5 + v) * a.nslots + e.slot] = Pib[v];
    }
    if (want_geo) {
      const size_t np = (size_t)N;
      const int d = e.f >> 1;
      const int64_t fc = hi ? e.cb + (e.cb - e.cn) : e.cb;   // the cell holding the face as its low face
      a.GF[(4 * d) * np + fc] = nb[0];
      a.GF[(4 * d + 1) * np + fc] = nb[1];
      a.GF[(4 * d + 2) * np + fc] = nb[2];
      a.GF[(4 * d + 3) * np + fc] = Ab;
      a.BG[e.slot] = bdlb;
      a.BG[(size_t)a.nslots + e.slot] = idlb;
#pragma unroll
      for (int r = 0; r < 6; ++r) a.IB[(size_t)r * a.nslots + e.slot] = bcb[r];
    }
  }
}

template <int LIM, int GEO>
__global__ void __launch_bounds__(256) k_adj_cell(AdjArgs a) {
  const double gm1 = a.gam - 1.0;
  const int64_t N = a.npool;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = a.nblocks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.blocks[mid].off <= c) lo = mid; else hi = mid - 1;
    }
    const BlockDev& B = a.blocks[lo];
    const int64_t r64 = c - B.off;
    if (r64 >= B.sz * (int64_t)(B.nk + 4)) continue;     // alignment padding between blocks
    const unsigned r = (unsigned)r64, sy = (unsigned)B.sy, rj = r / sy;
    const int pi = (int)(r - rj * sy) - 1, pj = (int)(rj % (unsigned)(B.nj + 4)) - 1,
              pk = (int)(rj / (unsigned)(B.nj + 4)) - 1;
    const int n3[3] = {B.ni, B.nj, B.nk};
    const int p3[3] = {pi, pj, pk};
    // cells outside the face-ghost cross (edges and corners) are never read
    int outside = 0;
    for (int d = 0; d < 3; ++d) outside += (p3[d] < 1 || p3[d] > n3[d]);
    double acc[5] = {0, 0, 0, 0, 0};
    double rlb[3] = {0, 0, 0}, dlb[3] = {0, 0, 0};
    if (outside <= 1) {
      const AdjGeo<GEO> G{a, B};
      for (int d = 0; d < 3; ++d) {
        const int64_t s = d == 0 ? 1 : (d == 1 ? (int64_t)B.sy : B.sz);
        const int qd = p3[d];
        for (int m = -1; m <= 1; ++m) {
          const int qm = qd + m;
          if (qm < 0 || qm > n3[d] + 1) continue;         // cells whose slope is never used
          if (outside == 1 && (p3[d] >= 1 && p3[d] <= n3[d])) {
            // a ghost cell of another direction: only its own slope along d could matter,
            // and ghost cells never need a slope across the ghost layer
            continue;
          }
          const int64_t cm = c + m * s;
          double Spb[5], Smb[5];
          bool any = false;
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            Spb[v] = a.W[(size_t)(10 * d + v) * N + cm];
            Smb[v] = a.W[(size_t)(10 * d + 5 + v) * N + cm];
            any = any || Spb[v] != 0.0 || Smb[v] != 0.0;
          }
          if (!any) continue;
          double Pm[5], P0[5], Pp[5];
          prim_load(a.Pp, N, cm - s, Pm);
          prim_load(a.Pp, N, cm, P0);
          prim_load(a.Pp, N, cm + s, Pp);
          double bm[5] = {0, 0, 0, 0, 0}, b0[5] = {0, 0, 0, 0, 0}, bp[5] = {0, 0, 0, 0, 0};
          double gb3[3] = {0.0, 0.0, 0.0};
          const bool sp_geo = GEO == 1 && a.GL && m <= 0;
          cell_states_adj<LIM>(Pm, P0, Pp, G.rl(d, cm, qm), G.rl(d, cm + s, qm + 1), G.hdl(d, cm, qm), Smb, Spb,
                               bm, b0, bp, sp_geo ? gb3 : nullptr);
          // rl(c) is rlm of c's own slope and rlp of the slope of c - e_d; hdl(c) = dl(c) / 2
          if (m == 0) { rlb[d] += gb3[0]; dlb[d] += 0.5 * gb3[2]; }
          else if (m == -1) rlb[d] += gb3[1];
          const double* src = m == -1 ? bp : (m == 0 ? b0 : bm);
#pragma unroll
          for (int v = 0; v < 5; ++v) acc[v] += src[v];
        }
      }
      // non-connection boundary faces naming this cell as boundary cell or neighbour
      if (outside == 0) {
        for (int f = 0; f < 6; ++f) {
          if (B.bct[f] == F_CONN) continue;
          const int d = f >> 1, d1 = (d + 1) % 3, d2 = (d + 2) % 3;
          const bool fh = f & 1;
          const int nd = n3[d];
          const int qb = fh ? nd : 1, qn = fh ? nd - 1 : 2;
          const int64_t slot = (int64_t)B.bcs[f] + (p3[d1] - 1) + (int64_t)n3[d1] * (p3[d2] - 1);
          if (p3[d] == qb) {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] += a.BB[(size_t)v * a.nslots + slot];
            if (GEO == 1 && a.GL) dlb[d] += 0.5 * a.BG[slot];                    // bdl = dl / 2
          }
          if (p3[d] == qn) {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] += a.BB[(size_t)(5 + v) * a.nslots + slot];
            if (GEO == 1 && a.GL) dlb[d] += 0.5 * a.BG[(size_t)a.nslots + slot];  // idl = dl / 2
          }
        }
      }
    }
    if (a.VPb && B.visc && outside == 0) {     // the viscous terms' part (k_adj_visc, on a second stream)
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += a.VPb[v * N + c];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) a.Pbar[v * N + c] = acc[v];
    if (GEO == 1 && a.GL) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a.GL[(size_t)d * N + c] = rlb[d];
        a.GL[(size_t)(3 + d) * N + c] = dlb[d];
      }
    }
  }
}
