// Reverse-mode (vector-Jacobian) counterparts of the per-cell and per-face
// 2D Euler numerics in device_numerics.hpp, for the discrete adjoint of the
// split RK2 step.
//
// Every function differentiates the forward expression piecewise: the
// branches (density / pressure floors, limiter cases, Davis wave-speed
// selection, HLLE upwind / subsonic branch, passive-flux upwinding) are
// decided on the primal values exactly as the forward decides them
// (euler1d.hpp:112-137,178-242; euler2d.hpp:20-64), and the derivative of
// the selected branch is propagated. Compiles for the host too (exact mode),
// where the CPU tests check each function against forward-mode dual numbers.
#pragma once
#include "device_numerics.hpp"

namespace NEST_NS {

// ---------------------------------------------------------------------------
// limiter: s = lim(l, r) per component; coefficients (cl, cr) with
// ds = cl*dl + cr*dr on the active branch
// ---------------------------------------------------------------------------
// minmod (euler1d.hpp:112-115): 0 if a*b <= 0, else the smaller magnitude
// (b on a tie).
__device__ __forceinline__ void minmod_coef(double a, double b, double& ca, double& cb) {
    const bool live = !(a * b <= 0.0);
    const bool pick_a = __builtin_fabs(a) < __builtin_fabs(b);
    ca = (live && pick_a) ? 1.0 : 0.0;
    cb = (live && !pick_a) ? 1.0 : 0.0;
}
template <int LIM>
__device__ __forceinline__ void limiter_coef(double l, double r, double& cl, double& cr) {
    if constexpr (LIM == LIM_MINMOD) {
        minmod_coef(l, r, cl, cr);
    } else if constexpr (LIM == LIM_MC) {
        // mc = minmod(2l, minmod(0.5(l+r), 2r))  (euler1d.hpp:117-121)
        const double c = 0.5 * (l + r);
        double a1, b1, a2, b2;
        minmod_coef(c, 2.0 * r, a1, b1);
        minmod_coef(2.0 * l, minmod(c, 2.0 * r), a2, b2);
        cl = 2.0 * a2 + 0.5 * b2 * a1;
        cr = b2 * (0.5 * a1 + 2.0 * b1);
    } else {
        cl = 0.5;
        cr = 0.5;
    }
}
struct coef4 {
    prim4 l, r;
};
template <int LIM>
__device__ __forceinline__ coef4 limiter_coef4(const prim4& dl, const prim4& dr) {
    coef4 c;
    limiter_coef<LIM>(dl.r, dr.r, c.l.r, c.r.r);
    limiter_coef<LIM>(dl.u, dr.u, c.l.u, c.r.u);
    limiter_coef<LIM>(dl.p, dr.p, c.l.p, c.r.p);
    limiter_coef<LIM>(dl.q, dr.q, c.l.q, c.r.q);
    return c;
}
// The limiter's branch of all 4 quantities, decided from (dl, dr) and applied
// later to the slope cotangent sb: dlb = cl*sb, drb = cr*sb. minmod keeps
// only which difference it picked (two lane masks per quantity, no
// floating-point coefficients); the other limiters keep the coefficients.
template <int LIM>
struct limsel4 {
    coef4 c;
    __device__ __forceinline__ void apply(const prim4& sb, prim4& dlb, prim4& drb) const {
        dlb = {c.l.r * sb.r, c.l.u * sb.u, c.l.p * sb.p, c.l.q * sb.q};
        drb = {c.r.r * sb.r, c.r.u * sb.u, c.r.p * sb.p, c.r.q * sb.q};
    }
};
template <>
struct limsel4<LIM_MINMOD> {
    bool l[4], r[4];
    __device__ __forceinline__ void apply(const prim4& sb, prim4& dlb, prim4& drb) const {
        dlb = {l[0] ? sb.r : 0.0, l[1] ? sb.u : 0.0, l[2] ? sb.p : 0.0, l[3] ? sb.q : 0.0};
        drb = {r[0] ? sb.r : 0.0, r[1] ? sb.u : 0.0, r[2] ? sb.p : 0.0, r[3] ? sb.q : 0.0};
    }
};
__device__ __forceinline__ void minmod_pick(double a, double b, bool& pa, bool& pb) {
    const bool live = !(a * b <= 0.0);
    const bool pick_a = __builtin_fabs(a) < __builtin_fabs(b);
    pa = live && pick_a;
    pb = live && !pick_a;
}
template <int LIM>
__device__ __forceinline__ limsel4<LIM> limiter_sel4(const prim4& dl, const prim4& dr) {
    limsel4<LIM> s;
    if constexpr (LIM == LIM_MINMOD) {
        minmod_pick(dl.r, dr.r, s.l[0], s.r[0]);
        minmod_pick(dl.u, dr.u, s.l[1], s.r[1]);
        minmod_pick(dl.p, dr.p, s.l[2], s.r[2]);
        minmod_pick(dl.q, dr.q, s.l[3], s.r[3]);
    } else {
        s.c = limiter_coef4<LIM>(dl, dr);
    }
    return s;
}

__device__ __forceinline__ prim4 mul4(const prim4& a, const prim4& b) {
    return {a.r * b.r, a.u * b.u, a.p * b.p, a.q * b.q};
}
__device__ __forceinline__ prim4 add4(const prim4& a, const prim4& b) {
    return {a.r + b.r, a.u + b.u, a.p + b.p, a.q + b.q};
}
__device__ __forceinline__ prim4 scale4(double s, const prim4& a) { return {s * a.r, s * a.u, s * a.p, s * a.q}; }
__device__ __forceinline__ prim4 zero4() { return {0.0, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ prim4 sel4(bool c, const prim4& a, const prim4& b) {
    return {c ? a.r : b.r, c ? a.u : b.u, c ? a.p : b.p, c ? a.q : b.q};
}

// ---------------------------------------------------------------------------
// get_prim_2d reverse (euler2d.hpp:44-52). W = (rho, u_n, p, q = m_t) as
// prim2d returns it; Wb its cotangent; Ub = cotangent of (rho, m_u, m_v, E).
// rho = max(r, 1e-10), p = max((g-1)(E - (m_u^2 + m_v^2)/(2 rho)), 1e-10):
// a floored quantity passes no derivative.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void prim2d_rev(const prim4& W, bool xsweep, const prim4& Wb, const eos_c& e,
                                           double Ub[4]) {
    const double ir = ddiv(1.0, W.r);
    const double un = W.u;
    const double ut = W.q * ir;
    const double gp = (W.p > 1e-10) ? e.gm1 * Wb.p : 0.0;
    const double mnb = Wb.u * ir - gp * un;
    const double mtb = Wb.q - gp * ut;
    const double rb = Wb.r + gp * (0.5 * (un * un + ut * ut)) - Wb.u * un * ir;
    Ub[0] = (W.r > 1e-10) ? rb : 0.0;
    Ub[1] = xsweep ? mnb : mtb;
    Ub[2] = xsweep ? mtb : mnb;
    Ub[3] = gp;
}

// ---------------------------------------------------------------------------
// Face flux reverse in the clamped form of the HLLE flux (the form the fast
// forward evaluates): with SL- = min(SL, 0), SR+ = max(SR, 0),
// inv = 1/(SR+ - SL-) and the weights wr = SR+ inv, wl = SL- inv,
// ws = SL- SR+ inv,
//   F_k = wr FL_k - wl FR_k + ws (UR_k - UL_k),  u_int = wr uL - wl uR,
// which is F_L when SL >= 0, F_R when SR <= 0 and the HLLE average between.
// Its derivative is the derivative of the branch the primal takes: in the
// upwind branches one weight is 0 and the other 1, and the speed cotangents
// vanish (SL- is constant for SL >= 0, SR+ for SR <= 0). With
// A = sum Fb_k FL_k, B = sum Fb_k FR_k, C = sum Fb_k (UR_k - UL_k) (the
// interface velocity counted in A and B with its cotangent):
//   d/dSL- = wr ((A - B) inv + C wr),  d/dSR+ = wl ((B - A) inv - C wl).
// One formula for every face: no per-branch selects. Evaluated with the
// numerics mode's division and sound speed (fast: c = (gamma p) y with
// y = rsq(gamma p rho), so dc/dp = gamma y / 2 and dc/drho = -c^2 y / 2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void face_flux_rev_clamped(const prim4& L, const prim4& R, const eos_c& e,
                                                      const double Fb[4], prim4& Lb, prim4& Rb) {
#if NEST_FAST
    const double gpL = e.gamma * L.p, gpR = e.gamma * R.p;
    const double yL = rsq(gpL * L.r), yR = rsq(gpR * R.r);
    const double cL = gpL * yL, cR = gpR * yR;
    const double cLp = 0.5 * e.gamma * yL, cRp = 0.5 * e.gamma * yR;
    const double cLr = -0.5 * cL * cL * yL, cRr = -0.5 * cR * cR * yR;
    const double igm1 = e.inv_gm1;
#else
    const double cL = sound_speed(L.r, L.p, e), cR = sound_speed(R.r, R.p, e);
    const double cLp = 0.5 * ddiv(cL, L.p), cRp = 0.5 * ddiv(cR, R.p);
    const double cLr = -0.5 * ddiv(cL, L.r), cRr = -0.5 * ddiv(cR, R.r);
    const double igm1 = 1.0 / e.gm1;
#endif
    const double aL = L.u - cL, aR = R.u - cR;
    const double bL = L.u + cL, bR = R.u + cR;
    const bool sl_from_r = aR < aL;  // SL = min(aL, aR), std::min order
    const bool sr_from_r = bL < bR;  // SR = max(bL, bR), std::max order
    const double SL = sl_from_r ? aR : aL;
    const double SR = sr_from_r ? bR : bL;
    const double SLm = dmin(SL, 0.0), SRp = dmax(SR, 0.0);
    const double inv = ddiv(1.0, SRp - SLm);
    const double wr = SRp * inv, wl = SLm * inv, ws = SLm * wr;
    // physical fluxes (r u, r u^2 + p, r H u) and jumps (dr, d(r u), dE),
    // r H = E + p = gamma/(gamma-1) p + r u^2 / 2
    const double g1 = e.gamma * igm1;
    const double mL = L.r * L.u, mR = R.r * R.u;
    const double kL = mL * L.u, kR = mR * R.u;
    const double HL = 0.5 * kL + g1 * L.p, HR = 0.5 * kR + g1 * R.p;
    const double FL1 = kL + L.p, FR1 = kR + R.p;
    const double FL2 = HL * L.u, FR2 = HR * R.u;
    const double dU0 = R.r - L.r, dU1 = mR - mL, dU2 = (HR - HL) - (R.p - L.p);
    // passive flux F3 = u_int * (u_int >= 0 ? qL : qR)  (euler2d.hpp:27-33)
    const double ui = wr * L.u - wl * R.u;
    const bool upw = ui >= 0.0;
    const double uib = Fb[3] * (upw ? L.q : R.q);
    Lb.q += upw ? Fb[3] * ui : 0.0;
    Rb.q += upw ? 0.0 : Fb[3] * ui;
    const double A = Fb[0] * mL + Fb[1] * FL1 + Fb[2] * FL2 + uib * L.u;
    const double B = Fb[0] * mR + Fb[1] * FR1 + Fb[2] * FR2 + uib * R.u;
    const double C = Fb[0] * dU0 + Fb[1] * dU1 + Fb[2] * dU2;
    const double SLb = (SL < 0.0) ? wr * ((A - B) * inv + C * wr) : 0.0;
    const double SRb = (SR > 0.0) ? wl * ((B - A) * inv - C * wl) : 0.0;
    // cotangents of FL_k (wr Fb_k), FR_k (-wl Fb_k) and the jumps (ws Fb_k)
    const double fl0 = wr * Fb[0], fl1 = wr * Fb[1], fl2 = wr * Fb[2];
    const double fr0 = -wl * Fb[0], fr1 = -wl * Fb[1], fr2 = -wl * Fb[2];
    const double du0 = ws * Fb[0], du1 = ws * Fb[1], du2 = ws * Fb[2];
    // left state: m = r u, k = m u, r H = k / 2 + g1 p
    const double hLb = fl2 * L.u - du2;
    const double kLb = fl1 + 0.5 * hLb;
    const double mLb = fl0 - du1;
    double uLb = fl2 * HL + mLb * L.r + 2.0 * kLb * mL + uib * wr;
    double rLb = L.u * (mLb + kLb * L.u) - du0;
    double pLb = fl1 + g1 * hLb + du2;
    // right state
    const double hRb = fr2 * R.u + du2;
    const double kRb = fr1 + 0.5 * hRb;
    const double mRb = fr0 + du1;
    double uRb = fr2 * HR + mRb * R.r + 2.0 * kRb * mR - uib * wl;
    double rRb = R.u * (mRb + kRb * R.u) + du0;
    double pRb = fr1 + g1 * hRb - du2;
    // Davis speeds and the sound speeds under them
    const double aLb = sl_from_r ? 0.0 : SLb, aRb = SLb - aLb;
    const double bLb = sr_from_r ? 0.0 : SRb, bRb = SRb - bLb;
    uLb += aLb + bLb;
    uRb += aRb + bRb;
    const double cLb = bLb - aLb, cRb = bRb - aRb;
    pLb += cLb * cLp;
    rLb += cLb * cLr;
    pRb += cRb * cRp;
    rRb += cRb * cRr;
    Lb.r += rLb;
    Lb.u += uLb;
    Lb.p += pLb;
    Rb.r += rRb;
    Rb.u += uRb;
    Rb.p += pRb;
}

// ---------------------------------------------------------------------------
// face flux reverse: face_flux<true> (HLLE with Davis speeds on (rho, u_n, p)
// plus the passive transverse-momentum flux). Fb = cotangent of F[0..3];
// accumulates the cotangents of the two face states into Lb, Rb.
// Fast numerics: the clamped form above (the fast forward's own form).
// Exact numerics: branch-free over the reference's branches: the subsonic
// (HLLE-average) terms are formed for every face (SR > SL always, since
// c > 0) and the branch the primal takes selects which terms are passed on,
// so a wave has no divergent paths here.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void face_flux_rev(const prim4& L, const prim4& R, const eos_c& e, const double Fb[4],
                                              prim4& Lb, prim4& Rb) {
#if NEST_FAST
    face_flux_rev_clamped(L, R, e, Fb, Lb, Rb);
    return;
#endif
    // ---- primal recompute, as face_flux<true>
#if NEST_FAST
    // c = (gamma p) y with y = rsq(gamma p rho), as sound_speed; y also gives
    // the derivatives without a division: dc/dp = gamma y / 2, dc/drho =
    // -c / (2 rho) = -c^2 y / 2
    const double gpL = e.gamma * L.p, gpR = e.gamma * R.p;
    const double yL = rsq(gpL * L.r), yR = rsq(gpR * R.r);
    const double cL = gpL * yL;
    const double cR = gpR * yR;
#else
    const double cL = sound_speed(L.r, L.p, e);
    const double cR = sound_speed(R.r, R.p, e);
#endif
    const double EL = total_energy(L.r, L.u, L.p, e);
    const double ER = total_energy(R.r, R.u, R.p, e);
    const double aL = L.u - cL, aR = R.u - cR;
    const double bL = L.u + cL, bR = R.u + cR;
    const bool sl_from_r = aR < aL;  // SL = min(aL, aR), std::min order
    const bool sr_from_r = bL < bR;  // SR = max(bL, bR), std::max order
    const double SL = sl_from_r ? aR : aL;
    const double SR = sr_from_r ? bR : bL;
    const bool left = SL >= 0.0;
    const bool right = !left && (SR <= 0.0);
    const bool mid = !left && !right;
    const double FL[3] = {L.r * L.u, L.r * L.u * L.u + L.p, (EL + L.p) * L.u};
    const double FR[3] = {R.r * R.u, R.r * R.u * R.u + R.p, (ER + R.p) * R.u};
#if NEST_FAST
    const double inv = rcp(SR - SL);
    const double um = (SR * L.u - SL * R.u) * inv;
#else
    const double inv = 1.0 / (SR - SL);
    const double um = (SR * L.u - SL * R.u) / (SR - SL);
#endif
    const double ui = left ? L.u : (right ? R.u : um);
    // passive flux F3 = ui * (ui >= 0 ? qL : qR)  (euler2d.hpp:27-33)
    const bool upw = ui >= 0.0;
    const double uib = Fb[3] * (upw ? L.q : R.q);
    Lb.q += upw ? Fb[3] * ui : 0.0;
    Rb.q += upw ? 0.0 : Fb[3] * ui;

    // subsonic branch: H_k = (SR FL_k - SL FR_k + SL SR (UR_k - UL_k)) / (SR - SL),
    // k = 0..2, U = (r, r u, E); interface velocity (SR uL - SL uR) / (SR - SL)
    const double dU[3] = {R.r - L.r, R.r * R.u - L.r * L.u, ER - EL};
    const double sLR = SL * SR;
    double sr_acc = 0.0, sl_acc = 0.0;
    double Fbl[3], Fbr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double H = (SR * FL[k] - SL * FR[k] + sLR * dU[k]) * inv;
        sr_acc += Fb[k] * (FL[k] + SL * dU[k] - H);
        sl_acc += Fb[k] * (SR * dU[k] - FR[k] + H);
        Fbl[k] = left ? Fb[k] : (mid ? Fb[k] * SR * inv : 0.0);
        Fbr[k] = right ? Fb[k] : (mid ? -Fb[k] * SL * inv : 0.0);
    }
    sr_acc += uib * (L.u - ui);
    sl_acc += uib * (ui - R.u);
    const double SRb = mid ? sr_acc * inv : 0.0;
    const double SLb = mid ? sl_acc * inv : 0.0;
    Lb.u += left ? uib : (mid ? uib * SR * inv : 0.0);
    Rb.u += right ? uib : (mid ? -uib * SL * inv : 0.0);
    // state jump (UR_k - UL_k) weighted by SL SR / (SR - SL)
    const double w = mid ? sLR * inv : 0.0;
    const double j0 = Fb[0] * w, j1 = Fb[1] * w, j2 = Fb[2] * w;
    Rb.r += j0 + j1 * R.u;
    Rb.u += j1 * R.r;
    Lb.r -= j0 + j1 * L.u;
    Lb.u -= j1 * L.r;
    double ERb = j2, ELb = -j2;
    // physical fluxes F(W) = (r u, r u^2 + p, (E + p) u)
    ELb += Fbl[2] * L.u;
    Lb.r += Fbl[0] * L.u + Fbl[1] * L.u * L.u;
    Lb.u += Fbl[0] * L.r + 2.0 * Fbl[1] * L.r * L.u + Fbl[2] * (EL + L.p);
    Lb.p += Fbl[1] + Fbl[2] * L.u;
    ERb += Fbr[2] * R.u;
    Rb.r += Fbr[0] * R.u + Fbr[1] * R.u * R.u;
    Rb.u += Fbr[0] * R.r + 2.0 * Fbr[1] * R.r * R.u + Fbr[2] * (ER + R.p);
    Rb.p += Fbr[1] + Fbr[2] * R.u;
    // total energy E = p/(g-1) + r u^2 / 2  (euler1d.hpp:47-51)
#if NEST_FAST
    const double igm1 = e.inv_gm1;
#else
    const double igm1 = 1.0 / e.gm1;
#endif
    Lb.p += ELb * igm1;
    Lb.r += ELb * 0.5 * L.u * L.u;
    Lb.u += ELb * L.r * L.u;
    Rb.p += ERb * igm1;
    Rb.r += ERb * 0.5 * R.u * R.u;
    Rb.u += ERb * R.r * R.u;
    // Davis speeds: SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR)
    Lb.u += (sl_from_r ? 0.0 : SLb) + (sr_from_r ? 0.0 : SRb);
    Rb.u += (sl_from_r ? SLb : 0.0) + (sr_from_r ? SRb : 0.0);
    const double cLb = (sr_from_r ? 0.0 : SRb) - (sl_from_r ? 0.0 : SLb);
    const double cRb = (sr_from_r ? SRb : 0.0) - (sl_from_r ? SLb : 0.0);
    // c = sqrt(gamma p / r): dc/dp = c / (2p), dc/dr = -c / (2r)  (p > 0 in 2D)
#if NEST_FAST
    Lb.p += cLb * (0.5 * e.gamma * yL);
    Lb.r -= cLb * (0.5 * cL * cL * yL);
    Rb.p += cRb * (0.5 * e.gamma * yR);
    Rb.r -= cRb * (0.5 * cR * cR * yR);
#else
    Lb.p += cLb * 0.5 * ddiv(cL, L.p);
    Lb.r -= cLb * 0.5 * ddiv(cL, L.r);
    Rb.p += cRb * 0.5 * ddiv(cR, R.p);
    Rb.r -= cRb * 0.5 * ddiv(cR, R.r);
#endif
}

}  // namespace NEST_NS
