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
// face flux reverse: face_flux<true> (HLLE with Davis speeds on (rho, u_n, p)
// plus the passive transverse-momentum flux). Fb = cotangent of F[0..3];
// accumulates the cotangents of the two face states into Lb, Rb.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void face_flux_rev(const prim4& L, const prim4& R, const eos_c& e, const double Fb[4],
                                              prim4& Lb, prim4& Rb) {
    // ---- primal recompute, as face_flux<true>
    const double cL = sound_speed(L.r, L.p, e);
    const double cR = sound_speed(R.r, R.p, e);
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
    const double FL[3] = {L.r * L.u, L.r * L.u * L.u + L.p, (EL + L.p) * L.u};
    const double FR[3] = {R.r * R.u, R.r * R.u * R.u + R.p, (ER + R.p) * R.u};

    double Fbl[3] = {0.0, 0.0, 0.0}, Fbr[3] = {0.0, 0.0, 0.0};
    double ELb = 0.0, ERb = 0.0, SLb = 0.0, SRb = 0.0;
    double ui, uib;
    if (left) {
        ui = L.u;
    } else if (right) {
        ui = R.u;
    } else {
        ui = ddiv(SR * L.u - SL * R.u, SR - SL);
    }
    // passive flux F3 = ui * (ui >= 0 ? qL : qR)  (euler2d.hpp:27-33)
    if (ui >= 0.0) {
        uib = Fb[3] * L.q;
        Lb.q += Fb[3] * ui;
    } else {
        uib = Fb[3] * R.q;
        Rb.q += Fb[3] * ui;
    }
    if (left) {
        Fbl[0] = Fb[0], Fbl[1] = Fb[1], Fbl[2] = Fb[2];
        Lb.u += uib;
    } else if (right) {
        Fbr[0] = Fb[0], Fbr[1] = Fb[1], Fbr[2] = Fb[2];
        Rb.u += uib;
    } else {
        // H_k = (SR FL_k - SL FR_k + SL SR (UR_k - UL_k)) / (SR - SL), k = 0..2,
        // U = (r, r u, E); interface velocity (SR uL - SL uR) / (SR - SL)
        const double inv = ddiv(1.0, SR - SL);
        const double dU[3] = {R.r - L.r, R.r * R.u - L.r * L.u, ER - EL};
        const double sLR = SL * SR;
        double sr_acc = 0.0, sl_acc = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double H = (SR * FL[k] - SL * FR[k] + sLR * dU[k]) * inv;
            sr_acc += Fb[k] * (FL[k] + SL * dU[k] - H);
            sl_acc += Fb[k] * (SR * dU[k] - FR[k] + H);
            Fbl[k] = Fb[k] * SR * inv;
            Fbr[k] = -Fb[k] * SL * inv;
        }
        sr_acc += uib * (L.u - ui);
        sl_acc += uib * (ui - R.u);
        SRb = sr_acc * inv;
        SLb = sl_acc * inv;
        Lb.u += uib * SR * inv;
        Rb.u -= uib * SL * inv;
        // state jump (UR_k - UL_k) weighted by SL SR / (SR - SL)
        const double j0 = Fb[0] * sLR * inv, j1 = Fb[1] * sLR * inv, j2 = Fb[2] * sLR * inv;
        Rb.r += j0 + j1 * R.u;
        Rb.u += j1 * R.r;
        ERb += j2;
        Lb.r -= j0 + j1 * L.u;
        Lb.u -= j1 * L.r;
        ELb -= j2;
    }
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
    const double igm1 = ddiv(1.0, e.gm1);
    Lb.p += ELb * igm1;
    Lb.r += ELb * 0.5 * L.u * L.u;
    Lb.u += ELb * L.r * L.u;
    Rb.p += ERb * igm1;
    Rb.r += ERb * 0.5 * R.u * R.u;
    Rb.u += ERb * R.r * R.u;
    // Davis speeds: SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR)
    double cLb = 0.0, cRb = 0.0;
    if (sl_from_r) {
        Rb.u += SLb;
        cRb -= SLb;
    } else {
        Lb.u += SLb;
        cLb -= SLb;
    }
    if (sr_from_r) {
        Rb.u += SRb;
        cRb += SRb;
    } else {
        Lb.u += SRb;
        cLb += SRb;
    }
    // c = sqrt(gamma p / r): dc/dp = c / (2p), dc/dr = -c / (2r)  (p > 0 in 2D)
    Lb.p += cLb * 0.5 * ddiv(cL, L.p);
    Lb.r -= cLb * 0.5 * ddiv(cL, L.r);
    Rb.p += cRb * 0.5 * ddiv(cR, R.p);
    Rb.r -= cRb * 0.5 * ddiv(cR, R.r);
}

}  // namespace NEST_NS
