// Device-side Euler numerics for gfx950 (FP64, wave64).
//
// Included by kernels.hip, which is compiled twice:
//   NEST_FAST=0 (-ffp-contract=off): every operation is the IEEE double
//     operation the reference CPU code performs, in its order
//     (examples/sod1d/euler1d.hpp:27-242, examples/sod2d/euler2d.hpp:20-64),
//     with correctly rounded division and square root, so results are
//     bitwise those of the CPU path.
//   NEST_FAST=1 (FMA contraction on, NaNs assumed absent, IEEE mode off):
//     divisions become reciprocal-multiplies (hardware v_rcp_f64 refined by
//     one Newton step), per-face quotients share one reciprocal, the EOS uses
//     1/(gamma-1) and the limiters hardware min / max. Differences from the
//     CPU path are a few ulp per operation; the per-step tolerance is 1e-12
//     relative.
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
// host compilation (CPU unit tests of the per-face math, exact mode only)
#define __device__
#define __forceinline__ inline
#endif

#include "nest_types.hpp"

#ifndef NEST_FAST
#error "define NEST_FAST to 0 or 1"
#endif
#if NEST_FAST && !defined(__HIPCC__)
#error "fast numerics use gfx950 builtins: device compilation only"
#endif

namespace NEST_NS {

enum : int { LIM_MINMOD = 0, LIM_MC = 1, LIM_NONE = 2 };

using eos_c = ::nest_eos;  // gamma, gm1 = gamma - 1.0, inv_gm1 = 1.0 / gm1 (fast mode only)

// std::min / std::max semantics (selects on a strict compare). Fast mode uses
// the hardware v_min_f64 / v_max_f64, which agree except for the sign of a
// zero result when the operands are zeros of opposite sign (every use below
// only compares such a result against 0 or takes it where it cannot be a zero).
#if NEST_FAST
__device__ __forceinline__ double dmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return __builtin_fmax(a, b); }
#else
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
#endif

#if NEST_FAST
// Newton-Raphson steps after the hardware reciprocal and reciprocal square
// root estimates (relative error 2^-24.3): one step leaves a relative error
// of about 2^-48 (3e-15), two about 1 ulp. One step keeps every fast-mode
// step within 1e-12 of the exact numerics and saves ~45 FP64 instructions
// per row of the one-launch step kernel (VALU-bound: -5% time).
#ifndef NEST_NR_STEPS
#define NEST_NR_STEPS 1
#endif
// 1/x: hardware estimate + NEST_NR_STEPS Newton-Raphson steps
__device__ __forceinline__ double rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int k = 0; k < NEST_NR_STEPS; ++k) {
        const double e = __builtin_fma(-x, y, 1.0);
        y = __builtin_fma(y, e, y);
    }
    return y;
}
__device__ __forceinline__ double ddiv(double a, double b) { return a * rcp(b); }
#else
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }
#endif
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }

// NaN and infinity tests on the bit pattern. The fast kernels are compiled
// with NaNs assumed absent (so min / max need no NaN quieting), under which
// x != x folds to false; these integer tests hold under any floating-point
// assumption.
__device__ __forceinline__ bool bits_nan(double x) {
    return ((unsigned long long)__builtin_bit_cast(long long, x) & 0x7fffffffffffffffull) > 0x7ff0000000000000ull;
}
__device__ __forceinline__ bool bits_inf(double x) {
    return ((unsigned long long)__builtin_bit_cast(long long, x) & 0x7fffffffffffffffull) == 0x7ff0000000000000ull;
}
// dt usable for a step: finite and positive
__device__ __forceinline__ bool dt_ok(double dt) { return !bits_nan(dt) && !bits_inf(dt) && dt > 0.0; }

// ---------------------------------------------------------------------------
// limiters — euler1d.hpp:112-137
// ---------------------------------------------------------------------------
// For a*b > 0 the reference's (|a| < |b|) ? a : b is the smaller magnitude
// with the common sign: copysign(min(|a|, |b|), a) returns the same bits
// (equal magnitudes give b's bits, as the reference does).
// Fast numerics take the median of (a, b, 0): the smaller magnitude when a
// and b share a sign, else 0 (the same value; a zero result may carry either
// sign). Four hardware min/max, no compare, no selects.
__device__ __forceinline__ double minmod(double a, double b) {
#if NEST_FAST
    return __builtin_fmax(__builtin_fmin(a, b), __builtin_fmin(__builtin_fmax(a, b), 0.0));
#else
    double r = (__builtin_fabs(a) < __builtin_fabs(b)) ? a : b;
    return (a * b <= 0.0) ? 0.0 : r;
#endif
}
template <int LIM>
__device__ __forceinline__ double limiter(double l, double r) {
    if constexpr (LIM == LIM_MINMOD) {
        return minmod(l, r);
    } else if constexpr (LIM == LIM_MC) {
        double c = 0.5 * (l + r);
        return minmod(2.0 * l, minmod(c, 2.0 * r));
    } else {
        return 0.5 * (l + r);
    }
}

// ---------------------------------------------------------------------------
// primitive state of one cell, 4 reconstructed quantities
// 2D: (rho, u_normal, p, q = transverse momentum), euler2d.hpp:36-64
// 1D: (rho, u, p, unused)
// ---------------------------------------------------------------------------
struct prim4 {
    double r, u, p, q;
};

// get_prim_2d — euler2d.hpp:44-52. mn = normal momentum, mt = transverse.
// The kinetic energy always uses u = rho_u/rho, v = rho_v/rho in (u*u + v*v)
// order, whichever direction is normal, as the reference does.
__device__ __forceinline__ prim4 prim2d(double r_in, double mu, double mv, double E, bool xsweep,
                                        const eos_c& e) {
    prim4 w;
    double rho = dmax(r_in, 1e-10);
#if NEST_FAST
    // rho (u^2 + v^2) = mu u + mv v: the same value with two fewer products
    double ir = rcp(rho);
    double u = mu * ir;
    double v = mv * ir;
    w.p = dmax(e.gm1 * __builtin_fma(-0.5, mu * u + mv * v, E), 1e-10);
#else
    double u = mu / rho;
    double v = mv / rho;
    double ke = 0.5 * rho * (u * u + v * v);
    w.p = dmax(e.gm1 * (E - ke), 1e-10);
#endif
    w.r = rho;
    w.u = xsweep ? u : v;
    w.q = xsweep ? mv : mu;
    return w;
}

// cons_to_prim — euler1d.hpp:69-84 with eos_t::pressure :31-44 (p >= 0 clamp,
// no density floor). The reference re-divides m/rho inside pressure(); the
// quotient is the same value.
__device__ __forceinline__ prim4 prim1d(double rho, double m, double E, const eos_c& e) {
    prim4 w;
    double u = ddiv(m, rho);
    double ke = 0.5 * rho * u * u;
    w.r = rho;
    w.u = u;
    w.p = dmax(e.gm1 * (E - ke), 0.0);
    w.q = 0.0;
    return w;
}

#if NEST_FAST
// 1/sqrt(x): hardware estimate + NEST_NR_STEPS Newton-Raphson steps
__device__ __forceinline__ double rsq(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
#pragma unroll
    for (int k = 0; k < NEST_NR_STEPS; ++k) y = y * __builtin_fma(-h, y * y, 1.5);
    return y;
}
#endif

// sound speed — euler1d.hpp:54-56: sqrt(gamma*p/rho). Fast mode uses
// sqrt(gamma*p/rho) = (gamma*p) * rsq(gamma*p*rho): one reciprocal square root
// instead of a division and a square root.
__device__ __forceinline__ double sound_speed(double rho, double p, const eos_c& e) {
#if NEST_FAST
    const double gp = e.gamma * p;
    return gp * rsq(gp * rho);
#else
    return dsqrt(ddiv(e.gamma * p, rho));
#endif
}
// 1D variant. The cell pressure is clamped to p >= 0 (euler1d.hpp:43), so p
// may be exactly 0, and a face state of the unlimited scheme may have p < 0
// together with rho < 0 (a positive quotient). Fast numerics use
// sqrt(gamma p / rho) = |gamma p| rsq(gamma p rho) for either sign of the
// pair, c = 0 for p = 0 (where the product form would give 0 * inf), and a
// NaN where the quotient is negative, as sqrt gives.
__device__ __forceinline__ double sound_speed_1d(double rho, double p, const eos_c& e) {
#if NEST_FAST
    const double gp = e.gamma * p;
    return gp != 0.0 ? __builtin_fabs(gp) * rsq(gp * rho) : 0.0;
#else
    return dsqrt(ddiv(e.gamma * p, rho));
#endif
}
// total energy — euler1d.hpp:47-51
__device__ __forceinline__ double total_energy(double rho, double u, double p, const eos_c& e) {
    double ke = 0.5 * rho * u * u;
#if NEST_FAST
    double ie = p * e.inv_gm1;
#else
    double ie = p / e.gm1;
#endif
    return ie + ke;
}

// HLLE flux with Davis speeds (euler1d.hpp:178-242) on (rho, u_n, p), plus the
// passive transverse-momentum flux upwinded by the HLLE-weighted interface
// velocity (euler2d.hpp:20-33, :115-122). The sound speeds are shared between
// the two (the reference recomputes the same values, :116-118).
// F[0..2] = HLLE (mass, normal momentum, energy); F[3] = passive flux.
//
// CLAMP (fast numerics): the three branches in one formula, with
// the speeds clamped as SL- = min(SL, 0), SR+ = max(SR, 0):
//   F = (SR+ F_L - SL- F_R + SL- SR+ (U_R - U_L)) / (SR+ - SL-)
// is F_L when SL >= 0 (SL- = 0, the quotient SR F_L / SR), F_R when SR <= 0,
// and the HLLE average otherwise; likewise the interface velocity. It
// replaces the branch selects by two min/max and differs from the selected
// value by the rounding of SR * (1/SR) (a few ulp). The denominator is
// positive in 2D: SR - SL >= c_L + c_R > 0, as the 2D floors keep rho, p > 0
// (1D: see the floor below).
template <bool PASSIVE, bool CLAMP = false>
__device__ __forceinline__ void face_flux(const prim4& L, const prim4& R, const eos_c& e, double F[4]) {
    double cL = PASSIVE ? sound_speed(L.r, L.p, e) : sound_speed_1d(L.r, L.p, e);
    double EL = total_energy(L.r, L.u, L.p, e);
    double cR = PASSIVE ? sound_speed(R.r, R.p, e) : sound_speed_1d(R.r, R.p, e);
    double ER = total_energy(R.r, R.u, R.p, e);
    double SL = dmin(L.u - cL, R.u - cR);
    double SR = dmax(L.u + cL, R.u + cR);
    double FL0 = L.r * L.u, FL1 = L.r * L.u * L.u + L.p, FL2 = (EL + L.p) * L.u;
    double FR0 = R.r * R.u, FR1 = R.r * R.u * R.u + R.p, FR2 = (ER + R.p) * R.u;
    double UL1 = L.r * L.u, UR1 = R.r * R.u;
#ifndef NEST_HLL_CLAMP
#define NEST_HLL_CLAMP 1
#endif
#if NEST_FAST && NEST_HLL_CLAMP
    if constexpr (CLAMP) {
        // the weights SR+/(SR+ - SL-), SL-/(SR+ - SL-) and their product
        // term formed once; enthalpy rho H = p gamma/(gamma-1) + rho u^2/2
        // gives the energy flux and E (= rho H - p) with one product less.
        // 1D allows c = 0 on both sides (p = 0): SR+ - SL- = 0 only when
        // u = c = 0 on both sides, where every flux is 0 and the floored
        // denominator gives exactly that.
        const double SLm = dmin(SL, 0.0), SRp = dmax(SR, 0.0);
        const double inv = PASSIVE ? rcp(SRp - SLm) : rcp(dmax(SRp - SLm, 1e-300));
        const double wr = SRp * inv, wl = SLm * inv, ws = SLm * wr;
        const double g1 = e.gamma * e.inv_gm1;
        const double mL = L.r * L.u, mR = R.r * R.u;
        const double kL = mL * L.u, kR = mR * R.u;
        const double HL = __builtin_fma(0.5, kL, L.p * g1), HR = __builtin_fma(0.5, kR, R.p * g1);
        F[0] = __builtin_fma(ws, R.r - L.r, __builtin_fma(-wl, mR, wr * mL));
        F[1] = __builtin_fma(ws, mR - mL, __builtin_fma(-wl, kR + R.p, wr * (kL + L.p)));
        F[2] = __builtin_fma(ws, (HR - HL) - (R.p - L.p), __builtin_fma(-wl, HR * R.u, wr * (HL * L.u)));
        if constexpr (PASSIVE) {
            const double ui = __builtin_fma(-wl, R.u, wr * L.u);
            F[3] = ui * ((ui >= 0.0) ? L.q : R.q);
        } else {
            F[3] = 0.0;
        }
        (void)EL;
        (void)ER;
        (void)FL0;
        (void)FL1;
        (void)FL2;
        (void)FR0;
        (void)FR1;
        (void)FR2;
        return;
    }
#endif
    double sLR = SL * SR;
#if NEST_FAST
    double inv = rcp(SR - SL);
    double H0 = (SR * FL0 - SL * FR0 + sLR * (R.r - L.r)) * inv;
    double H1 = (SR * FL1 - SL * FR1 + sLR * (UR1 - UL1)) * inv;
    double H2 = (SR * FL2 - SL * FR2 + sLR * (ER - EL)) * inv;
    double Hu = (SR * L.u - SL * R.u) * inv;
#else
    double d = SR - SL;
    double H0 = (SR * FL0 - SL * FR0 + sLR * (R.r - L.r)) / d;
    double H1 = (SR * FL1 - SL * FR1 + sLR * (UR1 - UL1)) / d;
    double H2 = (SR * FL2 - SL * FR2 + sLR * (ER - EL)) / d;
    double Hu = (SR * L.u - SL * R.u) / d;
#endif
    bool left = SL >= 0.0;
    bool right = !left && (SR <= 0.0);
    F[0] = left ? FL0 : (right ? FR0 : H0);
    F[1] = left ? FL1 : (right ? FR1 : H1);
    F[2] = left ? FL2 : (right ? FR2 : H2);
    if constexpr (PASSIVE) {
        double ui = left ? L.u : (right ? R.u : Hu);
        F[3] = (ui >= 0.0) ? ui * L.q : ui * R.q;
    } else {
        (void)Hu;
        F[3] = 0.0;
    }
}

#if defined(__HIPCC__)
// XCD-aware block order. The dispatcher hands block b of a launch to XCD
// b % 8 (gfx950: 8 XCDs, each with its own L2). Renumbering the blocks so
// that each XCD runs one contiguous run of the work list puts neighbouring
// strips (which share halo columns and partly filled cache lines) and the
// strip below each one (whose first rows are the last rows of this one) in
// the same L2. A bijection on [0, nb).
#ifndef NEST_XCD_SWIZZLE
#define NEST_XCD_SWIZZLE 1
#endif
__device__ __forceinline__ int xcd_block(int b, int nb) {
#if NEST_XCD_SWIZZLE
    constexpr int X = 8;
    const int q = nb / X, r = nb % X, x = b % X;
    return x * q + (x < r ? x : r) + b / X;
#else
    (void)nb;
    return b;
#endif
}

// Neighbour-lane values by DPP wave shifts (a VALU move, no LDS round trip):
// lane_up(x) on lane l is x of lane l-1 and lane_down(x) is x of lane l+1;
// lanes 0 / 63 keep their own value (the semantics of __shfl_up / __shfl_down
// with delta 1).
__device__ __forceinline__ double lane_up(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_down(double x) {
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(lo, lo, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// The same shifts with lane 0 (lane_up_z) / lane 63 (lane_down_z) receiving
// 0 instead of keeping its own value: no old operand, so no register copy
// before the DPP move. For stencils whose edge-lane results are not used.
__device__ __forceinline__ double lane_up_z(double x) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_down_z(double x) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x130 /* wave_shl:1 */, 0xf, 0xf, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
#endif

#if defined(__HIPCC__)
// ---------------------------------------------------------------------------
// global -> LDS row fetches (global_load_lds_dword) and the waits they need
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void* lds_void_ptr;

// s_waitcnt (gfx9 encoding: vmcnt[3:0] in bits 3:0 and vmcnt[5:4] in 15:14,
// expcnt in 6:4, lgkmcnt in 11:8). The signal fences keep the compiler from
// moving memory operations across the wait.
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void wait_lgkmcnt0() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
__device__ __forceinline__ void compiler_fence() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// A work item read by a whole wave (wave-uniform index), its fields marked
// wave-uniform: the compiler keeps them, and the march bounds derived from
// them, in scalar registers whatever else it infers.
__device__ __forceinline__ nest_item uniform_item(const nest_item& w) {
    return {__builtin_amdgcn_readfirstlane(w.patch), __builtin_amdgcn_readfirstlane(w.c0),
            __builtin_amdgcn_readfirstlane(w.r0), __builtin_amdgcn_readfirstlane(w.r1)};
}

__device__ __forceinline__ const char* shfl_cptr(const void* p, int src) {
    return reinterpret_cast<const char*>(__shfl(reinterpret_cast<long long>(p), src));
}

#endif

// slope of all 4 quantities from the left and right differences
template <int LIM>
__device__ __forceinline__ prim4 slope4(const prim4& dl, const prim4& dr) {
    return {limiter<LIM>(dl.r, dr.r), limiter<LIM>(dl.u, dr.u), limiter<LIM>(dl.p, dr.p),
            limiter<LIM>(dl.q, dr.q)};
}
__device__ __forceinline__ prim4 sub4(const prim4& a, const prim4& b) {
    return {a.r - b.r, a.u - b.u, a.p - b.p, a.q - b.q};
}
// W + 0.5*s (right face of a cell) / W - 0.5*s (left face), euler1d.hpp:166-167
__device__ __forceinline__ prim4 face_plus(const prim4& w, const prim4& s) {
    return {w.r + 0.5 * s.r, w.u + 0.5 * s.u, w.p + 0.5 * s.p, w.q + 0.5 * s.q};
}
__device__ __forceinline__ prim4 face_minus(const prim4& w, const prim4& s) {
    return {w.r - 0.5 * s.r, w.u - 0.5 * s.u, w.p - 0.5 * s.p, w.q - 0.5 * s.q};
}

}  // namespace NEST_NS
