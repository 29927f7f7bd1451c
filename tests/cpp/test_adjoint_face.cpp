// CPU check of the adjoint building blocks (nest_amd/csrc/adjoint_numerics.hpp,
// compiled for the host in exact mode) against forward-mode dual numbers run
// through the oracle's restatement of the reference numerics:
//   <Wb, J dU> == <J^T Wb, dU>   for get_prim_2d (both sweep directions)
//   <sb, ds>   == sb*(cl dl + cr dr)   for every limiter
//   <Fb, J dLR> == <Lb, dL> + <Rb, dR>   for the HLLE + passive face flux
//     (branch form, and the clamped form the fast numerics run)
// over random states that visit every branch (floors, upwind / subsonic
// HLLE, both Davis speed choices, both passive-flux signs). Exit status 0 iff
// every relative mismatch is below 1e-12.
#include <cmath>
#include <cstdint>
#include <cstdio>

#define NEST_FAST 0
#define NEST_NS nest_exact
#include "../../nest_amd/csrc/adjoint_numerics.hpp"
#include "../../oracle/dual.hpp"
#include "../../oracle/euler_oracle.hpp"

using nest_oracle::dual;
namespace ne = nest_exact;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand() {  // splitmix64 -> [0, 1)
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
static double urange(double a, double b) { return a + (b - a) * urand(); }

static int fails = 0;
static double worst = 0.0;
static void check(const char* what, int i, double lhs, double rhs, double scale) {
    const double err = std::fabs(lhs - rhs) / (scale > 0.0 ? scale : 1.0);
    if (err > worst) worst = err;
    if (!(err <= 1e-12)) {
        if (fails < 10) std::printf("FAIL %s case %d: %.17g vs %.17g (rel %.3g)\n", what, i, lhs, rhs, err);
        ++fails;
    }
}

static const ne::eos_c EOS{1.4, 1.4 - 1.0, 1.0 / (1.4 - 1.0)};

// ---- get_prim_2d -----------------------------------------------------------
static void test_prim(int n) {
    for (int i = 0; i < n; ++i) {
        double U[4];
        const int kind = i % 4;  // 0 regular, 1 floored density, 2 floored pressure, 3 both
        U[0] = (kind & 1) ? urange(-1e-3, 1e-11) : urange(0.1, 2.0);
        U[1] = urange(-1.0, 1.0) * ((kind & 1) ? 1e-12 : 1.0);
        U[2] = urange(-1.0, 1.0) * ((kind & 1) ? 1e-12 : 1.0);
        const double ke = (kind & 1) ? 0.0 : 0.5 * (U[1] * U[1] + U[2] * U[2]) / U[0];
        U[3] = (kind & 2) ? ke + urange(-0.5, 1e-12) : ke + urange(0.2, 3.0);
        for (int xs = 0; xs < 2; ++xs) {
            double dU[4], Wb[4];
            for (int k = 0; k < 4; ++k) dU[k] = urange(-1, 1), Wb[k] = urange(-1, 1);
            dual W[3], q;
            nest_oracle::get_prim_2d(dual(U[0], dU[0]), dual(U[1], dU[1]), dual(U[2], dU[2]), dual(U[3], dU[3]),
                                     1.4, xs != 0, W, q);
            const double lhs = Wb[0] * W[0].d + Wb[1] * W[1].d + Wb[2] * W[2].d + Wb[3] * q.d;
            const ne::prim4 Wp = ne::prim2d(U[0], U[1], U[2], U[3], xs != 0, EOS);
            double Ub[4];
            ne::prim2d_rev(Wp, xs != 0, {Wb[0], Wb[1], Wb[2], Wb[3]}, EOS, Ub);
            double rhs = 0.0, scale = 0.0;
            for (int k = 0; k < 4; ++k) rhs += Ub[k] * dU[k], scale += std::fabs(Ub[k] * dU[k]);
            check("prim2d", i, lhs, rhs, scale);
        }
    }
}

// ---- limiters --------------------------------------------------------------
template <int LIM>
static void test_limiter(int n, const char* name) {
    for (int i = 0; i < n; ++i) {
        double l = urange(-1, 1), r = urange(-1, 1);
        if (i % 5 == 0) r = 2.0 * l * urange(0.9, 1.1);  // near the MC branch switches
        const double dl = urange(-1, 1), dr = urange(-1, 1), sb = urange(-1, 1);
        const dual s = nest_oracle::apply_limiter(dual(l, dl), dual(r, dr), LIM);
        double cl, cr;
        ne::limiter_coef<LIM>(l, r, cl, cr);
        check(name, i, sb * s.d, sb * (cl * dl + cr * dr), std::fabs(sb) * (std::fabs(dl) + std::fabs(dr)) * 2.0);
        const double sv = ne::limiter<LIM>(l, r);
        if (sv != s.v) {
            std::printf("FAIL %s value case %d\n", name, i);
            ++fails;
        }
    }
}

// ---- face flux -------------------------------------------------------------
template <class T>
static void oracle_face(const T* L, const T* R, T* F) {
    // (rho, u_n, p) HLLE plus passive q flux, as one face of the oracle's sweep_cell
    T WL[3] = {L[0], L[1], L[2]}, WR[3] = {R[0], R[1], R[2]};
    nest_oracle::hlle_flux(WL, WR, F, 1.4);
    const T cL = nest_oracle::eos_sound_speed(L[0], L[2], 1.4);
    const T cR = nest_oracle::eos_sound_speed(R[0], R[2], 1.4);
    F[3] = nest_oracle::passive_flux(L[3], R[3], nest_oracle::interface_velocity(L[1], R[1], cL, cR));
}

static void test_face(int n) {
    int branch_count[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        double L[4], R[4];
        const int kind = i % 3;  // 0 subsonic, 1 supersonic to the right, 2 to the left
        const double shift = kind == 0 ? 0.0 : (kind == 1 ? 3.0 : -3.0);
        L[0] = urange(0.1, 2.0), L[1] = urange(-0.8, 0.8) + shift, L[2] = urange(0.1, 2.0), L[3] = urange(-1, 1);
        R[0] = urange(0.1, 2.0), R[1] = urange(-0.8, 0.8) + shift, R[2] = urange(0.1, 2.0), R[3] = urange(-1, 1);
        double dL[4], dR[4], Fb[4];
        for (int k = 0; k < 4; ++k) dL[k] = urange(-1, 1), dR[k] = urange(-1, 1), Fb[k] = urange(-1, 1);
        dual Ld[4], Rd[4], Fd[4];
        for (int k = 0; k < 4; ++k) Ld[k] = dual(L[k], dL[k]), Rd[k] = dual(R[k], dR[k]);
        oracle_face(Ld, Rd, Fd);
        double lhs = 0.0;
        for (int k = 0; k < 4; ++k) lhs += Fb[k] * Fd[k].d;
        // forward value agreement (exact mode is bitwise the reference numerics)
        double Fv[4];
        const ne::prim4 Lp{L[0], L[1], L[2], L[3]}, Rp{R[0], R[1], R[2], R[3]};
        ne::face_flux<true>(Lp, Rp, EOS, Fv);
        for (int k = 0; k < 4; ++k)
            if (Fv[k] != Fd[k].v) {
                std::printf("FAIL face value case %d comp %d: %.17g vs %.17g\n", i, k, Fv[k], Fd[k].v);
                ++fails;
            }
        ne::prim4 Lb = ne::zero4(), Rb = ne::zero4();
        ne::face_flux_rev(Lp, Rp, EOS, Fb, Lb, Rb);
        const double lb[4] = {Lb.r, Lb.u, Lb.p, Lb.q}, rb[4] = {Rb.r, Rb.u, Rb.p, Rb.q};
        double rhs = 0.0, scale = 0.0;
        for (int k = 0; k < 4; ++k) {
            rhs += lb[k] * dL[k] + rb[k] * dR[k];
            scale += std::fabs(lb[k] * dL[k]) + std::fabs(rb[k] * dR[k]);
        }
        check("face", i, lhs, rhs, scale);
        // the clamped-form reverse (what fast numerics run), same contract
        ne::prim4 Lc = ne::zero4(), Rc = ne::zero4();
        ne::face_flux_rev_clamped(Lp, Rp, EOS, Fb, Lc, Rc);
        const double lc[4] = {Lc.r, Lc.u, Lc.p, Lc.q}, rc[4] = {Rc.r, Rc.u, Rc.p, Rc.q};
        double rhs_c = 0.0, scale_c = 0.0;
        for (int k = 0; k < 4; ++k) {
            rhs_c += lc[k] * dL[k] + rc[k] * dR[k];
            scale_c += std::fabs(lc[k] * dL[k]) + std::fabs(rc[k] * dR[k]);
        }
        check("face clamped", i, lhs, rhs_c, scale_c);
        ++branch_count[kind];
    }
    std::printf("face cases: %d subsonic, %d right-upwind, %d left-upwind\n", branch_count[0], branch_count[1],
                branch_count[2]);
}

int main() {
    test_prim(4000);
    test_limiter<ne::LIM_MINMOD>(4000, "minmod");
    test_limiter<ne::LIM_MC>(4000, "mc");
    test_limiter<ne::LIM_NONE>(1000, "none");
    test_face(6000);
    std::printf("%s: worst relative mismatch %.3g, %d failure(s)\n", fails ? "FAIL" : "PASS", worst, fails);
    return fails ? 1 : 0;
}
