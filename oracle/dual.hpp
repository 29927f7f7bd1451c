// TEST INFRASTRUCTURE ONLY — forward-mode dual numbers for the oracle.
//
// dual = v + d*eps with eps^2 = 0. Instantiating the oracle numerics
// (euler_oracle.hpp) on this type gives the exact directional derivative
// (Jacobian-vector product) of the discrete RK2 step, which pins the hand-
// written adjoint kernels: <adjoint(w), v> must equal <w, jvp(v)>.
// Branches (limiter cases, HLLE wave-speed cases, floors) are taken on the
// value part, so the derivative is that of the branch the primal takes.
#pragma once
#include <cmath>

namespace nest_oracle {

struct dual {
    double v = 0.0, d = 0.0;
    dual() = default;
    dual(double x) : v(x), d(0.0) {}
    dual(double x, double dx) : v(x), d(dx) {}
};

inline double val(const dual& x) { return x.v; }
inline dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
inline dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
inline dual operator-(dual a) { return {-a.v, -a.d}; }
inline dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline dual operator/(dual a, dual b) {
    double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
inline dual operator+(double a, dual b) { return dual(a) + b; }
inline dual operator+(dual a, double b) { return a + dual(b); }
inline dual operator-(double a, dual b) { return dual(a) - b; }
inline dual operator-(dual a, double b) { return a - dual(b); }
inline dual operator*(double a, dual b) { return {a * b.v, a * b.d}; }
inline dual operator*(dual a, double b) { return {a.v * b, a.d * b}; }
inline dual operator/(dual a, double b) { return {a.v / b, a.d / b}; }
inline dual operator/(double a, dual b) { return dual(a) / b; }
inline dual nsqrt(dual a) {
    double s = std::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
inline dual nabs(dual a) { return a.v < 0.0 ? -a : a; }

}  // namespace nest_oracle
