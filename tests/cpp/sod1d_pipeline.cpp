// GPU driver test of exec::hip on 1D patch vectors (include/nest/
// backend_hip.hpp, hip::context<1>): a 1D Sod run written the way the
// reference's examples/sod1d/sod1d.cpp drives its patches (dt reduction,
// exchange, stage, exchange, stage; sod1d.cpp:219-272), through
// nest::pipeline on the MI355X backend.
//   1. per-stage pipeline: hip::cfl_reduce, exchange, rk2_stage<1>,
//      exchange, rk2_stage<2> (device launches), K steps;
//   2. the same with the dt from a user's host Reduce stage of the
//      reference's shape (value_type / init / reduce / finalize), whose
//      fold is a minimum: the stage runs on the host copies and its result
//      becomes the context's dt;
//   3. fused pipeline: fused_steps{K} (one launch per step, dt on the device);
//   4. the oracle (CPU restatement of the reference numerics), K steps.
// Cases: 4 periodic patches (build_periodic_mesh_1d, one device array over
// the whole domain) and one outflow patch. All must agree bit for bit
// (exact numerics), dt included. Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../oracle/euler_oracle.hpp"
#include "nest/backend_hip.hpp"

using namespace nest;

static constexpr double kGamma = 1.4;

// Sod tube over [0, 1] (sod1d.cpp:160-190), interiors only
static void sod(std::vector<patch_t<double, 1>>& patches, int n) {
    const double dx = 1.0 / n;
    for (auto& p : patches)
        for (auto i : p.interior) {
            const double x = (i[0] + 0.5) * dx;
            const double rho = x < 0.5 ? 1.0 : 0.125, pr = x < 0.5 ? 1.0 : 0.1;
            p(0, i) = rho;
            p(1, i) = 0.0;
            p(2, i) = pr / (kGamma - 1.0);
        }
}

static std::vector<patch_t<double, 1>> make(int n, int npatch) {
    std::vector<patch_t<double, 1>> v;
    if (npatch > 1) {
        v = build_periodic_mesh_1d<double>(index_space(ivec(0), uvec(static_cast<unsigned>(n))), npatch, 3, 2);
    } else {
        v.emplace_back(0, index_space(ivec(0), uvec(static_cast<unsigned>(n))), 3, 2);
    }
    sod(v, n);
    return v;
}

// A user's dt reduction of the reference's Reduce shape: per patch
// cfl*dx / max(|u|+c) over its interior, folded by min (bitwise the dt of
// the whole domain: correctly rounded division is monotone). finalize
// records it as the patch's time stamp.
struct user_dt {
    using value_type = double;
    double cfl, dx;
    static double init() { return 1e10; }
    double reduce(double acc, const patch_t<double, 1>& p) const {
        double smax = 0.0;
        for (auto i : p.interior) {
            const double U[3] = {p(0, i), p(1, i), p(2, i)};
            double W[3];
            nest_oracle::cons_to_prim(U, W, kGamma);
            smax = std::max(smax, std::fabs(W[1]) + nest_oracle::eos_sound_speed(W[0], W[2], kGamma));
        }
        return std::min(acc, cfl * dx / smax);
    }
    void finalize(double acc, patch_t<double, 1>& p) const { p.state.time = acc; }
};

static bool same(const std::vector<patch_t<double, 1>>& a, const std::vector<double>& U, int n, const char* what) {
    const int ne = n + 4;
    for (const auto& p : a)
        for (std::size_t v = 0; v < 3; ++v)
            for (auto i : p.interior)
                if (p(v, i) != U[v * ne + i[0] + 2]) {
                    std::printf("FAIL: %s: var %zu cell %d: %.17g vs %.17g\n", what, v, i[0], p(v, i),
                                U[v * ne + i[0] + 2]);
                    return false;
                }
    return true;
}

static bool run_case(int n, int npatch, boundary_t bc, int K) {
    const double dx = 1.0 / n, cfl = 0.4;
    const int obc = bc == boundary_t::periodic ? nest_oracle::BC_PERIODIC : nest_oracle::BC_OUTFLOW;
    bool ok = true;

    // 4. oracle
    std::vector<double> U(3 * static_cast<std::size_t>(n + 4), 0.0), U1, U2, W;
    {
        auto P = make(n, npatch);
        for (const auto& p : P)
            for (std::size_t v = 0; v < 3; ++v)
                for (auto i : p.interior) U[v * (n + 4) + i[0] + 2] = p(v, i);
    }
    std::vector<double> dts;
    for (int k = 0; k < K; ++k) {
        nest_oracle::fill_1d(U.data(), n, 2, 3, obc);
        const double dt = cfl * dx / nest_oracle::max_speed_1d(U.data(), n, 2, kGamma);
        dts.push_back(dt);
        nest_oracle::rk2_step_1d(U.data(), n, 2, obc, dt, dx, kGamma, 0, U1, U2, W);
    }

    // 1. per-stage pipeline, device dt
    auto A = make(n, npatch);
    hip::context ca(A, dx, kGamma, bc);
    ca.upload(A);
    auto step = pipeline(hip::cfl_reduce{cfl}, exchange_t{}, hip::rk2_stage<1>{}, exchange_t{}, hip::rk2_stage<2>{});
    for (int k = 0; k < K; ++k) {
        execute(step, A, exec::hip(ca));
        if (ca.dt() != dts[k]) {
            std::printf("FAIL: device dt of step %d: %.17g vs oracle %.17g\n", k, ca.dt(), dts[k]);
            ok = false;
        }
    }
    ca.download(A);
    ok = same(A, U, n, "per-stage pipeline") && ok;

    // 2. per-stage pipeline, dt from the user's host Reduce stage
    auto B = make(n, npatch);
    hip::context cb(B, dx, kGamma, bc);
    cb.upload(B);
    auto step_b = pipeline(user_dt{cfl, dx}, exchange_t{}, hip::rk2_stage<1>{}, exchange_t{}, hip::rk2_stage<2>{});
    for (int k = 0; k < K; ++k) {
        execute(step_b, B, exec::hip(cb));
        if (cb.dt() != dts[k] || B.front().state.time != dts[k]) {
            std::printf("FAIL: host-reduce dt of step %d: %.17g vs oracle %.17g\n", k, cb.dt(), dts[k]);
            ok = false;
        }
    }
    cb.download(B);
    ok = same(B, U, n, "host-reduce pipeline") && ok;

    // 3. fused steps
    auto C = make(n, npatch);
    hip::context cc(C, dx, kGamma, bc);
    cc.upload(C);
    auto fused = pipeline(hip::fused_steps{cfl, K / 2}, hip::fused_steps{cfl, K - K / 2});
    execute(fused, C, exec::hip(cc));
    cc.download(C);
    ok = same(C, U, n, "fused pipeline") && ok;
    // halo of the handed-out patches: what an exchange leaves there
    auto D = make(n, npatch);
    for (std::size_t q = 0; q < C.size(); ++q)
        for (std::size_t v = 0; v < 3; ++v)
            for (auto i : C[q].interior) D[q](v, i) = C[q](v, i);
    fill_all_halos(D);
    for (std::size_t q = 0; q < C.size() && npatch > 1; ++q)
        for (std::size_t v = 0; v < 3; ++v)
            for (auto i : C[q].extended)
                if (C[q](v, i) != D[q](v, i)) {
                    std::printf("FAIL: patch %zu halo cell %d\n", q, i[0]);
                    ok = false;
                }
    std::printf("%s: n=%d patches=%d bc=%s K=%d last dt %.17g\n", ok ? "ok" : "FAIL", n, npatch,
                bc == boundary_t::periodic ? "periodic" : "outflow", K, dts.back());
    return ok;
}

int main() {
    bool ok = true;
    ok = run_case(1000, 4, boundary_t::periodic, 25) && ok;
    ok = run_case(1000, 1, boundary_t::outflow, 25) && ok;
    ok = run_case(257, 3, boundary_t::periodic, 9) && ok;
    std::printf("%s\n", ok ? "PASS" : "FAIL");
    return ok ? 0 : 1;
}
