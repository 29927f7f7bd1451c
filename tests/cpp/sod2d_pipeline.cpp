// GPU driver test of exec::hip (include/nest/backend_hip.hpp): a 2D Sod run
// written the way the reference's examples/sod2d/sod2d.cpp drives patches,
// but through nest::pipeline on the MI355X backend.
//   1. per-patch pipeline: cfl_reduce, exchange, rk2_stage<1>, exchange,
//      rk2_stage<2> (one launch per patch per stage), K steps;
//   2. fused pipeline: fused_steps{K} (one launch set, dt on the device);
//   3. the per-patch pipeline with the dt from a user's host Reduce stage
//      of the reference's shape (value_type / init / reduce / finalize, fold
//      by min), run on host copies of the patches, its result becoming the
//      context's dt for the device stages;
//   4. the oracle (CPU restatement of the reference numerics), K steps.
// All three must agree bit for bit (exact numerics). Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../oracle/euler_oracle.hpp"
#include "nest/backend_hip.hpp"

using namespace nest;

static std::vector<patch_t<double, 2>> make_mesh(int nx, int ny, int npx, int npy) {
    auto patches = build_mesh_2d<double>(index_space(ivec(0, 0), uvec(nx, ny)), npx, npy, 4, 2, boundary_t::outflow,
                                         boundary_t::periodic);
    const double dx = 1.0 / nx, gamma = 1.4;
    for (auto& p : patches)
        for (auto idx : p.interior) {  // planar Sod (sod2d.cpp:266-290) with a deterministic density ripple
            double x = (idx[0] + 0.5) * dx;
            double rho = x < 0.5 ? 1.0 : 0.125, pr = x < 0.5 ? 1.0 : 0.1;
            rho += 1e-3 * std::sin(0.37 * idx[0] + 1.3 * idx[1]);
            double u = 0.0, v = 0.0;
            p(0, idx) = rho;
            p(1, idx) = rho * u;
            p(2, idx) = rho * v;
            p(3, idx) = pr / (gamma - 1.0) + 0.5 * rho * (u * u + v * v);
        }
    fill_all_halos(patches);
    return patches;
}

// compute_dt_2d (euler2d.hpp:242-271) as a user's Reduce stage: per patch
// cfl*min(min dx/(|u|+c), min dy/(|v|+c)), folded by min (bitwise the
// whole-mesh dt: multiplication by cfl > 0 is monotone).
struct user_dt {
    using value_type = double;
    double cfl, dx, dy;
    static double init() { return 1e10; }
    double reduce(double acc, const patch_t<double, 2>& p) const {
        double dtx = 1e10, dty = 1e10;
        nest_oracle::dt_2d_partial(p.state.conserved._data, static_cast<int>(shape(p.interior)[0]),
                                   static_cast<int>(shape(p.interior)[1]), 2, dx, dy, 1.4, dtx, dty);
        return std::min(acc, cfl * std::min(dtx, dty));
    }
    void finalize(double acc, patch_t<double, 2>& p) const { p.state.time = acc; }
};

static std::vector<double> flatten(const std::vector<patch_t<double, 2>>& patches) {
    std::vector<double> out;
    for (const auto& p : patches)
        out.insert(out.end(), p.state.conserved._data, p.state.conserved._data + 4 * size(p.extended));
    return out;
}

static bool same_interiors(const std::vector<patch_t<double, 2>>& a, const std::vector<patch_t<double, 2>>& b) {
    for (std::size_t q = 0; q < a.size(); ++q)
        for (std::size_t v = 0; v < 4; ++v)
            for (auto i : a[q].interior)
                if (a[q](v, i) != b[q](v, i)) {
                    std::printf("  patch %zu var %zu cell (%d,%d): %.17g vs %.17g\n", q, v, i[0], i[1], a[q](v, i),
                                b[q](v, i));
                    return false;
                }
    return true;
}

int main() {
    const int nx = 120, ny = 88, npx = 3, npy = 2, K = 12;
    const double dx = 1.0 / nx, dy = 1.0 / ny, cfl = 0.4;
    bool ok = true;

    // 1. per-patch pipeline under exec::hip
    auto A = make_mesh(nx, ny, npx, npy);
    hip::context ca(A, dx, dy, 1.4, boundary_t::outflow, boundary_t::periodic);
    ca.upload(A);
    auto step = pipeline(hip::cfl_reduce{cfl}, exchange_t{}, hip::rk2_stage<1>{}, exchange_t{},
                         hip::rk2_stage<2>{});
    std::vector<double> dts;
    for (int k = 0; k < K; ++k) {
        execute(step, A, exec::hip(ca));
        dts.push_back(ca.dt());
    }
    auto ex = pipeline(exchange_t{});
    execute(ex, A, exec::hip(ca));
    ca.download(A);

    // 2. fused pipeline under exec::hip
    auto B = make_mesh(nx, ny, npx, npy);
    hip::context cb(B, dx, dy, 1.4, boundary_t::outflow, boundary_t::periodic);
    cb.upload(B);
    auto fused = pipeline(hip::fused_steps{cfl, K, true});
    execute(fused, B, exec::hip(cb));
    cb.download(B);

    // 3. per-patch pipeline with a host Reduce stage for dt
    auto D = make_mesh(nx, ny, npx, npy);
    hip::context cd(D, dx, dy, 1.4, boundary_t::outflow, boundary_t::periodic);
    cd.upload(D);
    auto step_d = pipeline(user_dt{cfl, dx, dy}, exchange_t{}, hip::rk2_stage<1>{}, exchange_t{},
                           hip::rk2_stage<2>{});
    for (int k = 0; k < K; ++k) {
        execute(step_d, D, exec::hip(cd));
        if (cd.dt() != dts[k] || D.back().state.time != dts[k]) {
            std::printf("FAIL: host-reduce dt of step %d: %.17g vs device %.17g\n", k, cd.dt(), dts[k]);
            ok = false;
        }
    }
    execute(ex, D, exec::hip(cd));
    cd.download(D);

    // 4. oracle
    auto C = make_mesh(nx, ny, npx, npy);
    nest_oracle::mesh2d m(nx, ny, npx, npy, 2, nest_oracle::BC_OUTFLOW, nest_oracle::BC_PERIODIC);
    std::vector<double> U = flatten(C), U1, s;
    for (int k = 0; k < K; ++k) {
        double dt = nest_oracle::dt_mesh(U.data(), m, dx, dy, cfl, 1.4);
        if (dt != dts[k]) {
            std::printf("FAIL: dt of step %d: device %.17g oracle %.17g\n", k, dts[k], dt);
            ok = false;
        }
        nest_oracle::rk2_step_mesh(U.data(), m, dt, dx, dy, 1.4, 0, U1, s);
    }
    for (auto& p : C) std::copy(U.begin() + m.off[p.index], U.begin() + m.off[p.index] + 4 * size(p.extended),
                                p.state.conserved._data);

    if (!same_interiors(A, C)) {
        std::printf("FAIL: per-patch pipeline differs from the oracle\n");
        ok = false;
    }
    if (!same_interiors(D, C)) {
        std::printf("FAIL: host-reduce pipeline differs from the oracle\n");
        ok = false;
    }
    if (!same_interiors(B, C)) {
        std::printf("FAIL: fused pipeline differs from the oracle\n");
        ok = false;
    }
    std::printf("%s: %d steps, %zu patches, last dt %.17g\n", ok ? "PASS" : "FAIL", K, A.size(), dts.back());
    return ok ? 0 : 1;
}
