// exec::hip — the MI355X backend of nest's Pipeline, over the C ABI of
// include/nest_hip.h (libnest_hip.so).
//
// The reference declares exec::cuda without an executor (core.hpp:435,439)
// and runs pipelines only on exec::cpu / exec::omp (pipeline.hpp:78-151).
// This header adds one more tag, exec::hip_t, and execute_stage overloads
// for it; nothing of the cpu/omp paths changes.
//
// A hip::context mirrors a 2D tensor-product patch mesh (build_mesh_2d /
// build_periodic_mesh_2d, halo 2) in device memory with the identical
// field_t layout, so upload()/download() are plain copies per patch. Between
// them the state lives on the device: stages executed under exec::hip update
// the device mirror, not the host patches.
//
// Stage types understood by exec::hip:
//   hip::cfl_reduce{cfl}  Reduce: dt = compute_dt_2d (euler2d.hpp:242-271),
//                          blocking read-back (kept in the context)
//   anything that is not a hip:: stage, a Compute or a Reduce stage:
//                          Exchange of the current buffer (faces + corners,
//                          periodic / outflow), one device launch
//   hip::rk2_stage<1>     Compute, one launch per patch: U1 = S(U)
//   hip::rk2_stage<2>     Compute, one launch per patch: U = (U + S(U1)) / 2
//   hip::fused_steps{cfl, n}
//                          n whole SSP-RK2 steps in one launch set for all
//                          patches, dt on the device (halos gathered inside
//                          the stage kernels; no host round trip)
// So pipeline(cfl_reduce{}, exchange, rk2_stage<1>{}, exchange,
// rk2_stage<2>{}) is the plain per-patch drop-in of one RK2 step, and
// pipeline(fused_steps{}) the batched form. Generic host Compute/Reduce
// stages have no device form and are rejected at compile time.
#pragma once
#include <cmath>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../nest_hip.h"
#include "pipeline.hpp"

namespace nest {
namespace hip {

enum class numerics { exact = NEST_MODE_EXACT, fast = NEST_MODE_FAST };

inline void check(int rc, const char* what) {
    if (rc != NEST_OK) throw std::runtime_error(std::string("nest_hip: ") + what + ": " + nest_hip_last_error());
}

struct mesh_shape {
    int nx = 0, ny = 0, npx = 0, npy = 0;
};

// Recover (nx, ny, npx, npy) of a tensor-product patch vector and check that
// its partition is the subspace() one the device mirror uses.
inline mesh_shape infer_mesh(const std::vector<patch_t<double, 2>>& patches) {
    mesh_shape m;
    if (patches.empty()) throw std::invalid_argument("nest_hip: empty patch vector");
    std::vector<int> x0, y0;
    int nx = 0, ny = 0;
    for (const auto& p : patches) {
        if (p.halo_width != 2) throw std::invalid_argument("nest_hip: patches need halo_width 2");
        if (std::find(x0.begin(), x0.end(), start(p.interior)[0]) == x0.end()) {
            x0.push_back(start(p.interior)[0]);
            nx += static_cast<int>(shape(p.interior)[0]);
        }
        if (std::find(y0.begin(), y0.end(), start(p.interior)[1]) == y0.end()) {
            y0.push_back(start(p.interior)[1]);
            ny += static_cast<int>(shape(p.interior)[1]);
        }
    }
    m = {nx, ny, static_cast<int>(x0.size()), static_cast<int>(y0.size())};
    if (static_cast<int>(patches.size()) != m.npx * m.npy)
        throw std::invalid_argument("nest_hip: patches do not form a tensor-product mesh");
    auto dom = index_space(ivec(*std::min_element(x0.begin(), x0.end()), *std::min_element(y0.begin(), y0.end())),
                           uvec(static_cast<unsigned>(nx), static_cast<unsigned>(ny)));
    for (int py = 0; py < m.npy; ++py)
        for (int px = 0; px < m.npx; ++px) {
            const auto& p = patches[static_cast<std::size_t>(py * m.npx + px)];
            auto want = subspace(subspace(dom, m.npx, px, 0), m.npy, py, 1);
            if (p.index != py * m.npx + px || !(start(p.interior) == start(want)) ||
                !(shape(p.interior) == shape(want)))
                throw std::invalid_argument("nest_hip: patch partition differs from subspace() order");
        }
    return m;
}

class context {
  public:
    context(const std::vector<patch_t<double, 2>>& patches, double dx, double dy, double gamma = 1.4,
            boundary_t bcx = boundary_t::periodic, boundary_t bcy = boundary_t::periodic, int limiter = 0,
            numerics mode = numerics::exact)
        : shape_(infer_mesh(patches)) {
        auto bc = [](boundary_t b) {
            return b == boundary_t::outflow ? NEST_BC_OUTFLOW
                                            : (b == boundary_t::reflecting ? NEST_BC_REFLECTING : NEST_BC_PERIODIC);
        };
        check(nest_hip_mesh2d_create(shape_.nx, shape_.ny, shape_.npx, shape_.npy, dx, dy, gamma, bc(bcx), bc(bcy),
                                     limiter, static_cast<int>(mode), &sim_),
              "mesh2d_create");
    }
    ~context() { nest_hip_sim2d_destroy(sim_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    void upload(const std::vector<patch_t<double, 2>>& patches) {
        for (const auto& p : patches) check(nest_hip_sim2d_upload_patch(sim_, p.index, p.state.conserved._data), "upload");
        current_ = 0;
    }
    void download(std::vector<patch_t<double, 2>>& patches) {
        check(nest_hip_sim2d_synchronize(sim_), "synchronize");
        for (auto& p : patches) check(nest_hip_sim2d_download_patch(sim_, p.index, p.state.conserved._data), "download");
    }

    // stage implementations (called by execute_stage below)
    double reduce_dt(double cfl) {
        check(nest_hip_sim2d_compute_dt(sim_, cfl, &dt_), "compute_dt");
        return dt_;
    }
    void exchange() { check(nest_hip_sim2d_exchange(sim_, current_), "exchange"); }
    void stage(int s, int patch) {
        if (!(dt_ > 0.0)) throw std::logic_error("nest_hip: rk2_stage before a cfl_reduce");
        check(nest_hip_sim2d_stage(sim_, s, dt_, patch), "stage");
        current_ = s == 1 ? 1 : 0;
    }
    void fused(double cfl, int nsteps, bool graph) {
        if (!started_) {
            check(nest_hip_sim2d_start(sim_, cfl, INFINITY, 0.0), "start");
            started_ = true;
        }
        check(nest_hip_sim2d_advance(sim_, nsteps, graph ? 1 : 0), "advance");
        current_ = 0;
    }
    double dt() const { return dt_; }
    nest_sim2d* handle() { return sim_; }
    const mesh_shape& shape() const { return shape_; }

  private:
    nest_sim2d* sim_ = nullptr;
    mesh_shape shape_;
    double dt_ = 0.0;
    int current_ = 0;  // buffer holding the newest state: 0 = U, 1 = U1
    bool started_ = false;
};

struct cfl_reduce {
    double cfl = 0.4;
};
template <int S>
struct rk2_stage {
    static_assert(S == 1 || S == 2, "SSP-RK2 has stages 1 and 2");
};
struct fused_steps {
    double cfl = 0.4;
    int nsteps = 1;
    bool graph = true;
};

template <typename S> struct is_rk2_stage : std::false_type {};
template <int S> struct is_rk2_stage<rk2_stage<S>> : std::true_type {
    static constexpr int stage = S;
};
template <typename> inline constexpr bool always_false = false;

}  // namespace hip

namespace exec {
struct hip_t {
    hip::context* ctx = nullptr;
};
inline hip_t hip(hip::context& c) { return hip_t{&c}; }
}  // namespace exec

template <typename T, std::size_t Dim, typename Stage>
void execute_stage(Stage& stage, std::vector<patch_t<T, Dim>>& patches, exec::hip_t ex) {
    static_assert(std::is_same_v<T, double> && Dim == 2, "exec::hip runs 2D double patch meshes");
    using P = patch_t<T, Dim>;
    if (!ex.ctx) throw std::invalid_argument("exec::hip without a context");
    if constexpr (std::is_same_v<Stage, hip::cfl_reduce>) {
        ex.ctx->reduce_dt(stage.cfl);
    } else if constexpr (hip::is_rk2_stage<Stage>::value) {
        for (const auto& p : patches) ex.ctx->stage(hip::is_rk2_stage<Stage>::stage, p.index);
    } else if constexpr (std::is_same_v<Stage, hip::fused_steps>) {
        ex.ctx->fused(stage.cfl, stage.nsteps, stage.graph);
    } else if constexpr (detail::is_compute_stage<Stage, P>::value || detail::is_reduce_stage<Stage, P>::value) {
        static_assert(hip::always_false<Stage>,
                      "exec::hip runs nest::hip stage types; a host Compute/Reduce stage has no device form");
    } else {
        ex.ctx->exchange();
    }
}

}  // namespace nest
