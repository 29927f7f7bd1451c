// exec::hip — the MI355X backend of nest's Pipeline, over the C ABI of
// include/nest_hip.h (libnest_hip.so).
//
// The reference declares exec::cuda without an executor (core.hpp:435,439)
// and runs pipelines only on exec::cpu / exec::omp (pipeline.hpp:78-151).
// This header adds one more tag, exec::hip_t<Dim>, and its execute_stage;
// nothing of the cpu/omp paths changes.
//
// A hip::context<Dim> mirrors a patch vector in device memory in the
// identical field_t layout; upload() / download() move the state between
// the host patches and the device. Between them the state lives on the
// device: device stages update the device mirror, not the host patches.
//   hip::context<2>  a 2D tensor-product patch mesh (build_mesh_2d /
//                    build_periodic_mesh_2d order, halo 2, 4 variables)
//   hip::context<1>  a 1D patch vector tiling one contiguous domain
//                    (build_periodic_mesh_1d order, or one patch with an
//                    outflow / reflecting boundary), halo 2, 3 variables;
//                    the device runs it as one array over the whole domain
//                    (the full 5-point stencil everywhere: the reference's
//                    OpenMP/CUDA 1D form, benchmark_euler1d.cpp:113-165)
//
// Stages under exec::hip:
//   hip::cfl_reduce{cfl}   Reduce on the device: dt = compute_dt_2d
//                          (euler2d.hpp:242-271) / the 1D compute_dt
//                          (sod1d.cpp:97-128), blocking read-back, kept in
//                          the context for the rk2 stages that follow
//   hip::rk2_stage<1|2>    Compute on the device, one launch per patch (2D)
//                          or per domain (1D): U1 = S(U); U = (U + S(U1)) / 2
//   hip::fused_steps{cfl, n}  n whole SSP-RK2 steps, dt on the device, no
//                          host round trip (2D: halos gathered inside the
//                          stage kernels; 1D: one launch per step)
//   a host Reduce stage    (reference contract, pipeline.hpp:24-54): the
//                          device state is downloaded into the patches, the
//                          stage folds them on the host (exec::omp's fold:
//                          partials combined by min, pipeline.hpp:116) and
//                          finalizes every patch, the patches are uploaded
//                          again; a result convertible to double becomes the
//                          context's dt, so a user's own dt reduction drives
//                          the device rk2 stages
//   a host Compute stage   the same round trip around value(patch)
//   anything else          Exchange of the buffer holding the newest state
//                          (faces + corners; periodic / outflow /
//                          reflecting), on the device
// So pipeline(cfl_reduce{}, exchange, rk2_stage<1>{}, exchange,
// rk2_stage<2>{}) is the plain per-patch drop-in of one RK2 step, and
// pipeline(fused_steps{}) the batched form. Host stages move the whole state
// across PCIe twice; they serve user code between steps, not the step.
#pragma once
#include <algorithm>
#include <cmath>
#include <concepts>
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

inline int bc_code(boundary_t b) {
    return b == boundary_t::outflow ? NEST_BC_OUTFLOW
                                    : (b == boundary_t::reflecting ? NEST_BC_REFLECTING : NEST_BC_PERIODIC);
}

// ---- device-only stages ----------------------------------------------------
struct cfl_reduce {
    static constexpr bool device_only = true;
    double cfl = 0.4;
};
template <int S>
struct rk2_stage {
    static_assert(S == 1 || S == 2, "SSP-RK2 has stages 1 and 2");
    static constexpr bool device_only = true;
    static constexpr int stage = S;
};
struct fused_steps {
    static constexpr bool device_only = true;
    double cfl = 0.4;
    int nsteps = 1;
    bool graph = true;  // 2D: replay the captured step graph
};

template <class S>
concept rk2_stage_type = requires {
    requires S::device_only;
    { S::stage } -> std::convertible_to<int>;
};

struct mesh_shape {
    int nx = 0, ny = 0, npx = 0, npy = 0;
};

// Recover (nx, ny, npx, npy) of a tensor-product patch vector and check that
// its partition is the subspace() one the device mirror uses.
inline mesh_shape infer_mesh(const std::vector<patch_t<double, 2>>& patches) {
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
    mesh_shape m{nx, ny, static_cast<int>(x0.size()), static_cast<int>(y0.size())};
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

template <std::size_t Dim>
class context;

// ---- 2D: one mesh handle -----------------------------------------------------
template <>
class context<2> {
  public:
    context(const std::vector<patch_t<double, 2>>& patches, double dx, double dy, double gamma = 1.4,
            boundary_t bcx = boundary_t::periodic, boundary_t bcy = boundary_t::periodic, int limiter = 0,
            numerics mode = numerics::exact)
        : shape_(infer_mesh(patches)) {
        check(nest_hip_mesh2d_create(shape_.nx, shape_.ny, shape_.npx, shape_.npy, dx, dy, gamma, bc_code(bcx),
                                     bc_code(bcy), limiter, static_cast<int>(mode), &sim_),
              "mesh2d_create");
    }
    ~context() { nest_hip_sim2d_destroy(sim_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    void upload(const std::vector<patch_t<double, 2>>& patches) {
        for (const auto& p : patches)
            check(nest_hip_sim2d_upload_patch(sim_, p.index, p.state.conserved._data), "upload");
        current_ = 0;
        started_ = false;  // device dt state belongs to the replaced state
    }
    void download(std::vector<patch_t<double, 2>>& patches) {
        check(nest_hip_sim2d_synchronize(sim_), "synchronize");
        for (auto& p : patches)
            check(nest_hip_sim2d_download_patch(sim_, p.index, p.state.conserved._data), "download");
    }

    double reduce_dt(double cfl) {
        check(nest_hip_sim2d_compute_dt(sim_, cfl, &dt_), "compute_dt");
        return dt_;
    }
    void exchange() { check(nest_hip_sim2d_exchange(sim_, current_), "exchange"); }
    void stage(int s, const std::vector<patch_t<double, 2>>& patches) {
        if (!(dt_ > 0.0)) throw std::logic_error("nest_hip: rk2_stage before a dt reduction");
        for (const auto& p : patches) check(nest_hip_sim2d_stage(sim_, s, dt_, p.index), "stage");
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
    // the newest state must be a whole step's (U) for a host stage to see it
    void require_step_state() const {
        if (current_ != 0) throw std::logic_error("nest_hip: host stage between rk2_stage<1> and rk2_stage<2>");
    }
    void set_dt(double dt) { dt_ = dt; }
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

// ---- 1D: one solver handle over the whole domain ------------------------------
template <>
class context<1> {
  public:
    context(const std::vector<patch_t<double, 1>>& patches, double dx, double gamma = 1.4,
            boundary_t bc = boundary_t::periodic, int limiter = 0, numerics mode = numerics::exact) {
        if (patches.empty()) throw std::invalid_argument("nest_hip: empty patch vector");
        x0_ = start(patches.front().interior)[0];
        int next = x0_;
        for (std::size_t q = 0; q < patches.size(); ++q) {
            const auto& p = patches[q];
            if (p.halo_width != 2 || p.nvars() != 3)
                throw std::invalid_argument("nest_hip: 1D patches need halo_width 2 and 3 variables");
            if (p.index != static_cast<int>(q) || start(p.interior)[0] != next)
                throw std::invalid_argument("nest_hip: 1D patches must tile one domain in index order");
            next += static_cast<int>(shape(p.interior)[0]);
        }
        n_ = next - x0_;
        if (patches.size() > 1 && bc != boundary_t::periodic)
            throw std::invalid_argument("nest_hip: several 1D patches exchange periodically (build_periodic_mesh_1d)");
        check(nest_hip_solver1d_create(n_, dx, gamma, bc_code(bc), limiter, static_cast<int>(mode), &s_),
              "solver1d_create");
        buf_.assign(3 * static_cast<std::size_t>(n_ + 4), 0.0);
    }
    ~context() { nest_hip_solver1d_destroy(s_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;

    // interiors into one domain array (ghosts are set on the device)
    void upload(const std::vector<patch_t<double, 1>>& patches) {
        const std::size_t ne = static_cast<std::size_t>(n_ + 4);
        for (const auto& p : patches)
            for (std::size_t v = 0; v < 3; ++v)
                for (auto i : p.interior) buf_[v * ne + static_cast<std::size_t>(i[0] - x0_ + 2)] = p(v, i);
        check(nest_hip_solver1d_upload(s_, buf_.data()), "upload");
        check(nest_hip_solver1d_exchange(s_, 0), "exchange");
        current_ = 0;
        started_ = false;
    }
    // every patch's extended cells from the halo-complete domain array: the
    // values an exchange would have left in its halo
    void download(std::vector<patch_t<double, 1>>& patches) {
        check(nest_hip_solver1d_exchange(s_, 0), "exchange");
        check(nest_hip_solver1d_download(s_, buf_.data()), "download");
        const std::size_t ne = static_cast<std::size_t>(n_ + 4);
        for (auto& p : patches)
            for (std::size_t v = 0; v < 3; ++v)
                for (auto i : p.extended) p(v, i) = buf_[v * ne + static_cast<std::size_t>(i[0] - x0_ + 2)];
    }

    double reduce_dt(double cfl) {
        check(nest_hip_solver1d_compute_dt(s_, cfl, &dt_), "compute_dt");
        return dt_;
    }
    void exchange() { check(nest_hip_solver1d_exchange(s_, current_), "exchange"); }
    void stage(int s, const std::vector<patch_t<double, 1>>&) {
        if (!(dt_ > 0.0)) throw std::logic_error("nest_hip: rk2_stage before a dt reduction");
        check(nest_hip_solver1d_stage(s_, s, dt_), "stage");
        current_ = s == 1 ? 1 : 0;
    }
    void fused(double cfl, int nsteps, bool) {
        if (!started_) {
            check(nest_hip_solver1d_start(s_, cfl, INFINITY), "start");
            started_ = true;
        }
        check(nest_hip_solver1d_advance(s_, nsteps), "advance");
        current_ = 0;
    }
    void require_step_state() const {
        if (current_ != 0) throw std::logic_error("nest_hip: host stage between rk2_stage<1> and rk2_stage<2>");
    }
    void set_dt(double dt) { dt_ = dt; }
    double dt() const { return dt_; }
    int cells() const { return n_; }
    nest_solver1d* handle() { return s_; }

  private:
    nest_solver1d* s_ = nullptr;
    int x0_ = 0, n_ = 0;
    std::vector<double> buf_;
    double dt_ = 0.0;
    int current_ = 0;
    bool started_ = false;
};

// hip::context ctx(patches, ...) is the context of the patches' dimension
template <std::size_t Dim, class... Args>
context(const std::vector<patch_t<double, Dim>>&, Args...) -> context<Dim>;

// A user's Compute or Reduce stage under exec::hip: round trip of the state
// through the host patches (see the header comment).
template <std::size_t Dim, class Stage>
void run_host_stage(context<Dim>& c, Stage& stage, std::vector<patch_t<double, Dim>>& patches) {
    using P = patch_t<double, Dim>;
    using loop = host_loop<exec::omp_t>;
    c.require_step_state();
    c.download(patches);
    const std::size_t n = patches.size();
    if constexpr (compute_stage<Stage, P>) {
        loop::each(n, [&](std::size_t i) { patches[i] = stage.value(std::move(patches[i])); });
    } else {
        using V = typename Stage::value_type;
        const V acc =
            loop::fold(n, V(Stage::init()), [&](const V& a, std::size_t i) { return V(stage.reduce(a, patches[i])); });
        loop::each(n, [&](std::size_t i) { stage.finalize(acc, patches[i]); });
        if constexpr (std::convertible_to<V, double>) c.set_dt(static_cast<double>(acc));
    }
    c.upload(patches);
}

}  // namespace hip

namespace exec {
template <std::size_t Dim>
struct hip_t {
    hip::context<Dim>* ctx = nullptr;
};
template <std::size_t Dim>
hip_t<Dim> hip(hip::context<Dim>& c) {
    return hip_t<Dim>{&c};
}
}  // namespace exec

template <typename T, std::size_t Dim, typename Stage>
void execute_stage(Stage& stage, std::vector<patch_t<T, Dim>>& patches, exec::hip_t<Dim> ex) {
    static_assert(std::is_same_v<T, double>, "exec::hip runs double patches");
    using P = patch_t<T, Dim>;
    if (!ex.ctx) throw std::invalid_argument("exec::hip without a context");
    auto& c = *ex.ctx;
    if constexpr (std::is_same_v<Stage, hip::cfl_reduce>) {
        c.reduce_dt(stage.cfl);
    } else if constexpr (hip::rk2_stage_type<Stage>) {
        c.stage(Stage::stage, patches);
    } else if constexpr (std::is_same_v<Stage, hip::fused_steps>) {
        c.fused(stage.cfl, stage.nsteps, stage.graph);
    } else if constexpr (compute_stage<Stage, P> || reduce_stage<Stage, P>) {
        hip::run_host_stage(c, stage, patches);
    } else {
        static_assert(!device_only_stage<Stage>, "unknown device-only stage");
        c.exchange();
    }
}

}  // namespace nest
