// nest pipelines: a fixed sequence of stages run over a vector of patches.
//
// Public interface of the reference's include/nest/pipeline.hpp: pipeline()
// building a pipeline_t, execute(pipeline, patches[, exec]) and one
// execute_stage overload per execution tag (reference pipeline.hpp:62-151).
// The implementation is this library's own.
//
// The stage contract (reference pipeline.hpp:24-54) is written as C++20
// concepts over a stage type S and the patch type P:
//   compute_stage  s.value(P) -> P         the patch is moved in and the
//                                          result moved back in its place
//   reduce_stage   S::value_type, S::init(),
//                  s.reduce(acc, const P&) -> value_type   folds one patch,
//                  s.finalize(acc, P&)                     hands the result
//                                                          to every patch
//   anything else  an exchange: every patch's halo is filled from its
//                  exchange plan (the ghost cells of a patch are written from
//                  other patches' interiors only, so patches may be filled in
//                  any order or in parallel).
// A stage that is both compute- and reduce-shaped runs as compute, as in the
// reference.
//
// The host executors differ only in how a loop over independent patches
// runs and how a reduction's per-patch fold is split up; that policy is
// host_loop<E> below, and one execute_stage serves exec::cpu and exec::omp
// through it. Under exec::omp each thread folds its share of the patches
// from S::init() and the partial results are combined with min, which is
// the reference's reduction(min:acc) (pipeline.hpp:116): a Reduce stage
// whose fold is a minimum gets the same answer from either executor.
//
// Stages that only exist on the device (nest::hip::*, backend_hip.hpp)
// declare `static constexpr bool device_only = true` and are refused by the
// host executors at compile time instead of being taken for exchanges.
#pragma once
#include <algorithm>
#include <concepts>
#include <cstddef>
#include <tuple>
#include <utility>
#include <vector>

#include "patch.hpp"

namespace nest {

template <class S, class P>
concept compute_stage = requires(const S& s, P p) {
    { s.value(std::move(p)) } -> std::convertible_to<P>;
};

template <class S, class P>
concept reduce_stage = requires(const S& s, typename S::value_type acc, const P& in, P& out) {
    { S::init() } -> std::convertible_to<typename S::value_type>;
    { s.reduce(acc, in) } -> std::convertible_to<typename S::value_type>;
    s.finalize(acc, out);
};

template <class S>
concept device_only_stage = requires {
    requires S::device_only;
};

// An explicit exchange stage (any stage that is neither compute nor reduce
// exchanges halos; this one names the intent).
struct exchange_t {};

template <typename... Stages>
struct pipeline_t {
    std::tuple<Stages...> stages;
    explicit pipeline_t(Stages... s) : stages(std::move(s)...) {}
    static constexpr std::size_t size() { return sizeof...(Stages); }
};

template <typename... Stages>
pipeline_t<Stages...> pipeline(Stages... s) {
    return pipeline_t<Stages...>(std::move(s)...);
}

// How a host executor runs a loop over n independent patches (each) and a
// fold over them (fold: acc = step(acc, i) for every i, starting from init).
template <class E>
struct host_loop;

template <>
struct host_loop<exec::cpu_t> {
    template <class F>
    static void each(std::size_t n, F&& f) {
        for (std::size_t i = 0; i < n; ++i) f(i);
    }
    template <class V, class F>
    static V fold(std::size_t n, V init, F&& step) {
        for (std::size_t i = 0; i < n; ++i) init = step(init, i);
        return init;
    }
};

template <>
struct host_loop<exec::omp_t> {
    template <class F>
    static void each(std::size_t n, F&& f) {
        const long m = static_cast<long>(n);
#pragma omp parallel for schedule(dynamic)
        for (long i = 0; i < m; ++i) f(static_cast<std::size_t>(i));
    }
    // each thread folds its share from `init`; the partials meet by min
    template <class V, class F>
    static V fold(std::size_t n, const V& init, F&& step) {
        const long m = static_cast<long>(n);
        V result = init;
#pragma omp parallel
        {
            V part = init;
#pragma omp for schedule(static) nowait
            for (long i = 0; i < m; ++i) part = step(part, static_cast<std::size_t>(i));
#pragma omp critical(nest_pipeline_fold)
            result = std::min(result, part);
        }
        return result;
    }
};

template <class E>
concept host_executor = std::same_as<E, exec::cpu_t> || std::same_as<E, exec::omp_t>;

template <typename T, std::size_t Dim, typename Stage, host_executor E>
void execute_stage(Stage& stage, std::vector<patch_t<T, Dim>>& patches, E) {
    using P = patch_t<T, Dim>;
    using loop = host_loop<E>;
    static_assert(!device_only_stage<Stage>, "a device-only stage (nest::hip::*) needs exec::hip");
    const std::size_t n = patches.size();
    if constexpr (compute_stage<Stage, P>) {
        loop::each(n, [&](std::size_t i) { patches[i] = stage.value(std::move(patches[i])); });
    } else if constexpr (reduce_stage<Stage, P>) {
        using V = typename Stage::value_type;
        const V acc = loop::fold(n, V(Stage::init()),
                                 [&](const V& a, std::size_t i) { return V(stage.reduce(a, patches[i])); });
        loop::each(n, [&](std::size_t i) { stage.finalize(acc, patches[i]); });
    } else {
        loop::each(n, [&](std::size_t i) { fill_halos(patches[i], patches); });
    }
}

// Run the stages in order; each sees the patches as the previous one left them.
template <typename T, std::size_t Dim, typename... Stages, typename Exec>
void execute(pipeline_t<Stages...>& pl, std::vector<patch_t<T, Dim>>& patches, Exec exec) {
    [&]<std::size_t... I>(std::index_sequence<I...>) {
        (execute_stage<T, Dim>(std::get<I>(pl.stages), patches, exec), ...);
    }(std::index_sequence_for<Stages...>{});
}

template <typename T, std::size_t Dim, typename... Stages>
void execute(pipeline_t<Stages...>& pl, std::vector<patch_t<T, Dim>>& patches) {
    execute(pl, patches, exec::cpu);
}

}  // namespace nest
