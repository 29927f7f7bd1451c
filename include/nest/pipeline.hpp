// nest pipelines — a tuple of stages executed over a vector of patches.
//
// API-compatible with the reference's include/nest/pipeline.hpp (own
// implementation). Stage kinds are recognised by shape, as the reference
// does (pipeline.hpp:14-56):
//   Compute: stage.value(patch) -> patch
//   Reduce:  value_type, static init(), reduce(acc, const patch&) -> acc,
//            finalize(acc, patch&)
//   anything else: Exchange (fill every patch's halo from its plan).
// exec::cpu and exec::omp executors are here; exec::hip (the MI355X backend)
// is in backend_hip.hpp. The OpenMP Reduce combines per-thread accumulators
// with min, as the reference's reduction(min:acc) does (pipeline.hpp:116).
#pragma once
#include <algorithm>
#include <functional>
#include <limits>
#include <tuple>
#include <type_traits>
#include <vector>

#include "patch.hpp"

namespace nest {

namespace detail {
template <typename S, typename P, typename = void>
struct is_compute_stage : std::false_type {};
template <typename S, typename P>
struct is_compute_stage<S, P, std::void_t<decltype(std::declval<const S&>().value(std::declval<P>()))>>
    : std::true_type {};

template <typename S, typename P, typename = void>
struct is_reduce_stage : std::false_type {};
template <typename S, typename P>
struct is_reduce_stage<
    S, P,
    std::void_t<typename S::value_type, decltype(S::init()),
                decltype(std::declval<const S&>().reduce(std::declval<typename S::value_type>(),
                                                         std::declval<const P&>())),
                decltype(std::declval<const S&>().finalize(std::declval<typename S::value_type>(),
                                                           std::declval<P&>()))>> : std::true_type {};

template <typename S, typename P, typename = void>
struct is_exchange_stage : std::false_type {};
template <typename S, typename P>
struct is_exchange_stage<
    S, P,
    std::void_t<decltype(std::declval<const S&>().provides(std::declval<const P&>())),
                decltype(std::declval<const S&>().need(std::declval<P&>(),
                                                       std::declval<std::function<void(int, int)>>()))>>
    : std::true_type {};
}  // namespace detail

// An explicit Exchange stage (any stage that is neither Compute nor Reduce
// exchanges halos; this one names the intent).
struct exchange_t {};

template <typename... Stages>
struct pipeline_t {
    std::tuple<Stages...> stages;
    explicit pipeline_t(Stages... s) : stages(std::move(s)...) {}
};

template <typename... Stages>
pipeline_t<Stages...> pipeline(Stages... s) {
    return pipeline_t<Stages...>(std::move(s)...);
}

template <typename T, std::size_t Dim, typename Stage>
void execute_stage(Stage& stage, std::vector<patch_t<T, Dim>>& patches, exec::cpu_t) {
    using P = patch_t<T, Dim>;
    if constexpr (detail::is_compute_stage<Stage, P>::value) {
        for (auto& p : patches) p = stage.value(std::move(p));
    } else if constexpr (detail::is_reduce_stage<Stage, P>::value) {
        auto acc = Stage::init();
        for (const auto& p : patches) acc = stage.reduce(acc, p);
        for (auto& p : patches) stage.finalize(acc, p);
    } else {
        fill_all_halos(patches);
    }
}

template <typename T, std::size_t Dim, typename Stage>
void execute_stage(Stage& stage, std::vector<patch_t<T, Dim>>& patches, exec::omp_t) {
    using P = patch_t<T, Dim>;
    const long n = static_cast<long>(patches.size());
    if constexpr (detail::is_compute_stage<Stage, P>::value) {
#pragma omp parallel for schedule(dynamic)
        for (long i = 0; i < n; ++i) patches[i] = stage.value(std::move(patches[i]));
    } else if constexpr (detail::is_reduce_stage<Stage, P>::value) {
        auto acc = Stage::init();
#pragma omp parallel
        {
            auto mine = Stage::init();
#pragma omp for schedule(static) nowait
            for (long i = 0; i < n; ++i) mine = stage.reduce(mine, patches[i]);
#pragma omp critical(nest_reduce)
            acc = std::min(acc, mine);
        }
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) stage.finalize(acc, patches[i]);
    } else {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) fill_halos(patches[i], patches);
    }
}

template <typename T, std::size_t Dim, typename... Stages, typename Exec>
void execute(pipeline_t<Stages...>& pl, std::vector<patch_t<T, Dim>>& patches, Exec exec) {
    std::apply([&](auto&... s) { (execute_stage<T, Dim>(s, patches, exec), ...); }, pl.stages);
}

template <typename T, std::size_t Dim, typename... Stages>
void execute(pipeline_t<Stages...>& pl, std::vector<patch_t<T, Dim>>& patches) {
    execute(pl, patches, exec::cpu);
}

}  // namespace nest
