// CPU tests of the C++20 host API (include/nest/core.hpp, patch.hpp,
// pipeline.hpp): the reference's core known answers (tests/test_core.cpp
// :65-568 of the reference, same numbers, own test code) plus the
// extensions — 2D plans with corners and outflow, checked against the
// oracle's halo fill — and the cpu/omp pipeline executors.
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "../../oracle/euler_oracle.hpp"
#include "nest/pipeline.hpp"

using namespace nest;

static int g_fail = 0, g_run = 0;
static std::vector<std::pair<std::string, std::function<void()>>>& registry() {
    static std::vector<std::pair<std::string, std::function<void()>>> r;
    return r;
}
struct reg {
    reg(const char* n, std::function<void()> f) { registry().push_back({n, f}); }
};
#define TEST(name)                       \
    static void test_##name();           \
    static reg reg_##name(#name, test_##name); \
    static void test_##name()
#define EXPECT(c)                                                          \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::printf("  FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);     \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

TEST(vec_arithmetic) {
    auto a = ivec(1, 2, 3), b = ivec(4, 5, 6);
    EXPECT((a + b) == ivec(5, 7, 9));
    EXPECT((b - a) == ivec(3, 3, 3));
    EXPECT((a * 2) == ivec(2, 4, 6));
    EXPECT(product(uvec(2, 3, 4)) == 24u);
}

TEST(index_space_iteration_is_row_major) {
    auto s = index_space(ivec(0, 0), uvec(2, 3));
    std::vector<ivec_t<2>> got;
    for (auto i : s) got.push_back(i);
    std::vector<ivec_t<2>> want{ivec(0, 0), ivec(0, 1), ivec(0, 2), ivec(1, 0), ivec(1, 1), ivec(1, 2)};
    EXPECT(got == want);
    auto e = index_space(ivec(0, 0), uvec(0, 3));
    int n = 0;
    for (auto i : e) (void)i, ++n;
    EXPECT(n == 0);
}

TEST(subspace_partitions_with_remainder_first) {
    auto d = index_space(ivec(0), uvec(10u));
    EXPECT(start(subspace(d, 3, 0, 0))[0] == 0 && shape(subspace(d, 3, 0, 0))[0] == 4u);
    EXPECT(start(subspace(d, 3, 1, 0))[0] == 4 && shape(subspace(d, 3, 1, 0))[0] == 3u);
    EXPECT(start(subspace(d, 3, 2, 0))[0] == 7 && shape(subspace(d, 3, 2, 0))[0] == 3u);
}

TEST(linear_index_and_soa_offset) {
    auto s1 = index_space(ivec(-2), uvec(10u));
    EXPECT(linear_index(s1, ivec(-2)) == 0 && linear_index(s1, ivec(7)) == 9);
    auto s2 = index_space(ivec(0, 0), uvec(3, 4));  // last index fastest
    EXPECT(linear_index(s2, ivec(0, 1)) == 1 && linear_index(s2, ivec(1, 0)) == 4 && linear_index(s2, ivec(2, 3)) == 11);
    EXPECT(soa_offset(s2, 2, ivec(1, 1)) == 2 * 12 + 5);
}

TEST(md_view_strides_and_subview) {
    std::vector<double> d(12);
    for (int i = 0; i < 12; ++i) d[i] = i;
    md_view_t<double, 2> v(d.data(), index_space(ivec(0, 0), uvec(3, 4)));
    EXPECT(strides(v)[0] == 4u && strides(v)[1] == 1u && is_contiguous(v));
    EXPECT(v(ivec(2, 1)) == 9.0);
    auto sub = v.subview(index_space(ivec(1, 1), uvec(2, 2)));
    EXPECT(sub(ivec(1, 1)) == 5.0 && sub(ivec(2, 2)) == 10.0 && !is_contiguous(sub));
}

TEST(field_soa_layout) {
    field_t<double, 2> f(3, index_space(ivec(0, 0), uvec(2, 3)));
    EXPECT(f.nvars() == 3 && f.stride() == 6);
    f(1, ivec(1, 2)) = 7.0;
    EXPECT(f._data[6 + 5] == 7.0 && f.var_data(1)[5] == 7.0 && f[1](ivec(1, 2)) == 7.0);
    field_t<double, 2> g = std::move(f);
    EXPECT(g(1, ivec(1, 2)) == 7.0 && f._data == nullptr);
}

TEST(patch_with_halos) {
    patch_t<double, 1> p(0, index_space(ivec(0), uvec(8u)), 3, 2);
    EXPECT(start(p.extended)[0] == -2 && shape(p.extended)[0] == 12u && p.nvars() == 3);
    p(2, ivec(-2)) = 1.5;
    EXPECT(p.state.conserved._data[2 * 12] == 1.5);
}

TEST(periodic_1d_plan_and_exchange) {
    auto patches = build_periodic_mesh_1d<double>(index_space(ivec(0), uvec(12u)), 3, 1, 1);
    EXPECT(patches[0].exchange_plan.num_regions() == 2);
    EXPECT(patches[0].exchange_plan.regions[0].peer_index == 2 && patches[0].exchange_plan.regions[1].peer_index == 1);
    for (auto& p : patches)
        for (auto i : p.interior) p(0, i) = p.index;
    execute_all_exchanges(patches);
    EXPECT(patches[0](0, ivec(-1)) == 2.0 && patches[0](0, ivec(4)) == 1.0);
    EXPECT(patches[2](0, ivec(7)) == 1.0 && patches[2](0, ivec(12)) == 0.0);
}

TEST(periodic_1d_halo2) {
    auto patches = build_periodic_mesh_1d<double>(index_space(ivec(0), uvec(20u)), 2, 1, 2);
    for (auto& p : patches)
        for (auto i : p.interior) p(0, i) = i[0];
    execute_all_exchanges(patches);
    EXPECT(patches[0](0, ivec(-2)) == 18.0 && patches[0](0, ivec(-1)) == 19.0);
    EXPECT(patches[0](0, ivec(10)) == 10.0 && patches[0](0, ivec(11)) == 11.0);
}

TEST(periodic_2d_faces_match_wrap_formula) {
    const int nx = 6, ny = 4;
    auto patches = build_periodic_mesh_2d<double>(index_space(ivec(0, 0), uvec(nx, ny)), 2, 2, 1, 1);
    for (auto& p : patches)
        for (auto i : p.interior) p(0, i) = i[0] + 1000.0 * i[1];
    execute_all_exchanges(patches);
    auto wrap = [](int a, int n) { return ((a % n) + n) % n; };
    for (auto& p : patches) {
        EXPECT(p.exchange_plan.num_regions() == 4);  // faces only, as the reference
        for (int j = start(p.interior)[1]; j < upper(p.interior)[1]; ++j)
            for (int i : {start(p.interior)[0] - 1, upper(p.interior)[0]})
                EXPECT(p(0, ivec(i, j)) == wrap(i, nx) + 1000.0 * wrap(j, ny));
        for (int i = start(p.interior)[0]; i < upper(p.interior)[0]; ++i)
            for (int j : {start(p.interior)[1] - 1, upper(p.interior)[1]})
                EXPECT(p(0, ivec(i, j)) == wrap(i, nx) + 1000.0 * wrap(j, ny));
    }
}

TEST(mesh_2d_with_corners_matches_oracle_fill) {
    // faces + corners, periodic / outflow / reflecting per dimension, uneven
    // partitions (the oracle's bc codes: 0 periodic, 1 outflow, 2 reflecting)
    const boundary_t kinds[3] = {boundary_t::periodic, boundary_t::outflow, boundary_t::reflecting};
    for (int bx = 0; bx < 3; ++bx)
        for (int by = 0; by < 3; ++by) {
            const int nx = 13, ny = 9, npx = 3, npy = 2;
            auto bcx = kinds[bx];
            auto bcy = kinds[by];
            auto patches = build_mesh_2d<double>(index_space(ivec(0, 0), uvec(nx, ny)), npx, npy, 4, 2, bcx, bcy);
            nest_oracle::mesh2d m(nx, ny, npx, npy, 2, bx, by);
            std::vector<double> flat(m.total);
            std::mt19937_64 g(bx * 3 + by);
            std::uniform_real_distribution<double> U(-1, 1);
            for (auto& p : patches)
                for (std::size_t v = 0; v < 4; ++v)
                    for (auto i : p.interior) p(v, i) = U(g);
            for (auto& p : patches) {
                std::size_t n = 4 * size(p.extended);
                std::copy(p.state.conserved._data, p.state.conserved._data + n, flat.begin() + m.off[p.index]);
            }
            execute_all_exchanges(patches);
            nest_oracle::fill_mesh(flat.data(), m);
            for (auto& p : patches) {
                std::size_t n = 4 * size(p.extended);
                EXPECT(std::equal(p.state.conserved._data, p.state.conserved._data + n, flat.begin() + m.off[p.index]));
            }
        }
}

struct scale_stage {  // Compute
    double f;
    patch_t<double, 1> value(patch_t<double, 1> p) const {
        for (auto i : p.interior) p(0, i) *= f;
        return p;
    }
};
struct min_stage {  // Reduce (min of interior values), finalize writes time
    using value_type = double;
    static double init() { return 1e300; }
    double reduce(double acc, const patch_t<double, 1>& p) const {
        for (auto i : p.interior) acc = std::min(acc, p(0, i));
        return acc;
    }
    void finalize(double acc, patch_t<double, 1>& p) const { p.state.time = acc; }
};
struct exchange_stage {};

TEST(pipeline_cpu_and_omp) {
    for (int mode = 0; mode < 2; ++mode) {
        auto patches = build_periodic_mesh_1d<double>(index_space(ivec(0), uvec(16u)), 4, 1, 1);
        for (auto& p : patches)
            for (auto i : p.interior) p(0, i) = 1.0 + i[0];
        auto pl = pipeline(scale_stage{2.0}, exchange_stage{}, min_stage{});
        if (mode == 0) execute(pl, patches, exec::cpu);
        else execute(pl, patches, exec::omp);
        for (auto& p : patches) EXPECT(p.state.time == 2.0);
        EXPECT(patches[1](0, ivec(3)) == 8.0);   // halo from patch 0, scaled
        EXPECT(patches[0](0, ivec(-1)) == 32.0);  // periodic wrap from patch 3
    }
}

// The stage contract as concepts: classification of stage shapes.
struct not_a_stage_value {  // value() returning something that is not a patch: not compute
    int value(patch_t<double, 1>) const { return 0; }
};
struct device_tagged {
    static constexpr bool device_only = true;
};
static_assert(compute_stage<scale_stage, patch_t<double, 1>>);
static_assert(!reduce_stage<scale_stage, patch_t<double, 1>>);
static_assert(reduce_stage<min_stage, patch_t<double, 1>>);
static_assert(!compute_stage<min_stage, patch_t<double, 1>>);
static_assert(!compute_stage<exchange_stage, patch_t<double, 1>> && !reduce_stage<exchange_stage, patch_t<double, 1>>);
static_assert(!compute_stage<not_a_stage_value, patch_t<double, 1>>);
static_assert(!compute_stage<scale_stage, patch_t<double, 2>>);  // a 1D stage is no 2D stage
static_assert(device_only_stage<device_tagged> && !device_only_stage<exchange_stage>);
static_assert(host_executor<exec::cpu_t> && host_executor<exec::omp_t> && !host_executor<exec::cuda_t>);

struct append_stage {  // Compute that records the order stages ran in
    int tag;
    std::vector<int>* log;
    patch_t<double, 1> value(patch_t<double, 1> p) const {
        if (p.index == 0) log->push_back(tag);
        return p;
    }
};
struct count_reduce {  // Reduce whose fold is a minimum over a per-patch key
    using value_type = long;
    static long init() { return 1L << 40; }
    long reduce(long acc, const patch_t<double, 1>& p) const { return std::min(acc, 1000L - 7L * p.index); }
    void finalize(long acc, patch_t<double, 1>& p) const { p.state.time = static_cast<double>(acc); }
};

TEST(pipeline_stage_order_and_fold_equivalence) {
    auto patches = build_periodic_mesh_1d<double>(index_space(ivec(0), uvec(640u)), 64, 1, 2);
    std::vector<int> log;
    auto pl = pipeline(append_stage{1, &log}, append_stage{2, &log}, exchange_t{}, append_stage{3, &log});
    EXPECT(pl.size() == 4);
    execute(pl, patches);
    EXPECT((log == std::vector<int>{1, 2, 3}));
    for (int mode = 0; mode < 2; ++mode) {
        auto r = pipeline(count_reduce{});
        if (mode == 0) execute(r, patches, exec::cpu);
        else execute(r, patches, exec::omp);
        for (auto& p : patches) EXPECT(p.state.time == 1000.0 - 7.0 * 63);
    }
}

TEST(for_each_order_and_omp_coverage) {
    auto s = index_space(ivec(0), uvec(100u));
    std::vector<int> order;
    for_each(s, [&](ivec_t<1> i) { order.push_back(i[0]); }, exec::cpu);
    bool ok = order.size() == 100;
    for (int i = 0; ok && i < 100; ++i) ok = order[i] == i;
    EXPECT(ok);
    std::vector<int> hit(24, 0);
    for_each(index_space(ivec(0, 0), uvec(4, 6)), [&](ivec_t<2> i) { hit[i[0] * 6 + i[1]] += 1; }, exec::omp);
    for (int h : hit) EXPECT(h == 1);
}

int main() {
    for (auto& [name, f] : registry()) {
        int before = g_fail;
        f();
        ++g_run;
        std::printf("%s %s\n", g_fail == before ? "ok  " : "FAIL", name.c_str());
    }
    std::printf("%d tests, %d failed checks\n", g_run, g_fail);
    return g_fail ? 1 : 0;
}
