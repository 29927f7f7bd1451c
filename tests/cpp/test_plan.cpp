// CPU check of the device tables (nest_amd/csrc/plan.hpp): for meshes and
// batches of every shape the library builds, the halo-region table and the
// stage work list stay inside the buffers, and executing the region table
// reproduces the oracle's halo fill (faces + corners, periodic / outflow)
// bit for bit. Exit status 0 = pass.
#include <cstdio>
#include <random>
#include <vector>

#include "../../nest_amd/csrc/plan.hpp"
#include "../../oracle/euler_oracle.hpp"

static int failures = 0;
#define CHECK(c, ...)                        \
    do {                                     \
        if (!(c)) {                          \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");               \
            ++failures;                      \
        }                                    \
    } while (0)

static void check_mesh(int nx, int ny, int npx, int npy, int bcx, int bcy, int rows) {
    std::vector<nest_patch_desc> P;
    std::vector<nest_region> R;
    std::vector<nest_item> I;
    long long total = 0;
    nest_plan::build_mesh(nx, ny, npx, npy, bcx, bcy, 0, P, R, total);
    for (int p = 0; p < (int)P.size(); ++p) nest_plan::add_items(I, p, P[p].nx, P[p].ny, rows);
    CHECK(nest_plan::tables_in_bounds(P, R, I, total), "bounds %dx%d %dx%d bc %d%d", nx, ny, npx, npy, bcx, bcy);
    CHECK(R.size() == 8 * P.size(), "region count");
    // every interior cell is owned by exactly one work item
    std::vector<int> cover(total / 4, 0);
    for (const auto& it : I) {
        const auto& d = P[it.patch];
        for (int i = it.r0; i < it.r1; ++i)
            for (int lane = 2; lane < 62; ++lane) {
                int ec = it.c0 + lane;
                if (ec >= 2 && ec < d.ny + 2) cover[(d.off / 4) + (long long)(i + 2) * (d.ny + 4) + ec]++;
            }
    }
    for (const auto& d : P)
        for (int i = 0; i < d.nx; ++i)
            for (int j = 0; j < d.ny; ++j)
                CHECK(cover[(d.off / 4) + (long long)(i + 2) * (d.ny + 4) + j + 2] == 1, "coverage");
    // region execution == oracle fill
    nest_oracle::mesh2d m(nx, ny, npx, npy, 2, bcx, bcy);
    CHECK((long long)m.total == total, "total %lld vs %zu", total, m.total);
    std::mt19937_64 g(nx * 7919 + ny);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> a(total);
    for (auto& x : a) x = U(g);
    std::vector<double> b = a;
    nest_plan::execute_regions_host(a.data(), R, 4);
    nest_oracle::fill_mesh(b.data(), m);
    CHECK(a == b, "fill %dx%d %dx%d bc %d%d", nx, ny, npx, npy, bcx, bcy);
}

int main() {
    for (int bcx = 0; bcx < 2; ++bcx)
        for (int bcy = 0; bcy < 2; ++bcy) {
            check_mesh(96, 80, 3, 2, bcx, bcy, 64);
            check_mesh(70, 130, 2, 3, bcx, bcy, 17);
            check_mesh(40, 200, 1, 1, bcx, bcy, 64);
            check_mesh(16, 12, 2, 2, bcx, bcy, 3);
            check_mesh(9, 7, 4, 3, bcx, bcy, 5);  // uneven, minimum-size patches
            check_mesh(256, 256, 4, 4, bcx, bcy, 64);
        }
    // batch: independent single-patch problems appended to one buffer
    {
        std::vector<nest_patch_desc> P;
        std::vector<nest_region> R;
        std::vector<nest_item> I;
        long long total = 0;
        int sizes[] = {64, 81, 100, 130, 77, 512, 2, 3};
        for (int k = 0; k < 8; ++k) nest_plan::build_mesh(sizes[k], sizes[k] + 1, 1, 1, 1, 1, k, P, R, total);
        for (int p = 0; p < (int)P.size(); ++p) nest_plan::add_items(I, p, P[p].nx, P[p].ny, 64);
        CHECK(nest_plan::tables_in_bounds(P, R, I, total), "batch bounds");
        for (int k = 0; k < 8; ++k) CHECK(P[k].prob == k, "batch prob index");
        std::vector<double> a(total), b;
        std::mt19937_64 g(7);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (auto& x : a) x = U(g);
        b = a;
        nest_plan::execute_regions_host(a.data(), R, 4);
        for (int k = 0; k < 8; ++k) {
            nest_oracle::mesh2d m(sizes[k], sizes[k] + 1, 1, 1, 2, 1, 1);
            nest_oracle::fill_mesh(b.data() + P[k].off, m);
        }
        CHECK(a == b, "batch fill");
    }
    // full-size configurations: bounds only
    for (auto [n, np] : {std::pair{4096, 16}, std::pair{8192, 64}, std::pair{1024, 1}}) {
        std::vector<nest_patch_desc> P;
        std::vector<nest_region> R;
        std::vector<nest_item> I;
        long long total = 0;
        nest_plan::build_mesh(n, n, np, np, 1, 1, 0, P, R, total);
        for (int p = 0; p < (int)P.size(); ++p) nest_plan::add_items(I, p, P[p].nx, P[p].ny, 64);
        CHECK(nest_plan::tables_in_bounds(P, R, I, total), "full size %d", n);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
