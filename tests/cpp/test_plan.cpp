// CPU check of the device tables (nest_amd/csrc/plan.hpp): for meshes and
// batches of every shape the library builds,
//  * the halo-region table, the geometry / owner tables and both work lists
//    stay inside the buffers;
//  * executing the region table reproduces the oracle's halo fill (faces +
//    corners, periodic / outflow / reflecting) bit for bit;
//  * nest_gather_source (the addressing of the stage kernels' in-kernel halo
//    gather) names, for every extended cell of every patch, an interior cell
//    that holds the oracle's ghost value (negated where nest_gather_flip
//    says so);
//  * the per-patch strips and the global strips each write every interior
//    cell exactly once.
// Exit status 0 = pass.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../nest_amd/csrc/plan.hpp"
#include "../../oracle/euler_oracle.hpp"

static int failures = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            if (failures < 20) {               \
                std::printf("FAIL: " __VA_ARGS__); \
                std::printf("\n");             \
            }                                  \
            ++failures;                        \
        }                                      \
    } while (0)

using nest_plan::tables;

// output lanes of every item (2..61; the one-launch step's strips of 56
// columns: 4..59): which interior cells do they write?
static void check_coverage(const tables& t, const std::vector<nest_item>& I, bool gather, const char* what,
                           int lane0 = 2) {
    std::vector<int> cover(t.total / 4, 0);
    for (const auto& it : I) {
        const auto& pd = t.patches[it.patch];
        const auto& G = t.geoms[pd.prob];
        for (int lane = lane0; lane < 64 - lane0; ++lane) {
            long long off;
            int lc, nye;
            if (gather) {
                const int g = pd.y0 + it.c0 + lane - lane0;
                if (g >= G.NY) continue;
                int q, le;
                nest_gather_source(pd, G, t.itab.data(), 2, g, q, le, lc);
                off = t.patches[q].off;
                nye = t.patches[q].ny + 4;
            } else {
                const int ec = it.c0 + lane;
                if (ec < 2 || ec >= pd.ny + 2) continue;
                off = pd.off;
                lc = ec;
                nye = pd.ny + 4;
            }
            for (int i = it.r0; i < it.r1; ++i) cover[off / 4 + (long long)(i + 2) * nye + lc]++;
        }
    }
    for (const auto& d : t.patches)
        for (int i = 0; i < d.nx; ++i)
            for (int j = 0; j < d.ny; ++j)
                CHECK(cover[d.off / 4 + (long long)(i + 2) * (d.ny + 4) + j + 2] == 1, "%s coverage", what);
}

// The one-launch step's strips (56 output columns): every address its
// gather forms (lanes 0..63, U^n rows r0-4 .. r1+3) is an interior cell of
// a patch of the problem; lanes resolve their column within the home
// partition's rows; every lane stage 2 reads (2..61) that takes U1 from
// another lane (nest_step_lane_source) takes it from a lane stage 1
// computes (2..61) holding the mapped column.
static void check_step_items(const tables& t, const std::vector<nest_item>& FI, const char* what) {
    const int* T = t.itab.data();
    for (const auto& it : FI) {
        const auto& pd = t.patches[it.patch];
        const auto& G = t.geoms[pd.prob];
        for (int lane = 0; lane < 64; ++lane) {
            const int g = pd.y0 + it.c0 + lane - 4;
            int q, le, lc;
            nest_gather_source(pd, G, T, 2, g, q, le, lc);
            CHECK(q >= G.patch_base && q < G.patch_base + G.npx * G.npy && t.patches[q].nx == pd.nx &&
                      le == 2 && lc >= 2 && lc < t.patches[q].ny + 2,
                  "%s: lane column source", what);
            for (int i = it.r0 - 4; i <= it.r1 + 3; ++i) {
                nest_gather_source(pd, G, T, i + 2, g, q, le, lc);
                const auto& qd = t.patches[q];
                CHECK(q >= G.patch_base && q < G.patch_base + G.npx * G.npy && le >= 2 && le < qd.nx + 2 && lc >= 2 &&
                          lc < qd.ny + 2 && qd.off + 3LL * (qd.nx + 4) * (qd.ny + 4) + (long long)le * (qd.ny + 4) + lc < t.total,
                      "%s: row source", what);
            }
            if (lane >= 2 && lane < 62) {
                const int s = nest_step_lane_source(lane, g, G.NY, G.bcy);
                const int gs = pd.y0 + it.c0 + s - 4;
                const bool inside = g >= 0 && g < G.NY;
                // only lanes stage 2 reads on behalf of a stored output (g <= NY + 1) matter
                if (!inside && G.bcy != 0 && g >= -2 && g <= G.NY + 1)
                    CHECK(s >= 2 && s < 62 && gs == nest_map_index(g, G.NY, G.bcy), "%s: lane source", what);
                if (inside || G.bcy == 0) CHECK(s == lane, "%s: lane source identity", what);
            }
        }
    }
    check_coverage(t, FI, true, what, 4);
}

static void check_mesh(int nx, int ny, int npx, int npy, int bcx, int bcy, int rows) {
    tables t;
    nest_plan::build_mesh(t, nx, ny, npx, npy, bcx, bcy);
    std::vector<nest_item> I, GI;
    for (int p = 0; p < (int)t.patches.size(); ++p) nest_plan::add_patch_items(I, p, t.patches[p].nx, t.patches[p].ny, rows);
    nest_plan::add_gather_items(GI, t, 0, rows);
    CHECK(nest_plan::tables_in_bounds(t, I) && nest_plan::tables_in_bounds(t, GI), "bounds %dx%d %dx%d bc %d%d", nx,
          ny, npx, npy, bcx, bcy);
    CHECK(t.regions.size() == 8 * t.patches.size(), "region count");
    check_coverage(t, I, false, "patch strips");
    check_coverage(t, GI, true, "global strips");
    for (int r : {rows, 1, 4, 7}) {
        std::vector<nest_item> FI;
        nest_plan::add_gather_items(FI, t, 0, r, 56);
        CHECK(nest_plan::tables_in_bounds(t, FI), "step strip bounds");
        check_step_items(t, FI, "step strips");
    }
    // region execution == oracle fill
    nest_oracle::mesh2d m(nx, ny, npx, npy, 2, bcx, bcy);
    CHECK((long long)m.total == t.total, "total %lld vs %zu", t.total, m.total);
    std::mt19937_64 g(nx * 7919 + ny);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> a(t.total);
    for (auto& x : a) x = U(g);
    std::vector<double> b = a;
    nest_plan::execute_regions_host(a.data(), t.regions, 4);
    nest_oracle::fill_mesh(b.data(), m);
    CHECK(a == b, "fill %dx%d %dx%d bc %d%d", nx, ny, npx, npy, bcx, bcy);
    // gather addressing == oracle ghost values, every extended cell, and the
    // same for lanes far past the patch edge (global strips)
    const auto& G = t.geoms[0];
    for (int p = 0; p < (int)t.patches.size(); ++p) {
        const auto& d = t.patches[p];
        for (int e = 0; e < d.nx + 4; ++e)
            for (int ec = 0; ec < d.ny + 4; ++ec) {
                int q, le, lc;
                nest_gather_source(d, G, t.itab.data(), e, d.y0 + ec - 2, q, le, lc);
                const auto& qd = t.patches[q];
                bool inside = q >= 0 && q < (int)t.patches.size() && le >= 2 && le < qd.nx + 2 && lc >= 2 &&
                              lc < qd.ny + 2;
                CHECK(inside, "gather source not interior p%d e%d ec%d", p, e, ec);
                if (!inside) continue;
                const int fl = nest_gather_flip(d, G, e, d.y0 + ec - 2);
                for (int k = 0; k < 4; ++k) {
                    double got = b[qd.off + (long long)k * (qd.nx + 4) * (qd.ny + 4) + (long long)le * (qd.ny + 4) + lc];
                    if ((fl >> k) & 1) got = -got;
                    double want = b[d.off + (long long)k * (d.nx + 4) * (d.ny + 4) + (long long)e * (d.ny + 4) + ec];
                    CHECK(got == want, "gather value p%d e%d ec%d", p, e, ec);
                }
            }
        for (int e : {0, 1, 2, d.nx + 1, d.nx + 2, d.nx + 3})
            for (int gg = d.y0 - 2; gg < d.y0 + d.ny + 66; ++gg) {
                int q, le, lc;
                nest_gather_source(d, G, t.itab.data(), e, gg, q, le, lc);
                const auto& qd = t.patches[q];
                CHECK(le >= 2 && le < qd.nx + 2 && lc >= 2 && lc < qd.ny + 2, "far gather in bounds");
            }
    }
}

// The 1D stage kernel reads extended cell e of the input as interior cell
// nest_map_index(e - 2, n, bc), momentum negated past a reflecting wall; for
// every ghost that must be exactly what the oracle's fill leaves there, and
// cells past the array (a block's tail) must map inside it.
static void check_1d_gather() {
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> U(-1, 1);
    for (int n : {2, 3, 5, 17, 256, 1000})
        for (int bc = 0; bc < 3; ++bc) {
            const int ne = n + 4;
            std::vector<double> a(3 * ne);
            for (auto& x : a) x = U(g);
            std::vector<double> f = a;
            nest_oracle::fill_1d(f.data(), n, 2, 3, bc);
            for (int e = 0; e < ne + 300; ++e) {
                const int gi = e - 2;
                const int src = nest_map_index(gi, n, bc) + 2;
                CHECK(src >= 2 && src < n + 2, "1d gather n%d bc%d e%d in bounds", n, bc, e);
                if (e >= ne) continue;
                const bool neg = bc == 2 && (gi < 0 || gi >= n);
                for (int k = 0; k < 3; ++k) {
                    const double v = a[k * ne + src];
                    CHECK(f[k * ne + e] == (k == 1 && neg ? -v : v), "1d gather n%d bc%d e%d k%d", n, bc, e, k);
                }
            }
        }
}

// The fused 1D step's wall copies (nest_step1d_wall_source): for every size
// and block, every ghost position stage 2 reads gets a source, and every
// source is a tile position stage 1 computes ([2, 262)) holding an interior
// cell, never a position the copy loop itself writes.
static void check_step1d_wall() {
    for (int bc = 1; bc <= 2; ++bc)
        for (int n = 2; n <= 1100; ++n)
            for (int g0 = 0; g0 < n; g0 += 256) {
                const int gend = std::min(n, g0 + 256);
                for (int t = 2; t < 262; ++t) {
                    const int g = g0 - 4 + t;
                    const int ts = nest_step1d_wall_source(t, g0, n, bc);
                    const bool ghost = g < 0 || g >= n;
                    const bool read = g >= g0 - 2 && g < gend + 2;
                    CHECK((ts >= 0) == (ghost && read), "wall source presence n=%d g0=%d t=%d bc=%d", n, g0, t, bc);
                    if (ts < 0) continue;
                    const int gs = g0 - 4 + ts;
                    CHECK(ts >= 3 && ts < 262 && gs >= 0 && gs < n, "wall source range n=%d g0=%d t=%d bc=%d ts=%d", n,
                          g0, t, bc, ts);
                    CHECK(gs == nest_map_index(g, n, bc), "wall source cell n=%d t=%d", n, t);
                }
            }
}

// The 1D lane kernel's wall copies: in every 56-cell segment, every lane
// stage 2 reads (2..61) whose cell lies past an outflow / reflecting wall
// but within two cells of it takes U1 from a lane stage 1 computes (2..61)
// holding the mapped cell; lanes inside keep their own.
static void check_step1d_lanes() {
    for (int bc = 1; bc <= 2; ++bc)
        for (int n = 2; n <= 1100; ++n)
            for (int g0 = 0; g0 < n; g0 += 56)
                for (int lane = 2; lane < 62; ++lane) {
                    const int g = g0 - 4 + lane;
                    const int s = nest_step_lane_source(lane, g, n, bc);
                    if (g >= 0 && g < n) {
                        CHECK(s == lane, "lanes identity n=%d g0=%d lane=%d", n, g0, lane);
                    } else if (g >= -2 && g <= n + 1) {
                        CHECK(s >= 2 && s < 62 && g0 - 4 + s == nest_map_index(g, n, bc),
                              "lanes source n=%d g0=%d lane=%d bc=%d s=%d", n, g0, lane, bc, s);
                    }
                }
}

int main() {
    check_step1d_wall();
    check_step1d_lanes();
    check_1d_gather();
    for (int bcx = 0; bcx < 3; ++bcx)  // periodic, outflow, reflecting
        for (int bcy = 0; bcy < 3; ++bcy) {
            check_mesh(96, 80, 3, 2, bcx, bcy, 64);
            check_mesh(70, 130, 2, 3, bcx, bcy, 17);
            check_mesh(40, 200, 1, 1, bcx, bcy, 64);
            check_mesh(16, 12, 2, 2, bcx, bcy, 3);
            check_mesh(9, 7, 4, 3, bcx, bcy, 5);  // uneven, minimum-size patches
            check_mesh(256, 256, 4, 4, bcx, bcy, 64);
            check_mesh(128, 300, 2, 5, bcx, bcy, 64);  // strips crossing several patches
        }
    // batch: independent single-patch problems appended to one buffer
    {
        tables t;
        int sizes[] = {64, 81, 100, 130, 77, 512, 2, 3};
        for (int k = 0; k < 8; ++k) nest_plan::build_mesh(t, sizes[k], sizes[k] + 1, 1, 1, 1, 1);
        std::vector<nest_item> I, GI;
        for (int p = 0; p < (int)t.patches.size(); ++p)
            nest_plan::add_patch_items(I, p, t.patches[p].nx, t.patches[p].ny, 64);
        for (int k = 0; k < 8; ++k) nest_plan::add_gather_items(GI, t, k, 64);
        CHECK(nest_plan::tables_in_bounds(t, I) && nest_plan::tables_in_bounds(t, GI), "batch bounds");
        for (int k = 0; k < 8; ++k) CHECK(t.patches[k].prob == k && t.geoms[k].patch_base == k, "batch indices");
        check_coverage(t, I, false, "batch patch strips");
        check_coverage(t, GI, true, "batch global strips");
        std::vector<nest_item> FI;
        for (int k = 0; k < 8; ++k) nest_plan::add_gather_items(FI, t, k, 12, 56);
        CHECK(nest_plan::tables_in_bounds(t, FI), "batch step strip bounds");
        check_step_items(t, FI, "batch step strips");
        std::vector<double> a(t.total), b;
        std::mt19937_64 g(7);
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (auto& x : a) x = U(g);
        b = a;
        nest_plan::execute_regions_host(a.data(), t.regions, 4);
        for (int k = 0; k < 8; ++k) {
            nest_oracle::mesh2d m(sizes[k], sizes[k] + 1, 1, 1, 2, 1, 1);
            nest_oracle::fill_mesh(b.data() + t.patches[k].off, m);
        }
        CHECK(a == b, "batch fill");
    }
    // full-size configurations: bounds only
    for (auto [n, np] : {std::pair{4096, 16}, std::pair{8192, 64}, std::pair{1024, 1}}) {
        tables t;
        nest_plan::build_mesh(t, n, n, np, np, 1, 1);
        std::vector<nest_item> I, GI;
        for (int p = 0; p < (int)t.patches.size(); ++p)
            nest_plan::add_patch_items(I, p, t.patches[p].nx, t.patches[p].ny, 64);
        nest_plan::add_gather_items(GI, t, 0, 64);
        CHECK(nest_plan::tables_in_bounds(t, I) && nest_plan::tables_in_bounds(t, GI), "full size %d", n);
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "PASSED", failures);
    return failures ? 1 : 0;
}
