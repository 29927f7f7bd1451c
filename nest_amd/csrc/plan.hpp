// Host-side construction of the device tables of a 2D problem set: the
// tensor-product patch partition, the per-problem geometry and owner tables,
// the halo exchange plan (faces + corners, periodic / outflow / reflecting /
// patch-to-patch) and the stage work lists. Pure C++ (no HIP), so the exact
// tables the GPU consumes are also checked on the CPU (tests/cpp/
// test_plan.cpp runs them against the oracle's halo fill).
#pragma once
#include <algorithm>
#include <vector>

#include "nest_types.hpp"

namespace nest_plan {

enum { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTING = 2 };

// Tensor-product partition of one dimension with subspace() semantics
// (reference include/nest/core.hpp:184-199).
struct dim_part {
    std::vector<int> start, len;
};

inline dim_part partition(int total, int n) {
    dim_part d;
    const int base = total / n, extra = total % n;
    for (int p = 0; p < n; ++p) {
        d.start.push_back(p * base + std::min(p, extra));
        d.len.push_back(base + (p < extra ? 1 : 0));
    }
    return d;
}

// source of the ghost strip on `side` (-1, 0, +1) of partition p along one
// dimension: (source partition, first source extended index, broadcast,
// mirror). A reflecting wall mirrors the first / last h interior cells.
inline void dim_source(int side, int p, const dim_part& d, int bc, int& q, int& s, int& bcast, int& mirror) {
    const int np = (int)d.len.size();
    const int h = 2;
    bcast = 0;
    mirror = 0;
    if (side == 0) {
        q = p;
        s = h;
    } else if (side < 0) {
        if (p > 0 || bc == BC_PERIODIC) {
            q = (p - 1 + np) % np;
            s = d.len[q];  // = h + len - h: last h interior cells
        } else {
            q = p;
            s = h;
            bcast = bc == BC_OUTFLOW;
            mirror = bc == BC_REFLECTING;
        }
    } else {
        if (p < np - 1 || bc == BC_PERIODIC) {
            q = (p + 1) % np;
            s = h;
        } else {
            q = p;
            s = bc == BC_REFLECTING ? d.len[p] : h + d.len[p] - 1;
            bcast = bc == BC_OUTFLOW;
            mirror = bc == BC_REFLECTING;
        }
    }
}

// Halo regions (faces + corners) of every patch of one mesh; off[] holds the
// array offsets of the mesh's patches in local patch-index order (py*npx+px).
inline void add_mesh_regions(std::vector<nest_region>& R, const dim_part& X, const dim_part& Y, int bcx, int bcy,
                             const std::vector<long long>& off) {
    const int npx = (int)X.len.size(), npy = (int)Y.len.size();
    const int h = 2;
    for (int py = 0; py < npy; ++py)
        for (int px = 0; px < npx; ++px) {
            const int p = py * npx + px;
            const int lx = X.len[px], ly = Y.len[py];
            for (int sx = -1; sx <= 1; ++sx)
                for (int sy = -1; sy <= 1; ++sy) {
                    if (sx == 0 && sy == 0) continue;
                    nest_region g{};
                    int qx, qy;
                    dim_source(sx, px, X, bcx, qx, g.si, g.bi, g.mi);
                    dim_source(sy, py, Y, bcy, qy, g.sj, g.bj, g.mj);
                    g.neg = (g.mi ? 2 : 0) | (g.mj ? 4 : 0);  // normal momenta of mirrored directions
                    const int q = qy * npx + qx;
                    g.di = sx < 0 ? 0 : (sx == 0 ? h : h + lx);
                    g.ni = sx == 0 ? lx : h;
                    g.dj = sy < 0 ? 0 : (sy == 0 ? h : h + ly);
                    g.nj = sy == 0 ? ly : h;
                    g.dst_off = off[p];
                    g.src_off = off[q];
                    g.dst_nye = ly + 2 * h;
                    g.src_nye = Y.len[qy] + 2 * h;
                    g.dst_vs = (lx + 2 * h) * (ly + 2 * h);
                    g.src_vs = (X.len[qx] + 2 * h) * (Y.len[qy] + 2 * h);
                    R.push_back(g);
                }
        }
}

// All device tables of a problem set.
struct tables {
    std::vector<nest_patch_desc> patches;
    std::vector<nest_region> regions;
    std::vector<nest_prob_geom> geoms;
    std::vector<int> itab;  // owner / start tables of every problem (see nest_prob_geom)
    long long total = 0;    // doubles per buffer (U or U1)
};

// Append one problem's patch mesh (nx*ny cells on npx*npy patches): patch
// descriptors at consecutive buffer offsets, halo regions, geometry, owner
// tables.
inline void build_mesh(tables& t, int nx, int ny, int npx, int npy, int bcx, int bcy) {
    const int prob = (int)t.geoms.size();
    dim_part X = partition(nx, npx), Y = partition(ny, npy);
    nest_prob_geom G{};
    G.patch_base = (int)t.patches.size();
    G.npx = npx;
    G.npy = npy;
    G.NX = nx;
    G.NY = ny;
    G.bcx = bcx;
    G.bcy = bcy;
    G.tab = (int)t.itab.size();
    for (int px = 0; px < npx; ++px)
        for (int i = 0; i < X.len[px]; ++i) t.itab.push_back(px);
    for (int py = 0; py < npy; ++py)
        for (int j = 0; j < Y.len[py]; ++j) t.itab.push_back(py);
    for (int px = 0; px < npx; ++px) t.itab.push_back(X.start[px]);
    for (int py = 0; py < npy; ++py) t.itab.push_back(Y.start[py]);
    t.geoms.push_back(G);
    std::vector<long long> off;
    for (int py = 0; py < npy; ++py)
        for (int px = 0; px < npx; ++px) {
            off.push_back(t.total);
            t.patches.push_back({t.total, X.len[px], Y.len[py], prob, px, X.start[px], Y.start[py]});
            t.total += 4LL * (X.len[px] + 4) * (Y.len[py] + 4);
        }
    add_mesh_regions(t.regions, X, Y, bcx, bcy, off);
}

// Per-patch work list (halo read from the patch's own ghost cells): row
// chunks of `rows`, strips of 60 output columns over the patch.
inline void add_patch_items(std::vector<nest_item>& items, int patch, int nx, int ny, int rows) {
    for (int r0 = 0; r0 < nx; r0 += rows)
        for (int c0 = 0; c0 < ny; c0 += 60) items.push_back({patch, c0, r0, std::min(r0 + rows, nx)});
}

// Global-strip work list of one problem (halo gathered from the neighbours):
// for every dim-0 partition and row chunk, strips of `width` output columns
// (60 for the forward stage kernel) over
// the problem's whole dim-1 range, each anchored in the patch holding its
// first output column. A strip's lanes may run into the next patches.
inline void add_gather_items(std::vector<nest_item>& items, const tables& t, int prob, int rows, int width = 60) {
    const nest_prob_geom& G = t.geoms[prob];
    const int* T = t.itab.data() + G.tab;
    for (int px = 0; px < G.npx; ++px) {
        const int lx = t.patches[G.patch_base + px].nx;
        for (int r0 = 0; r0 < lx; r0 += rows)
            for (int gs = 0; gs < G.NY; gs += width) {
                const int qy = T[G.NX + gs];
                const int home = G.patch_base + qy * G.npx + px;
                const int c0 = gs - T[G.NX + G.NY + G.npx + qy];
                items.push_back({home, c0, r0, std::min(r0 + rows, lx)});
            }
    }
}

// Every table entry addresses memory inside a buffer of `total` doubles.
inline bool tables_in_bounds(const tables& t, const std::vector<nest_item>& I) {
    auto arr_ok = [&](long long off, long long vs) { return off >= 0 && off + 4 * vs <= t.total; };
    for (const auto& d : t.patches)
        if (d.nx < 2 || d.ny < 2 || d.prob < 0 || d.prob >= (int)t.geoms.size() ||
            !arr_ok(d.off, (long long)(d.nx + 4) * (d.ny + 4)))
            return false;
    for (const auto& G : t.geoms)
        if (G.tab < 0 || G.tab + G.NX + G.NY + G.npx + G.npy > (int)t.itab.size() || G.patch_base < 0 ||
            G.patch_base + G.npx * G.npy > (int)t.patches.size())
            return false;
    for (const auto& g : t.regions) {
        const int dnx = g.dst_vs / g.dst_nye, snx = g.src_vs / g.src_nye;
        bool ok = arr_ok(g.dst_off, g.dst_vs) && arr_ok(g.src_off, g.src_vs) && g.ni > 0 && g.nj > 0 &&
                  g.di >= 0 && g.di + g.ni <= dnx && g.dj >= 0 && g.dj + g.nj <= g.dst_nye && g.si >= 0 &&
                  g.si + (g.bi ? 1 : g.ni) <= snx && g.sj >= 0 && g.sj + (g.bj ? 1 : g.nj) <= g.src_nye;
        if (!ok) return false;
    }
    for (const auto& it : I) {
        if (it.patch < 0 || it.patch >= (int)t.patches.size()) return false;
        const auto& d = t.patches[it.patch];
        if (it.r0 < 0 || it.r0 >= it.r1 || it.r1 > d.nx || it.c0 < 0 || it.c0 >= d.ny) return false;
    }
    return true;
}

// Host execution of a region table (the same index arithmetic as the device
// fill kernel), for CPU checks of the plan.
inline void execute_regions_host(double* base, const std::vector<nest_region>& R, int nvars) {
    for (const auto& g : R)
        for (int k = 0; k < nvars; ++k)
            for (int ia = 0; ia < g.ni; ++ia)
                for (int jb = 0; jb < g.nj; ++jb) {
                    const int si = g.si + (g.bi ? 0 : (g.mi ? g.ni - 1 - ia : ia));
                    const int sj = g.sj + (g.bj ? 0 : (g.mj ? g.nj - 1 - jb : jb));
                    const double v = base[g.src_off + (long long)k * g.src_vs + (long long)si * g.src_nye + sj];
                    base[g.dst_off + (long long)k * g.dst_vs + (long long)(g.di + ia) * g.dst_nye + g.dj + jb] =
                        (g.neg >> k) & 1 ? -v : v;
                }
}

}  // namespace nest_plan
