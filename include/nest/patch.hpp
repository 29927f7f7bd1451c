// nest patches and halo exchange plans.
//
// API-compatible with the reference's include/nest/patch.hpp: halo_region_t /
// halo_exchange_plan_t (patch.hpp:14-37), patch_t (:43-81), the periodic mesh
// builders (:88-147, :207-326) and the plan executors (:153-204). Own
// implementation, plus three extensions the Euler step needs:
//   * physical boundaries: a region with peer_index < 0 is skipped, as in the
//     reference, unless its `bc` is boundary_t::outflow or ::reflecting; then
//     its source is this patch's own boundary cells, and a source extent of 1
//     along a dimension is broadcast over the destination (zero-gradient
//     ghosts);
//   * mirrored regions (reflecting walls): bit k of `mirror` walks the source
//     backwards along dimension k, bit v of `negate` flips the sign of
//     variable v (the wall-normal momentum);
//   * build_mesh_2d(): a 2D tensor-product mesh whose plans cover faces AND
//     corners, with periodic, outflow or reflecting boundaries per dimension.
//     Corners take the value of the global cell reached by mapping each
//     coordinate through its boundary condition — the halo a fused X-then-Y
//     stage needs.
#pragma once
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "core.hpp"

namespace nest {

enum class boundary_t { none, periodic, outflow, reflecting };

template <std::size_t Dim>
struct halo_region_t {
    int peer_index = -1;           // neighbour patch index (-1: physical boundary)
    index_space_t<Dim> src_space;  // source box (in the peer, or in this patch for outflow)
    index_space_t<Dim> dst_space;  // ghost box in this patch
    std::size_t copy_size = 0;     // elements per variable
    boundary_t bc = boundary_t::none;
    unsigned mirror = 0;           // bit k: source walked backwards along dimension k
    unsigned negate = 0;           // bit v: variable v changes sign (wall-normal momentum)
};

template <std::size_t Dim>
struct halo_exchange_plan_t {
    std::vector<halo_region_t<Dim>> regions;
    void add_region(int peer, index_space_t<Dim> src, index_space_t<Dim> dst,
                    boundary_t bc = boundary_t::none, unsigned mirror = 0, unsigned negate = 0) {
        regions.push_back(halo_region_t<Dim>{peer, src, dst, size(dst), bc, mirror, negate});
    }
    std::size_t num_regions() const { return regions.size(); }
};

template <typename T, std::size_t Dim>
struct patch_t {
    int index = 0;
    index_space_t<Dim> interior;
    index_space_t<Dim> extended;
    int halo_width = 1;
    state_t<T, Dim> state;
    halo_exchange_plan_t<Dim> exchange_plan;

    patch_t() = default;
    patch_t(int idx, index_space_t<Dim> in, std::size_t nvars, int halo = 1)
        : index(idx), interior(in), extended(expand(in, halo)), halo_width(halo), state(nvars, extended) {}

    md_view_t<T, Dim> operator[](std::size_t v) { return state[v]; }
    md_view_t<const T, Dim> operator[](std::size_t v) const { return state[v]; }
    T& operator()(std::size_t v, ivec_t<Dim> idx) { return state(v, idx); }
    const T& operator()(std::size_t v, ivec_t<Dim> idx) const { return state(v, idx); }
    std::size_t nvars() const { return state.nvars(); }
};

// ---------------------------------------------------------------------------
// plan executor
// ---------------------------------------------------------------------------
template <typename T, std::size_t Dim>
void execute_exchange(patch_t<T, Dim>& patch, const std::vector<patch_t<T, Dim>>& all) {
    for (const auto& r : patch.exchange_plan.regions) {
        const patch_t<T, Dim>* src;
        if (r.peer_index >= 0) {
            src = &all[static_cast<std::size_t>(r.peer_index)];
        } else if (r.bc == boundary_t::outflow || r.bc == boundary_t::reflecting) {
            src = &patch;
        } else {
            continue;  // physical boundary left to the caller (reference behaviour)
        }
        for (std::size_t v = 0; v < patch.nvars(); ++v) {
            if (Dim == 1 && r.peer_index >= 0 && size(r.src_space) == r.copy_size && !r.mirror && !r.negate) {
                std::memcpy(patch.state.conserved.var_data(v) + linear_index(patch.extended, start(r.dst_space)),
                            src->state.conserved.var_data(v) + linear_index(src->extended, start(r.src_space)),
                            r.copy_size * sizeof(T));
                continue;
            }
            const bool neg = (r.negate >> v) & 1u;
            for (auto d : r.dst_space) {
                ivec_t<Dim> s;
                for (std::size_t k = 0; k < Dim; ++k) {
                    const int o = d[k] - r.dst_space._start[k];
                    s[k] = r.src_space._start[k] +
                           (r.src_space._shape[k] == 1
                                ? 0
                                : ((r.mirror >> k) & 1u ? static_cast<int>(r.dst_space._shape[k]) - 1 - o : o));
                }
                const T x = (*src)(v, s);
                patch(v, d) = neg ? -x : x;
            }
        }
    }
}

template <typename T, std::size_t Dim>
void execute_all_exchanges(std::vector<patch_t<T, Dim>>& patches) {
    for (auto& p : patches) execute_exchange(p, patches);
}

template <std::size_t Dim> using neighbor_info_t = halo_region_t<Dim>;

template <typename T, std::size_t Dim>
void fill_halos(patch_t<T, Dim>& patch, const std::vector<patch_t<T, Dim>>& all) {
    execute_exchange(patch, all);
}
template <typename T, std::size_t Dim>
void fill_all_halos(std::vector<patch_t<T, Dim>>& patches) {
    execute_all_exchanges(patches);
}

// ---------------------------------------------------------------------------
// mesh builders
// ---------------------------------------------------------------------------
namespace detail {
// ghost strip on `side` (-1 low, +1 high) of partition p along one axis:
// source partition q and the first source cell along the axis, an outflow
// source (q = p, the boundary cell, broadcast), or a reflecting one (q = p,
// the `halo` cells next to the wall, walked backwards)
struct axis_source {
    int q;
    int first;
    bool outflow;
    bool mirror;
};
template <std::size_t Dim>
inline axis_source ghost_source(int side, int p, const std::vector<index_space_t<Dim>>& parts, int axis, int halo,
                                boundary_t bc) {
    const int n = static_cast<int>(parts.size());
    if (side < 0) {
        if (p > 0 || bc == boundary_t::periodic) {
            int q = (p - 1 + n) % n;
            return {q, upper(parts[q])[axis] - halo, false, false};
        }
        if (bc == boundary_t::reflecting) return {p, start(parts[p])[axis], false, true};
        return {p, start(parts[p])[axis], true, false};
    }
    if (p < n - 1 || bc == boundary_t::periodic) {
        int q = (p + 1) % n;
        return {q, start(parts[q])[axis], false, false};
    }
    if (bc == boundary_t::reflecting) return {p, upper(parts[p])[axis] - halo, false, true};
    return {p, upper(parts[p])[axis] - 1, true, false};
}
}  // namespace detail

// 1D periodic mesh (reference patch.hpp:88-147)
template <typename T, std::size_t Dim>
std::vector<patch_t<T, Dim>> build_periodic_mesh_1d(index_space_t<Dim> domain, int n_patches, std::size_t nvars,
                                                    int halo = 1) {
    static_assert(Dim == 1, "build_periodic_mesh_1d only supports 1D");
    std::vector<patch_t<T, Dim>> patches;
    patches.reserve(static_cast<std::size_t>(n_patches));
    for (int p = 0; p < n_patches; ++p) patches.emplace_back(p, subspace(domain, n_patches, p, 0), nvars, halo);
    for (int p = 0; p < n_patches; ++p) {
        auto& me = patches[p];
        const int l = (p - 1 + n_patches) % n_patches, r = (p + 1) % n_patches;
        auto w = uvec(static_cast<unsigned>(halo));
        me.exchange_plan.add_region(l, index_space(ivec(upper(patches[l].interior)[0] - halo), w),
                                    index_space(ivec(start(me.interior)[0] - halo), w));
        me.exchange_plan.add_region(r, index_space(ivec(start(patches[r].interior)[0]), w),
                                    index_space(ivec(upper(me.interior)[0]), w));
    }
    return patches;
}

// 2D tensor-product mesh, patch index = py*npx + px (reference
// patch.hpp:219-229). corners = false gives the reference's face-only plans
// (build_periodic_mesh_2d); corners = true adds the 4 corner blocks. A
// reflecting wall along dimension k negates variable 1 + k, the wall-normal
// momentum of the (rho, rho u, rho v, E) layout; it needs patches of at least
// `halo` cells along k.
template <typename T>
std::vector<patch_t<T, 2>> build_mesh_2d(index_space_t<2> domain, int npx, int npy, std::size_t nvars, int halo,
                                         boundary_t bcx, boundary_t bcy, bool corners = true) {
    std::vector<patch_t<T, 2>> patches;
    patches.reserve(static_cast<std::size_t>(npx * npy));
    std::vector<index_space_t<2>> xs, ys;
    for (int px = 0; px < npx; ++px) xs.push_back(subspace(domain, npx, px, 0));
    for (int py = 0; py < npy; ++py) ys.push_back(subspace(domain, npy, py, 1));
    for (int py = 0; py < npy; ++py)
        for (int px = 0; px < npx; ++px)
            patches.emplace_back(py * npx + px, subspace(xs[px], npy, py, 1), nvars, halo);
    for (int py = 0; py < npy; ++py)
        for (int px = 0; px < npx; ++px) {
            auto& me = patches[static_cast<std::size_t>(py * npx + px)];
            for (int sx = -1; sx <= 1; ++sx)
                for (int sy = -1; sy <= 1; ++sy) {
                    if ((sx == 0 && sy == 0) || (!corners && sx != 0 && sy != 0)) continue;
                    ivec_t<2> d0, s0;
                    uvec_t<2> dn, sn;
                    int qx = px, qy = py;
                    bool outflow = false;
                    unsigned mirror = 0, negate = 0;
                    // dim 0
                    if (sx == 0) {
                        d0[0] = s0[0] = start(me.interior)[0];
                        dn[0] = sn[0] = shape(me.interior)[0];
                    } else {
                        auto a = detail::ghost_source<2>(sx, px, xs, 0, halo, bcx);
                        qx = a.q;
                        d0[0] = sx < 0 ? start(me.interior)[0] - halo : upper(me.interior)[0];
                        dn[0] = static_cast<unsigned>(halo);
                        s0[0] = a.first;
                        sn[0] = a.outflow ? 1u : static_cast<unsigned>(halo);
                        outflow |= a.outflow;
                        if (a.mirror) mirror |= 1u, negate |= 2u;
                    }
                    // dim 1
                    if (sy == 0) {
                        d0[1] = s0[1] = start(me.interior)[1];
                        dn[1] = sn[1] = shape(me.interior)[1];
                    } else {
                        auto a = detail::ghost_source<2>(sy, py, ys, 1, halo, bcy);
                        qy = a.q;
                        d0[1] = sy < 0 ? start(me.interior)[1] - halo : upper(me.interior)[1];
                        dn[1] = static_cast<unsigned>(halo);
                        s0[1] = a.first;
                        sn[1] = a.outflow ? 1u : static_cast<unsigned>(halo);
                        outflow |= a.outflow;
                        if (a.mirror) mirror |= 2u, negate |= 4u;
                    }
                    const int q = qy * npx + qx;
                    // a region whose source is another patch (or itself through
                    // a periodic wrap) is an exchange; one that reads this
                    // patch's own boundary cells is a wall fill (outflow copy
                    // or reflecting mirror)
                    const bool wall = outflow || mirror;
                    const bool self_bc = wall && q == me.index;
                    me.exchange_plan.add_region(
                        self_bc ? -1 : q, index_space(s0, sn), index_space(d0, dn),
                        mirror ? boundary_t::reflecting : (outflow ? boundary_t::outflow : boundary_t::periodic),
                        mirror, negate);
                }
        }
    return patches;
}

// 2D periodic mesh with face-only plans (reference patch.hpp:207-326)
template <typename T>
std::vector<patch_t<T, 2>> build_periodic_mesh_2d(index_space_t<2> domain, int nx_patches, int ny_patches,
                                                  std::size_t nvars, int halo = 1) {
    return build_mesh_2d<T>(domain, nx_patches, ny_patches, nvars, halo, boundary_t::periodic,
                            boundary_t::periodic, false);
}

template <std::size_t Dim>
std::vector<patch_t<double, Dim>> build_periodic_mesh(index_space_t<Dim> domain, int n_patches, std::size_t nvars,
                                                      int halo = 1) {
    return build_periodic_mesh_1d<double, Dim>(domain, n_patches, nvars, halo);
}

}  // namespace nest
