// TEST INFRASTRUCTURE ONLY — C entry points around the reference's OWN
// numerics headers, compiled from where they lie under the reference tree
// (see oracle/Makefile, target `ref`) into oracle/_ref/libnest_ref.so.
// Nothing of the reference is copied here: this file only includes its
// headers and drives them, exactly as its examples do. It is used to write
// the golden vectors under tests/golden/ and, where the reference tree is
// present, to cross-check the oracle directly.
#include <cstring>
#include <vector>

#include "nest/core.hpp"
#include "nest/patch.hpp"
#include "sod2d/euler2d.hpp"

using namespace nest;
using namespace nest::euler1d;
using namespace nest::euler2d;

static limiter_t lim_of(int l) { return l == 1 ? limiter_t::mc : (l == 2 ? limiter_t::none : limiter_t::minmod); }

extern "C" {

double ref_eos_pressure(double r, double m, double e, double g) { return eos_t{g}.pressure(r, m, e); }
double ref_eos_total_energy(double r, double u, double p, double g) { return eos_t{g}.total_energy(r, u, p); }
double ref_eos_sound_speed(double r, double p, double g) { return eos_t{g}.sound_speed(r, p); }
void ref_cons_to_prim(const double* U, double* W, double g) { cons_to_prim(U, W, eos_t{g}); }
void ref_prim_to_cons(const double* W, double* U, double g) { prim_to_cons(W, U, eos_t{g}); }
double ref_minmod(double a, double b) { return minmod(a, b); }
double ref_mc(double a, double b) { return mc_limiter(a, b); }
double ref_limiter(double a, double b, int lim) { return apply_limiter(a, b, lim_of(lim)); }
void ref_plm_batch(int n, const double* Wm, const double* W0, const double* Wp, int lim, double* WL, double* WR) {
    for (int i = 0; i < n; ++i) {
        auto r = plm_reconstruct(Wm + 3 * i, W0 + 3 * i, Wp + 3 * i, lim_of(lim));
        std::memcpy(WL + 3 * i, r.W_L, sizeof(r.W_L));
        std::memcpy(WR + 3 * i, r.W_R, sizeof(r.W_R));
    }
}
void ref_hlle_batch(int n, const double* WL, const double* WR, double g, double* F) {
    eos_t eos{g};
    for (int i = 0; i < n; ++i) hlle_flux(WL + 3 * i, WR + 3 * i, F + 3 * i, eos);
}
double ref_interface_velocity(double uL, double uR, double cL, double cR) {
    return compute_interface_velocity(uL, uR, cL, cR);
}
double ref_passive_flux(double qL, double qR, double ui) { return passive_flux(qL, qR, ui); }

// euler_rhs_t (euler1d.hpp:279-397) on a single 1D patch: U arrays hold the
// extended space [-h, n+h) in field_t layout; writes the interior of Uout.
void ref_rhs_1d(const double* Uin, double* Uout, int n, int h, double dt, double dx, double g, int lim) {
    auto interior = index_space(ivec(0), uvec(static_cast<unsigned>(n)));
    auto ext = expand(interior, h);
    field_t<double, 1> fin(const_cast<double*>(Uin), NCONS, ext);
    field_t<double, 1> fout(Uout, NCONS, ext);
    eos_t eos{g};
    euler_rhs_t rhs(eos, dx, lim_of(lim));
    rhs(fin, fout, interior, dt);
}

// x_sweep_t / y_sweep_t (euler2d.hpp:67-239) on one 2D patch of nx*ny with
// halo h, arrays over the extended space; writes the interior of Uout.
void ref_sweep_2d(const double* Uin, double* Uout, int nx, int ny, int h, double dt, double dxy, double g,
                  int lim, int xsweep) {
    auto interior = index_space(ivec(0, 0), uvec(static_cast<unsigned>(nx), static_cast<unsigned>(ny)));
    auto ext = expand(interior, h);
    field_t<double, 2> fin(const_cast<double*>(Uin), 4, ext);
    field_t<double, 2> fout(Uout, 4, ext);
    eos_t eos{g};
    if (xsweep) {
        x_sweep_t s(eos, dxy, lim_of(lim));
        s(fin, fout, interior, dt);
    } else {
        y_sweep_t s(eos, dxy, lim_of(lim));
        s(fin, fout, interior, dt);
    }
}

// Reference periodic 2D mesh (build_periodic_mesh_2d, patch.hpp:207-326):
// loads every patch's extended array from U (patch-index order), then
//   mode 0: execute_all_exchanges (faces only)
//   mode 1: one Lie step X then Y with exchanges, as sod2d.cpp:328-356
// and writes the patches back. Returns compute_dt_2d (euler2d.hpp:242-271)
// of the input state.
double ref_periodic_mesh(double* U, int nx, int ny, int npx, int npy, int h, int mode, double dt, double dx,
                         double dy, double cfl, double g, int lim) {
    auto domain = index_space(ivec(0, 0), uvec(static_cast<unsigned>(nx), static_cast<unsigned>(ny)));
    auto patches = build_periodic_mesh_2d<double>(domain, npx, npy, 4, h);
    size_t off = 0;
    for (auto& p : patches) {
        size_t n = 4 * size(p.extended);
        std::memcpy(p.state.conserved._data, U + off, n * sizeof(double));
        off += n;
    }
    eos_t eos{g};
    double dt_ref = compute_dt_2d(patches, dx, dy, cfl, eos);
    if (mode == 0) {
        execute_all_exchanges(patches);
    } else {
        x_sweep_t xs(eos, dx, lim_of(lim));
        y_sweep_t ys(eos, dy, lim_of(lim));
        std::vector<field_t<double, 2>> work;
        for (auto& p : patches) work.emplace_back(4, p.interior, field_init_t::uninitialized);
        execute_all_exchanges(patches);
        for (size_t q = 0; q < patches.size(); ++q) {
            xs(patches[q].state.conserved, work[q], patches[q].interior, dt);
            for (int k = 0; k < 4; ++k)
                for (auto idx : patches[q].interior) patches[q](k, idx) = work[q](k, idx);
        }
        execute_all_exchanges(patches);
        for (size_t q = 0; q < patches.size(); ++q) {
            ys(patches[q].state.conserved, work[q], patches[q].interior, dt);
            for (int k = 0; k < 4; ++k)
                for (auto idx : patches[q].interior) patches[q](k, idx) = work[q](k, idx);
        }
    }
    off = 0;
    for (auto& p : patches) {
        size_t n = 4 * size(p.extended);
        std::memcpy(U + off, p.state.conserved._data, n * sizeof(double));
        off += n;
    }
    return dt_ref;
}

}  // extern "C"
