// TEST INFRASTRUCTURE ONLY — the parity oracle.
//
// A CPU restatement of nest's Euler finite-volume numerics, written from the
// reference's published algorithm, used only by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg, as the reference results that the HIP path is
// compared with. It is never linked into the
// product library (nest_amd/libnest_hip.so) and nothing in the product path
// calls it.
//
// Every function cites the reference file:line it restates. Evaluation order
// follows the reference source (C++ left-to-right association) so that, built
// with -ffp-contract=off and without fast-math, results are bitwise those of
// the reference CPU path. The scalar type T is a template parameter so the
// same code instantiates on forward-mode dual numbers (oracle/dual.hpp) for
// the adjoint check; the reference numerics are double-only.
//
// Parity pinning (see tests/test_oracle_*.py): every function here is checked
// against vectors produced by the reference's own headers compiled in
// oracle/_ref (committed under tests/golden/), against the reference's
// known-answer tests (tests/test_euler.cpp, tests/test_core.cpp), and against
// invariants (free-stream preservation, conservation, mirror symmetry). The
// reflecting wall, which the reference lacks, is pinned by the method of
// images (a walled box evolves as the mirror-doubled periodic box), closed-box
// conservation and brute-force image ghosts (tests/test_oracle_reflecting.py).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace nest_oracle {

// ---------------------------------------------------------------------------
// scalar helpers (overloaded for the dual type in dual.hpp)
// ---------------------------------------------------------------------------
inline double val(double x) { return x; }
inline double nsqrt(double x) { return std::sqrt(x); }
inline double nabs(double x) { return std::fabs(x); }
// std::min / std::max semantics: min(a,b) = (b<a)?b:a ; max(a,b) = (a<b)?b:a
template <class T> inline T nmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T> inline T nmax(T a, T b) { return (val(a) < val(b)) ? b : a; }

enum limiter_kind : int { LIM_MINMOD = 0, LIM_MC = 1, LIM_NONE = 2 };
// Boundary conditions. The reference has periodic ones only (its exchange
// plans, include/nest/patch.hpp:207-326); outflow and reflecting are
// extensions: outflow ghosts copy the boundary cell (zero gradient),
// reflecting ghosts are the mirror image of the interior with the normal
// momentum negated (a solid wall).
enum bc_kind : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTING = 2 };

// ---------------------------------------------------------------------------
// EOS — examples/sod1d/euler1d.hpp:27-62
// ---------------------------------------------------------------------------
template <class T>
inline T eos_pressure(T rho, T mom, T erg, double gamma) {  // euler1d.hpp:31-44
    T u = mom / rho;
    T ke = 0.5 * rho * u * u;
    T p = (gamma - 1.0) * (erg - ke);
    return nmax(p, T(0.0));
}
template <class T>
inline T eos_total_energy(T rho, T u, T p, double gamma) {  // euler1d.hpp:47-51
    T ke = 0.5 * rho * u * u;
    T ie = p / (gamma - 1.0);
    return ie + ke;
}
template <class T>
inline T eos_sound_speed(T rho, T p, double gamma) {  // euler1d.hpp:54-56
    return nsqrt(gamma * p / rho);
}

// cons_to_prim — euler1d.hpp:69-84 ; prim_to_cons — euler1d.hpp:87-100
template <class T>
inline void cons_to_prim(const T* U, T* W, double gamma) {
    T u = U[1] / U[0];
    T p = eos_pressure(U[0], U[1], U[2], gamma);
    W[0] = U[0];
    W[1] = u;
    W[2] = p;
}
template <class T>
inline void prim_to_cons(const T* W, T* U, double gamma) {
    U[0] = W[0];
    U[1] = W[0] * W[1];
    U[2] = eos_total_energy(W[0], W[1], W[2], gamma);
}

// ---------------------------------------------------------------------------
// limiters — euler1d.hpp:106-137
// ---------------------------------------------------------------------------
template <class T>
inline T minmod(T a, T b) {  // euler1d.hpp:112-115
    if (val(a * b) <= 0.0) return T(0.0);
    return (val(nabs(a)) < val(nabs(b))) ? a : b;
}
template <class T>
inline T mc_limiter(T a, T b) {  // euler1d.hpp:117-125
    T c = 0.5 * (a + b);
    return minmod(2.0 * a, minmod(c, 2.0 * b));
}
template <class T>
inline T apply_limiter(T l, T r, int lim) {  // euler1d.hpp:127-137
    switch (lim) {
        case LIM_MINMOD: return minmod(l, r);
        case LIM_MC: return mc_limiter(l, r);
        case LIM_NONE: return 0.5 * (l + r);
    }
    return T(0.0);
}

// PLM — euler1d.hpp:149-171. WL = left face (i-1/2) of cell 0, WR = right face.
template <class T>
inline void plm_reconstruct(const T* Wm, const T* W0, const T* Wp, int lim, T* WL, T* WR) {
    for (int k = 0; k < 3; ++k) {
        T dL = W0[k] - Wm[k];
        T dR = Wp[k] - W0[k];
        T s = apply_limiter(dL, dR, lim);
        WL[k] = W0[k] - 0.5 * s;
        WR[k] = W0[k] + 0.5 * s;
    }
}

// HLLE with Davis wave speeds — euler1d.hpp:178-242
template <class T>
inline void hlle_flux(const T* WL, const T* WR, T* F, double gamma) {
    T rL = WL[0], uL = WL[1], pL = WL[2];
    T cL = eos_sound_speed(rL, pL, gamma);
    T EL = eos_total_energy(rL, uL, pL, gamma);
    T rR = WR[0], uR = WR[1], pR = WR[2];
    T cR = eos_sound_speed(rR, pR, gamma);
    T ER = eos_total_energy(rR, uR, pR, gamma);
    T SL = nmin(uL - cL, uR - cR);
    T SR = nmax(uL + cL, uR + cR);
    T FL[3] = {rL * uL, rL * uL * uL + pL, (EL + pL) * uL};
    T FR[3] = {rR * uR, rR * uR * uR + pR, (ER + pR) * uR};
    T UL[3] = {rL, rL * uL, EL};
    T UR[3] = {rR, rR * uR, ER};
    if (val(SL) >= 0.0) {
        for (int k = 0; k < 3; ++k) F[k] = FL[k];
    } else if (val(SR) <= 0.0) {
        for (int k = 0; k < 3; ++k) F[k] = FR[k];
    } else {
        for (int k = 0; k < 3; ++k)
            F[k] = (SR * FL[k] - SL * FR[k] + SL * SR * (UR[k] - UL[k])) / (SR - SL);
    }
}

// ---------------------------------------------------------------------------
// 1D single-array stage with the full 5-point stencil.
// examples/benchmark/benchmark_euler1d.cpp:113-165 (OpenMP form) and
// tests/test_euler_cuda.cpp:109-158; limiter generalised through
// apply_limiter exactly as plm_reconstruct does (euler1d.hpp:157-168).
// Layout: U[var*(n+2h) + e], e = interior index + h.
// ---------------------------------------------------------------------------
template <class T>
void stage_1d(const T* Uin, T* Uout, int n, int h, double dt, double dx, double gamma, int lim,
              std::vector<T>& W) {
    const int ne = n + 2 * h;
    W.resize(3 * (size_t)ne);
    for (int e = 0; e < ne; ++e) {  // benchmark_euler1d.cpp:116-125
        T U[3] = {Uin[e], Uin[ne + e], Uin[2 * ne + e]};
        T w[3];
        cons_to_prim(U, w, gamma);
        for (int k = 0; k < 3; ++k) W[(size_t)k * ne + e] = w[k];
    }
#pragma omp parallel for schedule(static) if (n > 4096)
    for (int e = h; e < h + n; ++e) {  // benchmark_euler1d.cpp:131-164
        T WLm[3], WRm[3], WLp[3], WRp[3];
        for (int k = 0; k < 3; ++k) {
            const T* w = &W[(size_t)k * ne];
            T s_im1 = apply_limiter(w[e - 1] - w[e - 2], w[e] - w[e - 1], lim);
            T s_i = apply_limiter(w[e] - w[e - 1], w[e + 1] - w[e], lim);
            T s_ip1 = apply_limiter(w[e + 1] - w[e], w[e + 2] - w[e + 1], lim);
            WLm[k] = w[e - 1] + 0.5 * s_im1;
            WRm[k] = w[e] - 0.5 * s_i;
            WLp[k] = w[e] + 0.5 * s_i;
            WRp[k] = w[e + 1] - 0.5 * s_ip1;
        }
        T FL[3], FR[3];
        hlle_flux(WLm, WRm, FL, gamma);
        hlle_flux(WLp, WRp, FR, gamma);
        for (int k = 0; k < 3; ++k)
            Uout[(size_t)k * ne + e] = Uin[(size_t)k * ne + e] - (dt / dx) * (FR[k] - FL[k]);
    }
}

// ghost fill of a single 1D array: periodic as benchmark_euler1d.cpp:102-110,
// outflow (zero-gradient copy of the boundary cell) and reflecting (mirror
// image, momentum negated) as the build defines them.
template <class T>
void fill_1d(T* U, int n, int h, int nvars, int bc) {
    const int ne = n + 2 * h;
    for (int k = 0; k < nvars; ++k) {
        T* u = U + (size_t)k * ne;
        for (int g = 0; g < h; ++g) {
            if (bc == BC_PERIODIC) {
                u[g] = u[n + g];
                u[h + n + g] = u[h + g];
            } else if (bc == BC_REFLECTING) {
                // ghost h-1-j is the image of interior cell j (extended h+j)
                u[g] = k == 1 ? -u[2 * h - 1 - g] : u[2 * h - 1 - g];
                u[h + n + g] = k == 1 ? -u[h + n - 1 - g] : u[h + n - 1 - g];
            } else {
                u[g] = u[h];
                u[h + n + g] = u[h + n - 1];
            }
        }
    }
}

// max(|u|+c) over the interior — benchmark_euler1d.cpp:168-180, sod1d.cpp:97-128
template <class T>
double max_speed_1d(const T* U, int n, int h, double gamma) {
    const int ne = n + 2 * h;
    double ms = 0.0;
    for (int e = h; e < h + n; ++e) {
        T Uc[3] = {U[e], U[ne + e], U[2 * ne + e]};
        T W[3];
        cons_to_prim(Uc, W, gamma);
        T c = eos_sound_speed(W[0], W[2], gamma);
        double s = val(nabs(W[1]) + c);
        ms = std::max(ms, s);
    }
    return ms;
}

// 1D SSP-RK2 — benchmark_euler1d.cpp:188-209, sod1d.cpp:219-272
template <class T>
void rk2_step_1d(T* U, int n, int h, int bc, double dt, double dx, double gamma, int lim,
                 std::vector<T>& U1, std::vector<T>& U2, std::vector<T>& W) {
    const int ne = n + 2 * h;
    U1.assign(U, U + 3 * (size_t)ne);
    U2.assign(U, U + 3 * (size_t)ne);
    fill_1d(U, n, h, 3, bc);
    stage_1d(U, U1.data(), n, h, dt, dx, gamma, lim, W);
    fill_1d(U1.data(), n, h, 3, bc);
    stage_1d(U1.data(), U2.data(), n, h, dt, dx, gamma, lim, W);
    for (int k = 0; k < 3; ++k)
        for (int e = h; e < h + n; ++e) {
            size_t o = (size_t)k * ne + e;
            U[o] = 0.5 * (U[o] + U2[o]);
        }
}

// ---------------------------------------------------------------------------
// 2D numerics — examples/sod2d/euler2d.hpp
// Single ext array layout (field_t, last index fastest): (i = x = dim 0,
// j = y = dim 1), linear = i*(ny+2h) + j, var stride (nx+2h)*(ny+2h).
// ---------------------------------------------------------------------------

// get_prim_2d — euler2d.hpp:36-64. W = (rho, u_normal, p), q = transverse momentum.
template <class T>
inline void get_prim_2d(T r_in, T mu, T mv, T E, double gamma, bool xsweep, T* W, T& q) {
    T rho = nmax(r_in, T(1e-10));
    T u = mu / rho;
    T v = mv / rho;
    T ke = 0.5 * rho * (u * u + v * v);
    T p = nmax((gamma - 1.0) * (E - ke), T(1e-10));
    W[0] = rho;
    W[2] = p;
    if (xsweep) {
        W[1] = u;
        q = mv;
    } else {
        W[1] = v;
        q = mu;
    }
}

// compute_interface_velocity — euler2d.hpp:20-28
template <class T>
inline T interface_velocity(T uL, T uR, T cL, T cR) {
    T SL = nmin(uL - cL, uR - cR);
    T SR = nmax(uL + cL, uR + cR);
    if (val(SL) >= 0.0) return uL;
    if (val(SR) <= 0.0) return uR;
    return (SR * uL - SL * uR) / (SR - SL);
}
// passive_flux — euler2d.hpp:31-33
template <class T>
inline T passive_flux(T qL, T qR, T uint) {
    return (val(uint) >= 0.0) ? uint * qL : uint * qR;
}

// One directional sweep at one cell from its 5-cell stencil (cells m2..p2
// along the sweep direction). x_sweep_t::operator() euler2d.hpp:88-149 and
// y_sweep_t::operator() euler2d.hpp:176-236 are mirror images; this is the
// shared body. Ucell[4] is the centre cell's conserved state (rho, mu, mv, E);
// returns the updated state in out[4].
template <class T>
inline void sweep_cell(const T (*S)[4], double gamma, int lim, bool xsweep, double dtdx, T* out) {
    T W[5][3], q[5];
    for (int s = 0; s < 5; ++s) get_prim_2d(S[s][0], S[s][1], S[s][2], S[s][3], gamma, xsweep, W[s], q[s]);
    // left interface
    T a_L[3], a_R[3], b_L[3], b_R[3], c_L[3], c_R[3];
    plm_reconstruct(W[0], W[1], W[2], lim, a_L, a_R);  // recon_im1
    plm_reconstruct(W[1], W[2], W[3], lim, b_L, b_R);  // recon_i
    T FL[3];
    hlle_flux(a_R, b_L, FL, gamma);
    T cLL = eos_sound_speed(a_R[0], a_R[2], gamma);
    T cLR = eos_sound_speed(b_L[0], b_L[2], gamma);
    T uintL = interface_velocity(a_R[1], b_L[1], cLL, cLR);
    T qLL = q[1] + 0.5 * apply_limiter(q[1] - q[0], q[2] - q[1], lim);
    T qLR = q[2] - 0.5 * apply_limiter(q[2] - q[1], q[3] - q[2], lim);
    T GL = passive_flux(qLL, qLR, uintL);
    // right interface
    plm_reconstruct(W[2], W[3], W[4], lim, c_L, c_R);  // recon_ip1
    T FR[3];
    hlle_flux(b_R, c_L, FR, gamma);
    T cRL = eos_sound_speed(b_R[0], b_R[2], gamma);
    T cRR = eos_sound_speed(c_L[0], c_L[2], gamma);
    T uintR = interface_velocity(b_R[1], c_L[1], cRL, cRR);
    T qRL = q[2] + 0.5 * apply_limiter(q[2] - q[1], q[3] - q[2], lim);
    T qRR = q[3] - 0.5 * apply_limiter(q[3] - q[2], q[4] - q[3], lim);
    T GR = passive_flux(qRL, qRR, uintR);
    const T* U = S[2];
    out[0] = U[0] - dtdx * (FR[0] - FL[0]);
    if (xsweep) {
        out[1] = U[1] - dtdx * (FR[1] - FL[1]);
        out[2] = U[2] - dtdx * (GR - GL);
    } else {
        out[1] = U[1] - dtdx * (GR - GL);
        out[2] = U[2] - dtdx * (FR[1] - FL[1]);
    }
    out[3] = U[3] - dtdx * (FR[2] - FL[2]);
}

// x_sweep_t / y_sweep_t on one ext array, interior rows [r0, r1) of dim 0.
// Writes the interior of Uout only (as the reference writes a field over the
// interior, sod2d.cpp:313-340).
template <class T>
void sweep_2d_rows(const T* Uin, T* Uout, int nx, int ny, int h, double dt, double dxy, double gamma,
                   int lim, bool xsweep, int r0, int r1) {
    const int nye = ny + 2 * h;
    const size_t vs = (size_t)(nx + 2 * h) * nye;
    const double dtdx = dt / dxy;
    for (int i = r0; i < r1; ++i) {
        for (int j = 0; j < ny; ++j) {
            const int ei = i + h, ej = j + h;
            T S[5][4];
            for (int s = 0; s < 5; ++s) {
                int ii = xsweep ? ei + s - 2 : ei;
                int jj = xsweep ? ej : ej + s - 2;
                size_t o = (size_t)ii * nye + jj;
                for (int k = 0; k < 4; ++k) S[s][k] = Uin[k * vs + o];
            }
            T out[4];
            sweep_cell<T>(S, gamma, lim, xsweep, dtdx, out);
            size_t o = (size_t)ei * nye + ej;
            for (int k = 0; k < 4; ++k) Uout[k * vs + o] = out[k];
        }
    }
}

// compute_dt_2d — euler2d.hpp:242-271, per interior cell; returns (dt_x, dt_y)
// partial minima so meshes can fold them across patches.
template <class T>
void dt_2d_partial(const T* U, int nx, int ny, int h, double dx, double dy, double gamma, double& dtx,
                   double& dty) {
    const int nye = ny + 2 * h;
    const size_t vs = (size_t)(nx + 2 * h) * nye;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            size_t o = (size_t)(i + h) * nye + (j + h);
            double rho = std::max(val(U[o]), 1e-10);
            double u = val(U[vs + o]) / rho;
            double v = val(U[2 * vs + o]) / rho;
            double ke = 0.5 * rho * (u * u + v * v);
            double p = std::max((gamma - 1.0) * (val(U[3 * vs + o]) - ke), 1e-10);
            double c = std::sqrt(gamma * p / rho);
            double sx = std::fabs(u) + c;
            double sy = std::fabs(v) + c;
            if (sx > 0.0) dtx = std::min(dtx, dx / sx);
            if (sy > 0.0) dty = std::min(dty, dy / sy);
        }
}

// ---------------------------------------------------------------------------
// 2D patch mesh. Tensor-product partition with subspace() semantics
// (include/nest/core.hpp:184-199), patch index = py*npx + px
// (include/nest/patch.hpp:219-229). All patches' field_t arrays are
// concatenated in patch-index order.
// ---------------------------------------------------------------------------
struct mesh2d {
    int nx = 0, ny = 0, npx = 1, npy = 1, h = 2;
    int bcx = BC_OUTFLOW, bcy = BC_OUTFLOW;
    std::vector<int> x0, lx, y0, ly;  // per px / per py partition
    std::vector<size_t> off;          // element offset of each patch array
    std::vector<int> owner_x, owner_y;
    size_t total = 0;

    mesh2d() = default;
    mesh2d(int nx_, int ny_, int npx_, int npy_, int h_, int bcx_, int bcy_)
        : nx(nx_), ny(ny_), npx(npx_), npy(npy_), h(h_), bcx(bcx_), bcy(bcy_) {
        auto part = [](int total, int n, std::vector<int>& s, std::vector<int>& l) {
            int base = total / n, extra = total % n;
            for (int p = 0; p < n; ++p) {
                s.push_back(p * base + std::min(p, extra));
                l.push_back(base + (p < extra ? 1 : 0));
            }
        };
        part(nx, npx, x0, lx);
        part(ny, npy, y0, ly);
        owner_x.resize(nx);
        owner_y.resize(ny);
        for (int p = 0; p < npx; ++p)
            for (int i = 0; i < lx[p]; ++i) owner_x[x0[p] + i] = p;
        for (int p = 0; p < npy; ++p)
            for (int j = 0; j < ly[p]; ++j) owner_y[y0[p] + j] = p;
        for (int py = 0; py < npy; ++py)
            for (int px = 0; px < npx; ++px) {
                off.push_back(total);
                total += 4 * (size_t)(lx[px] + 2 * h) * (ly[py] + 2 * h);
            }
    }
    int npatch() const { return npx * npy; }
    int pxof(int p) const { return p % npx; }
    int pyof(int p) const { return p / npx; }
    static int mapc(int g, int n, int bc) {
        if (bc == BC_PERIODIC) return ((g % n) + n) % n;
        if (bc == BC_REFLECTING) return g < 0 ? -1 - g : (g >= n ? 2 * n - 1 - g : g);  // |g| within a halo
        return g < 0 ? 0 : (g >= n ? n - 1 : g);
    }
    // a coordinate past a reflecting wall (its normal momentum is negated)
    static bool mirrored(int g, int n, int bc) { return bc == BC_REFLECTING && (g < 0 || g >= n); }
};

// Halo fill over the whole mesh: every ghost cell (faces AND corners) takes the
// value of the global cell obtained by mapping each coordinate through its
// boundary condition (periodic wrap / outflow clamp / reflecting mirror, the
// last negating the normal momentum of each mirrored direction: rho u for x,
// rho v for y). On faces of a periodic
// mesh this is exactly include/nest/patch.hpp:153-190 driven by the plans of
// build_periodic_mesh_2d (patch.hpp:207-326); corners, outflow and reflecting
// walls extend it.
template <class T>
void fill_mesh(T* U, const mesh2d& m) {
    const int h = m.h;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < m.npatch(); ++p) {
        const int px = m.pxof(p), py = m.pyof(p);
        const int nxe = m.lx[px] + 2 * h, nye = m.ly[py] + 2 * h;
        const size_t vs = (size_t)nxe * nye;
        T* dst = U + m.off[p];
        for (int ei = 0; ei < nxe; ++ei)
            for (int ej = 0; ej < nye; ++ej) {
                bool inter = ei >= h && ei < nxe - h && ej >= h && ej < nye - h;
                if (inter) continue;
                const int ri = m.x0[px] + ei - h, rj = m.y0[py] + ej - h;
                int gi = m.mapc(ri, m.nx, m.bcx);
                int gj = m.mapc(rj, m.ny, m.bcy);
                const bool fx = m.mirrored(ri, m.nx, m.bcx), fy = m.mirrored(rj, m.ny, m.bcy);
                int qx = m.owner_x[gi], qy = m.owner_y[gj];
                int q = qy * m.npx + qx;
                const int qnye = m.ly[qy] + 2 * h;
                const size_t qvs = (size_t)(m.lx[qx] + 2 * h) * qnye;
                const T* src = U + m.off[q];
                size_t so = (size_t)(gi - m.x0[qx] + h) * qnye + (gj - m.y0[qy] + h);
                size_t d_o = (size_t)ei * nye + ej;
                for (int k = 0; k < 4; ++k) {
                    const T v = src[k * qvs + so];
                    dst[k * vs + d_o] = (k == 1 && fx) || (k == 2 && fy) ? -v : v;
                }
            }
    }
}

// One directional sweep over all patches, Uout interior <- sweep(Uin).
template <class T>
void sweep_mesh(const T* Uin, T* Uout, const mesh2d& m, double dt, double dxy, double gamma, int lim,
                bool xsweep) {
    // flatten (patch, row) so a handful of big patches still spreads over threads
    std::vector<std::pair<int, int>> rows;
    for (int p = 0; p < m.npatch(); ++p)
        for (int i = 0; i < m.lx[m.pxof(p)]; ++i) rows.push_back({p, i});
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)rows.size(); ++r) {
        int p = rows[r].first, i = rows[r].second;
        const int px = m.pxof(p), py = m.pyof(p);
        sweep_2d_rows(Uin + m.off[p], Uout + m.off[p], m.lx[px], m.ly[py], m.h, dt, dxy, gamma, lim,
                      xsweep, i, i + 1);
    }
}

template <class T>
void copy_interior(const T* src, T* dst, const mesh2d& m) {
#pragma omp parallel for schedule(static)
    for (int p = 0; p < m.npatch(); ++p) {
        const int px = m.pxof(p), py = m.pyof(p);
        const int nxe = m.lx[px] + 2 * m.h, nye = m.ly[py] + 2 * m.h;
        const size_t vs = (size_t)nxe * nye;
        for (int k = 0; k < 4; ++k)
            for (int ei = m.h; ei < nxe - m.h; ++ei)
                for (int ej = m.h; ej < nye - m.h; ++ej) {
                    size_t o = m.off[p] + k * vs + (size_t)ei * nye + ej;
                    dst[o] = src[o];
                }
    }
}

// Split stage S(U) = Y o fill o X o fill (U), in place on U, following the
// exchange/sweep/copy order of examples/sod2d/sod2d.cpp:328-356.
template <class T>
void split_stage_mesh(T* U, T* scratch, const mesh2d& m, double dt, double dx, double dy, double gamma,
                      int lim) {
    fill_mesh(U, m);
    sweep_mesh(U, scratch, m, dt, dx, gamma, lim, true);
    copy_interior(scratch, U, m);
    fill_mesh(U, m);
    sweep_mesh(U, scratch, m, dt, dy, gamma, lim, false);
    copy_interior(scratch, U, m);
}

// 2D SSP-RK2 (no 2D RK2 driver exists on the reference's CPU path):
// U1 = S(Un); U2 = S(U1); U^{n+1} = 0.5 (Un + U2)
// (the "2 stages x 2 sweeps" step of docs/PERFORMANCE.md:25-28 and
// src/backend_cuda/euler2d_cuda.cuh:610-630). The result's halos are refilled
// so the returned state is halo-complete.
template <class T>
void rk2_step_mesh(T* U, const mesh2d& m, double dt, double dx, double dy, double gamma, int lim,
                   std::vector<T>& U1, std::vector<T>& scratch) {
    U1.assign(U, U + m.total);
    scratch.assign(U, U + m.total);
    split_stage_mesh(U1.data(), scratch.data(), m, dt, dx, dy, gamma, lim);
    split_stage_mesh(U1.data(), scratch.data(), m, dt, dx, dy, gamma, lim);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < m.npatch(); ++p) {
        const int px = m.pxof(p), py = m.pyof(p);
        const int nxe = m.lx[px] + 2 * m.h, nye = m.ly[py] + 2 * m.h;
        const size_t vs = (size_t)nxe * nye;
        for (int k = 0; k < 4; ++k)
            for (int ei = m.h; ei < nxe - m.h; ++ei)
                for (int ej = m.h; ej < nye - m.h; ++ej) {
                    size_t o = m.off[p] + k * vs + (size_t)ei * nye + ej;
                    U[o] = 0.5 * (U[o] + U1[o]);
                }
    }
    fill_mesh(U, m);
}

// compute_dt_2d over the mesh — euler2d.hpp:242-271. Returns cfl*min(dt_x,dt_y).
template <class T>
double dt_mesh(const T* U, const mesh2d& m, double dx, double dy, double cfl, double gamma) {
    double dtx = 1e10, dty = 1e10;
    for (int p = 0; p < m.npatch(); ++p) {
        const int px = m.pxof(p), py = m.pyof(p);
        dt_2d_partial(U + m.off[p], m.lx[px], m.ly[py], m.h, dx, dy, gamma, dtx, dty);
    }
    return cfl * std::min(dtx, dty);
}

}  // namespace nest_oracle
