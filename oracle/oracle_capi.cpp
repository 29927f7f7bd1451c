// TEST INFRASTRUCTURE ONLY — C entry points of the parity oracle
// (oracle/liboracle.so), loaded by tests/ and bench.py's cpu_baseline through
// ctypes. Plain pointers and sizes; all arrays use the field_t SoA layout
// (include/nest/core.hpp:310-392 of the reference: data[var*size + linear],
// last index fastest).
#include <omp.h>

#include <cstring>
#include <vector>

#include "dual.hpp"
#include "euler_oracle.hpp"

using namespace nest_oracle;

extern "C" {

// OpenMP threads used by the mesh / 1D loops (the cpu_baseline sets this)
void or_set_threads(int n) { omp_set_num_threads(n); }
int or_get_threads(void) { return omp_get_max_threads(); }

// ---- scalar numerics (euler1d.hpp:27-242, euler2d.hpp:20-33) ----
double or_eos_pressure(double r, double m, double e, double g) { return eos_pressure(r, m, e, g); }
double or_eos_total_energy(double r, double u, double p, double g) { return eos_total_energy(r, u, p, g); }
double or_eos_sound_speed(double r, double p, double g) { return eos_sound_speed(r, p, g); }
void or_cons_to_prim(const double* U, double* W, double g) { cons_to_prim(U, W, g); }
void or_prim_to_cons(const double* W, double* U, double g) { prim_to_cons(W, U, g); }
double or_minmod(double a, double b) { return minmod(a, b); }
double or_mc(double a, double b) { return mc_limiter(a, b); }
double or_limiter(double a, double b, int lim) { return apply_limiter(a, b, lim); }
void or_plm(const double* Wm, const double* W0, const double* Wp, int lim, double* WL, double* WR) {
    plm_reconstruct(Wm, W0, Wp, lim, WL, WR);
}
void or_hlle(const double* WL, const double* WR, double g, double* F) { hlle_flux(WL, WR, F, g); }
double or_interface_velocity(double uL, double uR, double cL, double cR) {
    return interface_velocity(uL, uR, cL, cR);
}
double or_passive_flux(double qL, double qR, double ui) { return passive_flux(qL, qR, ui); }

// batched forms for brute-force comparisons: n independent cases
void or_hlle_batch(int n, const double* WL, const double* WR, double g, double* F) {
    for (int i = 0; i < n; ++i) hlle_flux(WL + 3 * i, WR + 3 * i, F + 3 * i, g);
}
void or_plm_batch(int n, const double* Wm, const double* W0, const double* Wp, int lim, double* WL,
                  double* WR) {
    for (int i = 0; i < n; ++i) plm_reconstruct(Wm + 3 * i, W0 + 3 * i, Wp + 3 * i, lim, WL + 3 * i, WR + 3 * i);
}
// sweep of one cell from a 5-cell stencil S[5][4]
void or_sweep_cell(const double* S, double g, int lim, int xsweep, double dtdx, double* out) {
    sweep_cell<double>(reinterpret_cast<const double(*)[4]>(S), g, lim, xsweep != 0, dtdx, out);
}

// ---- 1D single array ----
void or_fill_1d(double* U, int n, int h, int nvars, int bc) { fill_1d(U, n, h, nvars, bc); }
void or_stage_1d(const double* Uin, double* Uout, int n, int h, double dt, double dx, double g, int lim) {
    std::vector<double> W;
    stage_1d(Uin, Uout, n, h, dt, dx, g, lim, W);
}
double or_max_speed_1d(const double* U, int n, int h, double g) { return max_speed_1d(U, n, h, g); }
void or_rk2_step_1d(double* U, int n, int h, int bc, double dt, double dx, double g, int lim) {
    std::vector<double> U1, U2, W;
    rk2_step_1d(U, n, h, bc, dt, dx, g, lim, U1, U2, W);
}
// Run to t_end (dt = min(cfl*dx/max_speed, t_end - t), sod1d.cpp:219-222).
// dt_out (may be null) receives each step's dt. Returns the step count, or -1
// if max_steps was reached first.
int or_run_1d(double* U, int n, int h, int bc, double t_end, double cfl, double dx, double g, int lim,
              int max_steps, double* dt_out) {
    std::vector<double> U1, U2, W;
    double t = 0.0;
    int steps = 0;
    while (t < t_end) {
        if (steps >= max_steps) return -1;
        double ms = max_speed_1d(U, n, h, g);
        double dt = cfl * dx / ms;
        dt = std::min(dt, t_end - t);
        if (dt_out) dt_out[steps] = dt;
        rk2_step_1d(U, n, h, bc, dt, dx, g, lim, U1, U2, W);
        t += dt;
        ++steps;
    }
    fill_1d(U, n, h, 3, bc);
    return steps;
}

// ---- 2D single ext array ----
void or_sweep_2d(const double* Uin, double* Uout, int nx, int ny, int h, double dt, double dxy, double g,
                 int lim, int xsweep) {
    sweep_2d_rows(Uin, Uout, nx, ny, h, dt, dxy, g, lim, xsweep != 0, 0, nx);
}
void or_dt_2d_partial(const double* U, int nx, int ny, int h, double dx, double dy, double g, double* dtxy) {
    double dtx = 1e10, dty = 1e10;
    dt_2d_partial(U, nx, ny, h, dx, dy, g, dtx, dty);
    dtxy[0] = dtx;
    dtxy[1] = dty;
}

// ---- 2D patch mesh (all patches concatenated in patch-index order) ----
size_t or_mesh_size(int nx, int ny, int npx, int npy, int h) {
    return mesh2d(nx, ny, npx, npy, h, 0, 0).total;
}
void or_mesh_fill(double* U, int nx, int ny, int npx, int npy, int h, int bcx, int bcy) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    fill_mesh(U, m);
}
void or_mesh_sweep(double* U, int nx, int ny, int npx, int npy, int h, int bcx, int bcy, double dt,
                   double dxy, double g, int lim, int xsweep) {
    // fill, sweep into scratch, copy interior back (one Lie half-step)
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<double> s(U, U + m.total);
    fill_mesh(U, m);
    sweep_mesh(U, s.data(), m, dt, dxy, g, lim, xsweep != 0);
    copy_interior(s.data(), U, m);
}
void or_mesh_split_stage(double* U, int nx, int ny, int npx, int npy, int h, int bcx, int bcy, double dt,
                         double dx, double dy, double g, int lim) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<double> s(U, U + m.total);
    split_stage_mesh(U, s.data(), m, dt, dx, dy, g, lim);
    fill_mesh(U, m);
}
void or_mesh_rk2_step(double* U, int nx, int ny, int npx, int npy, int h, int bcx, int bcy, double dt,
                      double dx, double dy, double g, int lim) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<double> U1, s;
    rk2_step_mesh(U, m, dt, dx, dy, g, lim, U1, s);
}
double or_mesh_dt(const double* U, int nx, int ny, int npx, int npy, int h, double dx, double dy,
                  double cfl, double g) {
    mesh2d m(nx, ny, npx, npy, h, 0, 0);
    return dt_mesh(U, m, dx, dy, cfl, g);
}
// nsteps RK2 steps; dt from the state each step (compute_dt_2d), clamped to
// t_end - t as sod2d.cpp:319-320. fixed_dt > 0 holds dt fixed instead.
// dt_out (may be null) receives each dt. Returns steps taken.
int or_mesh_run(double* U, int nx, int ny, int npx, int npy, int h, int bcx, int bcy, int nsteps,
                double t_end, double cfl, double fixed_dt, double dx, double dy, double g, int lim,
                double* dt_out) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<double> U1, s;
    double t = 0.0;
    int k = 0;
    for (; k < nsteps && t < t_end; ++k) {
        double dt = fixed_dt > 0.0 ? fixed_dt : dt_mesh(U, m, dx, dy, cfl, g);
        dt = std::min(dt, t_end - t);
        if (dt_out) dt_out[k] = dt;
        rk2_step_mesh(U, m, dt, dx, dy, g, lim, U1, s);
        t += dt;
    }
    return k;
}
// Forward-mode derivative of K fixed-dt RK2 steps: (U, dU) -> (U_K, dU_K).
void or_mesh_rk2_jvp(double* U, double* dU, int nx, int ny, int npx, int npy, int h, int bcx, int bcy,
                     int K, double dt, double dx, double dy, double g, int lim) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<dual> X(m.total), U1, s;
    for (size_t i = 0; i < m.total; ++i) X[i] = dual(U[i], dU[i]);
    for (int k = 0; k < K; ++k) rk2_step_mesh(X.data(), m, dt, dx, dy, g, lim, U1, s);
    for (size_t i = 0; i < m.total; ++i) {
        U[i] = X[i].v;
        dU[i] = X[i].d;
    }
}

// Forward-mode derivative of one split stage S = Y o fill o X o fill.
void or_mesh_stage_jvp(double* U, double* dU, int nx, int ny, int npx, int npy, int h, int bcx, int bcy,
                       double dt, double dx, double dy, double g, int lim) {
    mesh2d m(nx, ny, npx, npy, h, bcx, bcy);
    std::vector<dual> X(m.total), s(m.total);
    for (size_t i = 0; i < m.total; ++i) X[i] = dual(U[i], dU[i]);
    s = X;
    split_stage_mesh(X.data(), s.data(), m, dt, dx, dy, g, lim);
    for (size_t i = 0; i < m.total; ++i) {
        U[i] = X[i].v;
        dU[i] = X[i].d;
    }
}

}  // extern "C"
