// Launch interface of the forward kernels (kernels.hip). The file is compiled
// once per numerics mode; each compilation defines the same launchers inside
// its own namespace (nest_exact / nest_fast), selected at run time by the
// host runtime (nest_hip.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "nest_types.hpp"

#define NEST_DECLARE_LAUNCHERS(NS)                                                                  \
    namespace NS {                                                                                  \
    void launch_stage2d(int stage, int lim, const stage_args& a, int grid, int block, hipStream_t s); \
    void launch_step2d(int lim, const stage_args& a, int grid, int block, hipStream_t s);           \
    int step2d_resident_blocks(int lim, int block);                                                 \
    void launch_speeds2d(const stage_args& a, int grid, int block, hipStream_t s);                  \
    void launch_fill_regions(double* base, const nest_region* R, int nreg, int nvars, hipStream_t s); \
    void launch_finalize(nest_prob_state* P, unsigned long long* slots, unsigned long long* slots_clr,  \
                         int nprob, int par, hipStream_t s);                                          \
    void launch_reset_speeds(nest_prob_state* P, unsigned long long* slots, int nprob, hipStream_t s); \
    void launch_stage1d(int stage, int lim, const stage1d_args& a, hipStream_t s);                  \
    void launch_step1d(int lim, const stage1d_args& a, hipStream_t s);                              \
    void prepare_step1d(int lim);                                                                   \
    void launch_speeds1d(const stage1d_args& a, hipStream_t s);                                     \
    void launch_fill1d(double* U, int n, int bc, const nest_prob_state* ps, hipStream_t s);         \
    void launch_adj_y(int lim, const adj_args& a, int grid, int block, hipStream_t s);              \
    void launch_adj_x(int lim, const adj_args& a, int grid, int block, hipStream_t s);              \
    void launch_pack(const double* src, double* dst, const nest_patch_desc* P, int npatch,          \
                     int max_nx, const nest_prob_geom* geoms, const long long* goff, int dir,        \
                     hipStream_t s);                                                                 \
    }

// arguments of the 2D stage / speed kernels
struct stage_args {
    double* U;   // U^n, updated in place to U^{n+1} by stage 2 unless Uo is set
    double* U1;  // first-stage state
    const nest_patch_desc* patches;
    const nest_item* items;
    int nitems;
    nest_prob_state* probs;
    nest_eos eos;
    double host_dt;  // > 0: use this dt for every problem (ignores done flags)
    // gather != 0: items are global strips and every cell (halo included) is
    // read from the patch that owns it (nest_gather_source); 0: items are
    // per-patch strips reading the patch's own ghost cells.
    int gather;
    const nest_prob_geom* geoms;
    const int* itab;
    double* Uo;  // stage 2 output (null: in place of U)
    long long total;  // doubles in each state buffer (all share the patch layout)
    unsigned long long* slots;  // stage 2: per-problem wave-speed slots (NEST_SPEED_SLOTS each)
    int kernel;  // stage kernel: 0 automatic, 1 register path, 2 LDS-ring path (state below 4 GiB)
    // one-launch step with the dt computed in the kernel (device dt): the
    // slot set the previous step wrote (read), the set the next step writes
    // (cleared by each problem's first work item), each problem's first work
    // item, the step's time parity. speeds2d: slots_rd, if set, receives the
    // speeds instead of the problems' smax.
    unsigned long long* slots_rd;
    unsigned long long* slots_clr;
    const int* first_item;
    int par;
};

// arguments of the two adjoint kernels of one split stage S = Y o X (the
// halo fills included): the Y kernel maps the cotangent g*G of S(U) to the
// cotangent Vb of the X-swept state; the X kernel maps Vb to
// out = X'(U)^T Vb + h*H. All buffers are in the state's patch layout; only
// interior cells are read or written.
struct adj_args {
    const double* U;  // stage input state
    const double* G;  // Y kernel: cotangent of the stage output
    double g;         // its scale
    double* Vb;       // Y kernel output, X kernel input
    const double* H;  // X kernel: added term (null: none); may alias out
    double h;
    double* out;      // X kernel output
    const nest_patch_desc* patches;
    const nest_item* items;
    int nitems;
    const nest_prob_state* probs;
    nest_eos eos;
    double dt;
    const nest_prob_geom* geoms;
    const int* itab;
    int outflow_x;  // some problem has outflow x boundaries
};

// arguments of the 1D kernels (one problem, one patch)
struct stage1d_args {
    double* U;
    double* U1;
    int n;
    nest_prob_state* ps;
    nest_eos eos;
    double host_dt;  // > 0: use this dt (ignores the done flag)
    int bc;          // ghost cells are read through this boundary condition
    unsigned long long* slots;  // NEST_SPEED_SLOTS speed slots the launch folds its speeds into (device dt)
    // fused 1D device-dt step only: the time-stepping state, the slot set the
    // previous step wrote (this step's dt), the set to clear for the next
    // step, and the parity of the step
    nest_run1d_state* run;
    const unsigned long long* slots_rd;
    unsigned long long* slots_clr;
    int par;
    // fused step kernel: lanes 0 the LDS-tile kernel (256 cells per block),
    // 1 a lane kernel: one cell per lane in 56-cell segments, or (pairs) two
    // cells per lane in 120-cell segments. waves: the lane kernel launch's
    // wave count (set by launch_step1d); wave w takes segments w, w + waves,
    // w + 2 waves, ...
    int lanes;
    int pairs;
    int waves;
};

NEST_DECLARE_LAUNCHERS(nest_exact)
NEST_DECLARE_LAUNCHERS(nest_fast)
