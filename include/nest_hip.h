/*
 * nest_hip.h — C ABI of the MI355X (gfx950) backend for nest's Euler
 * finite-volume step: fused PLM + HLLE + conservative update (1D and 2D),
 * SSP-RK2 stages, on-device patch halo fill and the CFL wave-speed
 * reduction, as hand-written HIP kernels.
 *
 * Conventions
 *  - Plain pointers and sizes only; no C++ or torch types cross this ABI.
 *  - Every array is in nest's field_t SoA layout (reference
 *    include/nest/core.hpp:206-221, :310-392): data[var*size(ext) + linear],
 *    linear index row-major with the LAST index contiguous, over the patch's
 *    extended space (interior + 2-cell halo on every side). 1D: 3 variables
 *    (rho, rho*u, E); 2D: 4 variables (rho, rho*u, rho*v, E) with u along
 *    dim 0 (x) and v along dim 1 (y, contiguous).
 *  - Multi-patch buffers hold every patch's array back to back in patch-index
 *    order (patch index = py*npx + px, reference patch.hpp:219-229).
 *  - Functions return NEST_OK (0) or an error code; nest_hip_last_error()
 *    gives a message (thread-local). No exception crosses the ABI.
 *  - A handle owns its device buffers and one HIP stream (default priority)
 *    unless the caller supplies a stream. A handle is not thread-safe;
 *    different handles may be used concurrently.
 *  - Numerics mode: NEST_MODE_EXACT reproduces the reference CPU arithmetic
 *    bit for bit (IEEE division/sqrt, no FMA contraction); NEST_MODE_FAST
 *    uses reciprocal-multiply and FMA (agrees to ~1e-15 relative per step).
 */
#ifndef NEST_HIP_H
#define NEST_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nest_status {
    NEST_OK = 0,
    NEST_ERR_INVALID = 1,     /* bad argument / handle */
    NEST_ERR_ALLOC = 2,       /* device allocation failed */
    NEST_ERR_HIP = 3,         /* HIP runtime error */
    NEST_ERR_NONFINITE = 4,   /* a problem produced a non-finite dt */
    NEST_ERR_UNSUPPORTED = 5  /* configuration not supported */
};
/* Physical boundaries. PERIODIC and OUTFLOW are the reference's two
   (periodic wrap, zero-gradient copy of the edge cell). REFLECTING has no
   counterpart in the reference: a solid wall whose ghost cells mirror the
   interior across the wall with the wall-normal momentum negated. The forward
   path supports it; the adjoint entry points reject it (NEST_ERR_UNSUPPORTED). */
enum nest_bc { NEST_BC_PERIODIC = 0, NEST_BC_OUTFLOW = 1, NEST_BC_REFLECTING = 2 };
/* limiter_t of reference examples/sod1d/euler1d.hpp:106-110 */
enum nest_limiter { NEST_LIM_MINMOD = 0, NEST_LIM_MC = 1, NEST_LIM_NONE = 2 };
enum nest_mode { NEST_MODE_EXACT = 0, NEST_MODE_FAST = 1 };

/* Message of the last failing call on this thread ("" if none). */
const char* nest_hip_last_error(void);
/* ABI version (major*100 + minor). */
int nest_hip_version(void);
/* Number of visible HIP devices (0 if none); does not create a context. */
int nest_hip_device_count(void);
/* dst[i] = src[i], i < n, on `stream` (null: the null stream): a plain
   device copy, 8 B per lane, grid-stride. No counterpart in the reference:
   it is the measured HBM ceiling the stage kernels are compared with, and
   the known-size kernel that calibrates the memory-traffic counters. */
int nest_hip_device_copy(const double* src, double* dst, long long n, void* stream);

/* ===================================================================== */
/* 1D single-patch solver. Replaces the CUDA solver object of reference   */
/* src/backend_cuda/euler1d_solver.cu:35-189 (behind euler1d_solver.hpp). */
/* ===================================================================== */
typedef struct nest_solver1d nest_solver1d;

/* n interior cells, halo 2, cell size dx, gamma, bc (periodic/outflow/reflecting),
   limiter, mode. Replaces euler1d_gpu_create (euler1d_solver.hpp:27). */
int nest_hip_solver1d_create(int n, double dx, double gamma, int bc, int limiter, int mode,
                             nest_solver1d** out);
int nest_hip_solver1d_destroy(nest_solver1d* s); /* euler1d_solver.hpp:32 */
/* host buffer of 3*(n+4) doubles, field_t layout (euler1d_solver.hpp:41,49) */
int nest_hip_solver1d_upload(nest_solver1d* s, const double* host);
int nest_hip_solver1d_download(nest_solver1d* s, double* host);
/* dt = cfl*dx/max(|u|+c) of the current state; blocking
   (euler1d_solver.hpp:58, euler1d_solver.cu:100-119). */
int nest_hip_solver1d_compute_dt(nest_solver1d* s, double cfl, double* dt);
/* One SSP-RK2 step with the given dt (euler1d_solver.hpp:66): two stage
   launches that read ghost cells through the boundary condition, then a
   ghost fill so the state handed out is halo-complete. */
int nest_hip_solver1d_rk2_step(nest_solver1d* s, double dt);
/* Run to t_final with dt = min(cfl*dx/max(|u|+c), t_final - t) computed on
   the device every step (euler1d_solver.hpp:76). Each step is one fused
   launch (dt from the previous step's wave speeds, U1 kept on chip, the
   state alternating between the handle's two buffers); steps are replayed
   48 at a time from a graph captured on the first call, and the host polls
   completion once per 48.
   The final state is in the handle's state buffer, halo-complete. *steps
   receives the step count. */
int nest_hip_solver1d_run(nest_solver1d* s, double t_final, double cfl, int* steps);
/* Use the caller's stream (hipStream_t, NULL = handle's own). */
int nest_hip_solver1d_set_stream(nest_solver1d* s, void* stream);
/* Which kernel takes the fused device-dt steps of run / advance: 0
   automatic (kind 3 from 2^19 cells, kind 2 below), 1
   the LDS-tile kernel (256 cells per block, U1 in LDS), 2 the lane kernel
   (each wave on its own segments of 56 cells, U1 in registers, no barrier),
   3 the lane kernel with two cells per lane (segments of 120 cells; n at
   most 178956966, else NEST_ERR_UNSUPPORTED). All compute the same arithmetic.
   No counterpart in the reference. */
int nest_hip_solver1d_set_step_kernel(nest_solver1d* s, int kind);
/* One SSP-RK2 stage with the given dt, the per-stage granularity of a
   Pipeline Compute stage (the reference's 1D loop runs the stages as
   separate passes, examples/sod1d/sod1d.cpp:219-272; benchmark_euler1d.cpp:
   188-209): stage 1 writes U1 = S(U); stage 2 writes U = (U + S(U1)) / 2.
   Ghost cells are read through the boundary condition. */
int nest_hip_solver1d_stage(nest_solver1d* s, int stage, double dt);
/* Ghost fill of U (buffer 0) or U1 (buffer 1) through the boundary
   condition: the Exchange stage of a 1D pipeline. */
int nest_hip_solver1d_exchange(nest_solver1d* s, int buffer);
/* Device-dt stepping without a host round trip, by step count: start()
   sets the CFL number, the end time (t_end = INFINITY for none) and the
   initial wave speeds; advance() enqueues nsteps fused steps (the kernel of
   nest_hip_solver1d_run; steps past t_end are no-ops) and leaves the
   state in U, halo-complete. Each step's dt comes from the state it steps:
   after upload, stage or rk2_step the next advance reduces the speeds of
   the current state again first. status() reads t, the last dt and the
   step count (blocking). */
int nest_hip_solver1d_start(nest_solver1d* s, double cfl, double t_end);
int nest_hip_solver1d_advance(nest_solver1d* s, int nsteps);
int nest_hip_solver1d_status(nest_solver1d* s, double* t, double* dt, int* steps, int* done);

/* One device-dt SSP-RK2 step of a 1D state that allocates nothing:
   U_out = step(U_in) (device buffers of 3*(n+4) doubles each, distinct;
   U_out comes back halo-complete), with dt = min(cfl*dx/max(|u|+c) of U_in,
   t_end - t) computed on the device and written to *dt_out (a device
   double; NULL: not written). It runs the kernels of nest_hip_solver1d_start
   + nest_hip_solver1d_advance(1) (the speeds of U_in, then the fused step
   kernel, then the ghost fill), bit for bit the same result, but keeps
   their speed slots, run state and parameters in a fixed device array of
   the library (25.6 KB per GPU, held until the process ends) instead of
   handle memory: no hipMalloc / hipFree and no host synchronisation, on
   `stream` (NULL: the null stream). Every call on one GPU shares that
   array, so calls that overlap on one GPU (several streams, several host
   threads) race; serialise them. n is limited to 2^28 cells (32-bit
   indexing); larger n returns NEST_ERR_INVALID. No counterpart in the
   reference: for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid allocation). */
int nest_hip_euler1d_step_noalloc(const double* U_in, double* U_out, int n, double dx, double gamma, int bc,
                                  int limiter, int mode, double cfl, double t, double t_end, double* dt_out,
                                  void* stream);

/* ---- reference-compatible entry points: exact signatures of reference
   src/backend_cuda/euler1d_solver.hpp:16-87 (periodic BC, minmod, halo 2,
   exact numerics). Errors print to stderr and return NULL / -1 / NaN
   instead of throwing through extern "C". ---- */
struct euler1d_gpu_t;
struct euler1d_gpu_t* euler1d_gpu_create(int n_interior, int n_halo, double dx, double gamma);
void euler1d_gpu_destroy(struct euler1d_gpu_t* solver);
void euler1d_gpu_upload(struct euler1d_gpu_t* solver, const double* h_data);
void euler1d_gpu_download(struct euler1d_gpu_t* solver, double* h_data);
double euler1d_gpu_compute_dt(struct euler1d_gpu_t* solver, double cfl);
void euler1d_gpu_rk2_step(struct euler1d_gpu_t* solver, double dt);
int euler1d_gpu_run(struct euler1d_gpu_t* solver, double t_final, double cfl);
double euler1d_gpu_run_timed(struct euler1d_gpu_t* solver, double t_final, double cfl, int* steps);

/* ===================================================================== */
/* 2D simulations: one patch mesh, or a batch of independent problems.    */
/* Replaces the 2D CPU driver loop of reference examples/sod2d/sod2d.cpp: */
/* 318-356 (exchange -> x_sweep_t -> exchange -> y_sweep_t, euler2d.hpp:  */
/* 67-239) and compute_dt_2d (euler2d.hpp:242-271), with the 2D step      */
/* defined as SSP-RK2 over the X-then-Y split stage.                      */
/* ===================================================================== */
typedef struct nest_sim2d nest_sim2d;

/* nx*ny cells on npx*npy patches (tensor-product partition with
   subspace() semantics, reference core.hpp:184-199, patch.hpp:207-229),
   halo 2, boundary conditions per dimension. Every patch must have at least
   2 interior cells per dimension. */
int nest_hip_mesh2d_create(int nx, int ny, int npx, int npy, double dx, double dy, double gamma, int bcx,
                           int bcy, int limiter, int mode, nest_sim2d** out);
/* nprob independent single-patch problems of nx[i]*ny[i] cells with cell
   sizes dx[i], dy[i], all sharing gamma, bc and limiter; stepped together. */
int nest_hip_batch2d_create(int nprob, const int* nx, const int* ny, const double* dx, const double* dy,
                            double gamma, int bc, int limiter, int mode, nest_sim2d** out);
int nest_hip_sim2d_destroy(nest_sim2d* s);
/* total doubles of the concatenated buffer, patch / problem / work-item counts */
int nest_hip_sim2d_info(const nest_sim2d* s, size_t* total, int* npatches, int* nprobs, int* nitems);
/* element offset, interior size and problem index of one patch */
int nest_hip_sim2d_patch_info(const nest_sim2d* s, int patch, long long* off, int* nx, int* ny, int* prob);
int nest_hip_sim2d_set_stream(nest_sim2d* s, void* stream);
/* Operate on caller-owned device buffers (each `total` doubles), e.g. torch
   tensors; NULL restores the handle's own buffer for that slot. */
int nest_hip_sim2d_bind(nest_sim2d* s, double* dev_U, double* dev_U1);
int nest_hip_sim2d_device_ptrs(nest_sim2d* s, double** dev_U, double** dev_U1);
/* Work decomposition: rows per work item (>= 1) and waves per block (1..4);
   0 keeps the current value. persistent != 0 caps the grid at the number of
   resident waves and loops over the work list. */
int nest_hip_sim2d_set_tiling(nest_sim2d* s, int rows, int waves_per_block, int persistent);
/* Which 2D stage kernel runs: 0 automatic (the LDS-ring kernel, which
   fetches rows straight into LDS and addresses the state buffers with 32-bit
   offsets, whenever a state buffer is below 4 GiB; else the register-path
   kernel), 1 register path, 2 LDS-ring path (NEST_ERR_UNSUPPORTED at 4 GiB
   and above). Both compute the same arithmetic. No counterpart in the
   reference: a backend choice, kept selectable so both paths stay tested. */
int nest_hip_sim2d_set_stage_kernel(nest_sim2d* s, int kind);
/* How a whole RK2 step of rk2_step / advance runs: 0 automatic (one launch
   per step, U1 kept on chip: the step kernel, whenever a state buffer is
   below 4 GiB and the stage kernel is not forced to the register path),
   1 two stage launches (stage 1 writes U1, stage 2 reads it). Both compute
   the same arithmetic. No counterpart in the reference: a backend choice,
   kept selectable so both paths stay tested. get_step_kernel(): 1 when the
   one-launch step runs, 0 otherwise. */
int nest_hip_sim2d_set_step_kernel(nest_sim2d* s, int kind);
int nest_hip_sim2d_get_step_kernel(const nest_sim2d* s, int* one_launch);
/* Tiling of one kernel of the device-dt / rk2_step path: stage 1 or 2 of
   the two-launch step, or 3 for the one-launch step kernel; 0 keeps the
   current value. get_tiling(): rows[2], waves_per_block[2] of the two stage
   kernels; get_step_tiling(): those of the one-launch step (either pointer
   may be NULL). */
int nest_hip_sim2d_set_stage_tiling(nest_sim2d* s, int stage, int rows, int waves_per_block);
/* Per-problem rows per work item of one kernel (stage 1, 2 or 3, as above;
   array of nprob; 0 = the kernel's default; NULL clears): lets a tuner pick
   the tile height per problem size class inside one launch. */
int nest_hip_sim2d_set_problem_rows(nest_sim2d* s, int stage, const int* rows);
int nest_hip_sim2d_get_tiling(const nest_sim2d* s, int* rows, int* waves_per_block);
int nest_hip_sim2d_get_step_tiling(const nest_sim2d* s, int* rows, int* waves_per_block);
/* host <-> device copies of the whole concatenated buffer U (blocking) */
int nest_hip_sim2d_upload(nest_sim2d* s, const double* host);
int nest_hip_sim2d_download(nest_sim2d* s, double* host);
int nest_hip_sim2d_upload_patch(nest_sim2d* s, int patch, const double* host);
int nest_hip_sim2d_download_patch(nest_sim2d* s, int patch, double* host);
/* Halo exchange of U: every ghost cell (faces and corners) of every patch
   from its neighbour or physical boundary (periodic wrap / outflow copy / wall mirror);
   one launch. Extends execute_all_exchanges (reference patch.hpp:153-190). */
int nest_hip_sim2d_fill_halos(nest_sim2d* s);
/* The same exchange on a chosen buffer: 0 = U, 1 = U1 (the first-stage
   state, whose halo a per-patch stage 2 reads). */
int nest_hip_sim2d_exchange(nest_sim2d* s, int buffer);
/* compute_dt_2d per problem into dt_host[nprob] (blocking). */
int nest_hip_sim2d_compute_dt(nest_sim2d* s, double cfl, double* dt_host);
/* One split stage with a host dt for every problem:
   stage 1: U1 = S(U); stage 2: U = 0.5*(U + S(U1)). Halos of the input must
   be filled. patch < 0: all patches in one launch; else that patch only. */
int nest_hip_sim2d_stage(nest_sim2d* s, int stage, double dt, int patch);
/* One full RK2 step with a host dt: the one-launch step kernel (into U1,
   then copied back to U) or stage 1 and stage 2 (set_step_kernel); either
   gathers its halo straight from the neighbour patches / boundary through
   the mesh's neighbour table (the same values an exchange would deliver),
   then one halo exchange leaves U halo-complete. */
int nest_hip_sim2d_rk2_step(nest_sim2d* s, double dt);
/* Device-resident time stepping. start(): t = 0, dt policy per problem
   (fixed_dt > 0 holds dt fixed, else CFL with cfl), stop at t_end (use
   INFINITY for a fixed step count), and reduces the initial wave speeds.
   advance(): nsteps steps of the one-launch step, which computes each
   step's dt in the kernel from the wave speeds the previous step left (the
   state alternating between U and U1; after an odd count it is copied back
   to U), or of [dt finalize, stage 1, stage 2] (set_step_kernel), halos
   gathered from neighbours as in rk2_step, enqueued without any host
   synchronisation, then one halo exchange of U; use_graph != 0 replays
   captured graphs of four steps (the last nsteps % 4 steps run outside
   them). A done problem passes its state through. Each step's dt comes
   from the state it steps: after upload, upload_patch, bind, device_ptrs
   (the caller may write through them), stage, pack into the patch layout
   or rollout, the next advance reduces the speeds of the current state
   again first. status(): per-problem t, dt, steps and done flags (arrays
   of nprob; any may be NULL), blocking. */
int nest_hip_sim2d_start(nest_sim2d* s, double cfl, double t_end, double fixed_dt);
int nest_hip_sim2d_advance(nest_sim2d* s, int nsteps, int use_graph);
int nest_hip_sim2d_status(nest_sim2d* s, double* t, double* dt, int* steps, int* done);
/* Plain per-patch drop-in path (what a patch-by-patch Pipeline executor
   does): host dt with a blocking read-back, then per patch one launch per
   stage. nsteps steps. */
int nest_hip_sim2d_advance_per_patch(nest_sim2d* s, int nsteps, double cfl);
int nest_hip_sim2d_synchronize(nest_sim2d* s);

/* ---- discrete adjoint of fixed-dt rollouts ---------------------------------
   No counterpart in the reference (it has no differentiation). These are the
   rollout_fwd / rollout_adj entry points behind the torch.autograd.Function
   (nest_amd/adjoint.py). The differentiated map is the library's RK2 step at
   a fixed dt, U1 = S(U), U' = 0.5 (U + S(U1)), S = Y o fill o X o fill of
   x_sweep_t / y_sweep_t (examples/sod2d/euler2d.hpp:67-239) with the halo
   fill (periodic wrap / zero-gradient copy) included. Derivatives are those
   of the branch the primal takes (limiter cases, Davis speeds, HLLE upwind or
   subsonic branch, floors). State and cotangent buffers are device pointers
   in the handle's patch layout (size = total from nest_hip_sim2d_info); only
   interior cells are read or written. Everything is enqueued on the handle's
   stream. */

/* Doubles in the problem-major interior layout [problem][var][x][y]. */
int nest_hip_sim2d_global_size(const nest_sim2d* s, long long* n);
/* direction 0: dst (interior layout) <- src (patch layout);
   direction 1: dst (patch layout, interior cells) <- src (interior layout). */
int nest_hip_sim2d_pack(nest_sim2d* s, const double* src, double* dst, int direction);
/* Work split of the two adjoint kernels (0 keeps a value; waves <= 4). */
int nest_hip_sim2d_set_adjoint_tiling(nest_sim2d* s, int rows_y, int waves_y, int rows_x, int waves_x);
/* out = S'(U)^T (g*G) + h*H for one split stage (H may be null, and may
   alias out; out may not alias U or G). */
int nest_hip_sim2d_adjoint_stage(nest_sim2d* s, const double* U, const double* G, double g, const double* H,
                                 double h, double* out, double dt);
/* Ubar <- (dU'/dU)^T Ubar for one RK2 step of size dt taken from state U0. */
int nest_hip_sim2d_rk2_vjp(nest_sim2d* s, const double* U0, double* Ubar, double dt);
/* nsteps fixed-dt RK2 steps of the handle's state in checkpoint intervals of
   `interval` steps (0: ceil(sqrt(nsteps))), the first interval taking the
   remainder; the first state of every interval is stored as its checkpoint
   in device buffers owned by the handle, and the final state's halos are
   refilled. With two or more intervals, the states of the last four
   intervals and their stage-1 states are kept too (in the four buffer sets
   the reverse sweep reads), so their reverse needs no recompute. interval <
   0 keeps every state and stage-1 state (2 nsteps states of device memory):
   the reverse then recomputes nothing. interval >= nsteps is a plain forward
   that keeps only the initial state. Buffers are kept by the handle (grown,
   never shrunk) until it is destroyed. */
int nest_hip_sim2d_rollout(nest_sim2d* s, int nsteps, double dt, int interval);
/* Reverse of the last rollout: Ubar (in: cotangent of the final state, out:
   cotangent of the initial state), recomputing each checkpoint interval's
   states (and their stage-1 states) from its checkpoint; the recompute of
   one interval runs on an auxiliary stream while the next is reversed.
   Memory: ceil(K/m) + 4 (2m - 1) + 3 states (2K + 3 with every state kept). */
int nest_hip_sim2d_rollout_adjoint(nest_sim2d* s, double* Ubar);
int nest_hip_sim2d_rollout_info(const nest_sim2d* s, int* nsteps, int* interval, double* dt);
/* advance() without a graph, with HIP events recorded on the handle's stream
   around each launch of the step; blocks at the end. ms[3] receives the
   summed durations of [dt finalize, stage 1, stage 2] over the nsteps steps
   (one-launch step: [0, step kernel, 0]), *total_ms the elapsed
   time of the whole sequence. */
int nest_hip_sim2d_advance_timed(nest_sim2d* s, int nsteps, double* ms, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif /* NEST_HIP_H */
