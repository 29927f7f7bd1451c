// Host runtime and C ABI (include/nest_hip.h) of the gfx950 Euler backend.
//
// Owns device buffers, the device-side tables (patches, work items, halo
// regions, per-problem time-stepping state) and the launch sequence of a
// step. Kernels live in kernels.hip (compiled once per numerics mode).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstddef>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "nest_hip.h"
#include "plan.hpp"

// Buffer sets of the rollout adjoint's checkpoint intervals: the rollout
// keeps its last NEST_ADJ_SETS intervals in them (their reverse recomputes
// nothing), and the recompute of an earlier interval (auxiliary stream) runs
// up to NEST_ADJ_SETS - 1 intervals ahead of the reverse sweep. Each set
// holds 2m - 1 states of an m-step interval. Four sets: at 1024^2, K = 200,
// m = 15, the checkpointed gradient takes 1.2 ms (3%) less than with two.
#ifndef NEST_ADJ_SETS
#define NEST_ADJ_SETS 4
#endif

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define NEST_HIP_TRY(call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(NEST_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// launcher table of one numerics mode
struct launch_table {
    void (*stage2d)(int, int, const stage_args&, int, int, hipStream_t);
    void (*step2d)(int, const stage_args&, int, int, hipStream_t);
    int (*step2d_blocks)(int, int);
    void (*speeds2d)(const stage_args&, int, int, hipStream_t);
    void (*fill_regions)(double*, const nest_region*, int, int, hipStream_t);
    void (*finalize)(nest_prob_state*, unsigned long long*, unsigned long long*, int, int, hipStream_t);
    void (*reset_speeds)(nest_prob_state*, unsigned long long*, int, hipStream_t);
    void (*stage1d)(int, int, const stage1d_args&, hipStream_t);
    void (*step1d)(int, const stage1d_args&, hipStream_t);
    void (*prepare_step1d)(int);
    void (*speeds1d)(const stage1d_args&, hipStream_t);
    void (*fill1d)(double*, int, int, const nest_prob_state*, hipStream_t);
    void (*adj_y)(int, const adj_args&, int, int, hipStream_t);
    void (*adj_x)(int, const adj_args&, int, int, hipStream_t);
    void (*pack)(const double*, double*, const nest_patch_desc*, int, int, const nest_prob_geom*, const long long*,
                 int, hipStream_t);
};

const launch_table kExact = {nest_exact::launch_stage2d,     nest_exact::launch_step2d,
                             nest_exact::step2d_resident_blocks, nest_exact::launch_speeds2d,
                             nest_exact::launch_fill_regions, nest_exact::launch_finalize,
                             nest_exact::launch_reset_speeds, nest_exact::launch_stage1d,
                             nest_exact::launch_step1d, nest_exact::prepare_step1d,       nest_exact::launch_speeds1d,
                             nest_exact::launch_fill1d,
                             nest_exact::launch_adj_y,        nest_exact::launch_adj_x,
                             nest_exact::launch_pack};
const launch_table kFast = {nest_fast::launch_stage2d,     nest_fast::launch_step2d,
                            nest_fast::step2d_resident_blocks, nest_fast::launch_speeds2d,
                            nest_fast::launch_fill_regions, nest_fast::launch_finalize,
                            nest_fast::launch_reset_speeds, nest_fast::launch_stage1d,
                            nest_fast::launch_step1d, nest_fast::prepare_step1d,       nest_fast::launch_speeds1d,
                            nest_fast::launch_fill1d,
                            nest_fast::launch_adj_y,        nest_fast::launch_adj_x,
                            nest_fast::launch_pack};

const launch_table& table(int mode) { return mode == NEST_MODE_FAST ? kFast : kExact; }

nest_eos make_eos(double gamma) { return {gamma, gamma - 1.0, 1.0 / (gamma - 1.0)}; }

double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// Host mirror of finalize_kernel's dt (used by the blocking compute_dt paths):
// compute_dt_2d (reference examples/sod2d/euler2d.hpp:249-270) and the 1D
// compute_dt (examples/sod1d/sod1d.cpp:127).
double host_dt_from_speeds(const nest_prob_state& s, double cfl) {
    const double sx = bits_to_double(s.smax[0]);
    const double sy = bits_to_double(s.smax[1]);
    if (s.dim == 1) return cfl * s.dx / sx;
    double dtx = 1e10, dty = 1e10;
    if (sx > 0.0) dtx = std::min(dtx, s.dx / sx);
    if (sy > 0.0) dty = std::min(dty, s.dy / sy);
    return cfl * std::min(dtx, dty);
}

int check_limiter_mode(int lim, int mode) {
    if (lim < 0 || lim > 2) return fail(NEST_ERR_INVALID, "limiter must be 0 (minmod), 1 (mc) or 2 (none)");
    if (mode != NEST_MODE_EXACT && mode != NEST_MODE_FAST) return fail(NEST_ERR_INVALID, "mode must be 0 or 1");
    return NEST_OK;
}

template <class T>
int dev_alloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) return NEST_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(NEST_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                        " bytes failed: " + hipGetErrorString(e));
    }
    return NEST_OK;
}

}  // namespace

// ===========================================================================
// 2D simulation object
// ===========================================================================
struct nest_sim2d {
    int mode = NEST_MODE_EXACT, lim = NEST_LIM_MINMOD;
    nest_eos eos{};
    int nprob = 0;
    std::vector<nest_prob_state> probs_h;
    nest_plan::tables tb;  // patches, halo regions, geometry, owner tables
    std::vector<nest_patch_desc>& patches_h = tb.patches;
    std::vector<nest_region>& regions_h = tb.regions;
    std::vector<nest_item> items_h;         // per-patch strips (own ghosts)
    std::vector<int> item_begin, item_end;  // per patch, into items_h
    std::vector<nest_item> gitems_h[2];     // global strips per stage (halo gathered from neighbours)
    std::vector<nest_item> fitems_h;        // global strips of the one-launch step (56 output columns)
    size_t total = 0;

    double* U_own = nullptr;
    double* U1_own = nullptr;
    double* U = nullptr;
    double* U1 = nullptr;
    nest_patch_desc* patches_d = nullptr;
    nest_region* regions_d = nullptr;
    nest_item* items_d = nullptr;
    nest_prob_state* probs_d = nullptr;
    // wave-speed slots of the device-dt steps: kSlotSets rotating sets of
    // NEST_SPEED_SLOTS per problem. Device step n reads set n % 4 (written by
    // step n - 1), writes set (n + 1) % 4 and clears set (n + 2) % 4 for
    // step n + 1. Four sets, so that a graph of four steps starts and ends
    // in the same phase.
    static constexpr int kSlotSets = 4;
    unsigned long long* slots_d = nullptr;
    int nstep_dev = 0;  // device-dt steps since start (the rotation phase and time parity)
    // the state changed outside a step since the speeds the next device-dt
    // step reads were reduced (upload, bind, stages, rollouts, writes through
    // device pointers): the next advance reduces them again first
    bool speeds_stale = false;
    nest_item* gitems_d[2] = {nullptr, nullptr};
    nest_item* fitems_d = nullptr;
    int* first_item_d = nullptr;  // per problem: its first work item in fitems
    nest_prob_geom* geoms_d = nullptr;
    int* itab_d = nullptr;

    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // captured device-dt step graphs of kGraphSteps steps, one per rotation
    // phase they start in
    static constexpr int kGraphSteps = 4;
    hipGraph_t graph[kSlotSets] = {};
    hipGraphExec_t graph_exec[kSlotSets] = {};

    int rows = 64;              // per-patch strips
    int wpb = 4;
    int grows[2] = {32, 48};    // global strips, rows per item, stage 1 / stage 2
    int gwpb[2] = {4, 4};       // waves per block, stage 1 / stage 2
    int frows = 32, fwpb = 4;   // the one-launch step kernel
    std::vector<int> prob_rows[3];  // optional per-problem override of grows / frows (0 = default)
    int persistent = 0;
    int resident_waves = 0;     // resident waves of the stage kernels (persistent launches)
    int fresident = 0;          // resident waves of the step kernel
    int ncu = 0;
    int stage_kernel = 0;  // 0 automatic, 1 register path, 2 LDS-ring path
    int step_kernel = 0;   // 0 automatic (one launch per step where it applies), 1 two stage launches

    // ---- discrete adjoint of fixed-dt rollouts
    std::vector<nest_item> aitems_h[2];  // adjoint strips: Y kernel (58 owned columns), X kernel (64)
    nest_item* aitems_d[2] = {nullptr, nullptr};
    int arows[2] = {32, 32};
    int awpb[2] = {4, 4};
    double* adj_U1 = nullptr;   // stage-1 state of the step being reversed
    double* adj_Vb = nullptr;   // cotangent of the X-swept intermediate
    double* adj_Ub1 = nullptr;  // cotangent of the stage-1 state
    std::vector<long long> goff_h;  // problem offsets in the problem-major interior layout
    long long gtotal = 0;
    long long* goff_d = nullptr;
    std::vector<double*> ckpt;  // rollout checkpoints (one per interval: its first state)
    // states of a checkpoint interval (kept by the rollout or recomputed)
    // and their stage-1 states, NEST_ADJ_SETS sets: the recompute of one
    // interval (auxiliary stream) overlaps the reverse sweep of a later one
    // (handle's stream)
    std::vector<double*> seg[NEST_ADJ_SETS], seg1[NEST_ADJ_SETS];
    hipStream_t aux_stream = nullptr;
    hipEvent_t ev_start = nullptr;
    std::vector<hipEvent_t> ev_rec_c, ev_rev_c;  // per interval: recompute done, reverse done
    std::vector<hipEvent_t> ev_step;             // per reversed step: its adjoint launches done
    double* roll_tmp = nullptr;  // discarded outputs of one-launch rollout steps
    int roll_K = 0, roll_m = 0;
    int roll_kept = 0;  // the last roll_kept intervals' states are in their buffer sets
    bool roll_one = false;  // the rollout took one-launch steps
    int roll_keep0 = 0;     // intervals the rollout kept (the last roll_keep0)
    double roll_dt = 0.0;

    ~nest_sim2d() { release(); }

    void drop_graph() {
        for (int k = 0; k < kSlotSets; ++k) {
            if (graph_exec[k]) (void)hipGraphExecDestroy(graph_exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            graph_exec[k] = nullptr;
            graph[k] = nullptr;
        }
    }
    void release() {
        drop_graph();
        for (void* p : {(void*)U_own, (void*)U1_own, (void*)patches_d, (void*)regions_d, (void*)items_d,
                        (void*)probs_d, (void*)slots_d, (void*)gitems_d[0], (void*)gitems_d[1], (void*)fitems_d,
                        (void*)first_item_d, (void*)geoms_d, (void*)itab_d,
                        (void*)aitems_d[0], (void*)aitems_d[1], (void*)adj_U1, (void*)adj_Vb, (void*)adj_Ub1,
                        (void*)goff_d, (void*)roll_tmp})
            if (p) (void)hipFree(p);
        roll_tmp = nullptr;
        for (double* p : ckpt) (void)hipFree(p);
        ckpt.clear();
        for (int k = 0; k < NEST_ADJ_SETS; ++k)
            for (auto* v : {&seg[k], &seg1[k]}) {
                for (double* p : *v) (void)hipFree(p);
                v->clear();
            }
        if (ev_start) (void)hipEventDestroy(ev_start);
        ev_start = nullptr;
        for (auto* v : {&ev_rec_c, &ev_rev_c, &ev_step}) {
            for (hipEvent_t e : *v) (void)hipEventDestroy(e);
            v->clear();
        }
        if (aux_stream) (void)hipStreamDestroy(aux_stream);
        aux_stream = nullptr;
        aitems_d[0] = aitems_d[1] = nullptr;
        adj_U1 = adj_Vb = adj_Ub1 = nullptr;
        goff_d = nullptr;
        roll_K = 0;
        roll_kept = 0;
        U_own = U1_own = U = U1 = nullptr;
        patches_d = nullptr;
        gitems_d[0] = gitems_d[1] = nullptr;
        fitems_d = nullptr;
        first_item_d = nullptr;
        geoms_d = nullptr;
        itab_d = nullptr;
        regions_d = nullptr;
        items_d = nullptr;
        probs_d = nullptr;
        slots_d = nullptr;
        if (own_stream) (void)hipStreamDestroy(own_stream);
        own_stream = nullptr;
    }

    const launch_table& L() const { return table(mode); }
    unsigned long long* slot_set(int k) const {
        return slots_d + (size_t)(k & (kSlotSets - 1)) * probs_h.size() * NEST_SPEED_SLOTS * NEST_SLOT_STRIDE;
    }

    // gather = 1 or 2: that stage's global strips (gitems) reading every cell,
    // halo included, from the patch that owns it (no exchange needed before
    // the stage); 0: per-patch strips (items) reading the patch's own ghosts.
    stage_args args(double host_dt, int item0, int nitems, int gather = 0) const {
        stage_args a{};
        a.U = U;
        a.U1 = U1;
        a.patches = patches_d;
        a.items = (gather ? gitems_d[gather - 1] : items_d) + item0;
        a.nitems = nitems;
        a.probs = probs_d;
        a.slots = slot_set(nstep_dev);  // host-dt stage 2: into the next device step's read set
        a.eos = eos;
        a.host_dt = host_dt;
        a.gather = gather;
        a.geoms = geoms_d;
        a.itab = itab_d;
        a.total = (long long)total;
        a.kernel = stage_kernel;
        return a;
    }
    // blocks for nitems work items at `waves` waves per block; persistent:
    // at most the resident waves (`resident`, default those of the stage
    // kernels), each wave looping over the work list
    int grid_for(int nitems, int waves, int resident = -1) const {
        int blocks = (nitems + waves - 1) / waves;
        const int res = resident < 0 ? resident_waves : resident;
        if (persistent && res > 0) blocks = std::min(blocks, std::max(1, res / waves));
        return std::max(blocks, 1);
    }
    // a whole RK2 step in one launch (step2d_kernel): chosen unless the
    // two-launch step or the register-path stage kernel is asked for; it
    // addresses the state buffers with 32-bit byte offsets
    bool one_launch() const {
        return step_kernel == 0 && stage_kernel != 1 && (unsigned long long)total * sizeof(double) < (1ULL << 32);
    }

    // Work list: per patch, row chunks of `rows`, strips of 60 output columns.
    int build_items() {
        items_h.clear();
        item_begin.assign(patches_h.size(), 0);
        item_end.assign(patches_h.size(), 0);
        for (size_t p = 0; p < patches_h.size(); ++p) {
            item_begin[p] = (int)items_h.size();
            nest_plan::add_patch_items(items_h, (int)p, patches_h[p].nx, patches_h[p].ny, rows);
            item_end[p] = (int)items_h.size();
        }
        for (int st = 0; st < 2; ++st) {
            gitems_h[st].clear();
            for (int q = 0; q < (int)tb.geoms.size(); ++q) {
                const int r = (int)prob_rows[st].size() == nprob && prob_rows[st][q] > 0 ? prob_rows[st][q] : grows[st];
                nest_plan::add_gather_items(gitems_h[st], tb, q, r);
            }
        }
        fitems_h.clear();
        std::vector<int> first_item_h(tb.geoms.size(), 0);
        for (int q = 0; q < (int)tb.geoms.size(); ++q) {
            const int r = (int)prob_rows[2].size() == nprob && prob_rows[2][q] > 0 ? prob_rows[2][q] : frows;
            first_item_h[q] = (int)fitems_h.size();
            nest_plan::add_gather_items(fitems_h, tb, q, r, 56);
        }
        for (int k = 0; k < 2; ++k) {
            aitems_h[k].clear();
            for (int q = 0; q < (int)tb.geoms.size(); ++q)
                nest_plan::add_gather_items(aitems_h[k], tb, q, arows[k], k == 0 ? 58 : 64);
        }
        if (int rc = validate_tables()) return rc;
        for (nest_item** d : {&items_d, &gitems_d[0], &gitems_d[1], &fitems_d, &aitems_d[0], &aitems_d[1]}) {
            if (*d) (void)hipFree(*d);
            *d = nullptr;
        }
        if (int rc = dev_alloc(&items_d, items_h.size())) return rc;
        NEST_HIP_TRY(hipMemcpy(items_d, items_h.data(), items_h.size() * sizeof(nest_item), hipMemcpyHostToDevice));
        for (int st = 0; st < 2; ++st) {
            if (int rc = dev_alloc(&gitems_d[st], gitems_h[st].size())) return rc;
            NEST_HIP_TRY(hipMemcpy(gitems_d[st], gitems_h[st].data(), gitems_h[st].size() * sizeof(nest_item),
                                   hipMemcpyHostToDevice));
            if (int rc = dev_alloc(&aitems_d[st], aitems_h[st].size())) return rc;
            NEST_HIP_TRY(hipMemcpy(aitems_d[st], aitems_h[st].data(), aitems_h[st].size() * sizeof(nest_item),
                                   hipMemcpyHostToDevice));
        }
        if (int rc = dev_alloc(&fitems_d, fitems_h.size())) return rc;
        NEST_HIP_TRY(hipMemcpy(fitems_d, fitems_h.data(), fitems_h.size() * sizeof(nest_item), hipMemcpyHostToDevice));
        if (first_item_d) (void)hipFree(first_item_d);
        first_item_d = nullptr;
        if (int rc = dev_alloc(&first_item_d, first_item_h.size())) return rc;
        NEST_HIP_TRY(hipMemcpy(first_item_d, first_item_h.data(), first_item_h.size() * sizeof(int),
                               hipMemcpyHostToDevice));
        fresident = L().step2d_blocks(lim, 64 * fwpb) * fwpb * ncu;
        drop_graph();
        return NEST_OK;
    }

    // Every device table entry must address memory inside the buffers: checked
    // on the host before anything is uploaded or launched.
    int validate_tables() const {
        if (!nest_plan::tables_in_bounds(tb, items_h) || !nest_plan::tables_in_bounds(tb, gitems_h[0]) ||
            !nest_plan::tables_in_bounds(tb, gitems_h[1]) || !nest_plan::tables_in_bounds(tb, fitems_h) ||
            !nest_plan::tables_in_bounds(tb, aitems_h[0]) ||
            !nest_plan::tables_in_bounds(tb, aitems_h[1]) ||
            (long long)total != tb.total)
            return fail(NEST_ERR_INVALID, "internal: a patch / halo region / work item addresses memory "
                                          "outside the buffers");
        return NEST_OK;
    }

    int init_device() {
        if (int rc = validate_tables()) return rc;
        if (int rc = dev_alloc(&U_own, total)) return rc;
        if (int rc = dev_alloc(&U1_own, total)) return rc;
        U = U_own;
        U1 = U1_own;
        NEST_HIP_TRY(hipMemset(U_own, 0, total * sizeof(double)));
        NEST_HIP_TRY(hipMemset(U1_own, 0, total * sizeof(double)));
        if (int rc = dev_alloc(&patches_d, patches_h.size())) return rc;
        if (int rc = dev_alloc(&regions_d, regions_h.size())) return rc;
        if (int rc = dev_alloc(&probs_d, probs_h.size())) return rc;
        const size_t nslot = kSlotSets * probs_h.size() * (size_t)NEST_SPEED_SLOTS * NEST_SLOT_STRIDE;
        if (int rc = dev_alloc(&slots_d, nslot)) return rc;
        NEST_HIP_TRY(hipMemset(slots_d, 0, nslot * sizeof(unsigned long long)));
        if (int rc = dev_alloc(&geoms_d, tb.geoms.size())) return rc;
        if (int rc = dev_alloc(&itab_d, tb.itab.size())) return rc;
        NEST_HIP_TRY(hipMemcpy(geoms_d, tb.geoms.data(), tb.geoms.size() * sizeof(nest_prob_geom),
                               hipMemcpyHostToDevice));
        NEST_HIP_TRY(hipMemcpy(itab_d, tb.itab.data(), tb.itab.size() * sizeof(int), hipMemcpyHostToDevice));
        NEST_HIP_TRY(hipMemcpy(patches_d, patches_h.data(), patches_h.size() * sizeof(nest_patch_desc),
                               hipMemcpyHostToDevice));
        if (!regions_h.empty())
            NEST_HIP_TRY(hipMemcpy(regions_d, regions_h.data(), regions_h.size() * sizeof(nest_region),
                                   hipMemcpyHostToDevice));
        NEST_HIP_TRY(hipMemcpy(probs_d, probs_h.data(), probs_h.size() * sizeof(nest_prob_state),
                               hipMemcpyHostToDevice));
        goff_h.assign(tb.geoms.size(), 0);
        gtotal = 0;
        for (size_t q = 0; q < tb.geoms.size(); ++q) {
            goff_h[q] = gtotal;
            gtotal += 4LL * tb.geoms[q].NX * tb.geoms[q].NY;
        }
        if (int rc = dev_alloc(&goff_d, goff_h.size())) return rc;
        NEST_HIP_TRY(hipMemcpy(goff_d, goff_h.data(), goff_h.size() * sizeof(long long), hipMemcpyHostToDevice));
        NEST_HIP_TRY(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
        stream = own_stream;
        int dev = 0, cus = 0;
        NEST_HIP_TRY(hipGetDevice(&dev));
        NEST_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ncu = cus;
        resident_waves = cus * 4 * 4;  // 4 SIMDs per CU, 4 waves/SIMD (the LDS-ring stage kernels)
        return build_items();
    }

    // ---- one launch of each step component ----
    void fill(double* base) { L().fill_regions(base, regions_d, (int)regions_h.size(), 4, stream); }
    // per-patch strips [item0, item0 + nitems) with a host or device dt
    void stage(int st, double host_dt, int item0, int nitems) {
        if (nitems <= 0) return;
        L().stage2d(st, lim, args(host_dt, item0, nitems, 0), grid_for(nitems, wpb), 64 * wpb, stream);
    }
    // all global strips of stage st (halo gathered in the kernel); stage 2
    // folds its speeds into `slots` (null: the default of args())
    void gstage(int st, double host_dt, unsigned long long* slots = nullptr) {
        const int n = (int)gitems_h[st - 1].size();
        const int w = gwpb[st - 1];
        stage_args a = args(host_dt, 0, n, st);
        if (slots) a.slots = slots;
        L().stage2d(st, lim, a, grid_for(n, w), 64 * w, stream);
    }
    void speeds(int item0, int nitems) {
        if (nitems <= 0) return;
        L().speeds2d(args(0.0, item0, nitems), grid_for(nitems, wpb), 64 * wpb, stream);
    }
    // the speeds of the current state into the read set of the next device
    // step
    void refresh_speeds() {
        const int n = (int)items_h.size();
        L().reset_speeds(probs_d, slot_set(nstep_dev), nprob, stream);
        stage_args a = args(0.0, 0, n);
        a.slots_rd = slot_set(nstep_dev);
        L().speeds2d(a, grid_for(n, wpb), 64 * wpb, stream);
        speeds_stale = false;
    }
    // one-launch step from src into dst (all global strips): the caller's
    // dt (host_dt > 0), or device step nstep_dev with its dt computed in the
    // kernel from the rotating slot sets
    // (u1: also the step's stage-1 state of every interior cell; on: a
    // stream other than the handle's)
    void fstep(double host_dt, const double* src, double* dst, double* u1 = nullptr, hipStream_t on = nullptr) {
        const int n = (int)fitems_h.size();
        stage_args a = args(host_dt, 0, n, 0);
        a.items = fitems_d;
        a.gather = 1;
        a.U = const_cast<double*>(src);
        a.U1 = u1;
        a.Uo = dst;
        if (!(host_dt > 0.0)) {
            a.slots_rd = slot_set(nstep_dev);
            a.slots = slot_set(nstep_dev + 1);
            a.slots_clr = slot_set(nstep_dev + 2);
            a.first_item = first_item_d;
            a.par = nstep_dev & 1;
        }
        L().step2d(lim, a, grid_for(n, fwpb, fresident), 64 * fwpb, on ? on : stream);
    }
    // dt of the two-launch step: fold the read set, clear the set the step
    // after next writes, schedule the step
    void finalize() {
        L().finalize(probs_d, slot_set(nstep_dev), slot_set(nstep_dev + 2), nprob, nstep_dev & 1, stream);
    }
    // Step k of a device-dt sequence (device step nstep_dev, which it
    // advances): the one-launch step, dt included, from U into U1 (even k)
    // or from U1 into U (odd k); or dt finalize and the two stages (stage 1
    // U -> U1, stage 2 U1, U -> U). Every kernel gathers its halo from the
    // neighbours (no exchange launches). Ghost cells of U are left stale;
    // callers refresh them with fill(U) when they hand U out.
    void device_step(int k) {
        if (one_launch()) {
            fstep(0.0, (k & 1) ? U1 : U, (k & 1) ? U : U1);
        } else {
            finalize();
            gstage(1, 0.0);
            gstage(2, 0.0, slot_set(nstep_dev + 1));
        }
        ++nstep_dev;
    }
    // after a device-dt sequence of nsteps one-launch steps the state is in
    // U1 when nsteps is odd: bring it back to U
    int settle(int nsteps) {
        if (one_launch() && (nsteps & 1))
            NEST_HIP_TRY(hipMemcpyAsync(U, U1, total * sizeof(double), hipMemcpyDeviceToDevice, stream));
        return NEST_OK;
    }

    // ---- adjoint and fixed-dt rollouts ----
    // forward stage st on an arbitrary state buffer X (stage 1: X -> X1;
    // stage 2: X <- 0.5 (X + S(X1)) in place), host dt
    // (stage 2 with Xo: Xo <- 0.5 (X + S(X1)), X left intact)
    void stage_on(int st, double* X, double* X1, double dt, double* Xo = nullptr, hipStream_t on = nullptr) {
        const int n = (int)gitems_h[st - 1].size();
        const int w = gwpb[st - 1];
        stage_args a = args(dt, 0, n, st);
        a.U = X;
        a.U1 = X1;
        a.Uo = Xo;
        L().stage2d(st, lim, a, grid_for(n, w), 64 * w, on ? on : stream);
    }
    int ensure_aux_stream() {
        if (!aux_stream) NEST_HIP_TRY(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
        if (!ev_start) NEST_HIP_TRY(hipEventCreateWithFlags(&ev_start, hipEventDisableTiming));
        return NEST_OK;
    }
    int ensure_interval_events(int n, int nsteps) {
        for (auto* v : {&ev_rec_c, &ev_rev_c})
            while ((int)v->size() < n) {
                hipEvent_t e = nullptr;
                NEST_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                v->push_back(e);
            }
        while ((int)ev_step.size() < nsteps) {
            hipEvent_t e = nullptr;
            NEST_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev_step.push_back(e);
        }
        return NEST_OK;
    }
    int ensure_adjoint_scratch() {
        // the adjoint kernels address every buffer with 32-bit element offsets
        if ((unsigned long long)total >= (1ULL << 32))
            return fail(NEST_ERR_UNSUPPORTED, "adjoint: state buffers of 2^32 elements or more");
        // their halo transpose folds periodic and outflow ghosts only
        for (const auto& G : tb.geoms)
            if (G.bcx == 2 || G.bcy == 2)
                return fail(NEST_ERR_UNSUPPORTED, "adjoint: reflecting boundaries are not differentiated");
        for (double** p : {&adj_U1, &adj_Vb, &adj_Ub1})
            if (!*p) {
                if (int rc = dev_alloc(p, total)) return rc;
                NEST_HIP_TRY(hipMemsetAsync(*p, 0, total * sizeof(double), stream));
            }
        return NEST_OK;
    }
    int grow(std::vector<double*>& v, size_t n) {
        while (v.size() < n) {
            double* p = nullptr;
            if (int rc = dev_alloc(&p, total)) return rc;
            v.push_back(p);
        }
        return NEST_OK;
    }
    // out = S'(Uin)^T (g G) + h H for one split stage S = Y o fill o X o fill
    void adj_stage(const double* Uin, const double* G, double g, const double* H, double h, double* out,
                   double dt) {
        adj_args a{};
        a.U = Uin;
        a.G = G;
        a.g = g;
        a.Vb = adj_Vb;
        a.H = H;
        a.h = h;
        a.out = out;
        a.patches = patches_d;
        a.probs = probs_d;
        a.eos = eos;
        a.dt = dt;
        a.geoms = geoms_d;
        a.itab = itab_d;
        a.outflow_x = 0;
        for (const auto& g : tb.geoms) a.outflow_x |= (g.bcx != 0);
        for (int k = 0; k < 2; ++k) {
            a.items = aitems_d[k];
            a.nitems = (int)aitems_h[k].size();
            const int w = awpb[k];
            (k == 0 ? L().adj_y : L().adj_x)(lim, a, grid_for(a.nitems, w), 64 * w, stream);
        }
    }
    // Ub <- (dU^{n+1}/dU^n)^T Ub for the fixed-dt RK2 step taken from U0:
    // U1 = S(U0), U^{n+1} = 0.5 (U0 + S(U1)), so
    // Ub1 = S'(U1)^T (0.5 Ub) and Ub <- 0.5 Ub + S'(U0)^T Ub1.
    void vjp_step(const double* U0, double* Ub, double dt) {
        stage_on(1, const_cast<double*>(U0), adj_U1, dt);
        vjp_step_u1(U0, adj_U1, Ub, dt);
    }
    // the same with the step's stage-1 state U1 = S(U0) already at hand
    void vjp_step_u1(const double* U0, const double* U1s, double* Ub, double dt) {
        adj_stage(U1s, Ub, 0.5, nullptr, 0.0, adj_Ub1, dt);
        adj_stage(U0, adj_Ub1, 1.0, Ub, 0.5, Ub, dt);
    }
};

namespace {

using namespace nest_plan;

nest_prob_state make_prob(double dx, double dy, int dim) {
    nest_prob_state s{};
    s.dt = 0.0;
    s.t = 0.0;
    s.t_end = std::numeric_limits<double>::infinity();
    s.cfl = 0.4;
    s.dx = dx;
    s.dy = dy;
    s.fixed_dt = 0.0;
    s.dim = dim;
    return s;
}

int valid_sim(const nest_sim2d* s) { return s ? NEST_OK : fail(NEST_ERR_INVALID, "null handle"); }

}  // namespace

extern "C" {

const char* nest_hip_last_error(void) { return g_err.c_str(); }
int nest_hip_version(void) { return 100; }
int nest_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

namespace {
__global__ __launch_bounds__(256) void device_copy_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                          long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" {

int nest_hip_device_copy(const double* src, double* dst, long long n, void* stream) {
    if (n < 0 || (n > 0 && (!src || !dst))) return fail(NEST_ERR_INVALID, "need n >= 0 and buffers");
    if (n == 0) return NEST_OK;
    const long long blocks = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(device_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, src, dst, n);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_mesh2d_create(int nx, int ny, int npx, int npy, double dx, double dy, double gamma, int bcx, int bcy,
                           int limiter, int mode, nest_sim2d** out) {
    if (!out) return fail(NEST_ERR_INVALID, "out is null");
    *out = nullptr;
    if (int rc = check_limiter_mode(limiter, mode)) return rc;
    if (nx < 2 || ny < 2 || npx < 1 || npy < 1 || nx / npx < 2 || ny / npy < 2)
        return fail(NEST_ERR_INVALID, "every patch needs at least 2 interior cells per dimension");
    if (bcx < 0 || bcx > 2 || bcy < 0 || bcy > 2)
        return fail(NEST_ERR_INVALID, "bc must be 0 (periodic), 1 (outflow) or 2 (reflecting)");
    if (!(dx > 0.0) || !(dy > 0.0) || !(gamma > 1.0)) return fail(NEST_ERR_INVALID, "need dx, dy > 0, gamma > 1");
    if ((double)(nx / npx + 5) * (ny / npy + 5) > 2.0e9) return fail(NEST_ERR_UNSUPPORTED, "patch too large");
    auto* s = new nest_sim2d();
    s->mode = mode;
    s->lim = limiter;
    s->eos = make_eos(gamma);
    s->nprob = 1;
    s->probs_h.push_back(make_prob(dx, dy, 2));
    build_mesh(s->tb, nx, ny, npx, npy, bcx, bcy);
    s->total = (size_t)s->tb.total;
    if (int rc = s->init_device()) {
        delete s;
        return rc;
    }
    *out = s;
    return NEST_OK;
}

int nest_hip_batch2d_create(int nprob, const int* nx, const int* ny, const double* dx, const double* dy,
                            double gamma, int bc, int limiter, int mode, nest_sim2d** out) {
    if (!out) return fail(NEST_ERR_INVALID, "out is null");
    *out = nullptr;
    if (int rc = check_limiter_mode(limiter, mode)) return rc;
    if (nprob < 1 || !nx || !ny || !dx || !dy) return fail(NEST_ERR_INVALID, "bad batch description");
    if (bc < 0 || bc > 2) return fail(NEST_ERR_INVALID, "bc must be 0 (periodic), 1 (outflow) or 2 (reflecting)");
    auto* s = new nest_sim2d();
    s->mode = mode;
    s->lim = limiter;
    s->eos = make_eos(gamma);
    s->nprob = nprob;
    for (int i = 0; i < nprob; ++i) {
        if (nx[i] < 2 || ny[i] < 2 || !(dx[i] > 0.0) || !(dy[i] > 0.0)) {
            delete s;
            return fail(NEST_ERR_INVALID, "problem " + std::to_string(i) + ": bad size or spacing");
        }
        s->probs_h.push_back(make_prob(dx[i], dy[i], 2));
        build_mesh(s->tb, nx[i], ny[i], 1, 1, bc, bc);
    }
    s->total = (size_t)s->tb.total;
    if (int rc = s->init_device()) {
        delete s;
        return rc;
    }
    *out = s;
    return NEST_OK;
}

int nest_hip_sim2d_destroy(nest_sim2d* s) {
    if (!s) return NEST_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    delete s;
    return NEST_OK;
}

int nest_hip_sim2d_info(const nest_sim2d* s, size_t* total, int* npatches, int* nprobs, int* nitems) {
    if (int rc = valid_sim(s)) return rc;
    if (total) *total = s->total;
    if (npatches) *npatches = (int)s->patches_h.size();
    if (nprobs) *nprobs = s->nprob;
    if (nitems) *nitems = (int)s->items_h.size();
    return NEST_OK;
}

int nest_hip_sim2d_patch_info(const nest_sim2d* s, int patch, long long* off, int* nx, int* ny, int* prob) {
    if (int rc = valid_sim(s)) return rc;
    if (patch < 0 || patch >= (int)s->patches_h.size()) return fail(NEST_ERR_INVALID, "patch out of range");
    const auto& d = s->patches_h[patch];
    if (off) *off = d.off;
    if (nx) *nx = d.nx;
    if (ny) *ny = d.ny;
    if (prob) *prob = d.prob;
    return NEST_OK;
}

int nest_hip_sim2d_set_stream(nest_sim2d* s, void* stream) {
    if (int rc = valid_sim(s)) return rc;
    s->stream = stream ? static_cast<hipStream_t>(stream) : s->own_stream;
    s->drop_graph();
    return NEST_OK;
}

int nest_hip_sim2d_bind(nest_sim2d* s, double* dev_U, double* dev_U1) {
    if (int rc = valid_sim(s)) return rc;
    s->U = dev_U ? dev_U : s->U_own;
    s->U1 = dev_U1 ? dev_U1 : s->U1_own;
    s->drop_graph();
    s->speeds_stale = true;
    return NEST_OK;
}

int nest_hip_sim2d_device_ptrs(nest_sim2d* s, double** dev_U, double** dev_U1) {
    if (int rc = valid_sim(s)) return rc;
    s->speeds_stale = true;  // the caller may write the state through them
    if (dev_U) *dev_U = s->U;
    if (dev_U1) *dev_U1 = s->U1;
    return NEST_OK;
}

int nest_hip_sim2d_set_tiling(nest_sim2d* s, int rows, int waves_per_block, int persistent) {
    if (int rc = valid_sim(s)) return rc;
    // the stage kernels are compiled with __launch_bounds__(256): at most 4 waves per block
    if (rows < 0 || waves_per_block < 0 || waves_per_block > 4)
        return fail(NEST_ERR_INVALID, "rows >= 1, 1 <= waves_per_block <= 4");
    if (waves_per_block) s->wpb = s->gwpb[0] = s->gwpb[1] = s->fwpb = waves_per_block;
    s->persistent = persistent;
    if (rows) s->rows = s->grows[0] = s->grows[1] = s->frows = rows;
    return s->build_items();
}

int nest_hip_sim2d_set_stage_kernel(nest_sim2d* s, int kind) {
    if (int rc = valid_sim(s)) return rc;
    if (kind < 0 || kind > 2) return fail(NEST_ERR_INVALID, "kind: 0 automatic, 1 register path, 2 LDS-ring path");
    if (kind == 2 && (long long)s->total * (long long)sizeof(double) >= (1LL << 32))
        return fail(NEST_ERR_UNSUPPORTED, "the LDS-ring stage kernel needs state buffers below 4 GiB");
    s->stage_kernel = kind;
    s->drop_graph();
    return NEST_OK;
}

int nest_hip_sim2d_set_step_kernel(nest_sim2d* s, int kind) {
    if (int rc = valid_sim(s)) return rc;
    if (kind != 0 && kind != 1) return fail(NEST_ERR_INVALID, "kind: 0 automatic (one launch per step), 1 two stage launches");
    s->step_kernel = kind;
    s->drop_graph();
    return NEST_OK;
}

int nest_hip_sim2d_get_step_kernel(const nest_sim2d* s, int* one_launch) {
    if (int rc = valid_sim(s)) return rc;
    if (one_launch) *one_launch = s->one_launch() ? 1 : 0;
    return NEST_OK;
}

int nest_hip_sim2d_set_stage_tiling(nest_sim2d* s, int stage, int rows, int waves_per_block) {
    if (int rc = valid_sim(s)) return rc;
    if (stage < 1 || stage > 3 || rows < 0 || waves_per_block < 0 || waves_per_block > 4)
        return fail(NEST_ERR_INVALID, "stage 1|2 (or 3: the one-launch step), rows >= 1, 1 <= waves_per_block <= 4");
    if (stage == 3) {
        if (waves_per_block) s->fwpb = waves_per_block;
        if (rows) s->frows = rows;
    } else {
        if (waves_per_block) s->gwpb[stage - 1] = waves_per_block;
        if (rows) s->grows[stage - 1] = rows;
    }
    return s->build_items();
}

int nest_hip_sim2d_get_step_tiling(const nest_sim2d* s, int* rows, int* waves_per_block) {
    if (int rc = valid_sim(s)) return rc;
    if (rows) *rows = s->frows;
    if (waves_per_block) *waves_per_block = s->fwpb;
    return NEST_OK;
}

int nest_hip_sim2d_set_problem_rows(nest_sim2d* s, int stage, const int* rows) {
    if (int rc = valid_sim(s)) return rc;
    if (stage < 1 || stage > 3) return fail(NEST_ERR_INVALID, "stage must be 1, 2 or 3");
    if (!rows) {
        s->prob_rows[stage - 1].clear();
    } else {
        for (int i = 0; i < s->nprob; ++i)
            if (rows[i] < 0) return fail(NEST_ERR_INVALID, "rows must be >= 0");
        s->prob_rows[stage - 1].assign(rows, rows + s->nprob);
    }
    return s->build_items();
}

int nest_hip_sim2d_get_tiling(const nest_sim2d* s, int* rows, int* waves_per_block) {
    if (int rc = valid_sim(s)) return rc;
    for (int st = 0; st < 2; ++st) {
        if (rows) rows[st] = s->grows[st];
        if (waves_per_block) waves_per_block[st] = s->gwpb[st];
    }
    return NEST_OK;
}

int nest_hip_sim2d_upload(nest_sim2d* s, const double* host) {
    if (int rc = valid_sim(s)) return rc;
    if (!host) return fail(NEST_ERR_INVALID, "host is null");
    NEST_HIP_TRY(hipMemcpyAsync(s->U, host, s->total * sizeof(double), hipMemcpyHostToDevice, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    s->speeds_stale = true;
    return NEST_OK;
}

int nest_hip_sim2d_download(nest_sim2d* s, double* host) {
    if (int rc = valid_sim(s)) return rc;
    if (!host) return fail(NEST_ERR_INVALID, "host is null");
    NEST_HIP_TRY(hipMemcpyAsync(host, s->U, s->total * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    return NEST_OK;
}

static size_t patch_elems(const nest_patch_desc& d) { return 4 * (size_t)(d.nx + 4) * (d.ny + 4); }

int nest_hip_sim2d_upload_patch(nest_sim2d* s, int patch, const double* host) {
    if (int rc = valid_sim(s)) return rc;
    if (!host || patch < 0 || patch >= (int)s->patches_h.size()) return fail(NEST_ERR_INVALID, "bad patch/host");
    const auto& d = s->patches_h[patch];
    NEST_HIP_TRY(hipMemcpyAsync(s->U + d.off, host, patch_elems(d) * sizeof(double), hipMemcpyHostToDevice, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    s->speeds_stale = true;
    return NEST_OK;
}

int nest_hip_sim2d_download_patch(nest_sim2d* s, int patch, double* host) {
    if (int rc = valid_sim(s)) return rc;
    if (!host || patch < 0 || patch >= (int)s->patches_h.size()) return fail(NEST_ERR_INVALID, "bad patch/host");
    const auto& d = s->patches_h[patch];
    NEST_HIP_TRY(hipMemcpyAsync(host, s->U + d.off, patch_elems(d) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    return NEST_OK;
}

int nest_hip_sim2d_fill_halos(nest_sim2d* s) {
    if (int rc = valid_sim(s)) return rc;
    s->fill(s->U);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_exchange(nest_sim2d* s, int buffer) {
    if (int rc = valid_sim(s)) return rc;
    if (buffer != 0 && buffer != 1) return fail(NEST_ERR_INVALID, "buffer must be 0 (U) or 1 (U1)");
    s->fill(buffer ? s->U1 : s->U);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_compute_dt(nest_sim2d* s, double cfl, double* dt_host) {
    if (int rc = valid_sim(s)) return rc;
    if (!dt_host) return fail(NEST_ERR_INVALID, "dt_host is null");
    // into smax (the device-dt slot sets keep the pending step's speeds)
    s->L().reset_speeds(s->probs_d, nullptr, s->nprob, s->stream);
    s->speeds(0, (int)s->items_h.size());
    NEST_HIP_TRY(hipGetLastError());
    NEST_HIP_TRY(hipMemcpyAsync(s->probs_h.data(), s->probs_d, s->probs_h.size() * sizeof(nest_prob_state),
                                hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->nprob; ++i) dt_host[i] = host_dt_from_speeds(s->probs_h[i], cfl);
    return NEST_OK;
}

int nest_hip_sim2d_stage(nest_sim2d* s, int stage, double dt, int patch) {
    if (int rc = valid_sim(s)) return rc;
    if ((stage != 1 && stage != 2) || !(dt > 0.0)) return fail(NEST_ERR_INVALID, "stage must be 1|2, dt > 0");
    if (patch >= (int)s->patches_h.size()) return fail(NEST_ERR_INVALID, "patch out of range");
    if (patch < 0) s->stage(stage, dt, 0, (int)s->items_h.size());
    else s->stage(stage, dt, s->item_begin[patch], s->item_end[patch] - s->item_begin[patch]);
    s->speeds_stale = true;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_rk2_step(nest_sim2d* s, double dt) {
    if (int rc = valid_sim(s)) return rc;
    if (!(dt > 0.0)) return fail(NEST_ERR_INVALID, "dt must be > 0");
    // the step leaves the new state's speeds in the next device-dt step's
    // read set (cleared first, so they are exactly the new state's)
    s->L().reset_speeds(s->probs_d, s->slot_set(s->nstep_dev), s->nprob, s->stream);
    if (s->one_launch()) {
        s->fstep(dt, s->U, s->U1);
        NEST_HIP_TRY(hipMemcpyAsync(s->U, s->U1, s->total * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    } else {
        s->gstage(1, dt);
        s->gstage(2, dt);
    }
    s->fill(s->U);
    s->speeds_stale = false;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_start(nest_sim2d* s, double cfl, double t_end, double fixed_dt) {
    if (int rc = valid_sim(s)) return rc;
    if (!(cfl > 0.0) && !(fixed_dt > 0.0)) return fail(NEST_ERR_INVALID, "need cfl > 0 or fixed_dt > 0");
    for (auto& p : s->probs_h) {
        p.t = 0.0;
        p.dt = 0.0;
        p.t_end = t_end;
        p.cfl = cfl;
        p.fixed_dt = fixed_dt;
        p.steps = 0;
        p.done = 0;
        p.smax[0] = p.smax[1] = 0ull;
        p.tn[0] = p.tn[1] = 0.0;
    }
    NEST_HIP_TRY(hipMemcpyAsync(s->probs_d, s->probs_h.data(), s->probs_h.size() * sizeof(nest_prob_state),
                                hipMemcpyHostToDevice, s->stream));
    // every slot set may hold maxima of earlier steps; the speeds of the
    // state go to the read set of device step 0
    s->nstep_dev = 0;
    for (int k = 0; k < nest_sim2d::kSlotSets; ++k) s->L().reset_speeds(s->probs_d, s->slot_set(k), s->nprob, s->stream);
    s->refresh_speeds();
    NEST_HIP_TRY(hipGetLastError());
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    return NEST_OK;
}

int nest_hip_sim2d_advance(nest_sim2d* s, int nsteps, int use_graph) {
    if (int rc = valid_sim(s)) return rc;
    if (nsteps < 0) return fail(NEST_ERR_INVALID, "nsteps < 0");
    if (nsteps > 0 && s->speeds_stale) s->refresh_speeds();
    if (!use_graph) {
        for (int k = 0; k < nsteps; ++k) s->device_step(k);
    } else {
        // a graph holds four steps (an even count: one-launch steps end in U;
        // the slot-set rotation and time parity return to the phase it
        // started in), one graph per starting phase; the last nsteps % 4
        // steps run outside it
        const int per = nest_sim2d::kGraphSteps;
        const int ph = s->nstep_dev & (nest_sim2d::kSlotSets - 1);
        if (nsteps >= per && !s->graph_exec[ph]) {
            const int n0 = s->nstep_dev;
            NEST_HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < per; ++k) s->device_step(k);
            s->nstep_dev = n0;  // captured, not run
            hipError_t e = hipStreamEndCapture(s->stream, &s->graph[ph]);
            if (e != hipSuccess) return fail(NEST_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            NEST_HIP_TRY(hipGraphInstantiate(&s->graph_exec[ph], s->graph[ph], nullptr, nullptr, 0));
        }
        for (int r = 0; r < nsteps / per; ++r) {
            NEST_HIP_TRY(hipGraphLaunch(s->graph_exec[ph], s->stream));
            s->nstep_dev += per;
        }
        for (int k = nsteps / per * per; k < nsteps; ++k) s->device_step(k);
    }
    if (int rc = s->settle(nsteps)) return rc;
    if (nsteps > 0) s->fill(s->U);  // the fused steps leave U's ghost cells stale
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_status(nest_sim2d* s, double* t, double* dt, int* steps, int* done) {
    if (int rc = valid_sim(s)) return rc;
    NEST_HIP_TRY(hipMemcpyAsync(s->probs_h.data(), s->probs_d, s->probs_h.size() * sizeof(nest_prob_state),
                                hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    for (int i = 0; i < s->nprob; ++i) {
        const auto& p = s->probs_h[i];
        if (t) t[i] = p.t;
        if (dt) dt[i] = p.dt;
        if (steps) steps[i] = p.steps;
        if (done) done[i] = p.done;
    }
    for (const auto& p : s->probs_h)
        if (p.done == 2) return fail(NEST_ERR_NONFINITE, "a problem produced a non-finite or non-positive dt");
    return NEST_OK;
}

int nest_hip_sim2d_advance_per_patch(nest_sim2d* s, int nsteps, double cfl) {
    if (int rc = valid_sim(s)) return rc;
    const int np = (int)s->patches_h.size();
    std::vector<double> dts(s->nprob);
    for (int k = 0; k < nsteps; ++k) {
        // Reduce: wave speeds patch by patch, blocking read-back, host dt
        // (the read set of the next device step is cleared: stage 2 below
        // leaves the new state's speeds there)
        s->L().reset_speeds(s->probs_d, s->slot_set(s->nstep_dev), s->nprob, s->stream);
        for (int p = 0; p < np; ++p) s->speeds(s->item_begin[p], s->item_end[p] - s->item_begin[p]);
        NEST_HIP_TRY(hipMemcpyAsync(s->probs_h.data(), s->probs_d, s->probs_h.size() * sizeof(nest_prob_state),
                                    hipMemcpyDeviceToHost, s->stream));
        NEST_HIP_TRY(hipStreamSynchronize(s->stream));
        for (int i = 0; i < s->nprob; ++i) {
            dts[i] = host_dt_from_speeds(s->probs_h[i], cfl);
            if (!(dts[i] > 0.0) || !std::isfinite(dts[i])) return fail(NEST_ERR_NONFINITE, "non-finite dt");
        }
        // Exchange, Compute (stage 1), Exchange, Compute (stage 2), per patch
        s->fill(s->U);
        for (int p = 0; p < np; ++p)
            s->stage(1, dts[s->patches_h[p].prob], s->item_begin[p], s->item_end[p] - s->item_begin[p]);
        s->fill(s->U1);
        for (int p = 0; p < np; ++p)
            s->stage(2, dts[s->patches_h[p].prob], s->item_begin[p], s->item_end[p] - s->item_begin[p]);
        NEST_HIP_TRY(hipGetLastError());
        s->speeds_stale = false;  // stage 2 left the new state's speeds in the cleared read set
    }
    return NEST_OK;
}

int nest_hip_sim2d_advance_timed(nest_sim2d* s, int nsteps, double* ms, double* total_ms) {
    if (int rc = valid_sim(s)) return rc;
    if (nsteps < 1 || !ms) return fail(NEST_ERR_INVALID, "nsteps >= 1 and ms[3] required");
    std::vector<hipEvent_t> ev(4 * (size_t)nsteps, nullptr);
    int rc = NEST_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) rc = fail(NEST_ERR_HIP, "hipEventCreate failed");
    if (!rc) {
        if (s->speeds_stale) s->refresh_speeds();
        const bool one = s->one_launch();
        for (int k = 0; k < nsteps; ++k) {  // the launches of device_step(k), with events between
            hipEvent_t* e = &ev[4 * (size_t)k];
            (void)hipEventRecord(e[0], s->stream);
            if (!one) s->finalize();
            (void)hipEventRecord(e[1], s->stream);
            if (one) {
                s->fstep(0.0, (k & 1) ? s->U1 : s->U, (k & 1) ? s->U : s->U1);
                (void)hipEventRecord(e[2], s->stream);
            } else {
                s->gstage(1, 0.0);
                (void)hipEventRecord(e[2], s->stream);
                s->gstage(2, 0.0, s->slot_set(s->nstep_dev + 1));
            }
            (void)hipEventRecord(e[3], s->stream);
            ++s->nstep_dev;
        }
        if (int r2 = s->settle(nsteps)) rc = r2;
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = fail(NEST_ERR_HIP, std::string("advance_timed: ") + hipGetErrorString(e));
    }
    if (!rc) {
        for (int c = 0; c < 3; ++c) ms[c] = 0.0;
        for (int k = 0; k < nsteps; ++k)
            for (int c = 0; c < 3; ++c) {
                float t = 0.f;
                (void)hipEventElapsedTime(&t, ev[4 * (size_t)k + c], ev[4 * (size_t)k + c + 1]);
                ms[c] += t;
            }
        if (total_ms) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ev.front(), ev.back());
            *total_ms = t;
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (!rc) s->fill(s->U);  // as advance(): hand out a halo-complete U
    return rc;
}

int nest_hip_sim2d_synchronize(nest_sim2d* s) {
    if (int rc = valid_sim(s)) return rc;
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

// ---- discrete adjoint ------------------------------------------------------

int nest_hip_sim2d_global_size(const nest_sim2d* s, long long* n) {
    if (int rc = valid_sim(s)) return rc;
    if (!n) return fail(NEST_ERR_INVALID, "n is null");
    *n = s->gtotal;
    return NEST_OK;
}

int nest_hip_sim2d_pack(nest_sim2d* s, const double* src, double* dst, int direction) {
    if (int rc = valid_sim(s)) return rc;
    if (!src || !dst) return fail(NEST_ERR_INVALID, "null buffer");
    if (direction != 0 && direction != 1) return fail(NEST_ERR_INVALID, "direction must be 0 or 1");
    int max_nx = 0;
    for (const auto& d : s->patches_h) max_nx = std::max(max_nx, d.nx);
    s->L().pack(src, dst, s->patches_d, (int)s->patches_h.size(), max_nx, s->geoms_d, s->goff_d, direction,
                s->stream);
    if (direction == 1) s->speeds_stale = true;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_set_adjoint_tiling(nest_sim2d* s, int rows_y, int waves_y, int rows_x, int waves_x) {
    if (int rc = valid_sim(s)) return rc;
    if (rows_y < 0 || rows_x < 0 || waves_y < 0 || waves_y > 4 || waves_x < 0 || waves_x > 4)
        return fail(NEST_ERR_INVALID, "rows >= 1, 1 <= waves_per_block <= 4");
    if (rows_y) s->arows[0] = rows_y;
    if (rows_x) s->arows[1] = rows_x;
    if (waves_y) s->awpb[0] = waves_y;
    if (waves_x) s->awpb[1] = waves_x;
    return s->build_items();
}

int nest_hip_sim2d_adjoint_stage(nest_sim2d* s, const double* U, const double* G, double g, const double* H,
                                 double h, double* out, double dt) {
    if (int rc = valid_sim(s)) return rc;
    if (!U || !G || !out) return fail(NEST_ERR_INVALID, "null buffer");
    if (!(dt > 0.0)) return fail(NEST_ERR_INVALID, "dt must be > 0");
    if (out == G || out == U) return fail(NEST_ERR_INVALID, "out may alias H only");
    if (int rc = s->ensure_adjoint_scratch()) return rc;
    s->adj_stage(U, G, g, H, h, out, dt);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_rk2_vjp(nest_sim2d* s, const double* U0, double* Ubar, double dt) {
    if (int rc = valid_sim(s)) return rc;
    if (!U0 || !Ubar) return fail(NEST_ERR_INVALID, "null buffer");
    if (U0 == Ubar) return fail(NEST_ERR_INVALID, "U0 and Ubar must differ");
    if (!(dt > 0.0)) return fail(NEST_ERR_INVALID, "dt must be > 0");
    if (int rc = s->ensure_adjoint_scratch()) return rc;
    s->vjp_step(U0, Ubar, dt);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

}  // extern "C"

namespace {
// Interval layout of a rollout of K steps with interval m: nck = ceil(K / m)
// intervals; interval 0 takes the remainder (1..m steps) and every later one
// m steps, so that the intervals the forward keeps (the last ones) are whole.
// Interval c holds steps [interval_start(c), interval_start(c + 1)).
int interval_start(int K, int m, int nck, int c) { return c <= 0 ? 0 : (c >= nck ? K : K - (nck - c) * m); }
}  // namespace

extern "C" {

int nest_hip_sim2d_rollout(nest_sim2d* s, int nsteps, double dt, int interval) {
    if (int rc = valid_sim(s)) return rc;
    if (nsteps < 1 || !(dt > 0.0)) return fail(NEST_ERR_INVALID, "nsteps >= 1 and dt > 0");
    // interval < 0 keeps every state: one interval of nsteps, kept whole
    const bool all = interval < 0;
    const int m = all ? nsteps : (interval > 0 ? interval : std::max(1, (int)std::ceil(std::sqrt((double)nsteps))));
    const int nck = (nsteps + m - 1) / m;
    if (int rc = s->grow(s->ckpt, (size_t)nck)) return rc;
    // With two or more intervals (or every state kept), the states of the
    // last min(NEST_ADJ_SETS, nck) intervals (and their stage-1 states) are
    // written straight into the buffer sets their reverse reads, so the
    // reverse sweep recomputes none of them. A single interval of a plain
    // forward (interval >= nsteps) keeps nothing: its reverse recomputes it
    // from the initial state.
    const int nsets = std::min(NEST_ADJ_SETS, nck);
    const int nkeep = all ? 1 : (nck >= 2 ? nsets : 0);
    for (int c = nck - nkeep; c < nck; ++c) {
        if (int rc = s->grow(s->seg[c % nsets], (size_t)std::max(0, m - 1))) return rc;
        if (int rc = s->grow(s->seg1[c % nsets], (size_t)m)) return rc;
    }
    // checkpoint c = first state of interval c: the initial state is copied;
    // every later one is written by the stage 2 before it (out of place) and
    // read from there by the next step
    NEST_HIP_TRY(hipMemcpyAsync(s->ckpt[0], s->U, s->total * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
    // One launch per step (out of place) where the step kernel applies; it
    // also writes the stage-1 state, kept for the reverse sweep or (interval
    // not kept) into a discarded buffer.
    // The reverse sweep's recompute of a single interval runs the same
    // kernel instance, so it reproduces the forward's states bit for bit.
    // Intervals recomputed alongside the reverse of later ones run the two
    // stage kernels instead, which share the GPU with the adjoint kernels
    // better (the step kernel's 227 VGPRs and 16 KB of LDS per wave crowd
    // them out: at 1024^2, K = 200 in 15-step intervals the gradient took
    // 41.2 against 40.7 ms). Their fast arithmetic rounds differently from
    // the step kernel's, so those intervals are reversed about states within
    // round-off of the forward's (exact numerics: bitwise the same).
    const bool one = s->one_launch();
    if (one && !s->roll_tmp)
        if (int rc = dev_alloc(&s->roll_tmp, s->total)) return rc;
    for (int c = 0, k = 0; c < nck; ++c) {
        const int len = interval_start(nsteps, m, nck, c + 1) - interval_start(nsteps, m, nck, c);
        const bool kept = c >= nck - nkeep;
        const int set = c % nsets;
        for (int i = 0; i < len; ++i, ++k) {
            // the state after this step: the next checkpoint, the final
            // state (U), or (kept interval) state i + 1 in the set
            double* next = i + 1 < len ? (kept ? s->seg[set][i] : s->U) : (c + 1 < nck ? s->ckpt[c + 1] : s->U);
            double* src = i == 0 ? (k > 0 ? s->ckpt[c] : s->U) : (kept ? s->seg[set][i - 1] : s->U);
            double* u1 = kept ? s->seg1[set][i] : s->U1;
            if (one) {
                // every interval starts from its checkpoint (ckpt[0]: the
                // copy of U); the intermediate states of an interval that is
                // not kept alternate between U1 and U, phased so that its
                // last step reads U1: state j goes to U1 when len - 2 - j is
                // even
                auto inter = [&](int j) { return ((len - 2 - j) & 1) ? s->U : s->U1; };
                if (i == 0) src = s->ckpt[c];
                else if (!kept) src = inter(i - 1);
                if (!kept && i + 1 < len) next = inter(i);
                s->fstep(dt, src, next, kept ? u1 : s->roll_tmp);
            } else {
                s->stage_on(1, src, u1, dt);
                s->stage_on(2, src, u1, dt, next == src ? nullptr : next);
            }
        }
    }
    s->fill(s->U);
    s->speeds_stale = true;
    s->roll_kept = nkeep;
    s->roll_one = one;
    s->roll_keep0 = nkeep;
    s->roll_K = nsteps;
    s->roll_m = m;
    s->roll_dt = dt;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_rollout_adjoint(nest_sim2d* s, double* Ubar) {
    if (int rc = valid_sim(s)) return rc;
    if (!Ubar) return fail(NEST_ERR_INVALID, "Ubar is null");
    if (s->roll_K < 1) return fail(NEST_ERR_INVALID, "no rollout recorded: call nest_hip_sim2d_rollout first");
    if (int rc = s->ensure_adjoint_scratch()) return rc;
    const int K = s->roll_K, m = s->roll_m;
    const double dt = s->roll_dt;
    const int nck = (K + m - 1) / m;
    if (int rc = s->ensure_aux_stream()) return rc;
    // S buffer sets: the recompute runs up to S - 1 intervals ahead of the
    // reverse sweep
    const int nsets = std::min(NEST_ADJ_SETS, nck);
    for (int set = 0; set < nsets; ++set) {
        if (int rc = s->grow(s->seg[set], (size_t)std::max(0, m - 1))) return rc;
        if (int rc = s->grow(s->seg1[set], (size_t)m)) return rc;
    }
    // one event per interval and direction, and one per reversed step (an
    // event is never recorded again while a wait on its previous record may
    // still be pending)
    if (int rc = s->ensure_interval_events(nck, K)) return rc;
    // states of interval c: the checkpoint, then the recomputed (or kept)
    // states in buffer set c % S, with the stage-1 state of each
    auto state = [&](int c, int i) { return i == 0 ? s->ckpt[c] : s->seg[c % nsets][i - 1]; };
    auto length = [&](int c) { return interval_start(K, m, nck, c + 1) - interval_start(K, m, nck, c); };
    // the last roll_kept intervals were kept by the rollout
    const int nkept = s->roll_kept;
    auto kept = [&](int c) { return c >= nck - nkept; };
    // One-launch steps (the output of the interval's last step discarded)
    // where the forward took them and the recompute must reproduce its states
    // bit for bit: a single interval, and the intervals the forward kept (a
    // second reverse of the same rollout recomputes them after the first
    // reverse's recomputes overwrote their buffer sets). The two stage
    // kernels for the other intervals, which the forward did not keep.
    auto one_at = [&](int r) { return s->roll_one && (nck == 1 || r >= nck - s->roll_keep0); };
    if (s->roll_one && !s->roll_tmp)
        if (int rc = dev_alloc(&s->roll_tmp, s->total)) return rc;
    // step j of the recompute of interval r (auxiliary stream): its stage-1
    // state, and the next state unless j is the interval's last step (whose
    // successor is the next checkpoint)
    auto recompute_step = [&](int r, int j) {
        const int len = length(r), set = r % nsets;
        if (one_at(r)) {
            s->fstep(dt, state(r, j), j + 1 < len ? state(r, j + 1) : s->roll_tmp, s->seg1[set][j], s->aux_stream);
            return;
        }
        s->stage_on(1, state(r, j), s->seg1[set][j], dt, nullptr, s->aux_stream);
        if (j + 1 < len) s->stage_on(2, state(r, j), s->seg1[set][j], dt, state(r, j + 1), s->aux_stream);
    };
    // the whole recompute of interval r at once, after the reverse of
    // interval r + S (the last to read its buffer set)
    auto recompute = [&](int r) -> hipError_t {
        if (r + nsets <= nck - 1) {
            hipError_t e = hipStreamWaitEvent(s->aux_stream, s->ev_rev_c[r + nsets], 0);
            if (e != hipSuccess) return e;
        }
        for (int j = 0; j < length(r); ++j) recompute_step(r, j);
        return hipEventRecord(s->ev_rec_c[r], s->aux_stream);
    };
    // everything enqueued before (the rollout and its checkpoints) comes first
    NEST_HIP_TRY(hipEventRecord(s->ev_start, s->stream));
    NEST_HIP_TRY(hipStreamWaitEvent(s->aux_stream, s->ev_start, 0));
    // the intervals the reverse needs before any recompute can be spread out
    for (int c = nck - 1; c >= std::max(0, nck - (nsets - 1)); --c)
        if (!kept(c)) NEST_HIP_TRY(recompute(c));
    // Then, during the reverse of interval c (handle's stream), the
    // recompute of interval r = c - (S - 1) (auxiliary stream), whose buffer
    // set the reverse of c + 1 was the last to read: its step j waits for
    // the j-th reversed step of c, so the auxiliary stream never holds more
    // than one step's launches ready. (Released all at once, an interval's
    // recompute ran alone for ~0.4 ms at each interval boundary, the reverse
    // sweep stalled behind it.) The reverse of r, S - 1 intervals later,
    // waits for the recompute's last step.
    bool recomputed = false;
    int ks = 0;  // reversed steps so far (the index of their events)
    for (int c = nck - 1; c >= 0; --c) {
        const int set = c % nsets, len = length(c);
        const int r = c - (nsets - 1);
        // one buffer set: the interval is recomputed whole before its reverse
        if (nsets == 1 && !kept(c)) NEST_HIP_TRY(recompute(c));
        const bool spread = nsets > 1 && r >= 0 && !kept(r);
        if (spread && c + 1 <= nck - 1) NEST_HIP_TRY(hipStreamWaitEvent(s->aux_stream, s->ev_rev_c[c + 1], 0));
        if (!kept(c)) {
            NEST_HIP_TRY(hipStreamWaitEvent(s->stream, s->ev_rec_c[c], 0));
            recomputed = true;
        }
        const int rlen = spread ? length(r) : 0;
        for (int i = len - 1, k = 0; i >= 0; --i, ++k) {
            s->vjp_step_u1(state(c, i), s->seg1[set][i], Ubar, dt);
            if (k < rlen) {
                NEST_HIP_TRY(hipEventRecord(s->ev_step[ks], s->stream));
                NEST_HIP_TRY(hipStreamWaitEvent(s->aux_stream, s->ev_step[ks], 0));
                ++ks;
                recompute_step(r, k);
            }
        }
        if (spread) {
            for (int j = len; j < rlen; ++j) recompute_step(r, j);
            NEST_HIP_TRY(hipEventRecord(s->ev_rec_c[r], s->aux_stream));
        }
        NEST_HIP_TRY(hipEventRecord(s->ev_rev_c[c], s->stream));
    }
    // A recompute overwrites the buffer set of a kept interval: a further
    // reverse sweep of this rollout (a second backward of the same forward)
    // must then recompute every interval.
    if (recomputed) s->roll_kept = 0;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_sim2d_rollout_info(const nest_sim2d* s, int* nsteps, int* interval, double* dt) {
    if (int rc = valid_sim(s)) return rc;
    if (nsteps) *nsteps = s->roll_K;
    if (interval) *interval = s->roll_m;
    if (dt) *dt = s->roll_dt;
    return NEST_OK;
}

}  // extern "C"

// ===========================================================================
// 1D solver
// ===========================================================================
// The pair kernel addresses the state (3 (n + 4) doubles) with 32-bit byte
// offsets.
static bool step1d_pairs_fit(int n) { return 24ULL * ((unsigned long long)n + 4) < (1ULL << 32); }
// The fused 1D step kernel for n cells: kind 1 the LDS-tile kernel; kind 2
// the lane kernel (one cell per lane, 56-cell segments), kind 3 the pair
// kernel (two cells per lane, 120-cell segments), each with no more waves
// than stay resident, every wave looping over its share of the segments.
// Kind 0 automatic: the pair kernel from 2^19 cells (where it fits), the
// lane kernel below. Measured per step (fast / exact numerics): 2^24 cells
// 186.9 / 322.4 us against 204.6 / 350.8 us for the lane kernel, 2^19 12.0 /
// 17.1 against 13.1 / 18.4, 2^18 8.5 / 11.7 against 8.6 / 11.3, 1e5 6.2 /
// 6.9 against 6.3 / 6.8, 1000 5.8 / 6.7 against 5.3 / 5.8 (the lane kernel:
// 7.1 against 10.3 us at 1e5 and 262 against 406 us at 2^24 for the tile
// kernel).
static void step1d_kernel_args(int n, int kind, stage1d_args& a) {
    if (kind == 0) kind = n >= (1 << 19) && step1d_pairs_fit(n) ? 3 : 2;
    a.lanes = kind != 1;
    a.pairs = kind == 3;
}
struct nest_solver1d {
    int n = 0, bc = 0, lim = 0, mode = 0;
    nest_eos eos{};
    double dx = 0.0;
    double* U = nullptr;
    double* U1 = nullptr;
    nest_prob_state* ps_d = nullptr;
    nest_prob_state ps_h{};
    unsigned long long* slots_d = nullptr;  // 3 rotating sets of speed slots (device dt)
    nest_run1d_state* run_d = nullptr;      // time-stepping state of run()
    nest_run1d_state run_h{};
    hipStream_t own_stream = nullptr, stream = nullptr;
    int k_next = 0;  // index of the next fused step of start() / advance() (buffer and slot-set parity)
    // the state changed outside a fused step since the speeds the next one
    // reads were reduced (upload, caller-dt stages or steps): the next
    // advance reduces them again first
    bool speeds_stale = false;
    int step_kernel = 0;  // fused step kernel: 0 automatic, 1 LDS tile, 2 lanes, 3 lane pairs
    hipGraph_t graph = nullptr;  // kBatch device-dt steps, replayed by run()
    hipGraphExec_t graph_exec = nullptr;
    static constexpr int kBatch = 48;  // a multiple of 2 (buffers, time slots) and 3 (slot sets)
    static constexpr size_t kSet = (size_t)NEST_SPEED_SLOTS * NEST_SLOT_STRIDE;

    ~nest_solver1d() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (U) (void)hipFree(U);
        if (U1) (void)hipFree(U1);
        if (ps_d) (void)hipFree(ps_d);
        if (slots_d) (void)hipFree(slots_d);
        if (run_d) (void)hipFree(run_d);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
    const launch_table& L() const { return table(mode); }
    stage1d_args args(double host_dt) const {
        return {U, U1, n, ps_d, eos, host_dt, bc, nullptr, nullptr, nullptr, nullptr, 0, 0};
    }
    // One RK2 step with the caller's dt, in place: the two stage kernels
    // (they read ghost cells through the boundary condition, so no fill runs
    // between them).
    void step(double host_dt) {
        L().stage1d(1, lim, args(host_dt), stream);
        L().stage1d(2, lim, args(host_dt), stream);
    }
    // Device-dt RK2 step k, out of place, one launch: the fused step kernel
    // from the state in U (even k) or U1 (odd k) into the other. It takes dt
    // from slot set (k + 2) % 3 (written by step k - 1), folds its speeds into
    // set k % 3 and clears set (k + 1) % 3 for step k + 1.
    void device_step(int k) {
        stage1d_args a = args(0.0);
        if (k & 1) std::swap(a.U, a.U1);
        a.par = k & 1;
        a.run = run_d;
        a.slots = slots_d + (size_t)(k % 3) * kSet;
        a.slots_rd = slots_d + (size_t)((k + 2) % 3) * kSet;
        a.slots_clr = slots_d + (size_t)((k + 1) % 3) * kSet;
        step1d_kernel_args(n, step_kernel, a);
        L().step1d(lim, a, stream);
    }
    // the speeds of the current state (in U) into the read set of step k_next
    void refresh_speeds() {
        unsigned long long* set = slots_d + (size_t)((k_next + 2) % 3) * kSet;
        L().reset_speeds(ps_d, set, 1, stream);
        stage1d_args a = args(0.0);
        a.slots = set;
        L().speeds1d(a, stream);
        speeds_stale = false;
    }
    int ensure_graph() {
        if (graph_exec) return NEST_OK;
        NEST_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < kBatch; ++k) device_step(k);  // kBatch is even: the state ends in U
        hipError_t e = hipStreamEndCapture(stream, &graph);
        if (e != hipSuccess) return fail(NEST_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        NEST_HIP_TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        return NEST_OK;
    }
};

extern "C" {

int nest_hip_solver1d_create(int n, double dx, double gamma, int bc, int limiter, int mode, nest_solver1d** out) {
    if (!out) return fail(NEST_ERR_INVALID, "out is null");
    *out = nullptr;
    if (int rc = check_limiter_mode(limiter, mode)) return rc;
    if (n < 2 || !(dx > 0.0) || !(gamma > 1.0) || bc < 0 || bc > 2)
        return fail(NEST_ERR_INVALID, "need n >= 2, dx > 0, gamma > 1, bc in {0,1,2}");
    auto* s = new nest_solver1d();
    s->n = n;
    s->bc = bc;
    s->lim = limiter;
    s->mode = mode;
    s->eos = make_eos(gamma);
    s->dx = dx;
    s->ps_h = make_prob(dx, 1.0, 1);
    const size_t total = 3 * (size_t)(n + 4);
    int rc = dev_alloc(&s->U, total);
    if (!rc) rc = dev_alloc(&s->U1, total);
    if (!rc) rc = dev_alloc(&s->ps_d, 1);
    if (!rc) rc = dev_alloc(&s->slots_d, 3 * nest_solver1d::kSet);
    if (!rc) rc = dev_alloc(&s->run_d, 1);
    if (rc) {
        delete s;
        return rc;
    }
    if (hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemset(s->U1, 0, total * sizeof(double)) != hipSuccess ||
        hipMemset(s->slots_d, 0, 3 * nest_solver1d::kSet * sizeof(unsigned long long)) != hipSuccess ||
        hipMemcpy(s->ps_d, &s->ps_h, sizeof(nest_prob_state), hipMemcpyHostToDevice) != hipSuccess) {
        delete s;
        return fail(NEST_ERR_HIP, "stream/memory setup failed");
    }
    s->stream = s->own_stream;
    s->L().prepare_step1d(limiter);  // before any graph capture of the step
    *out = s;
    return NEST_OK;
}

int nest_hip_solver1d_destroy(nest_solver1d* s) {
    if (s && s->stream) (void)hipStreamSynchronize(s->stream);
    delete s;
    return NEST_OK;
}

int nest_hip_solver1d_set_stream(nest_solver1d* s, void* stream) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    s->stream = stream ? static_cast<hipStream_t>(stream) : s->own_stream;
    return NEST_OK;
}

int nest_hip_solver1d_upload(nest_solver1d* s, const double* host) {
    if (!s || !host) return fail(NEST_ERR_INVALID, "null handle or host");
    NEST_HIP_TRY(hipMemcpyAsync(s->U, host, 3 * (size_t)(s->n + 4) * sizeof(double), hipMemcpyHostToDevice, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    s->speeds_stale = true;
    return NEST_OK;
}

int nest_hip_solver1d_download(nest_solver1d* s, double* host) {
    if (!s || !host) return fail(NEST_ERR_INVALID, "null handle or host");
    NEST_HIP_TRY(hipMemcpyAsync(host, s->U, 3 * (size_t)(s->n + 4) * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    return NEST_OK;
}

int nest_hip_solver1d_compute_dt(nest_solver1d* s, double cfl, double* dt) {
    if (!s || !dt) return fail(NEST_ERR_INVALID, "null handle or dt");
    s->L().reset_speeds(s->ps_d, nullptr, 1, s->stream);
    s->L().speeds1d(s->args(0.0), s->stream);
    NEST_HIP_TRY(hipGetLastError());
    NEST_HIP_TRY(hipMemcpyAsync(&s->ps_h, s->ps_d, sizeof(nest_prob_state), hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    *dt = host_dt_from_speeds(s->ps_h, cfl);
    return NEST_OK;
}

int nest_hip_solver1d_rk2_step(nest_solver1d* s, double dt) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (!(dt > 0.0)) return fail(NEST_ERR_INVALID, "dt must be > 0");
    s->step(dt);
    s->L().fill1d(s->U, s->n, s->bc, nullptr, s->stream);  // hand out a halo-complete state
    s->speeds_stale = true;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

// Time-stepping state of the fused device-dt steps: CFL number, end time,
// step 0 next, and the speeds of the current state in slot set 2, which the
// first step reads its dt from (sets 0 and 1 start clear).
static int solver1d_begin(nest_solver1d* s, double cfl, double t_end) {
    nest_prob_state p = make_prob(s->dx, 1.0, 1);
    p.t_end = t_end;
    p.cfl = cfl;
    s->ps_h = p;
    NEST_HIP_TRY(hipMemcpyAsync(s->ps_d, &s->ps_h, sizeof(nest_prob_state), hipMemcpyHostToDevice, s->stream));
    s->run_h = nest_run1d_state{};
    NEST_HIP_TRY(hipMemcpyAsync(s->run_d, &s->run_h, sizeof(nest_run1d_state), hipMemcpyHostToDevice, s->stream));
    for (int set = 0; set < 3; ++set)
        s->L().reset_speeds(s->ps_d, s->slots_d + (size_t)set * nest_solver1d::kSet, 1, s->stream);
    stage1d_args a0 = s->args(0.0);
    a0.slots = s->slots_d + 2 * nest_solver1d::kSet;
    s->L().speeds1d(a0, s->stream);
    s->k_next = 0;
    s->speeds_stale = false;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_solver1d_run(nest_solver1d* s, double t_final, double cfl, int* steps) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (!(cfl > 0.0)) return fail(NEST_ERR_INVALID, "cfl must be > 0");
    if (int rc = solver1d_begin(s, cfl, t_final)) return rc;
    if (int rc = s->ensure_graph()) return rc;
    for (;;) {
        // steps past t_end are no-ops (every launch sees the done flag)
        NEST_HIP_TRY(hipGraphLaunch(s->graph_exec, s->stream));
        NEST_HIP_TRY(hipGetLastError());
        NEST_HIP_TRY(hipMemcpyAsync(&s->run_h, s->run_d, sizeof(nest_run1d_state), hipMemcpyDeviceToHost, s->stream));
        NEST_HIP_TRY(hipStreamSynchronize(s->stream));
        if (s->run_h.done) break;
    }
    s->ps_h.steps = s->run_h.steps;
    s->ps_h.done = s->run_h.done;
    s->ps_h.t = s->run_h.t[s->run_h.steps & 1];
    // launches alternate U -> U1 -> U (a step past the end passes the state
    // through) and every graph holds an even number of them: the state is in U
    s->L().fill1d(s->U, s->n, s->bc, nullptr, s->stream);
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    if (steps) *steps = s->ps_h.steps;
    if (s->ps_h.done == 2) return fail(NEST_ERR_NONFINITE, "non-finite dt");
    return NEST_OK;
}

int nest_hip_solver1d_set_step_kernel(nest_solver1d* s, int kind) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (kind < 0 || kind > 3) return fail(NEST_ERR_INVALID, "kind: 0 automatic, 1 LDS tile, 2 lanes, 3 lane pairs");
    if (kind == 3 && !step1d_pairs_fit(s->n))
        return fail(NEST_ERR_UNSUPPORTED, "the lane-pair kernel addresses the state with 32-bit byte offsets: n <= 178956966");
    s->step_kernel = kind;
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
    return NEST_OK;
}

int nest_hip_solver1d_stage(nest_solver1d* s, int stage, double dt) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if ((stage != 1 && stage != 2) || !(dt > 0.0)) return fail(NEST_ERR_INVALID, "stage must be 1|2, dt > 0");
    s->L().stage1d(stage, s->lim, s->args(dt), s->stream);
    s->speeds_stale = true;
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_solver1d_exchange(nest_solver1d* s, int buffer) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (buffer != 0 && buffer != 1) return fail(NEST_ERR_INVALID, "buffer must be 0 (U) or 1 (U1)");
    s->L().fill1d(buffer ? s->U1 : s->U, s->n, s->bc, nullptr, s->stream);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_solver1d_start(nest_solver1d* s, double cfl, double t_end) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (!(cfl > 0.0)) return fail(NEST_ERR_INVALID, "cfl must be > 0");
    if (int rc = solver1d_begin(s, cfl, t_end)) return rc;
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    return NEST_OK;
}

int nest_hip_solver1d_advance(nest_solver1d* s, int nsteps) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    if (nsteps < 0) return fail(NEST_ERR_INVALID, "nsteps < 0");
    const size_t bytes = 3 * (size_t)(s->n + 4) * sizeof(double);
    if (nsteps > 0 && s->speeds_stale) s->refresh_speeds();
    // fused step k reads the buffer of parity k (U for even k, U1 for odd k);
    // the handle's state is in U between calls
    if (s->k_next & 1)
        NEST_HIP_TRY(hipMemcpyAsync(s->U1, s->U, bytes, hipMemcpyDeviceToDevice, s->stream));
    for (int i = 0; i < nsteps; ++i) s->device_step(s->k_next++);
    if (s->k_next & 1)
        NEST_HIP_TRY(hipMemcpyAsync(s->U, s->U1, bytes, hipMemcpyDeviceToDevice, s->stream));
    s->L().fill1d(s->U, s->n, s->bc, nullptr, s->stream);
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

int nest_hip_solver1d_status(nest_solver1d* s, double* t, double* dt, int* steps, int* done) {
    if (!s) return fail(NEST_ERR_INVALID, "null handle");
    NEST_HIP_TRY(hipMemcpyAsync(&s->run_h, s->run_d, sizeof(nest_run1d_state), hipMemcpyDeviceToHost, s->stream));
    NEST_HIP_TRY(hipStreamSynchronize(s->stream));
    // step k reads the time of slot k & 1 and the last step taken wrote
    // slot steps & 1
    if (t) *t = s->run_h.t[s->run_h.steps & 1];
    if (dt) *dt = s->run_h.dt;
    if (steps) *steps = s->run_h.steps;
    if (done) *done = s->run_h.done;
    if (s->run_h.done == 2) return fail(NEST_ERR_NONFINITE, "non-finite dt");
    return NEST_OK;
}

// ---- one 1D step that allocates nothing -----------------------------------
// The speed slots, run state and problem parameters of
// nest_hip_solver1d_start + nest_hip_solver1d_advance live, for this entry
// point, in a fixed device array of the library instead of handle memory:
//   [0, 128)            nest_prob_state (cfl, t_end, dx of the call)
//   [128, 256)          nest_run1d_state (t, dt, step count, done flag)
//   [1024, 1024 + 3 S)  three speed-slot sets of S = NEST_SPEED_SLOTS *
//                       NEST_SLOT_STRIDE u64, laid out as the handle's
// Every GPU holds its own copy; calls that overlap on one GPU race on it.
namespace {
constexpr size_t kNoallocSet = (size_t)NEST_SPEED_SLOTS * NEST_SLOT_STRIDE;
constexpr size_t kNoallocRun = 128, kNoallocSlots = 1024;
constexpr size_t kNoallocBytes = kNoallocSlots + 3 * kNoallocSet * sizeof(unsigned long long);
static_assert(sizeof(nest_prob_state) <= kNoallocRun && kNoallocRun + sizeof(nest_run1d_state) <= kNoallocSlots,
              "scratch layout");
// the kernels index the state with 32-bit ints: 3 (n + 4) < 2^31
constexpr int kNoallocMaxCells = 1 << 28;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_dev[kMaxDevices] = {};

// device address of this GPU's copy of `scratch`, looked up once per device
// (with that device current); the array's name is a host-side stand-in
int scratch_address(unsigned char** out) {
    int dev = -1;
    NEST_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(NEST_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NEST_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_dev[dev]);
    return NEST_OK;
}

// the state start() leaves: the call's parameters, time t, step 0, and all
// three slot sets clear (a kernel, not a memset: callers may capture it)
__global__ __launch_bounds__(256) void noalloc1d_init_kernel(unsigned char* base, nest_prob_state ps, double t) {
    unsigned long long* slots = reinterpret_cast<unsigned long long*>(base + kNoallocSlots);
    for (size_t i = threadIdx.x; i < 3 * kNoallocSet; i += blockDim.x) slots[i] = 0ull;
    if (threadIdx.x == 0) {
        *reinterpret_cast<nest_prob_state*>(base) = ps;
        nest_run1d_state r{};
        r.t[0] = t;
        r.t[1] = t;
        *reinterpret_cast<nest_run1d_state*>(base + kNoallocRun) = r;
    }
}
}  // namespace

int nest_hip_euler1d_step_noalloc(const double* U_in, double* U_out, int n, double dx, double gamma, int bc,
                                  int limiter, int mode, double cfl, double t, double t_end, double* dt_out,
                                  void* stream) {
    if (!U_in || !U_out || U_in == U_out) return fail(NEST_ERR_INVALID, "need distinct U_in and U_out");
    if (n < 2 || n > kNoallocMaxCells)
        return fail(NEST_ERR_INVALID, "n must be in [2, " + std::to_string(kNoallocMaxCells) + "] (32-bit indexing)");
    if (int rc = check_limiter_mode(limiter, mode)) return rc;
    if (!(dx > 0.0) || !(gamma > 1.0) || bc < 0 || bc > 2 || !(cfl > 0.0))
        return fail(NEST_ERR_INVALID, "need dx > 0, gamma > 1, bc in {0,1,2}, cfl > 0");
    unsigned char* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const launch_table& L = table(mode);
    nest_prob_state ps = make_prob(dx, 1.0, 1);
    ps.t_end = t_end;
    ps.cfl = cfl;
    hipLaunchKernelGGL(noalloc1d_init_kernel, 1, 256, 0, st, base, ps, t);
    unsigned long long* slots = reinterpret_cast<unsigned long long*>(base + kNoallocSlots);
    stage1d_args a{};
    a.U = const_cast<double*>(U_in);
    a.U1 = U_out;
    a.n = n;
    a.ps = reinterpret_cast<nest_prob_state*>(base);
    a.eos = make_eos(gamma);
    a.host_dt = 0.0;
    a.bc = bc;
    // speeds of U_in into slot set 2, which step 0 reads its dt from
    stage1d_args a0 = a;
    a0.slots = slots + 2 * kNoallocSet;
    L.speeds1d(a0, st);
    // step 0 of nest_solver1d::device_step: U_in -> U_out
    nest_run1d_state* run = reinterpret_cast<nest_run1d_state*>(base + kNoallocRun);
    a.run = run;
    a.par = 0;
    a.slots = slots;
    a.slots_rd = slots + 2 * kNoallocSet;
    a.slots_clr = slots + kNoallocSet;
    step1d_kernel_args(n, 0, a);  // the solver's automatic choice
    L.step1d(limiter, a, st);
    L.fill1d(U_out, n, bc, nullptr, st);  // hand out a halo-complete state
    if (dt_out)
        NEST_HIP_TRY(hipMemcpyAsync(dt_out, reinterpret_cast<unsigned char*>(run) + offsetof(nest_run1d_state, dt),
                                    sizeof(double), hipMemcpyDeviceToDevice, st));
    NEST_HIP_TRY(hipGetLastError());
    return NEST_OK;
}

// ---- reference-compatible wrappers (reference euler1d_solver.hpp:16-87) ----
struct euler1d_gpu_t;

static nest_solver1d* as1d(euler1d_gpu_t* p) { return reinterpret_cast<nest_solver1d*>(p); }
static void report(const char* where) { std::fprintf(stderr, "nest_hip: %s: %s\n", where, g_err.c_str()); }

euler1d_gpu_t* euler1d_gpu_create(int n_interior, int n_halo, double dx, double gamma) {
    if (n_halo != 2) {
        fail(NEST_ERR_UNSUPPORTED, "n_halo must be 2 (PLM stencil)");
        report("euler1d_gpu_create");
        return nullptr;
    }
    nest_solver1d* s = nullptr;
    if (nest_hip_solver1d_create(n_interior, dx, gamma, NEST_BC_PERIODIC, NEST_LIM_MINMOD, NEST_MODE_EXACT, &s)) {
        report("euler1d_gpu_create");
        return nullptr;
    }
    return reinterpret_cast<euler1d_gpu_t*>(s);
}
void euler1d_gpu_destroy(euler1d_gpu_t* solver) { nest_hip_solver1d_destroy(as1d(solver)); }
void euler1d_gpu_upload(euler1d_gpu_t* solver, const double* h) {
    if (nest_hip_solver1d_upload(as1d(solver), h)) report("euler1d_gpu_upload");
}
void euler1d_gpu_download(euler1d_gpu_t* solver, double* h) {
    if (nest_hip_solver1d_download(as1d(solver), h)) report("euler1d_gpu_download");
}
double euler1d_gpu_compute_dt(euler1d_gpu_t* solver, double cfl) {
    double dt = std::numeric_limits<double>::quiet_NaN();
    nest_solver1d* s = as1d(solver);
    // the reference reduces over the state as uploaded (its halos are refilled
    // only inside rk2_step); the interior-only reduction is unaffected.
    if (nest_hip_solver1d_compute_dt(s, cfl, &dt)) report("euler1d_gpu_compute_dt");
    return dt;
}
void euler1d_gpu_rk2_step(euler1d_gpu_t* solver, double dt) {
    if (nest_hip_solver1d_rk2_step(as1d(solver), dt)) report("euler1d_gpu_rk2_step");
}
int euler1d_gpu_run(euler1d_gpu_t* solver, double t_final, double cfl) {
    int steps = -1;
    if (nest_hip_solver1d_run(as1d(solver), t_final, cfl, &steps)) {
        report("euler1d_gpu_run");
        return -1;
    }
    return steps;
}
double euler1d_gpu_run_timed(euler1d_gpu_t* solver, double t_final, double cfl, int* steps) {
    nest_solver1d* s = as1d(solver);
    if (!s) return std::numeric_limits<double>::quiet_NaN();
    (void)hipStreamSynchronize(s->stream);
    auto t0 = std::chrono::high_resolution_clock::now();
    int k = euler1d_gpu_run(solver, t_final, cfl);
    (void)hipStreamSynchronize(s->stream);
    auto t1 = std::chrono::high_resolution_clock::now();
    if (steps) *steps = k;
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

}  // extern "C"
