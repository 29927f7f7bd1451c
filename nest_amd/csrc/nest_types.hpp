// Plain structs shared by the host runtime (nest_hip.cpp), the host-side plan
// builder (plan.hpp) and the device kernels (kernels.hip). Everything here is
// trivially copyable and lives in device memory as flat tables.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define NEST_HD __host__ __device__
#else
#define NEST_HD
#endif

// One rectangular patch of a problem: its field_t array (extended space,
// 2-cell halo, variable-major, last index contiguous) starts at element
// offset `off` of the problem-set buffers U / U1.
struct nest_patch_desc {
    long long off;  // element offset of the patch array in U and U1
    int nx;         // interior cells along dim 0 (x, strided)
    int ny;         // interior cells along dim 1 (y, contiguous); 1D: unused
    int prob;       // index of the independent problem the patch belongs to
    int px;         // partition index of the patch along dim 0 within its problem
    int x0, y0;     // global index of the first interior cell (problem coordinates)
};

// Static geometry of one problem: an npx x npy patch grid over NX x NY cells.
// The problem-set int table holds, from `tab`: owner_x[NX] (partition index
// of every global row), owner_y[NY], then the partition starts x0[npx],
// y0[npy]. Patch (qx, qy) is patch_base + qy*npx + qx.
struct nest_prob_geom {
    int patch_base, npx, npy;
    int NX, NY;
    int bcx, bcy;  // 0 periodic, 1 outflow (zero-gradient), 2 reflecting wall
    int tab;
};

// Wave-speed maxima of the 2D stage-2 kernels: every work item folds its
// maxima into one of NEST_SPEED_SLOTS per-problem slots (item index modulo
// the slot count), each slot on its own 128-byte line, so concurrent items
// do not queue on one address; the dt finalize folds the slots (one lane
// each) and clears them. Layout: slots[(prob * NEST_SPEED_SLOTS + k) *
// NEST_SLOT_STRIDE + {0: |u|+c, 1: |v|+c}], u64 bit patterns.
constexpr int NEST_SPEED_SLOTS = 64;
constexpr int NEST_SLOT_STRIDE = 16;

// Per-problem time-stepping state, resident on the device so a whole step
// (dt reduction included) runs without a host round trip.
struct nest_prob_state {
    double dt;        // dt of the step in flight
    double t;         // simulated time at the end of the step in flight
    double t_end;     // stop time (inf for a fixed step count)
    double cfl;       // CFL number
    double dx, dy;    // cell sizes (dy unused in 1D)
    double fixed_dt;  // > 0: use this dt every step instead of the CFL dt
    unsigned long long smax[2];  // max wave speeds (|u|+c, |v|+c) the blocking compute_dt paths read back, as bit patterns
    int steps;        // steps started
    int done;         // 1: t reached t_end; 2: non-finite dt (error)
    int dim;          // 1 or 2 (dt formula)
    int pad;
    // t at the end of the last device-dt step, by step parity: step k reads
    // tn[k & 1] and writes tn[(k + 1) & 1], so the one-launch step, whose
    // every wave computes the step's dt, never reads what its own launch
    // writes
    double tn[2];
};

// Time-stepping state of the 1D solver's fused device-dt loop. Step k reads
// the time of slot k & 1 and block 0 of its launch writes slot (k + 1) & 1,
// so no block of a launch reads what another block of it writes. steps counts
// the steps taken; done is 1 once t reached t_end, 2 on a non-finite dt.
struct nest_run1d_state {
    double t[2];
    double dt;
    int steps;
    int done;
};

// One halo region of the device exchange plan: copies an ni x nj box of all
// variables from a source patch array into the ghost cells of a destination
// patch array. Coordinates are extended-space local indices. A broadcast flag
// repeats the single source row (bi) / column (bj): that is the zero-gradient
// outflow ghost fill. Periodic and patch-to-patch regions are translations.
// A mirror flag runs the source box backwards along that dimension and the
// negation mask flips the sign of the listed variables (bit k: variable k):
// that is the reflecting-wall fill (mirror image, normal momentum negated).
struct nest_region {
    long long dst_off, src_off;  // patch array offsets (elements)
    int dst_nye, src_nye;        // extended row lengths (dim 1)
    int dst_vs, src_vs;          // variable strides
    int di, dj, ni, nj;          // destination box
    int si, sj;                  // source box start
    int bi, bj;                  // broadcast flags
    int mi, mj;                  // mirror flags
    int neg;                     // variables negated (bit mask)
};

// gamma-law EOS constants: gamma, gm1 = gamma - 1.0 (evaluated as the
// reference does), inv_gm1 = 1.0 / gm1 (fast numerics only)
struct nest_eos {
    double gamma, gm1, inv_gm1;
};

// A unit of stage work: a strip of 64 consecutive extended columns starting
// at extended column c0 of patch `patch` (outputs for the 60 middle lanes),
// interior rows [r0, r1) of that patch. In gather mode a strip may run past
// the patch's last column into its dim-1 neighbours (global strips).
struct nest_item {
    int patch, c0, r0, r1;
};

// Global index mapped through a boundary condition: periodic wrap,
// zero-gradient (outflow) clamp, or reflecting-wall mirror (ghost cell -1-k
// is the image of cell k, n+k of n-1-k). Indices past the mirror image of
// the whole line (only lanes far past the domain reach them, and their
// values are not used) are clamped into range.
NEST_HD inline int nest_map_index(int g, int n, int bc) {
    if (bc == 0) {
        int r = g % n;
        return r < 0 ? r + n : r;
    }
    if (bc == 2) g = g < 0 ? -1 - g : (g >= n ? 2 * n - 1 - g : g);
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

// The fused 1D step keeps U1 in a tile of 264 positions (position t holds
// cell g0 - 4 + t of a block whose output cells are g0 .. g0 + 255) and sets
// U1 past an outflow / reflecting wall from the block's own boundary cells.
// For position t: the tile position whose U1 it copies, or -1 when t needs
// no copy (an interior cell, or a position stage 2 does not read: stage 2
// reads g0 - 2 .. gend + 1, gend = min(n, g0 + 256)). Every source lies in
// [3, 262) and is an interior cell.
NEST_HD inline int nest_step1d_wall_source(int t, int g0, int n, int bc) {
    const int g = g0 - 4 + t;
    const int gend = n < g0 + 256 ? n : g0 + 256;
    if (bc == 0 || (g >= 0 && g < n) || g < g0 - 2 || g >= gend + 2) return -1;
    return nest_map_index(g, n, bc) - (g0 - 4);
}

// The one-launch 2D step keeps U1 of a strip's 64 lanes (lane l holds global
// column gs + l - 4) in registers; past an outflow / reflecting y edge a
// lane takes U1 from the lane of its mapped column. The lane a lane of
// global column g reads (itself inside the domain; clamped into the wave
// for lanes far past the edge, whose values are never read).
NEST_HD inline int nest_step_lane_source(int lane, int g, int ny, int bcy) {
    if (bcy == 0 || (g >= 0 && g < ny)) return lane;
    const int s = lane + (nest_map_index(g, ny, bcy) - g);
    return s < 0 ? 0 : (s > 63 ? 63 : s);
}

// Variables a gathered value negates relative to its source cell: bit 1
// (rho u) past a reflecting x wall, bit 2 (rho v) past a reflecting y wall.
NEST_HD inline int nest_gather_flip(const nest_patch_desc& pd, const nest_prob_geom& G, int e, int g) {
    const int gr = pd.x0 + e - 2;
    return (G.bcx == 2 && (gr < 0 || gr >= G.NX) ? 2 : 0) | (G.bcy == 2 && (g < 0 || g >= G.NY) ? 4 : 0);
}

// Where the value of extended row e of patch p, at global column g of its
// problem, lives: patch q, extended row le, extended column lc of q (always
// an interior cell of q). Ghost rows/columns resolve to the neighbour patch
// holding that global cell, or to the boundary cell (outflow) / wrapped cell
// (periodic) — faces and corners alike. Host plan checks and the device
// gather use this same arithmetic.
NEST_HD inline void nest_gather_source(const nest_patch_desc& pd, const nest_prob_geom& G, const int* T, int e,
                                       int g, int& q, int& le, int& lc) {
    const int gc = nest_map_index(g, G.NY, G.bcy);
    const int qy = T[G.tab + G.NX + gc];
    lc = gc - T[G.tab + G.NX + G.NY + G.npx + qy] + 2;
    const int gr = nest_map_index(pd.x0 + e - 2, G.NX, G.bcx);
    const int qx = T[G.tab + gr];
    le = gr - T[G.tab + G.NX + G.NY + qx] + 2;
    q = G.patch_base + qy * G.npx + qx;
}
