// Plain structs shared by the host runtime (nest_hip.cpp) and the device
// kernels (kernels.hip). Everything here is trivially copyable and lives in
// device memory as flat tables.
#pragma once
#include <cstdint>

// One rectangular patch of a problem: its field_t array (extended space,
// 2-cell halo, variable-major, last index contiguous) starts at element
// offset `off` of the problem buffers U / U1.
struct nest_patch_desc {
    long long off;  // element offset of the patch array in U and U1
    int nx;         // interior cells along dim 0 (x, strided)
    int ny;         // interior cells along dim 1 (y, contiguous); 1D: unused
    int prob;       // index of the independent problem the patch belongs to
    int pad;
};

// Per-problem time-stepping state, resident on the device so a whole step
// (dt reduction included) runs without a host round trip.
struct nest_prob_state {
    double dt;        // dt of the step in flight
    double t;         // simulated time at the end of the step in flight
    double t_end;     // stop time (inf for a fixed step count)
    double cfl;       // CFL number
    double dx, dy;    // cell sizes (dy unused in 1D)
    double fixed_dt;  // > 0: use this dt every step instead of the CFL dt
    unsigned long long smax[2];  // max wave speed (|u|+c, |v|+c) of the newest state, as bit patterns
    int steps;        // steps started
    int done;         // 1: t reached t_end; 2: non-finite dt (error)
    int dim;          // 1 or 2 (dt formula)
    int pad;
};

// One halo region of the device exchange plan: copies an ni x nj box of all
// variables from a source patch array into the ghost cells of a destination
// patch array. Coordinates are extended-space local indices. A broadcast flag
// repeats the single source row (bi) / column (bj): that is the zero-gradient
// outflow ghost fill. Periodic and patch-to-patch regions are translations.
struct nest_region {
    long long dst_off, src_off;  // patch array offsets (elements)
    int dst_nye, src_nye;        // extended row lengths (dim 1)
    int dst_vs, src_vs;          // variable strides
    int di, dj, ni, nj;          // destination box
    int si, sj;                  // source box start
    int bi, bj;                  // broadcast flags
};

// gamma-law EOS constants: gamma, gm1 = gamma - 1.0 (evaluated as the
// reference does), inv_gm1 = 1.0 / gm1 (fast numerics only)
struct nest_eos {
    double gamma, gm1, inv_gm1;
};

// A unit of stage work: patch `patch`, a strip of 64 extended columns starting
// at c0 (outputs for the 60 middle ones), interior rows [r0, r1).
struct nest_item {
    int patch, c0, r0, r1;
};
