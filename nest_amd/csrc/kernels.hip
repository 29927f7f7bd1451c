// Forward kernels of the Euler finite-volume step for gfx950 (wave64, FP64).
//
// Compiled twice into the library: NEST_FAST=0 / namespace nest_exact
// (-ffp-contract=off, IEEE division and sqrt: bitwise the CPU numerics) and
// NEST_FAST=1 / namespace nest_fast (FMA contraction, reciprocal division).
//
// 2D stage kernels ("strip march"): one wave owns a strip of 64 consecutive
// extended columns (dim 1, contiguous in memory) of one patch and marches
// down interior rows [r0, r1) of dim 0. Per row it loads one new row of the
// conserved state (4 variables x 512 B), and:
//   * X-sweep (along the march): primitives, differences, slopes and the
//     face flux are carried in registers as a sliding window, so every
//     cell's primitives and every x-face flux are computed exactly once;
//   * Y-sweep (across lanes) on the freshly X-swept row: neighbour
//     primitives, differences, face states and fluxes move between lanes by
//     wave shuffles; lanes 2..61 own outputs, lanes 0,1,62,63 are the
//     stencil halo (the X-sweep of the halo columns is exactly what the
//     split scheme's halo exchange would have delivered);
//   * epilogue: stage 1 stores U1; stage 2 stores 0.5*(U^n + S(U1)) in place
//     of U^n and folds |u|+c, |v|+c of the new state into one of the
//     problem's wave-speed slots (u64 bit-pattern atomicMax: exact,
//     order-free); the dt finalize folds the slots.
// Two forms: stage2d_dma_kernel (rows fetched straight into a per-wave LDS
// ring, used whenever a state buffer is below 4 GiB) and stage2d_kernel
// (rows loaded into registers, with an LDS ring for the centre rows; 64-bit
// addressing, used above that size). Nothing but the state itself touches
// HBM.
#include <atomic>

#include "device_numerics.hpp"
#include "kernels.hpp"

// Minimum waves per SIMD requested for stage 1 / stage 2 of the stage kernel
// (register budget 512/N per lane); 1 = the compiler's choice.
#ifndef NEST_STAGE1_WAVES
#define NEST_STAGE1_WAVES 1
#endif
#ifndef NEST_STAGE2_WAVES
#define NEST_STAGE2_WAVES 1
#endif
// Stage 1's strip march unrolled by two with alternating row buffers (1) or
// one row buffer loaded a step ahead (0).
#ifndef NEST_STAGE1_UNROLL
#define NEST_STAGE1_UNROLL 1
#endif
// Software pipeline of the march (1): the X step of row i + 1 is issued
// together with the Y step of row i (they do not depend on each other), and
// the Y step has no branch before its stores.
#ifndef NEST_STAGE_PIPE
#define NEST_STAGE_PIPE 0
#endif
#define NEST_STAGE_OCCUPANCY \
    __attribute__((amdgpu_waves_per_eu(STAGE == 1 ? NEST_STAGE1_WAVES : NEST_STAGE2_WAVES, 8)))

namespace NEST_NS {

__device__ __forceinline__ double ld(const double* p) { return *p; }

// x with its sign bit XORed with the sign bit of `mask` (mask 0: x itself;
// 0x80000000: -x, bit for bit)
__device__ __forceinline__ double xor_sign(double x, int mask) {
    return __hiloint2double(__double2hiint(x) ^ mask, __double2loint(x));
}

__device__ __forceinline__ prim4 shfl_up4(const prim4& a) {
    return {lane_up(a.r), lane_up(a.u), lane_up(a.p), lane_up(a.q)};
}
__device__ __forceinline__ prim4 shfl_down4(const prim4& a) {
    return {lane_down(a.r), lane_down(a.u), lane_down(a.p), lane_down(a.q)};
}

__device__ __forceinline__ prim4 shfl_up4_z(const prim4& a) {
    return {lane_up_z(a.r), lane_up_z(a.u), lane_up_z(a.p), lane_up_z(a.q)};
}
__device__ __forceinline__ prim4 shfl_down4_z(const prim4& a) {
    return {lane_down_z(a.r), lane_down_z(a.u), lane_down_z(a.p), lane_down_z(a.q)};
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o);
        v = v > w ? v : w;
    }
    return v;
}

// Slot value (bit pattern) of a wave speed. Fast numerics floor density and
// pressure with hardware max, which drops a NaN operand, so a NaN or infinity
// in a cell's density or energy need not reach its speed: such a cell
// reports the all-ones pattern, which ranks above every speed in the slots'
// u64 maxima and reads back as a NaN (the run then stops with a non-finite
// dt, as it does in exact numerics, where the floors pass a NaN on). Tested
// on the exponent bits, which no floating-point assumption of the compiler
// can fold away. Momenta reach the speeds without a floor.
__device__ __forceinline__ unsigned long long speed_slot_bits(double s, double rho, double E) {
    unsigned long long b = (unsigned long long)__double_as_longlong(s);
#if NEST_FAST
    const unsigned hr = (unsigned)__double2hiint(rho) & 0x7fffffffu;
    const unsigned hE = (unsigned)__double2hiint(E) & 0x7fffffffu;
    b = (hr > hE ? hr : hE) >= 0x7ff00000u ? ~0ull : b;
#else
    (void)rho;
    (void)E;
#endif
    return b;
}

// |u|+c and |v|+c of one cell, as compute_dt_2d (euler2d.hpp:254-263)
__device__ __forceinline__ void cell_speeds2d(double r, double mu, double mv, double E, const eos_c& e,
                                              double& sx, double& sy) {
    double rho = dmax(r, 1e-10);
    double u = ddiv(mu, rho);
    double v = ddiv(mv, rho);
#if NEST_FAST
    double p = dmax(e.gm1 * __builtin_fma(-0.5, mu * u + mv * v, E), 1e-10);
#else
    double ke = 0.5 * rho * (u * u + v * v);
    double p = dmax(e.gm1 * (E - ke), 1e-10);
#endif
    double c = sound_speed(rho, p, e);
    sx = __builtin_fabs(u) + c;
    sy = __builtin_fabs(v) + c;
}

template <int LIM, int STAGE>
__global__ __launch_bounds__(256) NEST_STAGE_OCCUPANCY void stage2d_kernel(stage_args a) {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    // Per wave: a ring of the last 4 loaded rows (4 variables x 64 lanes each)
    // in LDS. A row enters the X window 3 iterations before it is the centre
    // row whose conserved state the update needs; keeping it here instead of
    // re-reading it from memory halves the stage's HBM reads (the L2 does not
    // hold rows that long at full occupancy). Each lane touches only its own
    // column, so no barrier is needed.
    __shared__ double ring_all[4 * 4 * 4 * 64];
    double* const ring = ring_all + (threadIdx.x >> 6) * (4 * 4 * 64) + lane;
    auto ring_put = [&](int e, const double* u) {
        double* q = ring + (e & 3) * (4 * 64);
        q[0] = u[0];
        q[64] = u[1];
        q[128] = u[2];
        q[192] = u[3];
    };
    auto ring_get = [&](int e, double* u) {
        const double* q = ring + (e & 3) * (4 * 64);
        u[0] = q[0];
        u[1] = q[64];
        u[2] = q[128];
        u[3] = q[192];
    };
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6); it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        nest_prob_state* ps = a.probs + pd.prob;
        if (!(a.host_dt > 0.0) && ps->done) continue;
        const double dt = a.host_dt > 0.0 ? a.host_dt : ps->dt;
        const double dtdx = dt / ps->dx;
        const double dtdy = dt / ps->dy;

        const double* Ubuf = STAGE == 1 ? a.U : a.U1;
        // stage 2 writes 0.5 (U^n + S(U1)) to a.Uo when given (U^n is then
        // left intact), else in place of U^n; U^n is read at the output cell
        // through the wave-uniform offset u0_off from the output pointer
        double* Obuf = STAGE == 1 ? a.U1 : (a.Uo ? a.Uo : a.U);
        const ptrdiff_t u0_off = a.U - Obuf;
        // Per lane: where this lane's column lives for interior rows (base,
        // row length, variable stride) and where its output goes.
        const double* base;
        double* out;
        int snye, svs;
        bool store;
        int g = 0;  // global column (gather mode)
        nest_prob_geom G{};
        if (a.gather) {
            // global strip: lane -> global column g -> owning patch (q) and its
            // local column; outputs for lanes 2..61 inside the domain
            G = a.geoms[pd.prob];
            g = pd.y0 + w.c0 + lane - 2;
            int q, le, lc;
            nest_gather_source(pd, G, a.itab, 2, g, q, le, lc);
            const nest_patch_desc qd = a.patches[q];
            snye = qd.ny + 4;
            svs = (qd.nx + 4) * snye;
            base = Ubuf + qd.off + lc;
            out = Obuf + qd.off + lc;
            store = lane >= 2 && lane < 62 && g < G.NY;
        } else {
            // per-patch strip over the patch's own extended columns (ghosts filled)
            const int nye = pd.ny + 4;
            const int ec = w.c0 + lane;
            const int ecl = ec < nye ? ec : nye - 1;  // lanes past the patch: any in-bounds column
            snye = nye;
            svs = (pd.nx + 4) * nye;
            base = Ubuf + pd.off + ecl;
            out = Obuf + pd.off + ecl;
            store = lane >= 2 && lane < 62 && ec >= 2 && ec < pd.ny + 2;
        }
        const int vs = svs;  // variable stride of the output patch

        auto load_row = [&](int e, double* u) {
            const double* p;
            int s;
            if (a.gather && (e < 2 || e >= pd.nx + 2)) {  // ghost rows: wave-uniform slow path
                int q, le, lc;
                nest_gather_source(pd, G, a.itab, e, g, q, le, lc);
                const nest_patch_desc qd = a.patches[q];
                const int qnye = qd.ny + 4;
                s = (qd.nx + 4) * qnye;
                p = Ubuf + qd.off + (size_t)le * qnye + lc;
            } else {
                p = base + (size_t)e * snye;
                s = svs;
            }
            u[0] = ld(p);
            u[1] = ld(p + s);
            u[2] = ld(p + 2 * (size_t)s);
            u[3] = ld(p + 3 * (size_t)s);
            if (a.gather) {  // past a reflecting wall: the normal momentum is negated
                const int fl = nest_gather_flip(pd, G, e, g);
                if (fl & 2) u[1] = -u[1];
                if (fl & 4) u[2] = -u[2];
            }
        };

        // ---- X window warm-up: rows i = r0-2 .. r0+1 (extended rows r0 .. r0+3).
        // Carried down the march (registers): W1 = W(i+1), dW1 = W(i+1) - W(i),
        // FL = W(i) + 0.5*slope(i) (left state of face i+1/2) and Fm = F(i-1/2).
        prim4 W1, dW1, FL;
        double Fm[4];
        {
            double A[4];
            load_row(w.r0, A);
            prim4 Wm2 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            load_row(w.r0 + 1, A);
            prim4 Wm1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            load_row(w.r0 + 2, A);
            ring_put(w.r0 + 2, A);
            prim4 W0 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            load_row(w.r0 + 3, A);
            ring_put(w.r0 + 3, A);
            W1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            prim4 d0 = sub4(W0, Wm1);
            dW1 = sub4(W1, W0);
            prim4 s_m1 = slope4<LIM>(sub4(Wm1, Wm2), d0);
            prim4 s0 = slope4<LIM>(d0, dW1);
            face_flux<true, true>(face_plus(Wm1, s_m1), face_minus(W0, s0), a.eos, Fm);
            FL = face_plus(W0, s0);
        }

        unsigned long long smx = 0, smy = 0;
        double UA[4], UB[4];  // rows e = r0+4 (and r0+5, stage 1) for the first X steps
        load_row(w.r0 + 4, UA);
        if constexpr (STAGE == 1 && NEST_STAGE1_UNROLL) load_row(min(w.r0 + 5, pd.nx + 3), UB);
        // X step of row j: consumes the window and the row loaded for it (Ub,
        // extended row j+4), refills Ub with the row two steps ahead (two row
        // buffers alternate, so each load has two steps to arrive), writes
        // the X-swept state of row j to ux (x_sweep_t update,
        // euler2d.hpp:144-148)
        auto xstep = [&](int j, double* ux, double* Ub, int ahead) {
            double Uc[4];
            ring_get(j + 2, Uc);
            ring_put(j + 4, Ub);
            prim4 W2 = prim2d(Ub[0], Ub[1], Ub[2], Ub[3], true, a.eos);
            load_row(min(j + 4 + ahead, pd.nx + 3), Ub);
            prim4 dW2 = sub4(W2, W1);
            prim4 s1 = slope4<LIM>(dW1, dW2);
            double Fp[4];
            face_flux<true, true>(FL, face_minus(W1, s1), a.eos, Fp);
            FL = face_plus(W1, s1);
            ux[0] = Uc[0] - dtdx * (Fp[0] - Fm[0]);
            ux[1] = Uc[1] - dtdx * (Fp[1] - Fm[1]);
            ux[2] = Uc[2] - dtdx * (Fp[3] - Fm[3]);
            ux[3] = Uc[3] - dtdx * (Fp[2] - Fm[2]);
            W1 = W2;
            dW1 = dW2;
#pragma unroll
            for (int k = 0; k < 4; ++k) Fm[k] = Fp[k];
        };
        // Y step of row i on its X-swept state (euler2d.hpp:176-236), RK2
        // combine and wave speeds (stage 2), store
        auto ystep = [&](int i, const double* ux) {
            double* po = out + (size_t)(i + 2) * snye;
            double U0[4];  // U^n at the output cell (stage 2 combine), issued early
            if constexpr (STAGE == 2) {
                // (every lane's po is an in-bounds cell: lanes that store
                // nothing load a value they do not use)
                if (NEST_STAGE_PIPE || store) {
                    const double* pu = po + u0_off;
                    U0[0] = pu[0];
                    U0[1] = pu[vs];
                    U0[2] = pu[2 * (size_t)vs];
                    U0[3] = pu[3 * (size_t)vs];
                }
            }
            prim4 Y = prim2d(ux[0], ux[1], ux[2], ux[3], false, a.eos);
            prim4 dl = sub4(Y, shfl_up4(Y));
            prim4 dr = shfl_down4(dl);
            prim4 sy = slope4<LIM>(dl, dr);
            prim4 Rf = shfl_down4(face_minus(Y, sy));
            double G[4];
            face_flux<true, true>(face_plus(Y, sy), Rf, a.eos, G);
            double Gm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Gm[k] = lane_up(G[k]);
            double o0 = ux[0] - dtdy * (G[0] - Gm[0]);
            double o1 = ux[1] - dtdy * (G[3] - Gm[3]);
            double o2 = ux[2] - dtdy * (G[1] - Gm[1]);
            double o3 = ux[3] - dtdy * (G[2] - Gm[2]);
            if constexpr (STAGE == 2) {
                // SSP-RK2 combine, U^{n+1} = 0.5*(U^n + U2)
                if (NEST_STAGE_PIPE || store) {
                    o0 = 0.5 * (U0[0] + o0);
                    o1 = 0.5 * (U0[1] + o1);
                    o2 = 0.5 * (U0[2] + o2);
                    o3 = 0.5 * (U0[3] + o3);
                    double sx, sy2;
                    cell_speeds2d(o0, o1, o2, o3, a.eos, sx, sy2);
                    unsigned long long bx = speed_slot_bits(sx, o0, o3);
                    unsigned long long by = (unsigned long long)__double_as_longlong(sy2);
                    smx = (store && bx > smx) ? bx : smx;
                    smy = (store && by > smy) ? by : smy;
                }
            }
            if (store) {
                po[0] = o0;
                po[vs] = o1;
                po[2 * (size_t)vs] = o2;
                po[3 * (size_t)vs] = o3;
            }
        };
        if constexpr (NEST_STAGE_PIPE && STAGE == 1) {
            // two row buffers alternate (as below); X step of row i + 1 with
            // the Y step of row i
            double Vc[4], Vn[4];
            xstep(w.r0, Vc, UA, 2);
            int i = w.r0;
            for (; i + 2 < w.r1; i += 2) {
                xstep(i + 1, Vn, UB, 2);
                ystep(i, Vc);
                xstep(i + 2, Vc, UA, 2);
                ystep(i + 1, Vn);
            }
            if (i + 1 < w.r1) {
                xstep(i + 1, Vn, UB, 2);
                ystep(i, Vc);
                ystep(i + 1, Vn);
            } else {
                ystep(i, Vc);
            }
        } else if constexpr (NEST_STAGE_PIPE) {
            double Vc[4];
            xstep(w.r0, Vc, UA, 1);
            int i = w.r0;
            for (; i + 1 < w.r1; ++i) {
                double Vn[4];
                xstep(i + 1, Vn, UA, 1);
                ystep(i, Vc);
#pragma unroll
                for (int k = 0; k < 4; ++k) Vc[k] = Vn[k];
            }
            ystep(i, Vc);
        } else if constexpr (STAGE == 1 && NEST_STAGE1_UNROLL) {
            // the march, unrolled by two so that the two row buffers
            // alternate without register copies (a copy would wait for the
            // load at once): each row load has two steps to arrive
            for (int i = w.r0; i < w.r1; i += 2) {
                double V[4];
                xstep(i, V, UA, 2);
                ystep(i, V);
                if (i + 1 < w.r1) {
                    xstep(i + 1, V, UB, 2);
                    ystep(i + 1, V);
                }
            }
        } else {
            // one row buffer, loaded one step ahead (stage 2 holds more state,
            // U^n and the wave-speed maxima: this keeps it at 3 waves per SIMD)
            for (int i = w.r0; i < w.r1; ++i) {
                double V[4];
                xstep(i, V, UA, 1);
                ystep(i, V);
            }
        }
        if constexpr (STAGE == 2) {
            smx = wave_max_u64(smx);
            smy = wave_max_u64(smy);
            if (lane == 0) {
                unsigned long long* sl =
                    a.slots + ((size_t)pd.prob * NEST_SPEED_SLOTS + (it & (NEST_SPEED_SLOTS - 1))) * NEST_SLOT_STRIDE;
                atomicMax(sl, smx);
                atomicMax(sl + 1, smy);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Stage kernel with the row loads going straight into LDS.
//
// Same strip march and arithmetic as stage2d_kernel, but every row of the
// input state (and, in stage 2, every row of U^n at the output cells) is
// fetched with global_load_lds_dword into a per-wave LDS ring instead of
// into registers. Nothing loaded is carried in VGPRs across the march, so no
// register copy at the loop back-edge waits on a load in flight, and each
// row is requested two march steps before it is used (stage 2's U^n row one
// step before), with explicit vmcnt waits that count only the memory
// operations known to be issued after the awaited one.
//
// Dword view of a row: the 64 doubles of one variable are 512 contiguous
// bytes of LDS, filled by two instructions of 64 lanes x 4 B. Lane l of the
// first fetches half (l & 1) of the strip's double (l >> 1), of the second
// half (l & 1) of double 32 + (l >> 1). The source addresses are those the
// register path would use for those doubles (shuffled from the owning lane),
// so every access stays inside the state buffers.
//
// Per wave, LDS: 4 row slots x 4 variables x 64 doubles (8 KB); stage 2 adds
// one slot for the U^n row (2 KB). Dynamic shared memory: waves per block x
// that (4 waves of stage 2 per SIMD: 16 x 10 KB = the CU's 160 KB).
// ---------------------------------------------------------------------------
constexpr int kRowSlot = 4 * 64;  // doubles per ring slot (4 variables x 64 columns)
constexpr int kWaveLds1 = 4 * kRowSlot;                 // stage 1: 4 row slots
constexpr int kWaveLds2 = 4 * kRowSlot + kRowSlot;  // stage 2: + 1 U^n slot

// 4 waves per SIMD: both stages fit 128 VGPRs, and stage 2's LDS (10 KB per
// wave) fills the CU's 160 KB exactly
template <int LIM, int STAGE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void stage2d_dma_kernel(stage_args a) {
    extern __shared__ double lds_dyn[];
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    // the wave's index within the block, as a wave-uniform value: the work
    // item, its patch and rows, and the LDS slot addresses then live in
    // scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int kWave = STAGE == 1 ? kWaveLds1 : kWaveLds2;
    double* const ring = lds_dyn + wave * kWave;  // row slot of extended row e: (e & 3)
    double* const u0slot = ring + 4 * kRowSlot;   // stage 2: U^n row of the output row being marched
    const int hl = lane >> 1;
    const int part = (lane & 1) * 4;
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + wave; it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        nest_prob_state* ps = a.probs + pd.prob;
        if (!(a.host_dt > 0.0) && ps->done) continue;
        const double dt = a.host_dt > 0.0 ? a.host_dt : ps->dt;
        const double dtdx = dt / ps->dx;
        const double dtdy = dt / ps->dy;

        const double* Ubuf = STAGE == 1 ? a.U : a.U1;
        double* Obuf = STAGE == 1 ? a.U1 : (a.Uo ? a.Uo : a.U);
        const char* const Ub8 = reinterpret_cast<const char*>(Ubuf);
        const char* const U08 = reinterpret_cast<const char*>(a.U);  // stage 2: U^n at the output cells
        char* const Ob8 = reinterpret_cast<char*>(Obuf);
        // Per lane: byte offset (in every state buffer alike: they share the
        // patch layout and are below 4 GiB) of this lane's column at extended
        // row 0, its row length and variable stride in bytes.
        unsigned coff;
        int snye, svs;
        bool store;
        int g = 0;
        nest_prob_geom G{};
        if (a.gather) {
            G = a.geoms[pd.prob];
            g = pd.y0 + w.c0 + lane - 2;
            int q, le, lc;
            nest_gather_source(pd, G, a.itab, 2, g, q, le, lc);
            const nest_patch_desc qd = a.patches[q];
            snye = (qd.ny + 4) * 8;
            svs = (qd.nx + 4) * (qd.ny + 4) * 8;
            coff = (unsigned)((qd.off + lc) * 8);
            store = lane >= 2 && lane < 62 && g < G.NY;
        } else {
            const int nye = pd.ny + 4;
            const int ec = w.c0 + lane;
            const int ecl = ec < nye ? ec : nye - 1;
            snye = nye * 8;
            svs = (pd.nx + 4) * nye * 8;
            coff = (unsigned)((pd.off + ecl) * 8);
            store = lane >= 2 && lane < 62 && ec >= 2 && ec < pd.ny + 2;
        }
        // the dword view's source columns: strip doubles hl and 32 + hl
        const unsigned bA = (unsigned)__shfl((int)coff, hl) + part;
        const unsigned bB = (unsigned)__shfl((int)coff, 32 + hl) + part;
        const int rsA = __shfl(snye, hl), rsB = __shfl(snye, 32 + hl);
        const int vsA = __shfl(svs, hl), vsB = __shfl(svs, 32 + hl);

        // row min(e, last extended row) of the input state into ring slot
        // (e & 3): 8 memory operations
        const int elast = pd.nx + 3;
        auto issue_row = [&](int e_slot) {
            const int e = min(e_slot, elast);
            unsigned oA, oB;
            int sA, sB;
            if (a.gather && (e < 2 || e >= pd.nx + 2)) {  // ghost rows: wave-uniform slow path
                int q, le, lc;
                nest_gather_source(pd, G, a.itab, e, g, q, le, lc);
                const nest_patch_desc qd = a.patches[q];
                const int qnye = qd.ny + 4;
                const unsigned o = (unsigned)((qd.off + (long long)le * qnye + lc) * 8);
                const int s8 = (qd.nx + 4) * qnye * 8;
                oA = (unsigned)__shfl((int)o, hl) + part;
                oB = (unsigned)__shfl((int)o, 32 + hl) + part;
                sA = __shfl(s8, hl);
                sB = __shfl(s8, 32 + hl);
            } else {
                oA = bA + (unsigned)e * (unsigned)rsA;
                oB = bB + (unsigned)e * (unsigned)rsB;
                sA = vsA;
                sB = vsB;
            }
            double* slot = ring + (e_slot & 3) * kRowSlot;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                __builtin_amdgcn_global_load_lds(Ub8 + (oA + (unsigned)k * (unsigned)sA), (lds_void_ptr)(slot + k * 64), 4, 0,
                                                 0);
                __builtin_amdgcn_global_load_lds(Ub8 + (oB + (unsigned)k * (unsigned)sB), (lds_void_ptr)(slot + k * 64 + 32),
                                                 4, 0, 0);
            }
        };
        // stage 2: U^n of output row min(i, r1 - 1) (extended row + 2) into
        // the U^n slot: 8 memory operations
        auto issue_u0 = [&](int i_want) {
            const int i = min(i_want, w.r1 - 1);
            const unsigned oA = bA + (unsigned)(i + 2) * (unsigned)rsA;
            const unsigned oB = bB + (unsigned)(i + 2) * (unsigned)rsB;
            double* slot = u0slot;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                __builtin_amdgcn_global_load_lds(U08 + (oA + (unsigned)k * (unsigned)vsA), (lds_void_ptr)(slot + k * 64), 4, 0,
                                                 0);
                __builtin_amdgcn_global_load_lds(U08 + (oB + (unsigned)k * (unsigned)vsB), (lds_void_ptr)(slot + k * 64 + 32),
                                                 4, 0, 0);
            }
        };
        // past a reflecting wall the normal momentum is negated (gather
        // mode; per-patch strips read ghost cells the plan filled): rho v on
        // lanes past a y wall, rho u on rows past an x wall
        // (negation is a flip of the sign bit: XOR of the high word with a
        // mask that is 0 where nothing is negated)
        const int sign_v = a.gather && (nest_gather_flip(pd, G, 2, g) & 4) ? (int)0x80000000 : 0;
        auto read_row = [&](int e, double* u) {
            const double* q = ring + (e & 3) * kRowSlot + lane;
            const int sign_u = a.gather && G.bcx == 2 && (nest_gather_flip(pd, G, e, g) & 2) ? (int)0x80000000 : 0;
            u[0] = q[0];
            u[1] = xor_sign(q[64], sign_u);
            u[2] = xor_sign(q[128], sign_v);
            u[3] = q[192];
        };

        // ---- prologue: rows r0 .. r0+3 (X window warm-up), then the first
        // two prefetches (rows r0+4, r0+5) and stage 2's first U^n row
        issue_row(w.r0);
        issue_row(w.r0 + 1);
        issue_row(w.r0 + 2);
        issue_row(w.r0 + 3);
        wait_vmcnt<0>();
        prim4 W1, dW1, FL;
        double Fm[4];
        {
            double A[4];
            read_row(w.r0, A);
            prim4 Wm2 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(w.r0 + 1, A);
            prim4 Wm1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(w.r0 + 2, A);
            prim4 W0 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(w.r0 + 3, A);
            W1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            wait_lgkmcnt0();  // slots of rows r0, r0+1 are refilled next
            issue_row(w.r0 + 4);
            compiler_fence();
            issue_row(w.r0 + 5);
            if constexpr (STAGE == 2) {
                compiler_fence();
                issue_u0(w.r0);
            }
            prim4 d0 = sub4(W0, Wm1);
            dW1 = sub4(W1, W0);
            prim4 s_m1 = slope4<LIM>(sub4(Wm1, Wm2), d0);
            prim4 s0 = slope4<LIM>(d0, dW1);
            face_flux<true, true>(face_plus(Wm1, s_m1), face_minus(W0, s0), a.eos, Fm);
            FL = face_plus(W0, s0);
        }

        unsigned long long smx = 0, smy = 0;
        for (int i = w.r0; i < w.r1; ++i) {
            // Row i+4 was issued two steps ago (prologue for i < r0+2); issued
            // after it for sure: the whole previous step's fetches (row i+5,
            // and in stage 2 the U^n rows issued after it), or the
            // prologue's later groups.
            if constexpr (STAGE == 1) wait_vmcnt<8>();
            else wait_vmcnt<16>();
            double Uc[4], Ub[4];
            read_row(i + 2, Uc);
            read_row(i + 4, Ub);
            wait_lgkmcnt0();  // the centre row's slot is refilled next
            issue_row(i + 6);

            // X step of row i (x_sweep_t update, euler2d.hpp:144-148)
            double ux[4];
            {
                prim4 W2 = prim2d(Ub[0], Ub[1], Ub[2], Ub[3], true, a.eos);
                prim4 dW2 = sub4(W2, W1);
                prim4 s1 = slope4<LIM>(dW1, dW2);
                double Fp[4];
                face_flux<true, true>(FL, face_minus(W1, s1), a.eos, Fp);
                FL = face_plus(W1, s1);
                ux[0] = Uc[0] - dtdx * (Fp[0] - Fm[0]);
                ux[1] = Uc[1] - dtdx * (Fp[1] - Fm[1]);
                ux[2] = Uc[2] - dtdx * (Fp[3] - Fm[3]);
                ux[3] = Uc[3] - dtdx * (Fp[2] - Fm[2]);
                W1 = W2;
                dW1 = dW2;
#pragma unroll
                for (int k = 0; k < 4; ++k) Fm[k] = Fp[k];
            }
            // Y step of row i (euler2d.hpp:176-236). The values lanes 0 and
            // 63 receive from outside the wave (zero) reach only lanes 0, 1,
            // 62 and 63, which store nothing.
            prim4 Y = prim2d(ux[0], ux[1], ux[2], ux[3], false, a.eos);
            prim4 dl = sub4(Y, shfl_up4_z(Y));
            prim4 dr = shfl_down4_z(dl);
            prim4 sy = slope4<LIM>(dl, dr);
            prim4 Rf = shfl_down4_z(face_minus(Y, sy));
            double Gf[4];
            face_flux<true, true>(face_plus(Y, sy), Rf, a.eos, Gf);
            double Gm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Gm[k] = lane_up_z(Gf[k]);
            double o0 = ux[0] - dtdy * (Gf[0] - Gm[0]);
            double o1 = ux[1] - dtdy * (Gf[3] - Gm[3]);
            double o2 = ux[2] - dtdy * (Gf[1] - Gm[1]);
            double o3 = ux[3] - dtdy * (Gf[2] - Gm[2]);
            if constexpr (STAGE == 2) {
                // U^n row i was issued at the end of the previous step
                // (prologue for i = r0); issued after it for sure: this
                // step's row fetch
                wait_vmcnt<8>();
                const double* q = u0slot + lane;
                const double z0 = q[0], z1 = q[64], z2 = q[128], z3 = q[192];
                wait_lgkmcnt0();  // the slot is refilled next
                issue_u0(i + 1);
                // SSP-RK2 combine, U^{n+1} = 0.5*(U^n + U2)
                o0 = 0.5 * (z0 + o0);
                o1 = 0.5 * (z1 + o1);
                o2 = 0.5 * (z2 + o2);
                o3 = 0.5 * (z3 + o3);
                double sx, sy2;
                cell_speeds2d(o0, o1, o2, o3, a.eos, sx, sy2);
                unsigned long long bx = speed_slot_bits(sx, o0, o3);
                unsigned long long by = (unsigned long long)__double_as_longlong(sy2);
                smx = (store && bx > smx) ? bx : smx;
                smy = (store && by > smy) ? by : smy;
            }
            if (store) {
                const unsigned o = coff + (unsigned)(i + 2) * (unsigned)snye;
                *reinterpret_cast<double*>(Ob8 + o) = o0;
                *reinterpret_cast<double*>(Ob8 + (o + (unsigned)svs)) = o1;
                *reinterpret_cast<double*>(Ob8 + (o + 2u * (unsigned)svs)) = o2;
                *reinterpret_cast<double*>(Ob8 + (o + 3u * (unsigned)svs)) = o3;
            }
        }
        // the last steps' prefetches must land before the ring is reused by
        // the next work item (or the wave ends)
        wait_vmcnt<0>();
        if constexpr (STAGE == 2) {
            smx = wave_max_u64(smx);
            smy = wave_max_u64(smy);
            if (lane == 0) {
                unsigned long long* sl =
                    a.slots + ((size_t)pd.prob * NEST_SPEED_SLOTS + (it & (NEST_SPEED_SLOTS - 1))) * NEST_SLOT_STRIDE;
                atomicMax(sl, smx);
                atomicMax(sl + 1, smy);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// One whole SSP-RK2 step per launch: U^{n+1} = 0.5 (U^n + S(S(U^n))), with
// U1 = S(U^n) kept on chip (it never touches HBM).
//
// One wave owns a strip of 64 consecutive global columns (lane l holds
// column gs + l - 4; lanes 4..59 produce output, 56 columns) and marches
// down rows r0 .. r1 - 1 of one dim-0 partition. Iteration k runs two
// independent chains, so the compiler can interleave them:
//   * stage 1 of row k: the X-then-Y split stage of stage2d_dma_kernel on
//     U^n, giving U1(k) on lanes 2..61;
//   * stage 2 of row k - 3: the same split stage on U1, from a queue of the
//     last three U1 rows in registers (U1(k-3) is the centre row, U1(k-1)
//     the row entering the X window), then the RK2 combine with U^n(k-3),
//     the wave speeds of the new state and the store.
// U^n rows are fetched straight into a per-wave LDS ring of 8 rows (interior
// row i in slot i & 7), two iterations before stage 1 needs them; the ring
// still holds row k - 3 for the combine.
//
// Every U1 value stage 2 reads is the one the split scheme's exchange of U1
// would deliver (the fill between the stages, sod2d.cpp:318-356): inside the
// domain and across periodic edges stage 1 computes it directly from the
// gathered U^n (a cell past a periodic edge sees the stencil of the cell it
// wraps to, so it is that cell's U1 bit for bit); past an outflow or
// reflecting edge it is copied from the boundary cells (lanes: a shuffle
// from the lane of the mapped column; rows: the queue entry of the mapped
// row), with the wall-normal momentum negated at a reflecting wall. Corners
// take both mappings, as the plan's corner regions do.
//
// The output goes out of place (a.Uo): neighbouring strips read U^n four
// columns and rows into this one. A problem already done (device dt) copies
// its state through, so the two buffers stay interchangeable.
// LDS: 8 rows x 4 variables x 64 doubles = 16 KB per wave.
// ---------------------------------------------------------------------------
// dt of the next step of one problem at time t from the max wave speeds
// (bit patterns of |u|+c, |v|+c) of its newest state (CFL or fixed, clamped
// to t_end): returns 0 (step), 1 (t_end reached) or 2 (non-finite or
// non-positive dt, left in dt).
__device__ inline int step_dt(const nest_prob_state& s, double t, unsigned long long bx, unsigned long long by,
                              double& dt) {
    if (!(t < s.t_end)) return 1;
    const double sx = __longlong_as_double((long long)bx);
    const double sy = __longlong_as_double((long long)by);
    if (s.fixed_dt > 0.0) {
        dt = s.fixed_dt;
    } else if (s.dim == 1) {
        dt = s.cfl * s.dx / sx;
    } else {
        double dtx = 1e10, dty = 1e10;
        if (sx > 0.0) dtx = dmin(dtx, s.dx / sx);
        if (sy > 0.0) dty = dmin(dty, s.dy / sy);
        dt = s.cfl * dmin(dtx, dty);
    }
    dt = dmin(dt, s.t_end - t);
    if (!dt_ok(dt) || bits_nan(sx) || bits_nan(sy)) return 2;
    return 0;
}
constexpr int kStepRing = 8;
constexpr int kWaveLdsStep = kStepRing * kRowSlot;
// 1: the scheduler keeps stage 1 and stage 2 of an iteration apart
#ifndef NEST_STEP_SCHED_BARRIER
#define NEST_STEP_SCHED_BARRIER 1
#endif

// U1OUT: the stage-1 state of the cells the launch updates is also written
// to a.U1 (a rollout that the reverse sweep differentiates keeps it).
template <int LIM, bool U1OUT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 4))) void step2d_kernel(stage_args a) {
    extern __shared__ double lds_dyn[];
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double* const ring = lds_dyn + wave * kWaveLdsStep;
    const int hl = lane >> 1;
    const int part = (lane & 1) * 4;
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + wave; it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        nest_prob_state* ps = a.probs + pd.prob;
        const nest_prob_geom G = a.geoms[pd.prob];
        const char* const Ub8 = reinterpret_cast<const char*>(a.U);
        char* const Ob8 = reinterpret_cast<char*>(a.Uo);
        // lane -> global column g -> the patch holding it in the home
        // partition's rows, its local column, row length and variable stride
        // (byte offsets: every state buffer is below 4 GiB)
        const int gs = pd.y0 + w.c0;
        const int g = gs + lane - 4;
        unsigned coff;
        int snye, svs;
        {
            int q, le, lc;
            nest_gather_source(pd, G, a.itab, 2, g, q, le, lc);
            const nest_patch_desc qd = a.patches[q];
            snye = (qd.ny + 4) * 8;
            svs = (qd.nx + 4) * (qd.ny + 4) * 8;
            coff = (unsigned)((qd.off + lc) * 8);
        }
        const bool store = lane >= 4 && lane < 60 && g < G.NY;
        const int r0 = w.r0, r1 = w.r1;
        // The step's dt: the caller's, or (device dt) computed by every work
        // item from the speed slots the previous step wrote, with
        // finalize_kernel's arithmetic (-1: the problem was already done).
        // The problem's first work item records the step (dt, t, step count
        // or done flag) and clears the slot set the next step writes. No
        // wave reads what another wave of the launch writes: the time comes
        // from tn[par], the record goes to tn[par ^ 1], and a done flag set
        // here gives every item the same outcome (pass the state through).
        double dt = a.host_dt;
        bool pass = false;
        if (!(a.host_dt > 0.0)) {
            const unsigned long long* sr = a.slots_rd + ((size_t)pd.prob * NEST_SPEED_SLOTS + lane) * NEST_SLOT_STRIDE;
            const unsigned long long bx = wave_max_u64(sr[0]);
            const unsigned long long by = wave_max_u64(sr[1]);
            const double ctin = ps->tn[a.par];
            double cdt = 0.0;
            const int cdone = ps->done ? -1 : step_dt(*ps, ctin, bx, by, cdt);
            if (it == a.first_item[pd.prob]) {
                unsigned long long* sc = a.slots_clr + ((size_t)pd.prob * NEST_SPEED_SLOTS + lane) * NEST_SLOT_STRIDE;
                sc[0] = 0ull;
                sc[1] = 0ull;
                if (lane == 0) {
                    if (cdone == 0) {
                        ps->dt = cdt;
                        ps->t = ctin + cdt;
                        ps->tn[a.par ^ 1] = ctin + cdt;
                        ps->steps += 1;
                    } else if (cdone > 0) {
                        ps->done = cdone;
                        if (cdone == 2) ps->dt = cdt;
                    }
                }
            }
            dt = cdt;
            pass = cdone != 0;
        }
        if (pass) {
            // finished problem: its state passes through to the output buffer
            if (store) {
                for (int i = r0; i < r1; ++i) {
                    const unsigned o = coff + (unsigned)(i + 2) * (unsigned)snye;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const unsigned ov = o + (unsigned)v * (unsigned)svs;
                        *reinterpret_cast<double*>(Ob8 + ov) = *reinterpret_cast<const double*>(Ub8 + ov);
                    }
                }
            }
            continue;
        }
        const double dtdx = dt / ps->dx;
        const double dtdy = dt / ps->dy;

        // the dword view's source columns: strip doubles hl and 32 + hl
        const unsigned bA = (unsigned)__shfl((int)coff, hl) + part;
        const unsigned bB = (unsigned)__shfl((int)coff, 32 + hl) + part;
        const int rsA = __shfl(snye, hl), rsB = __shfl(snye, 32 + hl);
        const int vsA = __shfl(svs, hl), vsB = __shfl(svs, 32 + hl);

        // interior row i of U^n into ring slot (i & 7): 8 memory operations.
        // Rows outside the home patch resolve through the owner tables and
        // the boundary condition (wave-uniform slow path).
        auto issue_row = [&](int i) {
            const int e = i + 2;
            unsigned oA, oB;
            int sA, sB;
            if (e < 2 || e >= pd.nx + 2) {
                int q, le, lc;
                nest_gather_source(pd, G, a.itab, e, g, q, le, lc);
                const nest_patch_desc qd = a.patches[q];
                const int qnye = qd.ny + 4;
                const unsigned o = (unsigned)((qd.off + (long long)le * qnye + lc) * 8);
                const int s8 = (qd.nx + 4) * qnye * 8;
                oA = (unsigned)__shfl((int)o, hl) + part;
                oB = (unsigned)__shfl((int)o, 32 + hl) + part;
                sA = __shfl(s8, hl);
                sB = __shfl(s8, 32 + hl);
            } else {
                oA = bA + (unsigned)e * (unsigned)rsA;
                oB = bB + (unsigned)e * (unsigned)rsB;
                sA = vsA;
                sB = vsB;
            }
            double* slot = ring + (i & (kStepRing - 1)) * kRowSlot;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                __builtin_amdgcn_global_load_lds(Ub8 + (oA + (unsigned)k * (unsigned)sA), (lds_void_ptr)(slot + k * 64), 4, 0,
                                                 0);
                __builtin_amdgcn_global_load_lds(Ub8 + (oB + (unsigned)k * (unsigned)sB), (lds_void_ptr)(slot + k * 64 + 32),
                                                 4, 0, 0);
            }
        };
        // past a reflecting wall the wall-normal momentum is negated: rho v
        // on lanes past a y wall, rho u on rows past an x wall
        const int sign_v = (nest_gather_flip(pd, G, 2, g) & 4) ? (int)0x80000000 : 0;
        auto read_row = [&](int i, double* u) {
            const double* q = ring + (i & (kStepRing - 1)) * kRowSlot + lane;
            const int sign_u = G.bcx == 2 && (nest_gather_flip(pd, G, i + 2, g) & 2) ? (int)0x80000000 : 0;
            u[0] = q[0];
            u[1] = xor_sign(q[64], sign_u);
            u[2] = xor_sign(q[128], sign_v);
            u[3] = q[192];
        };
        // U^n of output row j (interior cells: nothing negated)
        auto read_u0 = [&](int j, double* u) {
            const double* q = ring + (j & (kStepRing - 1)) * kRowSlot + lane;
            u[0] = q[0];
            u[1] = q[64];
            u[2] = q[128];
            u[3] = q[192];
        };

        // ---- stage 1: X window over U^n, carried down the march: W1 =
        // W(k+1), dW1 = W(k+1) - W(k), FL = left state of face k+1/2, Fm =
        // F(k-1/2). Split stage of row k from centre Uc = U^n(k) and the new
        // row Ub = U^n(k+2) -> V = U1(k) (euler2d.hpp:67-239).
        prim4 W1, dW1, FL;
        double Fm[4];
        auto stage1 = [&](const double* Uc, const double* Ub, double* V) {
            double ux[4];
            {
                prim4 W2 = prim2d(Ub[0], Ub[1], Ub[2], Ub[3], true, a.eos);
                prim4 dW2 = sub4(W2, W1);
                prim4 s1 = slope4<LIM>(dW1, dW2);
                double Fp[4];
                face_flux<true, true>(FL, face_minus(W1, s1), a.eos, Fp);
                FL = face_plus(W1, s1);
                ux[0] = Uc[0] - dtdx * (Fp[0] - Fm[0]);
                ux[1] = Uc[1] - dtdx * (Fp[1] - Fm[1]);
                ux[2] = Uc[2] - dtdx * (Fp[3] - Fm[3]);
                ux[3] = Uc[3] - dtdx * (Fp[2] - Fm[2]);
                W1 = W2;
                dW1 = dW2;
#pragma unroll
                for (int k = 0; k < 4; ++k) Fm[k] = Fp[k];
            }
            prim4 Y = prim2d(ux[0], ux[1], ux[2], ux[3], false, a.eos);
            prim4 dl = sub4(Y, shfl_up4_z(Y));
            prim4 dr = shfl_down4_z(dl);
            prim4 sy = slope4<LIM>(dl, dr);
            prim4 Rf = shfl_down4_z(face_minus(Y, sy));
            double Gf[4];
            face_flux<true, true>(face_plus(Y, sy), Rf, a.eos, Gf);
            double Gm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Gm[k] = lane_up_z(Gf[k]);
            V[0] = ux[0] - dtdy * (Gf[0] - Gm[0]);
            V[1] = ux[1] - dtdy * (Gf[3] - Gm[3]);
            V[2] = ux[2] - dtdy * (Gf[1] - Gm[1]);
            V[3] = ux[3] - dtdy * (Gf[2] - Gm[2]);
        };
        // ---- stage 2: X window over U1 (same carry). Split stage of row j
        // from centre Qc = U1(j) and the new row Qn = U1(j+2), RK2 combine
        // with z = U^n(j), wave speeds, store.
        prim4 W1b, dW1b, FLb;
        double Fmb[4];
        unsigned long long smx = 0, smy = 0;
        auto stage2 = [&](int j, const double* Qc, const double* Qn) {
            double ux[4];
            {
                prim4 W2 = prim2d(Qn[0], Qn[1], Qn[2], Qn[3], true, a.eos);
                prim4 dW2 = sub4(W2, W1b);
                prim4 s1 = slope4<LIM>(dW1b, dW2);
                double Fp[4];
                face_flux<true, true>(FLb, face_minus(W1b, s1), a.eos, Fp);
                FLb = face_plus(W1b, s1);
                ux[0] = Qc[0] - dtdx * (Fp[0] - Fmb[0]);
                ux[1] = Qc[1] - dtdx * (Fp[1] - Fmb[1]);
                ux[2] = Qc[2] - dtdx * (Fp[3] - Fmb[3]);
                ux[3] = Qc[3] - dtdx * (Fp[2] - Fmb[2]);
                W1b = W2;
                dW1b = dW2;
#pragma unroll
                for (int k = 0; k < 4; ++k) Fmb[k] = Fp[k];
            }
            prim4 Y = prim2d(ux[0], ux[1], ux[2], ux[3], false, a.eos);
            prim4 dl = sub4(Y, shfl_up4_z(Y));
            prim4 dr = shfl_down4_z(dl);
            prim4 sy = slope4<LIM>(dl, dr);
            prim4 Rf = shfl_down4_z(face_minus(Y, sy));
            double Gf[4];
            face_flux<true, true>(face_plus(Y, sy), Rf, a.eos, Gf);
            double Gm[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) Gm[k] = lane_up_z(Gf[k]);
            double o0 = ux[0] - dtdy * (Gf[0] - Gm[0]);
            double o1 = ux[1] - dtdy * (Gf[3] - Gm[3]);
            double o2 = ux[2] - dtdy * (Gf[1] - Gm[1]);
            double o3 = ux[3] - dtdy * (Gf[2] - Gm[2]);
            // SSP-RK2 combine, U^{n+1} = 0.5*(U^n + U2), U^n from the ring
            double z[4];
            read_u0(j, z);
            o0 = 0.5 * (z[0] + o0);
            o1 = 0.5 * (z[1] + o1);
            o2 = 0.5 * (z[2] + o2);
            o3 = 0.5 * (z[3] + o3);
            double sx, sy2;
            cell_speeds2d(o0, o1, o2, o3, a.eos, sx, sy2);
            const unsigned long long bx = speed_slot_bits(sx, o0, o3);
            const unsigned long long by = (unsigned long long)__double_as_longlong(sy2);
            smx = (store && bx > smx) ? bx : smx;
            smy = (store && by > smy) ? by : smy;
            if (store) {
                const unsigned o = coff + (unsigned)(j + 2) * (unsigned)snye;
                *reinterpret_cast<double*>(Ob8 + o) = o0;
                *reinterpret_cast<double*>(Ob8 + (o + (unsigned)svs)) = o1;
                *reinterpret_cast<double*>(Ob8 + (o + 2u * (unsigned)svs)) = o2;
                *reinterpret_cast<double*>(Ob8 + (o + 3u * (unsigned)svs)) = o3;
            }
        };

        // U1 queue: at the start of iteration k, Q0 = U1(k-3), Q1 = U1(k-2),
        // Q2 = U1(k-1)
        double Q0[4], Q1[4], Q2[4];
        // V = U1(k) on lanes past an outflow / reflecting y edge: the value
        // of the lane of the mapped column (lanes inside the domain keep
        // their own; lanes far past the edge receive values never read)
        const bool yfix = G.bcy != 0 && (gs - 4 < 0 || gs + 59 >= G.NY);
        const int ysrc = nest_step_lane_source(lane, g, G.NY, G.bcy);
        const int yneg = G.bcy == 2 && (g < 0 || g >= G.NY) ? (int)0x80000000 : 0;
        auto fix_lanes = [&](double* V) {
            if (!yfix) return;
#pragma unroll
            for (int v = 0; v < 4; ++v) V[v] = __shfl(V[v], ysrc);
            V[2] = xor_sign(V[2], yneg);
        };
        // V = U1(k) on a row past the bottom x edge (outflow / reflecting):
        // the queue row of the mapped row (k-1, k-2 or k-3). Rows further out
        // are never read (stage 2 reads U1 down to global row NX + 1).
        const int xneg = G.bcx == 2 ? (int)0x80000000 : 0;
        auto fix_bottom = [&](int k, double* V) {
            const int gk = pd.x0 + k;
            if (G.bcx == 0 || gk < G.NX) return;
            const int d = gk - nest_map_index(gk, G.NX, G.bcx);
            if (d < 1 || d > 3) return;
#pragma unroll
            for (int v = 0; v < 4; ++v) V[v] = d == 1 ? Q2[v] : (d == 2 ? Q1[v] : Q0[v]);
            V[1] = xor_sign(V[1], xneg);
        };
        // U1(k) of the owned cells of row k, before the edge fixes (which
        // change only lanes and rows outside the domain)
        auto put_u1 = [&](int k, const double* V) {
            if constexpr (U1OUT) {
                if (k < r0 || k >= r1 || !store) return;
                char* const Q8 = reinterpret_cast<char*>(a.U1);
                const unsigned o = coff + (unsigned)(k + 2) * (unsigned)snye;
                *reinterpret_cast<double*>(Q8 + o) = V[0];
                *reinterpret_cast<double*>(Q8 + (o + (unsigned)svs)) = V[1];
                *reinterpret_cast<double*>(Q8 + (o + 2u * (unsigned)svs)) = V[2];
                *reinterpret_cast<double*>(Q8 + (o + 3u * (unsigned)svs)) = V[3];
            } else {
                (void)k;
                (void)V;
            }
        };
        auto rotate = [&](const double* V) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                Q0[v] = Q1[v];
                Q1[v] = Q2[v];
                Q2[v] = V[v];
            }
        };
        // Row fetches: U^n rows r0-4 .. r1+3 are needed (stage 1 of rows
        // r0-2 .. r1+1). Iteration k fetches row k+4 (while that is needed)
        // and waits for row k+2, fetched two iterations earlier; issued after
        // it for sure: row k+3, when k+3 <= r1+3.
        auto wait_row = [&](int k) {
            if (k <= r1) wait_vmcnt<8>();
            else wait_vmcnt<0>();
        };
        auto fetch_ahead = [&](int k) {
            if (k + 4 <= r1 + 3) issue_row(k + 4);
        };

        // ---- prologue: rows r0-4 .. r0-1 warm the stage-1 X window up;
        // rows r0 and r0+1 are the new rows of the first two iterations
        issue_row(r0 - 4);
        issue_row(r0 - 3);
        issue_row(r0 - 2);
        issue_row(r0 - 1);
        wait_vmcnt<0>();
        {
            double A[4];
            read_row(r0 - 4, A);
            prim4 Wm2 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(r0 - 3, A);
            prim4 Wm1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(r0 - 2, A);
            prim4 W0 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            read_row(r0 - 1, A);
            W1 = prim2d(A[0], A[1], A[2], A[3], true, a.eos);
            wait_lgkmcnt0();
            issue_row(r0);
            compiler_fence();
            issue_row(r0 + 1);
            prim4 d0 = sub4(W0, Wm1);
            dW1 = sub4(W1, W0);
            prim4 s_m1 = slope4<LIM>(sub4(Wm1, Wm2), d0);
            prim4 s0 = slope4<LIM>(d0, dW1);
            face_flux<true, true>(face_plus(Wm1, s_m1), face_minus(W0, s0), a.eos, Fm);
            FL = face_plus(W0, s0);
        }
        // ---- stage 1 of rows r0-2 .. r0+2; at r0+1 the stage-2 X window
        // warms up on U1 rows r0-2 .. r0+1
        for (int k = r0 - 2; k <= r0 + 2; ++k) {
            wait_row(k);
            double Uc[4], Ub[4], V[4];
            read_row(k, Uc);
            read_row(k + 2, Ub);
            wait_lgkmcnt0();
            fetch_ahead(k);
            stage1(Uc, Ub, V);
            put_u1(k, V);
            fix_lanes(V);
            fix_bottom(k, V);
            if (k == r0 + 1) {
                // U1 rows past the top x edge (outflow / reflecting), when
                // the chunk starts at global row 0: row -1 (Q1) <- row 0
                // (Q2); row -2 (Q0) <- row 0 (outflow) or row 1 (V,
                // reflecting)
                if (G.bcx != 0 && pd.x0 + r0 == 0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        Q0[v] = G.bcx == 2 ? V[v] : Q2[v];
                        Q1[v] = Q2[v];
                    }
                    Q0[1] = xor_sign(Q0[1], xneg);
                    Q1[1] = xor_sign(Q1[1], xneg);
                }
                prim4 Wm2 = prim2d(Q0[0], Q0[1], Q0[2], Q0[3], true, a.eos);
                prim4 Wm1 = prim2d(Q1[0], Q1[1], Q1[2], Q1[3], true, a.eos);
                prim4 W0 = prim2d(Q2[0], Q2[1], Q2[2], Q2[3], true, a.eos);
                W1b = prim2d(V[0], V[1], V[2], V[3], true, a.eos);
                prim4 d0 = sub4(W0, Wm1);
                dW1b = sub4(W1b, W0);
                prim4 s_m1 = slope4<LIM>(sub4(Wm1, Wm2), d0);
                prim4 s0 = slope4<LIM>(d0, dW1b);
                face_flux<true, true>(face_plus(Wm1, s_m1), face_minus(W0, s0), a.eos, Fmb);
                FLb = face_plus(W0, s0);
            }
            rotate(V);
        }
        // ---- main march: stage 1 of row k with stage 2 of row k-3
        for (int k = r0 + 3; k <= r1 + 1; ++k) {
            wait_row(k);
            double Uc[4], Ub[4], V[4];
            read_row(k, Uc);
            read_row(k + 2, Ub);
            wait_lgkmcnt0();
            fetch_ahead(k);
            stage1(Uc, Ub, V);
            put_u1(k, V);
#if NEST_STEP_SCHED_BARRIER
            __builtin_amdgcn_sched_barrier(0);
#endif
            stage2(k - 3, Q0, Q2);
            fix_lanes(V);
            fix_bottom(k, V);
            rotate(V);
        }
        // ---- last output row r1-1 (its new U1 row r1+1 is Q2)
        stage2(r1 - 1, Q0, Q2);
        // every fetch has landed before the ring is reused (or the wave ends)
        wait_vmcnt<0>();
        smx = wave_max_u64(smx);
        smy = wave_max_u64(smy);
        if (lane == 0) {
            unsigned long long* sl =
                a.slots + ((size_t)pd.prob * NEST_SPEED_SLOTS + (it & (NEST_SPEED_SLOTS - 1))) * NEST_SLOT_STRIDE;
            atomicMax(sl, smx);
            atomicMax(sl + 1, smy);
        }
    }
}

// Wave speeds of the current state (first step of a run): same work items.
__global__ __launch_bounds__(256) void speeds2d_kernel(stage_args a) {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + (threadIdx.x >> 6); it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        nest_prob_state* ps = a.probs + pd.prob;
        const int nye = pd.ny + 4;
        const int vs = (pd.nx + 4) * nye;
        const int ec = w.c0 + lane;
        const bool own = lane >= 2 && lane < 62 && ec >= 2 && ec < pd.ny + 2;
        unsigned long long smx = 0, smy = 0;
        if (own) {
            const double* p0 = a.U + pd.off + ec;
            for (int i = w.r0; i < w.r1; ++i) {
                const double* p = p0 + (size_t)(i + 2) * nye;
                double sx, sy;
                const double r = p[0], E = p[3 * (size_t)vs];
                cell_speeds2d(r, p[vs], p[2 * (size_t)vs], E, a.eos, sx, sy);
                unsigned long long bx = speed_slot_bits(sx, r, E);
                unsigned long long by = (unsigned long long)__double_as_longlong(sy);
                smx = bx > smx ? bx : smx;
                smy = by > smy ? by : smy;
            }
        }
        smx = wave_max_u64(smx);
        smy = wave_max_u64(smy);
        if (lane == 0) {
            unsigned long long* sl =
                a.slots_rd ? a.slots_rd + ((size_t)pd.prob * NEST_SPEED_SLOTS + (it & (NEST_SPEED_SLOTS - 1))) *
                                              NEST_SLOT_STRIDE
                           : &ps->smax[0];
            atomicMax(sl, smx);
            atomicMax(sl + 1, smy);
        }
    }
}

// Device exchange-plan executor: every region of every patch in one launch.
__global__ __launch_bounds__(256) void fill_regions_kernel(double* base, const nest_region* R, int nreg,
                                                           int nvars) {
    for (int r = blockIdx.x; r < nreg; r += gridDim.x) {
        const nest_region g = R[r];
        const int n = g.ni * g.nj;
        for (int t = threadIdx.x; t < n * nvars; t += blockDim.x) {
            const int k = t / n;
            const int c = t - k * n;
            const int ia = c / g.nj;
            const int jb = c - ia * g.nj;
            const int si = g.si + (g.bi ? 0 : (g.mi ? g.ni - 1 - ia : ia));
            const int sj = g.sj + (g.bj ? 0 : (g.mj ? g.nj - 1 - jb : jb));
            const double v = base[g.src_off + (long long)k * g.src_vs + (long long)si * g.src_nye + sj];
            base[g.dst_off + (long long)k * g.dst_vs + (long long)(g.di + ia) * g.dst_nye + g.dj + jb] =
                (g.neg >> k) & 1 ? -v : v;
        }
    }
}

// dt for the next step of every problem from the reduced wave speeds:
// 2D compute_dt_2d (euler2d.hpp:249-270): cfl*min(min dx/sx, min dy/sy),
//    and min over cells of dx/s equals dx/max s (correctly rounded division
//    is monotone), init 1e10, cells with s == 0 skipped;
// 1D compute_dt (sod1d.cpp:127): (cfl*dx)/max(|u|+c);
// then dt = min(dt, t_end - t) (sod1d.cpp:222, sod2d.cpp:320).
// Schedule the next step (parity par) of one problem from the max wave
// speeds of its newest state: set dt, advance t (and tn[par ^ 1]) and the
// step count, or flag the problem done (t_end reached: 1; non-finite dt: 2).
__device__ inline void schedule_step(nest_prob_state& s, unsigned long long bx, unsigned long long by, int par) {
    if (s.done) return;
    double dt = 0.0;
    const int d = step_dt(s, s.t, bx, by, dt);
    if (d) {
        s.done = d;
        if (d == 2) s.dt = dt;
        return;
    }
    s.dt = dt;
    s.t = s.t + dt;
    s.tn[par ^ 1] = s.t;
    s.steps += 1;
}

// Step scheduling of the two-launch step: folds and clears the slot set the
// previous step wrote (slots), clears the set the next step writes
// (slots_clr, may be null) and schedules step parity par.
__global__ __launch_bounds__(64) void finalize_kernel(nest_prob_state* P, unsigned long long* slots,
                                                      unsigned long long* slots_clr, int nprob, int par) {
    const int i = blockIdx.x;  // one wave per problem
    const int lane = threadIdx.x;
    if (i >= nprob) return;
    unsigned long long vx = 0, vy = 0;
    if (slots) {  // lane k: slot k
        unsigned long long* sl = slots + ((size_t)i * NEST_SPEED_SLOTS + lane) * NEST_SLOT_STRIDE;
        vx = sl[0];
        vy = sl[1];
        sl[0] = 0ull;
        sl[1] = 0ull;
    }
    if (slots_clr) {
        unsigned long long* sc = slots_clr + ((size_t)i * NEST_SPEED_SLOTS + lane) * NEST_SLOT_STRIDE;
        sc[0] = 0ull;
        sc[1] = 0ull;
    }
    vx = wave_max_u64(vx);
    vy = wave_max_u64(vy);
    if (lane != 0) return;
    schedule_step(P[i], vx, vy, par);
}

__global__ __launch_bounds__(64) void reset_speeds_kernel(nest_prob_state* P, unsigned long long* slots, int nprob) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= nprob) return;
    if (slots) {
        unsigned long long* sl = slots + ((size_t)i * NEST_SPEED_SLOTS + lane) * NEST_SLOT_STRIDE;
        sl[0] = 0ull;
        sl[1] = 0ull;
    }
    if (lane == 0) {
        P[i].smax[0] = 0ull;
        P[i].smax[1] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// 1D (single patch, n interior cells, halo 2, 3 variables)
// Block of 256 threads owns 256 consecutive interior cells; primitives,
// slopes and the 257 face fluxes of the block are staged in LDS so each is
// computed once (benchmark_euler1d.cpp:113-165 numerics). Ghost cells are
// read through the boundary condition (periodic wrap, outflow clamp,
// reflecting mirror with the momentum negated): the values a ghost fill
// would leave there, so no fill launch runs between stages.
// ---------------------------------------------------------------------------
template <int LIM, int STAGE>
__global__ __launch_bounds__(256) void stage1d_kernel(stage1d_args a) {
    __shared__ prim4 W[260];
    __shared__ prim4 S[258];
    __shared__ double F[3][257];
    if (!(a.host_dt > 0.0) && a.ps->done) return;
    const int n = a.n, ne = n + 4;
    const int base = 2 + blockIdx.x * 256;  // first extended cell of the block
    const double* in = STAGE == 1 ? a.U : a.U1;
    double* out = STAGE == 1 ? a.U1 : a.U;
    for (int t = threadIdx.x; t < 260; t += 256) {
        const int g = base - 4 + t;  // interior index (ghosts and the block tail outside [0, n))
        const int src = nest_map_index(g, n, a.bc) + 2;
        const double m = in[ne + src];
        W[t] = prim1d(in[src], a.bc == 2 && (g < 0 || g >= n) ? -m : m, in[2 * ne + src], a.eos);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 258; t += 256) {  // slopes of cells base-1+t
        S[t] = slope4<LIM>(sub4(W[t + 1], W[t]), sub4(W[t + 2], W[t + 1]));
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 257; t += 256) {  // face between cells base-1+t and base+t
        double f[4];
        face_flux<false, true>(face_plus(W[t + 1], S[t]), face_minus(W[t + 2], S[t + 1]), a.eos, f);
        F[0][t] = f[0];
        F[1][t] = f[1];
        F[2][t] = f[2];
    }
    __syncthreads();
    const int e = base + threadIdx.x;
    unsigned long long sm = 0;
    if (e < n + 2) {
        const double dt = a.host_dt > 0.0 ? a.host_dt : a.ps->dt;
        const double dtdx = dt / a.ps->dx;
        const int t = threadIdx.x;
        double o[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = in[k * ne + e] - dtdx * (F[k][t + 1] - F[k][t]);
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int k = 0; k < 3; ++k) o[k] = 0.5 * (out[k * ne + e] + o[k]);
            prim4 w = prim1d(o[0], o[1], o[2], a.eos);
            double s = __builtin_fabs(w.u) + sound_speed_1d(w.r, w.p, a.eos);
            sm = speed_slot_bits(s, o[0], o[2]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k * ne + e] = o[k];
    }
    if constexpr (STAGE == 2) {
        if (a.host_dt > 0.0 || !a.slots) return;  // the caller chose dt: no speeds to fold
        // Device dt: the block's maximum goes to one of NEST_SPEED_SLOTS
        // slots (spread so that blocks do not queue on one address);
        // finalize_kernel folds and clears them and schedules the next step.
        __shared__ unsigned long long wmax[4];
        sm = wave_max_u64(sm);
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = sm;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long b = wmax[0];
            for (int w = 1; w < 4; ++w) b = b > wmax[w] ? b : wmax[w];
            atomicMax(a.slots + (size_t)(blockIdx.x & (NEST_SPEED_SLOTS - 1)) * NEST_SLOT_STRIDE, b);
        }
    }
}

// One whole RK2 step of the 1D solver per launch, out of place: U^{n+1}
// (into a.U1) from U^n (a.U). A block owns 256 output cells g0 .. g0+255 and
// stages U^n over g0-4 .. g0+259 (ghosts read through the boundary
// condition) in LDS; it computes stage 1 at g0-2 .. g0+257, so U1 never
// leaves the block, then stage 2 at its own cells. The numerics and their
// order are those of the two stage kernels, cell for cell: stage 1 at a
// position past a periodic end sees the stencil of the cell it wraps to, so
// it equals that cell's U1 bit for bit; past an outflow or reflecting wall
// U1 is set from the block's own boundary cells, as the ghost fill of U1
// would set it. Every block takes the step's dt from the speeds the
// previous step left in the slots (the arithmetic of schedule_step), block 0
// records it, and the block's maximum wave speed of U^{n+1} goes to the next
// slot set: one launch per step.
// Speed slots the fused 1D step uses of each set (a power of two, at most
// NEST_SPEED_SLOTS): every block reads all of them for its dt, so fewer lines
// to read, while the blocks' atomics stay spread over this many addresses.
constexpr int kStepSlots = 16;
static_assert((kStepSlots & (kStepSlots - 1)) == 0 && kStepSlots <= NEST_SPEED_SLOTS, "slot count");

template <int LIM>
__global__ __launch_bounds__(256) void step1d_kernel(stage1d_args a) {
    constexpr int T = 264;  // tile position t holds interior cell g0 - 4 + t
    // one array per variable (8-byte lane stride: no LDS bank conflicts);
    // the passive component of prim4 is zero in 1D
    __shared__ double C[3][T];  // conserved U^n, then U1 at t in [2, 262)
    __shared__ double Wv[3][T];  // primitives rho, u, p
    __shared__ double Sv[3][T];  // their limited slopes
    __shared__ double F[3][T];  // F[k][t]: face between tile positions t and t + 1
    auto W = [&](int t) { return prim4{Wv[0][t], Wv[1][t], Wv[2][t], 0.0}; };
    auto S = [&](int t) { return prim4{Sv[0][t], Sv[1][t], Sv[2][t], 0.0}; };
    auto put = [](double (*A)[T], int t, const prim4& w) {
        A[0][t] = w.r;
        A[1][t] = w.u;
        A[2][t] = w.p;
    };
    __shared__ unsigned long long wmax[4];
    __shared__ double s_dt;
    __shared__ int s_go;
    const int n = a.n, ne = n + 4;
    const int g0 = blockIdx.x * 256;
    const double* in = a.U;
    double* out = a.U1;
    // Issued ahead of the tile loads, used after them: the run's flag, this
    // step's time and the previous step's speeds. A finished or failed run
    // makes every later step a no-op. (Block 0 of this launch may set the
    // flag while others read it; it does so only when every block finds the
    // same from the same inputs.)
    const int done_in = a.run->done;
    const double t_in = a.run->t[a.par];
    const unsigned long long v_in =
        threadIdx.x < kStepSlots ? a.slots_rd[(size_t)threadIdx.x * NEST_SLOT_STRIDE] : 0ull;
    for (int t = threadIdx.x; t < T; t += 256) {
        const int g = g0 - 4 + t;
        const int src = nest_map_index(g, n, a.bc) + 2;
        const double m = in[ne + src];
        C[0][t] = in[src];
        C[1][t] = a.bc == 2 && (g < 0 || g >= n) ? -m : m;
        C[2][t] = in[2 * ne + src];
        put(Wv, t, prim1d(C[0][t], C[1][t], C[2][t], a.eos));
    }
    // the step's dt (its read overlaps the tile loads above)
    if (threadIdx.x < 64) {
        const unsigned long long v = wave_max_u64(v_in);
        if (threadIdx.x == 0) {
            const nest_prob_state& P = *a.ps;
            const double t = t_in;
            const double sx = __longlong_as_double((long long)v);
            int done = 0;
            double dtn = 0.0;
            if (done_in) {
                done = -1;  // already finished: leave the run's state alone
            } else if (!(t < P.t_end)) {
                done = 1;
            } else {
                dtn = P.fixed_dt > 0.0 ? P.fixed_dt : P.cfl * P.dx / sx;
                dtn = dmin(dtn, P.t_end - t);
                if (!dt_ok(dtn) || bits_nan(sx)) done = 2;
            }
            s_go = !done;
            s_dt = dtn;
            if (blockIdx.x == 0 && done >= 0) {
                if (done) {
                    a.run->done = done;
                    if (done == 2) a.run->dt = dtn;
                } else {
                    a.run->dt = dtn;
                    a.run->t[a.par ^ 1] = t + dtn;
                    a.run->steps += 1;
                }
            }
        }
        if (blockIdx.x == 0 && threadIdx.x < kStepSlots)
            a.slots_clr[(size_t)threadIdx.x * NEST_SLOT_STRIDE] = 0ull;  // for the next step
    }
    __syncthreads();
    if (!s_go) {
        // a step past the end (or after a failure) passes the state through,
        // so after step k the state is in the buffer of parity k + 1
        const int g = g0 + threadIdx.x;
        if (g < n) {
#pragma unroll
            for (int k = 0; k < 3; ++k) out[k * ne + g + 2] = in[k * ne + g + 2];
        }
        return;
    }
    const double dt = s_dt;
    const double dtdx = dt / a.ps->dx;
    // ---- stage 1 at t in [2, 262)
    for (int t = threadIdx.x + 1; t < T - 1; t += 256) put(Sv, t, slope4<LIM>(sub4(W(t), W(t - 1)), sub4(W(t + 1), W(t))));
    __syncthreads();
    for (int t = threadIdx.x + 1; t < T - 2; t += 256) {
        double f[4];
        face_flux<false, true>(face_plus(W(t), S(t)), face_minus(W(t + 1), S(t + 1)), a.eos, f);
        F[0][t] = f[0];
        F[1][t] = f[1];
        F[2][t] = f[2];
    }
    __syncthreads();
    double u1[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int t = threadIdx.x + 2 + 256 * r;
        if (t < T - 2) {
#pragma unroll
            for (int k = 0; k < 3; ++k) u1[r][k] = C[k][t] - dtdx * (F[k][t] - F[k][t - 1]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int t = threadIdx.x + 2 + 256 * r;
        if (t < T - 2) {
#pragma unroll
            for (int k = 0; k < 3; ++k) C[k][t] = u1[r][k];
        }
    }
    __syncthreads();
    if (a.bc != 0) {  // U1 past an outflow / reflecting wall, from the block's own cells
        // only the ghost positions stage 2 reads; each maps to cell 0 / 1 or
        // n - 1 / n - 2, inside the tile (nest_step1d_wall_source)
        for (int t = threadIdx.x + 2; t < T - 2; t += 256) {
            const int ts = nest_step1d_wall_source(t, g0, n, a.bc);
            if (ts < 0) continue;
            const bool neg = a.bc == 2;
            C[0][t] = C[0][ts];
            C[1][t] = neg ? -C[1][ts] : C[1][ts];
            C[2][t] = C[2][ts];
        }
        __syncthreads();
    }
    // ---- stage 2 at t in [4, 260)
    for (int t = threadIdx.x + 2; t < T - 2; t += 256) put(Wv, t, prim1d(C[0][t], C[1][t], C[2][t], a.eos));
    __syncthreads();
    for (int t = threadIdx.x + 3; t < T - 3; t += 256) put(Sv, t, slope4<LIM>(sub4(W(t), W(t - 1)), sub4(W(t + 1), W(t))));
    __syncthreads();
    for (int t = threadIdx.x + 3; t < T - 4; t += 256) {
        double f[4];
        face_flux<false, true>(face_plus(W(t), S(t)), face_minus(W(t + 1), S(t + 1)), a.eos, f);
        F[0][t] = f[0];
        F[1][t] = f[1];
        F[2][t] = f[2];
    }
    __syncthreads();
    const int g = g0 + threadIdx.x;
    unsigned long long sm = 0;
    if (g < n) {
        const int t = threadIdx.x + 4, e = g + 2;
        double o[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = C[k][t] - dtdx * (F[k][t] - F[k][t - 1]);
#pragma unroll
        for (int k = 0; k < 3; ++k) o[k] = 0.5 * (in[k * ne + e] + o[k]);
        prim4 w = prim1d(o[0], o[1], o[2], a.eos);
        double sp = __builtin_fabs(w.u) + sound_speed_1d(w.r, w.p, a.eos);
        sm = speed_slot_bits(sp, o[0], o[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k * ne + e] = o[k];
    }
    sm = wave_max_u64(sm);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = sm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wmax[0];
        for (int w = 1; w < 4; ++w) b = b > wmax[w] ? b : wmax[w];
        atomicMax(a.slots + (size_t)(blockIdx.x & (kStepSlots - 1)) * NEST_SLOT_STRIDE, b);
    }
}

// The device-dt step's dt (finalize_kernel's arithmetic) from the speed
// slots the previous step wrote, the same in every wave. Wave 0 records the
// step (dt, time, step count, or the done flag) and clears the slot set of
// the next step. Returns 0 (take the step), 1 / 2 (the run ends here: end
// time reached / no usable dt) or -1 (the run had already ended).
__device__ __forceinline__ int step1d_schedule(const stage1d_args& a, int wid, int lane, double& dt) {
    const int done_in = a.run->done;
    const double t_in = a.run->t[a.par];
    const unsigned long long v_in = lane < kStepSlots ? a.slots_rd[(size_t)lane * NEST_SLOT_STRIDE] : 0ull;
    const unsigned long long v = wave_max_u64(v_in);
    const nest_prob_state& P = *a.ps;
    const double sx = __longlong_as_double((long long)v);
    int done = 0;
    dt = 0.0;
    if (done_in) {
        done = -1;  // already finished: leave the run's state alone
    } else if (!(t_in < P.t_end)) {
        done = 1;
    } else {
        dt = P.fixed_dt > 0.0 ? P.fixed_dt : P.cfl * P.dx / sx;
        dt = dmin(dt, P.t_end - t_in);
        if (!dt_ok(dt) || bits_nan(sx)) done = 2;
    }
    if (wid == 0) {
        if (lane == 0 && done >= 0) {
            if (done) {
                a.run->done = done;
                if (done == 2) a.run->dt = dt;
            } else {
                a.run->dt = dt;
                a.run->t[a.par ^ 1] = t_in + dt;
                a.run->steps += 1;
            }
        }
        if (lane < kStepSlots) a.slots_clr[(size_t)lane * NEST_SLOT_STRIDE] = 0ull;  // for the next step
    }
    return done;
}

// the wave's speed maximum meets the block's in LDS and reaches the slots
// with one atomic per block
__device__ __forceinline__ void step1d_fold_speeds(const stage1d_args& a, unsigned long long sm,
                                                   unsigned long long* wmax) {
    sm = wave_max_u64(sm);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = sm;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long b = wmax[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = b > wmax[w] ? b : wmax[w];
        atomicMax(a.slots + (size_t)(blockIdx.x & (kStepSlots - 1)) * NEST_SLOT_STRIDE, b);
    }
}

// The same whole 1D RK2 step with no LDS tile and no barrier inside the
// step: every wave works on its own. Lane l of a segment holds cell
// g0 - 4 + l; neighbour values move by DPP wave shifts (the Y-sweep
// machinery of the 2D kernels on 3 variables): stage 1 is valid on lanes
// 2..61, stage 2 on lanes 4..59, so a segment produces 56 cells. Past an
// outflow / reflecting wall, U1 comes from the lane of the mapped cell
// (nest_step_lane_source), momentum negated at a reflecting wall; past a
// periodic end stage 1 sees the stencil of the cell it wraps to. The launch
// holds no more waves than the GPU keeps resident; wave w takes segments w,
// w + W, w + 2W, ... (W = a.waves), so every wave runs to the end of the
// launch with the same number of segments, give or take one, and reads dt
// once. Arithmetic and order per cell are those of step1d_kernel.
template <int LIM>
__global__ __launch_bounds__(256) void step1d_lanes_kernel(stage1d_args a) {
    __shared__ unsigned long long wmax[4];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wib);
    const int n = a.n, ne = n + 4;
    const double* in = a.U;
    double* out = a.U1;
    // U^n of segment sg's lanes (cell g0 - 4 + lane through the boundary map;
    // momentum not yet negated past a reflecting wall)
    const int nseg = (n + 55) / 56, W = a.waves;
    double p0 = 0.0, pm = 0.0, p2 = 0.0;
    auto fetch = [&](int sg) {
        if (sg < nseg) {
            const int src = nest_map_index(sg * 56 - 4 + lane, n, a.bc) + 2;
            p0 = in[src];
            pm = in[ne + src];
            p2 = in[2 * ne + src];
        }
    };
    // the first segment's loads go out before the dt arithmetic, which they
    // do not depend on
    fetch(wid);
    double dt;
    const bool go = step1d_schedule(a, wid, lane, dt) == 0;
    const double dtdx = dt / a.ps->dx;
    unsigned long long sm = 0;
    for (int sg = wid; sg < nseg; sg += W) {
        const int g0 = sg * 56;
        const int g = g0 - 4 + lane;
        const bool store = lane >= 4 && lane < 60 && g < n;
        const double m = pm;
        double c0 = p0, c2 = p2;
        fetch(sg + W);  // the next segment's loads, in flight during this one
        if (!go) {
            // a step past the end passes the state through (stored lanes
            // hold interior cells: their loads are the cells themselves)
            if (store) {
                const int e = g + 2;
                out[e] = c0;
                out[ne + e] = m;
                out[2 * ne + e] = c2;
            }
            continue;
        }
        const bool outside = g < 0 || g >= n;
        double c1 = a.bc == 2 && outside ? -m : m;
        const double z0 = c0, z1 = c1, z2 = c2;  // U^n of this lane's cell (the combine)
        // one stage on the lanes: c <- c - dtdx (F(l+1/2) - F(l-1/2))
        auto stage = [&](double& u0, double& u1, double& u2) {
            prim4 W = prim1d(u0, u1, u2, a.eos);
            const double d0 = W.r - lane_up_z(W.r), d1 = W.u - lane_up_z(W.u), d2 = W.p - lane_up_z(W.p);
            const prim4 dl{d0, d1, d2, 0.0};
            const prim4 dr{lane_down_z(d0), lane_down_z(d1), lane_down_z(d2), 0.0};
            const prim4 S = slope4<LIM>(dl, dr);
            const prim4 Rm = face_minus(W, S);
            const prim4 Rf{lane_down_z(Rm.r), lane_down_z(Rm.u), lane_down_z(Rm.p), 0.0};
            double f[4];
            face_flux<false, true>(face_plus(W, S), Rf, a.eos, f);
            u0 = u0 - dtdx * (f[0] - lane_up_z(f[0]));
            u1 = u1 - dtdx * (f[1] - lane_up_z(f[1]));
            u2 = u2 - dtdx * (f[2] - lane_up_z(f[2]));
        };
        stage(c0, c1, c2);
        if (a.bc != 0 && (g0 - 4 < 0 || g0 + 59 >= n)) {
            // U1 past an outflow / reflecting wall from the lane of the mapped cell
            const int sl = nest_step_lane_source(lane, g, n, a.bc);
            c0 = __shfl(c0, sl);
            c1 = __shfl(c1, sl);
            c2 = __shfl(c2, sl);
            if (a.bc == 2 && outside) c1 = -c1;
        }
        stage(c0, c1, c2);
        if (store) {
            const double o0 = 0.5 * (z0 + c0), o1 = 0.5 * (z1 + c1), o2 = 0.5 * (z2 + c2);
            const prim4 w = prim1d(o0, o1, o2, a.eos);
            const double sp = __builtin_fabs(w.u) + sound_speed_1d(w.r, w.p, a.eos);
            const unsigned long long b = speed_slot_bits(sp, o0, o2);
            sm = b > sm ? b : sm;
            const int e = g + 2;
            out[e] = o0;
            out[ne + e] = o1;
            out[2 * ne + e] = o2;
        }
    }
    if (!go) return;
    step1d_fold_speeds(a, sm, wmax);
}

// The lane step with two cells per lane: lane l of a segment holds cells
// g0 - 4 + 2l (A) and g0 - 3 + 2l (B). A cell's left neighbour is the B
// cell of the lane before (A) or the lane's own A cell (B), so each stage
// moves one value per quantity between lanes where the one-cell kernel
// moves two, and the face between A and B is computed on the lane. Stage 1
// is valid on lanes 1..62, stage 2 on lanes 2..61: a segment of 128 cells
// produces 120. Past an outflow / reflecting wall, U1 comes from the lane
// and half of the mapped cell, momentum negated at a reflecting wall; past
// a periodic end stage 1 sees the stencil of the cell it wraps to. Waves
// take segments as in step1d_lanes_kernel. WIDE (state buffers and the
// variable stride 16-byte aligned: n even): a lane moves its two cells with
// one 16-byte access per variable (a wave's access spans whole cache lines,
// where two 8-byte accesses each touch every other 8 bytes of twice as
// many), and segments with a cell past the domain load cell by cell through
// the boundary map. Arithmetic and order per cell are those of
// step1d_kernel.
typedef double nest_d2 __attribute__((ext_vector_type(2)));
template <int LIM, bool WIDE>
__global__ __launch_bounds__(256) void step1d_pairs_kernel(stage1d_args a) {
    __shared__ unsigned long long wmax[4];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wib);
    const int n = a.n;
    // byte offsets from the state buffers (3 (n + 4) doubles: below 4 GiB
    // wherever this kernel is chosen)
    const char* const in8 = reinterpret_cast<const char*>(a.U);
    char* const out8 = reinterpret_cast<char*>(a.U1);
    const unsigned vs = (unsigned)(n + 4) * 8u;
    auto ld = [&](unsigned o) { return *reinterpret_cast<const double*>(in8 + o); };
    auto st = [&](unsigned o, double x) { *reinterpret_cast<double*>(out8 + o) = x; };
    // U^n of segment sg's cells A and B (through the boundary map; momentum
    // not yet negated past a reflecting wall)
    const int nseg = (n + 119) / 120, W = a.waves;
    double pA[3] = {0.0, 0.0, 0.0}, pB[3] = {0.0, 0.0, 0.0};
    auto fetch = [&](int sg) {
        if (sg < nseg) {
            const int g0 = sg * 120;
            const int gA = g0 - 4 + 2 * lane;
            if (WIDE && g0 >= 4 && g0 + 124 <= n) {
                const unsigned oA = (unsigned)(gA + 2) * 8u;
#pragma unroll
                for (int v = 0; v < 3; ++v) {
                    const nest_d2 x = *reinterpret_cast<const nest_d2*>(in8 + (oA + (unsigned)v * vs));
                    pA[v] = x.x;
                    pB[v] = x.y;
                }
                return;
            }
            const unsigned oA = (unsigned)(nest_map_index(gA, n, a.bc) + 2) * 8u;
            const unsigned oB = (unsigned)(nest_map_index(gA + 1, n, a.bc) + 2) * 8u;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                pA[v] = ld(oA + (unsigned)v * vs);
                pB[v] = ld(oB + (unsigned)v * vs);
            }
        }
    };
    fetch(wid);
    double dt;
    const bool go = step1d_schedule(a, wid, lane, dt) == 0;
    const double dtdx = dt / a.ps->dx;
    unsigned long long sm = 0;
    for (int sg = wid; sg < nseg; sg += W) {
        const int g0 = sg * 120;
        const int gA = g0 - 4 + 2 * lane, gB = gA + 1;
        const bool mid = lane >= 2 && lane < 62;
        const bool storeA = mid && gA < n, storeB = mid && gB < n;
        double A[3], B[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            A[v] = pA[v];
            B[v] = pB[v];
        }
        fetch(sg + W);  // the next segment's loads, in flight during this one
        const unsigned eA = (unsigned)(gA + 2) * 8u;  // cell A's byte offset (stored lanes)
        // both cells of a lane (n even: storeA == storeB)
        auto st2 = [&](unsigned o, double x, double y) {
            if constexpr (WIDE) {
                nest_d2 q;
                q.x = x;
                q.y = y;
                *reinterpret_cast<nest_d2*>(out8 + o) = q;
            } else {
                if (storeA) st(o, x);
                if (storeB) st(o + 8u, y);
            }
        };
        if (!go) {
            // a step past the end passes the state through
            if (storeA || storeB) {
#pragma unroll
                for (int v = 0; v < 3; ++v) st2(eA + (unsigned)v * vs, A[v], B[v]);
            }
            continue;
        }
        const bool outA = gA < 0 || gA >= n, outB = gB < 0 || gB >= n;
        if (a.bc == 2) {
            if (outA) A[1] = -A[1];
            if (outB) B[1] = -B[1];
        }
        const double zA[3] = {A[0], A[1], A[2]}, zB[3] = {B[0], B[1], B[2]};  // U^n (the combine)
        // one stage: c <- c - dtdx (F(c+1/2) - F(c-1/2)) on both cells
        auto stage = [&](double* a_, double* b_) {
            const prim4 WA = prim1d(a_[0], a_[1], a_[2], a.eos);
            const prim4 WB = prim1d(b_[0], b_[1], b_[2], a.eos);
            // left differences: A - (B of the lane before), B - A
            const prim4 dA{WA.r - lane_up_z(WB.r), WA.u - lane_up_z(WB.u), WA.p - lane_up_z(WB.p), 0.0};
            const prim4 dB{WB.r - WA.r, WB.u - WA.u, WB.p - WA.p, 0.0};
            const prim4 SA = slope4<LIM>(dA, dB);
            const prim4 SB = slope4<LIM>(dB, prim4{lane_down_z(dA.r), lane_down_z(dA.u), lane_down_z(dA.p), 0.0});
            const prim4 RA = face_minus(WA, SA);
            double fA[4], fB[4];
            face_flux<false, true>(face_plus(WA, SA), face_minus(WB, SB), a.eos, fA);  // face A | B
            face_flux<false, true>(face_plus(WB, SB), prim4{lane_down_z(RA.r), lane_down_z(RA.u), lane_down_z(RA.p), 0.0},
                                   a.eos, fB);  // face B | A of the lane after
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                a_[v] = a_[v] - dtdx * (fA[v] - lane_up_z(fB[v]));
                b_[v] = b_[v] - dtdx * (fB[v] - fA[v]);
            }
        };
        stage(A, B);
        if (a.bc != 0 && (g0 - 4 < 0 || g0 + 123 >= n)) {
            // U1 past an outflow / reflecting wall from the lane and half of
            // the mapped cell (clamped into the segment for cells far past
            // the wall, whose values stage 2 never reads)
            auto pos = [&](int g, int own) {
                if (g >= 0 && g < n) return own;
                const int q = nest_map_index(g, n, a.bc) - (g0 - 4);
                return q < 0 ? 0 : (q > 127 ? 127 : q);
            };
            const int qA = pos(gA, 2 * lane), qB = pos(gB, 2 * lane + 1);
            double nA[3], nB[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const double xa = __shfl(A[v], qA >> 1), xb = __shfl(B[v], qA >> 1);
                const double ya = __shfl(A[v], qB >> 1), yb = __shfl(B[v], qB >> 1);
                nA[v] = (qA & 1) ? xb : xa;
                nB[v] = (qB & 1) ? yb : ya;
            }
            if (a.bc == 2) {
                if (outA) nA[1] = -nA[1];
                if (outB) nB[1] = -nB[1];
            }
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                A[v] = nA[v];
                B[v] = nB[v];
            }
        }
        stage(A, B);
        if (storeA || storeB) {
            double oA[3], oB[3];
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                oA[v] = 0.5 * (zA[v] + A[v]);
                oB[v] = 0.5 * (zB[v] + B[v]);
            }
            auto speed = [&](const double* o, bool own) {
                const prim4 w = prim1d(o[0], o[1], o[2], a.eos);
                const double sp = __builtin_fabs(w.u) + sound_speed_1d(w.r, w.p, a.eos);
                const unsigned long long b = speed_slot_bits(sp, o[0], o[2]);
                sm = own && b > sm ? b : sm;
            };
            speed(oA, storeA);
            speed(oB, storeB);
#pragma unroll
            for (int v = 0; v < 3; ++v) st2(eA + (unsigned)v * vs, oA[v], oB[v]);
        }
    }
    if (!go) return;
    step1d_fold_speeds(a, sm, wmax);
}

__global__ __launch_bounds__(256) void speeds1d_kernel(stage1d_args a) {
    __shared__ unsigned long long wmax[4];
    const int n = a.n, ne = n + 4;
    const int e = 2 + blockIdx.x * 256 + threadIdx.x;
    unsigned long long sm = 0;
    if (e < n + 2) {
        const double r = a.U[e], E = a.U[2 * ne + e];
        prim4 w = prim1d(r, a.U[ne + e], E, a.eos);
        double s = __builtin_fabs(w.u) + sound_speed_1d(w.r, w.p, a.eos);
        sm = speed_slot_bits(s, r, E);
    }
    sm = wave_max_u64(sm);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = sm;
    __syncthreads();
    if (threadIdx.x == 0) {  // one atomic per block: into slot 0 of a slot set, or smax
        unsigned long long b = wmax[0];
        for (int w = 1; w < 4; ++w) b = b > wmax[w] ? b : wmax[w];
        atomicMax(a.slots ? a.slots : &a.ps->smax[0], b);
    }
}

// periodic (benchmark_euler1d.cpp:102-110), outflow or reflecting (mirror
// image, momentum negated) ghost fill, 3 variables
__global__ void fill1d_kernel(double* U, int n, int bc, const nest_prob_state* ps) {
    if (ps && ps->done) return;
    const int t = threadIdx.x;  // 3 vars x 2 ghosts x 2 sides = 12 threads
    if (t >= 12) return;
    const int ne = n + 4;
    const int k = t / 4, g = (t / 2) & 1, side = t & 1;
    double* u = U + k * ne;
    if (bc == 0) {
        if (side == 0) u[g] = u[n + g];
        else u[2 + n + g] = u[2 + g];
    } else if (bc == 2) {
        const double v = side == 0 ? u[3 - g] : u[n + 1 - g];
        const double r = k == 1 ? -v : v;
        if (side == 0) u[g] = r;
        else u[2 + n + g] = r;
    } else {
        if (side == 0) u[g] = u[2];
        else u[2 + n + g] = u[n + 1];
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
#ifndef NEST_STAGE_DMA
#define NEST_STAGE_DMA 1
#endif
template <int STAGE>
static void stage2d_dispatch(int lim, const stage_args& a, int grid, int block, hipStream_t s) {
#if NEST_STAGE_DMA
    // the LDS-ring kernel addresses the state buffers with 32-bit byte offsets
    if (a.kernel != 1 && a.total > 0 && a.total * (long long)sizeof(double) < (1LL << 32)) {
    const size_t lds = (size_t)(block / 64) * (STAGE == 1 ? kWaveLds1 : kWaveLds2) * sizeof(double);
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((stage2d_dma_kernel<LIM_MC, STAGE>), grid, block, lds, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((stage2d_dma_kernel<LIM_NONE, STAGE>), grid, block, lds, s, a); break;
        default: hipLaunchKernelGGL((stage2d_dma_kernel<LIM_MINMOD, STAGE>), grid, block, lds, s, a); break;
    }
    return;
    }
#endif
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((stage2d_kernel<LIM_MC, STAGE>), grid, block, 0, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((stage2d_kernel<LIM_NONE, STAGE>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((stage2d_kernel<LIM_MINMOD, STAGE>), grid, block, 0, s, a); break;
    }
}

void launch_stage2d(int stage, int lim, const stage_args& a, int grid, int block, hipStream_t s) {
    if (stage == 1) stage2d_dispatch<1>(lim, a, grid, block, s);
    else stage2d_dispatch<2>(lim, a, grid, block, s);
}

template <bool U1OUT>
void step2d_dispatch(int lim, const stage_args& a, int grid, int block, size_t lds, hipStream_t s) {
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((step2d_kernel<LIM_MC, U1OUT>), grid, block, lds, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((step2d_kernel<LIM_NONE, U1OUT>), grid, block, lds, s, a); break;
        default: hipLaunchKernelGGL((step2d_kernel<LIM_MINMOD, U1OUT>), grid, block, lds, s, a); break;
    }
}
// a.U1 set: the launch also writes the stage-1 state there
void launch_step2d(int lim, const stage_args& a, int grid, int block, hipStream_t s) {
    const size_t lds = (size_t)(block / 64) * kWaveLdsStep * sizeof(double);
    if (a.U1) step2d_dispatch<true>(lim, a, grid, block, lds, s);
    else step2d_dispatch<false>(lim, a, grid, block, lds, s);
}

int step2d_resident_blocks(int lim, int block) {
    const size_t lds = (size_t)(block / 64) * kWaveLdsStep * sizeof(double);
    int nb = 0;
    hipError_t e;
    switch (lim) {
        case LIM_MC: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step2d_kernel<LIM_MC, false>, block, lds); break;
        case LIM_NONE: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step2d_kernel<LIM_NONE, false>, block, lds); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step2d_kernel<LIM_MINMOD, false>, block, lds); break;
    }
    return e == hipSuccess ? nb : 0;
}

void launch_speeds2d(const stage_args& a, int grid, int block, hipStream_t s) {
    hipLaunchKernelGGL(speeds2d_kernel, grid, block, 0, s, a);
}

void launch_fill_regions(double* base, const nest_region* R, int nreg, int nvars, hipStream_t s) {
    if (nreg <= 0) return;
    hipLaunchKernelGGL(fill_regions_kernel, nreg, 256, 0, s, base, R, nreg, nvars);
}

void launch_finalize(nest_prob_state* P, unsigned long long* slots, unsigned long long* slots_clr, int nprob,
                     int par, hipStream_t s) {
    if (nprob <= 0) return;
    hipLaunchKernelGGL(finalize_kernel, nprob, NEST_SPEED_SLOTS, 0, s, P, slots, slots_clr, nprob, par);
}

void launch_reset_speeds(nest_prob_state* P, unsigned long long* slots, int nprob, hipStream_t s) {
    if (nprob <= 0) return;
    hipLaunchKernelGGL(reset_speeds_kernel, nprob, NEST_SPEED_SLOTS, 0, s, P, slots, nprob);
}

template <int STAGE>
static void stage1d_dispatch(int lim, const stage1d_args& a, int grid, hipStream_t s) {
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((stage1d_kernel<LIM_MC, STAGE>), grid, 256, 0, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((stage1d_kernel<LIM_NONE, STAGE>), grid, 256, 0, s, a); break;
        default: hipLaunchKernelGGL((stage1d_kernel<LIM_MINMOD, STAGE>), grid, 256, 0, s, a); break;
    }
}

void launch_stage1d(int stage, int lim, const stage1d_args& a, hipStream_t s) {
    const int grid = (a.n + 255) / 256;
    if (stage == 1) stage1d_dispatch<1>(lim, a, grid, s);
    else stage1d_dispatch<2>(lim, a, grid, s);
}

// Resident blocks (256 threads) of kernel K on the current device: its
// occupancy per compute unit times the compute units, queried once per
// device (0 if a query fails).
template <void (*K)(stage1d_args)>
static int resident_blocks_1d() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int v = cached[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, K, 256, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    v = per_cu * cus;
    cached[dev].store(v, std::memory_order_relaxed);
    return v;
}

// one lane kernel launch over nseg segments: at most the resident blocks of
// 4 waves, every wave looping over its share of the segments
template <void (*K)(stage1d_args)>
static void launch_lanes(stage1d_args a, int nseg, hipStream_t s) {
    const int need = (nseg + 3) / 4;
    const int res = resident_blocks_1d<K>();
    const int grid = std::max(1, res > 0 ? std::min(need, res) : need);
    a.waves = 4 * grid;
    hipLaunchKernelGGL(K, grid, 256, 0, s, a);
}

template <int LIM>
static void lanes_dispatch(const stage1d_args& a, hipStream_t s) {
    if (!a.pairs) {
        launch_lanes<step1d_lanes_kernel<LIM>>(a, (a.n + 55) / 56, s);
        return;
    }
    const int nseg = (a.n + 119) / 120;
    if ((((uintptr_t)a.U | (uintptr_t)a.U1) & 15) == 0 && (a.n & 1) == 0)
        launch_lanes<step1d_pairs_kernel<LIM, true>>(a, nseg, s);
    else
        launch_lanes<step1d_pairs_kernel<LIM, false>>(a, nseg, s);
}

// The lane kernels' resident counts of limiter lim on the current device,
// queried ahead (host queries only, no launch), so that launches inside a
// later graph capture make none.
void prepare_step1d(int lim) {
    switch (lim) {
        case LIM_MC:
            (void)resident_blocks_1d<step1d_lanes_kernel<LIM_MC>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_MC, true>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_MC, false>>();
            break;
        case LIM_NONE:
            (void)resident_blocks_1d<step1d_lanes_kernel<LIM_NONE>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_NONE, true>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_NONE, false>>();
            break;
        default:
            (void)resident_blocks_1d<step1d_lanes_kernel<LIM_MINMOD>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_MINMOD, true>>();
            (void)resident_blocks_1d<step1d_pairs_kernel<LIM_MINMOD, false>>();
            break;
    }
}

void launch_step1d(int lim, const stage1d_args& a, hipStream_t s) {
    if (a.lanes) {
        switch (lim) {
            case LIM_MC: lanes_dispatch<LIM_MC>(a, s); break;
            case LIM_NONE: lanes_dispatch<LIM_NONE>(a, s); break;
            default: lanes_dispatch<LIM_MINMOD>(a, s); break;
        }
        return;
    }
    const int grid = (a.n + 255) / 256;
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((step1d_kernel<LIM_MC>), grid, 256, 0, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((step1d_kernel<LIM_NONE>), grid, 256, 0, s, a); break;
        default: hipLaunchKernelGGL((step1d_kernel<LIM_MINMOD>), grid, 256, 0, s, a); break;
    }
}

void launch_speeds1d(const stage1d_args& a, hipStream_t s) {
    hipLaunchKernelGGL(speeds1d_kernel, (a.n + 255) / 256, 256, 0, s, a);
}

void launch_fill1d(double* U, int n, int bc, const nest_prob_state* ps, hipStream_t s) {
    hipLaunchKernelGGL(fill1d_kernel, 1, 64, 0, s, U, n, bc, ps);
}

}  // namespace NEST_NS
