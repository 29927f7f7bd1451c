// Discrete adjoint of one split stage of the 2D Euler step, S = Y o fill o
// X o fill, for gfx950 (wave64, FP64). Compiled once per numerics mode, like
// the forward kernels.
//
// The halo fill is part of the differentiated map: every ghost cell is a
// copy of the interior cell the boundary condition maps it to (periodic wrap
// or zero-gradient clamp), so along each grid line a sweep is a function of
// the interior cells only. Periodic lines are rings and need nothing more;
// on an outflow line the two ghost positions past each end are copies of the
// end cell, and the cotangents they receive are folded back into that cell.
//
// Y kernel (one wave per strip of 64 extended columns, lanes 3..60 own
// output, marching interior rows like the forward stage kernel):
//   * recomputes the X-swept row V(i) with the forward's sliding register
//     window (the X-sweep of the halo columns is what the fill would copy),
//     one row ahead of the Y reverse (software pipeline);
//   * Y-sweep reverse across lanes: cotangent of each y-face flux from the
//     stage-output cotangent, reverse HLLE + passive flux, slopes, limiter
//     and primitives, all neighbour terms moved by wave shuffles;
//   * writes Vb(i) = g*G(i) + Y'(V(i))^T g*G(i) for the owned columns.
// X kernel (one wave per strip of 64 columns, every lane owns its column;
// the X-sweep is a per-column recurrence, so no lane exchange): marches
// positions t along x with the primitives W(t-3..t) in registers; at each
// step it reverses the x-face t-3/2 (its flux cotangent comes from Vb rows
// t-2 and t-1), completes the slope cotangent of cell t-2 and reverses its
// limiter, which completes the primitive cotangent of cell t-3; cell t-3 is
// then emitted as out = Vb + (dW/dU)^T Wb (+ h*H).
#include "adjoint_numerics.hpp"
#include "kernels.hpp"

// Minimum waves per SIMD requested for the Y and X kernels (register budget
// 512/N per lane). Buffers are addressed by 32-bit element offsets (below
// 2^32 elements, checked by the host), which keeps the register counts down.
// Y kernel: 2 (with the cotangent row loaded one iteration ahead, the fast
// minmod / none instances need 168+ VGPRs, past the 3-wave budget). X
// kernel: 2 (its interior runs two march positions per trip; one position
// per trip fits 3 waves in the periodic fast instances but measured 4%
// slower overall).
#ifndef NEST_ADJY_WAVES
#define NEST_ADJY_WAVES(LIM) 2
#endif
#ifndef NEST_ADJX_WAVES
#define NEST_ADJX_WAVES 2
#endif

namespace NEST_NS {

__device__ __forceinline__ prim4 shup4(const prim4& a, int d) {
    return {__shfl_up(a.r, d), __shfl_up(a.u, d), __shfl_up(a.p, d), __shfl_up(a.q, d)};
}
__device__ __forceinline__ prim4 shdn4(const prim4& a, int d) {
    return {__shfl_down(a.r, d), __shfl_down(a.u, d), __shfl_down(a.p, d), __shfl_down(a.q, d)};
}
// delta 1 by zero-fill DPP wave shifts: lane 0 (up1) / lane 63 (dn1)
// receive 0. In the Y-sweep reverse those values reach only lane 0's and
// lanes 62-63's own faces, which are zeroed before use (has_face), and the
// limiter reverse of lanes 0 and 63, whose slope cotangent is then 0; owned
// lanes 3..60 and the outflow fold (which reads lanes >= 1) never see them.
__device__ __forceinline__ prim4 up1(const prim4& a) {
    return {lane_up_z(a.r), lane_up_z(a.u), lane_up_z(a.p), lane_up_z(a.q)};
}
__device__ __forceinline__ prim4 dn1(const prim4& a) {
    return {lane_down_z(a.r), lane_down_z(a.u), lane_down_z(a.p), lane_down_z(a.q)};
}

// ---------------------------------------------------------------------------
// Y kernel
// ---------------------------------------------------------------------------
template <int LIM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NEST_ADJY_WAVES(LIM), 8))) void adj_y_kernel(
    adj_args a) {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    // wave-uniform wave index: the work item and its tables live in scalar
    // registers
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // per-wave LDS ring of the last 4 loaded rows: the centre row of the X
    // update comes from here instead of a second read from memory (as in the
    // forward stage kernel); each lane touches only its own column
    __shared__ double ring_all[4 * 4 * 4 * 64];
    double* const ring = ring_all + wave * (4 * 4 * 64) + lane;
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
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + wave; it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        const nest_prob_state& ps = a.probs[pd.prob];
        const double dtdx = a.dt / ps.dx;
        const double dtdy = a.dt / ps.dy;
        const nest_prob_geom G = a.geoms[pd.prob];
        // lane -> global column (unmapped position along y) -> owning patch
        const int gs = pd.y0 + w.c0;
        const int gcol = gs + lane - 3;
        int q, le, lc;
        nest_gather_source(pd, G, a.itab, 2, gcol, q, le, lc);
        const nest_patch_desc qd = a.patches[q];
        const int snye = qd.ny + 4;
        const int svs = (qd.nx + 4) * snye;
        // this lane's column at extended row 0, as an element offset (a state
        // buffer stays far below 2^32 elements), and its row offset within
        const unsigned col = (unsigned)(qd.off + lc);
        const bool ring = G.bcy == 0;
        const bool real = ring || (gcol >= 0 && gcol < G.NY);  // outflow ghosts are not outputs
        const double greal = real ? a.g : 0.0;                   // the cotangent's scale on this lane
        const bool own = lane >= 3 && lane < 61 && gcol < G.NY;
        const bool edge = !ring && (gs == 0 || gs + 58 >= G.NY);  // wave-uniform

        // interior row i of buffer B at this lane's column
        auto load_int = [&](const double* B, int i, double* u) {
            const double* p = B + (col + (unsigned)(i + 2) * (unsigned)snye);
            u[0] = p[0];
            u[1] = p[svs];
            u[2] = p[2 * (long long)svs];
            u[3] = p[3 * (long long)svs];
        };
        // Y-sweep reverse of row i across lanes, from its X-swept state v and
        // the stage-output cotangent A (unscaled); writes Vb(i)
        auto yrev = [&](int i, const double* v, double* A) {
#pragma unroll
            for (int k = 0; k < 4; ++k) A[k] = greal * A[k];
            const prim4 Y = prim2d(v[0], v[1], v[2], v[3], false, a.eos);
            const prim4 dl = sub4(Y, up1(Y));
            const prim4 dr = dn1(dl);
            const prim4 sy = slope4<LIM>(dl, dr);
            const limsel4<LIM> cf = limiter_sel4<LIM>(dl, dr);
            const prim4 Lf = face_plus(Y, sy);
            const prim4 Rf = dn1(face_minus(Y, sy));
            // face between this lane and the next: flux cotangent from the
            // output cotangents of the two cells (y_sweep_t update order:
            // rho <- G0, m_u <- G3, m_v <- G1, E <- G2)
            double An[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) An[k] = lane_down_z(A[k]);
            double Fb[4];
            Fb[0] = dtdy * (An[0] - A[0]);
            Fb[1] = dtdy * (An[2] - A[2]);
            Fb[2] = dtdy * (An[3] - A[3]);
            Fb[3] = dtdy * (An[1] - A[1]);
            // every lane reverses its face (branch-free); lanes 0, 62, 63
            // have no face of the strip and drop theirs
            prim4 Lb = zero4(), Rb = zero4();
            face_flux_rev(Lf, Rf, a.eos, Fb, Lb, Rb);
            const bool has_face = lane >= 1 && lane <= 61;
            Lb = sel4(has_face, Lb, zero4());
            Rb = sel4(has_face, Rb, zero4());
            const prim4 Rbm = up1(Rb);  // right-state cotangent of the face below this cell
            const prim4 sb = scale4(0.5, sub4(Lb, Rbm));
            prim4 dlb, drb;
            cf.apply(sb, dlb, drb);
            prim4 Wb = add4(add4(Lb, Rbm), sub4(dlb, drb));
            Wb = sub4(Wb, dn1(dlb));
            Wb = add4(Wb, up1(drb));
            if (edge) {
                // outflow: ghost positions -1, -2 are copies of column 0 and
                // NY, NY+1 of column NY-1 (same primitives, so the
                // cotangents add before the primitive reverse)
                const prim4 u1 = up1(Wb), u2 = shup4(Wb, 2);
                const prim4 d1 = dn1(Wb), d2 = shdn4(Wb, 2);
                if (gcol == 0) Wb = add4(Wb, add4(u1, u2));
                if (gcol == G.NY - 1) Wb = add4(Wb, add4(d1, d2));
            }
            double Ub[4];
            prim2d_rev(Y, false, Wb, a.eos, Ub);
            if (own) {
                double* po = a.Vb + (col + (unsigned)(i + 2) * (unsigned)snye);
                po[0] = A[0] + Ub[0];
                po[svs] = A[1] + Ub[1];
                po[2 * (long long)svs] = A[2] + Ub[2];
                po[3 * (long long)svs] = A[3] + Ub[3];
            }
        };

        auto load_row = [&](int e, double* u) {
            const double* p;
            int s;
            if (e < 2 || e >= pd.nx + 2) {  // ghost rows: wave-uniform slow path
                int q2, le2, lc2;
                nest_gather_source(pd, G, a.itab, e, gcol, q2, le2, lc2);
                const nest_patch_desc qd2 = a.patches[q2];
                const int qnye = qd2.ny + 4;
                s = (qd2.nx + 4) * qnye;
                p = a.U + qd2.off + (long long)le2 * qnye + lc2;
            } else {
                p = a.U + (col + (unsigned)e * (unsigned)snye);
                s = svs;
            }
            u[0] = p[0];
            u[1] = p[s];
            u[2] = p[2 * (long long)s];
            u[3] = p[3 * (long long)s];
        };

        // X window warm-up (as the forward stage kernel; the same flux
        // arithmetic as the forward kernels of this numerics mode)
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
        double Un[4];
        load_row(w.r0 + 4, Un);
        // X-swept row i (forward recompute), advancing the window
        auto xstep = [&](int i, double* v) {
            double Unx[4], Uc[4];
            load_row(min(i + 5, pd.nx + 3), Unx);
            ring_get(i + 2, Uc);
            ring_put(i + 4, Un);
            prim4 W2 = prim2d(Un[0], Un[1], Un[2], Un[3], true, a.eos);
            prim4 dW2 = sub4(W2, W1);
            prim4 s1 = slope4<LIM>(dW1, dW2);
            double Fp[4];
            face_flux<true, true>(FL, face_minus(W1, s1), a.eos, Fp);
            FL = face_plus(W1, s1);
            v[0] = Uc[0] - dtdx * (Fp[0] - Fm[0]);
            v[1] = Uc[1] - dtdx * (Fp[1] - Fm[1]);
            v[2] = Uc[2] - dtdx * (Fp[3] - Fm[3]);
            v[3] = Uc[3] - dtdx * (Fp[2] - Fm[2]);
            W1 = W2;
            dW1 = dW2;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                Fm[k] = Fp[k];
                Un[k] = Unx[k];
            }
        };
        // software pipeline: the X-sweep of row i + 1 is issued with the
        // Y-sweep reverse of row i (the two do not depend on each other);
        // past the strip's last row it computes a row that is discarded
        // (its loads are clamped to the patch)
        // the cotangent row of the next iteration is loaded one iteration
        // ahead (unconditionally: row r1 is at most the patch's first ghost
        // row, in bounds)
        double vc[4], Gn[4];
        load_int(a.G, w.r0, Gn);
        xstep(w.r0, vc);
        for (int i = w.r0; i < w.r1; ++i) {
            double A[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) A[k] = Gn[k];
            load_int(a.G, i + 1, Gn);
            double vn[4];
            xstep(i + 1, vn);
            yrev(i, vc, A);
#pragma unroll
            for (int k = 0; k < 4; ++k) vc[k] = vn[k];
        }
    }
}

// ---------------------------------------------------------------------------
// X kernel
// ---------------------------------------------------------------------------
// Two waves per SIMD: the register budget (256 per lane) holds the march's
// carried state and two interleaved face reverses without spilling.
// OUTX: some problem of the launch has outflow x boundaries (the ghost-row
// fold is compiled in only then).
template <int LIM, bool OUTX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NEST_ADJX_WAVES, 8))) void adj_x_kernel(adj_args a) {
    const int lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int nw = gridDim.x * wpb;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int it = xcd_block(blockIdx.x, gridDim.x) * wpb + wave; it < a.nitems; it += nw) {
        const nest_item w = uniform_item(a.items[it]);
        const nest_patch_desc pd = a.patches[w.patch];
        const nest_prob_state& ps = a.probs[pd.prob];
        const double dtdx = a.dt / ps.dx;
        const nest_prob_geom G = a.geoms[pd.prob];
        const int gcol = pd.y0 + w.c0 + lane;
        const bool own = gcol < G.NY;
        const int gc = own ? gcol : G.NY - 1;  // lanes past the domain: any in-bounds column
        int q, le, lc;
        nest_gather_source(pd, G, a.itab, 2, gc, q, le, lc);
        const nest_patch_desc qd = a.patches[q];
        const int snye = qd.ny + 4;
        const int svs = (qd.nx + 4) * snye;
        const unsigned col = (unsigned)(qd.off + lc);  // element offset (buffers < 2^32 elements)
        const bool ring = G.bcx == 0;
        const bool lo = OUTX && !ring && pd.x0 + w.r0 == 0;     // outflow end at row 0
        const bool hi = OUTX && !ring && pd.x0 + w.r1 == G.NX;  // outflow end at row NX-1

        // value of buffer B at position t (local row of the home patch, any
        // integer; rows outside the patch resolve through the boundary map)
        auto load_pos = [&](const double* B, int t, double* u) {
            const int e = t + 2;
            const double* p;
            int s;
            if (e < 2 || e >= pd.nx + 2) {
                int q2, le2, lc2;
                nest_gather_source(pd, G, a.itab, e, gc, q2, le2, lc2);
                const nest_patch_desc qd2 = a.patches[q2];
                const int qnye = qd2.ny + 4;
                s = (qd2.nx + 4) * qnye;
                p = B + qd2.off + (long long)le2 * qnye + lc2;
            } else {
                p = B + (col + (unsigned)e * (unsigned)snye);
                s = svs;
            }
            u[0] = p[0];
            u[1] = p[s];
            u[2] = p[2 * (long long)s];
            u[3] = p[3 * (long long)s];
        };
        auto interior = [&](int t) {
            const int gr = pd.x0 + t;
            return ring || (gr >= 0 && gr < G.NX);
        };
        // out(row) = Vb(row) + (dW/dU)^T Wbar (+ h H(row)), in two halves: the
        // loads of Vb(row) and H(row) (emit_load; Vb(row) was loaded two
        // positions earlier, so it is cache-resident), issued before the
        // position's face reverse, and the sum and store (emit_store)
        // (loads without branches, so that the compiler's wait counts stay
        // exact: without H the second half re-reads Vb and is not used)
        const double* const Hs = a.H ? a.H : a.Vb;
        auto emit_load = [&](int row, double* vb, double* hh) {
            const unsigned o = col + (unsigned)(row + 2) * (unsigned)snye;
#pragma unroll
            for (int k = 0; k < 4; ++k) vb[k] = a.Vb[o + (unsigned)k * (unsigned)svs];
#pragma unroll
            for (int k = 0; k < 4; ++k) hh[k] = Hs[o + (unsigned)k * (unsigned)svs];
        };
        auto emit_store = [&](int row, const prim4& Wp, const prim4& Wbar, const double* vb, const double* hh) {
            double Ub[4];
            prim2d_rev(Wp, true, Wbar, a.eos, Ub);
            const unsigned o = col + (unsigned)(row + 2) * (unsigned)snye;
            double r[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) r[k] = vb[k] + Ub[k];
            if (a.H) {
#pragma unroll
                for (int k = 0; k < 4; ++k) r[k] += a.h * hh[k];
            }
            if (own) {
#pragma unroll
                for (int k = 0; k < 4; ++k) a.out[o + (unsigned)k * (unsigned)svs] = r[k];
            }
        };
        auto emit = [&](int row, const prim4& Wp, const prim4& Wbar) {
            double vb[4], hh[4];
            emit_load(row, vb, hh);
            emit_store(row, Wp, Wbar, vb, hh);
        };

        // The first face that reaches an emitted cell is r0 - 3/2 (through
        // the slope of cell r0 - 1), reversed at t = r0: the march starts
        // there with the window W(r0-3 .. r0-1) and zero cotangents (faces
        // before it touch only cells below r0; at an outflow end the faces
        // past the ghosts carry zero cotangent). Cell r1 - 1 completes at
        // t = r1 + 2, the ghosts past an outflow end two steps later.
        const int te = w.r1 + 3 + (hi ? 2 : 0);
        double u[4];
        load_pos(a.U, w.r0 - 3, u);
        prim4 W3 = prim2d(u[0], u[1], u[2], u[3], true, a.eos);
        load_pos(a.U, w.r0 - 2, u);
        prim4 W2 = prim2d(u[0], u[1], u[2], u[3], true, a.eos);
        load_pos(a.U, w.r0 - 1, u);
        prim4 W1 = prim2d(u[0], u[1], u[2], u[3], true, a.eos);
        double Vb2[4] = {0.0, 0.0, 0.0, 0.0};
        if (interior(w.r0 - 2)) load_pos(a.Vb, w.r0 - 2, Vb2);
        prim4 Wb3 = zero4(), Wb2 = zero4(), sb2 = zero4(), acc = zero4();
        prim4 s2 = slope4<LIM>(sub4(W2, W3), sub4(W1, W2));  // slope of cell t-2, carried
        // the operands of march position t, U(t) and Vb(t-1), are loaded one
        // position ahead (software pipeline: they land while the previous
        // position's face is reversed). The loads are unconditional (any
        // position resolves through the boundary map, in bounds); Vb of a
        // ghost row past an outflow end is then taken as 0.
        double nu[4], nvb[4];
        auto prefetch = [&](int t) {
            load_pos(a.U, t, nu);
            load_pos(a.Vb, t - 1, nvb);
            if (OUTX) {
                const bool in = interior(t - 1);
#pragma unroll
                for (int k = 0; k < 4; ++k) nvb[k] = in ? nvb[k] : 0.0;
            }
        };
        prefetch(w.r0);
        // one march position: face t - 3/2 reversed, the cotangent of cell
        // t - 3 completed in Wb3
        auto body = [&](int t) {
            double uu[4], Vb1[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                uu[k] = nu[k];
                Vb1[k] = nvb[k];
            }
            prefetch(t + 1);
            const prim4 Wt = prim2d(uu[0], uu[1], uu[2], uu[3], true, a.eos);
            const prim4 d3 = sub4(W2, W3), d2 = sub4(W1, W2), d1 = sub4(Wt, W1);
            const prim4 s1 = slope4<LIM>(d2, d1);
            const limsel4<LIM> cf = limiter_sel4<LIM>(d3, d2);
            // face t-3/2 (x_sweep_t update order: rho <- F0, m_u <- F1, m_v <- F3, E <- F2)
            double Fb[4];
            Fb[0] = dtdx * (Vb1[0] - Vb2[0]);
            Fb[1] = dtdx * (Vb1[1] - Vb2[1]);
            Fb[2] = dtdx * (Vb1[3] - Vb2[3]);
            Fb[3] = dtdx * (Vb1[2] - Vb2[2]);
            prim4 Lb = zero4(), Rb = zero4();
            face_flux_rev(face_plus(W2, s2), face_minus(W1, s1), a.eos, Fb, Lb, Rb);
            Wb2 = add4(Wb2, Lb);
            sb2 = add4(sb2, scale4(0.5, Lb));  // slope cotangent of cell t-2 complete
            prim4 Wb1 = Rb;
            const prim4 sb1 = scale4(-0.5, Rb);
            prim4 dlb, drb;
            cf.apply(sb2, dlb, drb);
            Wb2 = add4(Wb2, sub4(dlb, drb));
            Wb3 = sub4(Wb3, dlb);  // primitive cotangent of cell t-3 complete
            Wb1 = add4(Wb1, drb);
            return [&, Wt, Wb1, sb1, s1, Vb1]() {  // the shift to position t + 1
                W3 = W2;
                W2 = W1;
                W1 = Wt;
#pragma unroll
                for (int k = 0; k < 4; ++k) Vb2[k] = Vb1[k];
                Wb3 = Wb2;
                Wb2 = Wb1;
                sb2 = sb1;
                s2 = s1;
            };
        };
        // any position, with the outflow ends: ghost cotangents fold into
        // the end cells (row 0 gets those of positions -2, -1; row NX-1 those
        // of NX, NX+1, emitted once the last of them is complete)
        auto slow = [&](int t) {
            auto shift = body(t);
            const int p = t - 3;
            if (p < w.r0) {
                if (lo) acc = add4(acc, Wb3);
            } else if (p < w.r1) {
                prim4 val = Wb3;
                if (lo && p == 0) {
                    val = add4(val, acc);
                    acc = zero4();
                }
                if (hi && p == w.r1 - 1)
                    acc = val;
                else
                    emit(p, W3, val);
            } else {
                acc = add4(acc, Wb3);
                if (p == w.r1 + 1) emit(w.r1 - 1, W3, acc);  // W3: ghost copy of row NX-1
            }
            shift();
        };
        // the interior of the strip: emit every position (no end cases), two
        // positions per trip so that the two face reverses, which do not
        // depend on each other, can be interleaved
        auto fast = [&](int t) {
            double vb[4], hh[4];
            emit_load(t - 3, vb, hh);
            auto shift = body(t);
            emit_store(t - 3, W3, Wb3, vb, hh);
            shift();
        };
        const int ta = w.r0 + 3 + (lo ? 1 : 0);  // first position of the interior part
        const int tz = w.r1 + 3 - (hi ? 1 : 0);  // its end
        int t = w.r0;
        for (; t < ta && t < te; ++t) slow(t);
        for (; t + 1 < tz; t += 2) {
            fast(t);
            fast(t + 1);
        }
        for (; t < tz; ++t) fast(t);
        for (; t < te; ++t) slow(t);
    }
}

// Copies between the patch layout and the problem-major interior layout
// [problem][var][x][y] (problem q starts at goff[q]): dir 0 patch -> global,
// dir 1 global -> patch (interior cells only; ghosts untouched). Block
// (patch, variable, z) copies rows [PACK_ROWS z, PACK_ROWS (z+1)) of the
// patch, threads along y (both sides contiguous).
constexpr int PACK_ROWS = 4;
__global__ __launch_bounds__(256) void pack_kernel(const double* src, double* dst, const nest_patch_desc* P,
                                                   const nest_prob_geom* geoms, const long long* goff, int dir) {
    const nest_patch_desc d = P[blockIdx.x];
    const int k = blockIdx.y;
    const int x0 = blockIdx.z * PACK_ROWS;
    if (x0 >= d.nx) return;
    const nest_prob_geom G = geoms[d.prob];
    const int nye = d.ny + 4;
    const long long vs = (long long)(d.nx + 4) * nye;
    const long long gbase = goff[d.prob] + (long long)k * G.NX * G.NY;
    const int x1 = min(x0 + PACK_ROWS, d.nx);
    for (int ix = x0; ix < x1; ++ix) {
        const long long ps = d.off + k * vs + (long long)(ix + 2) * nye + 2;
        const long long gi = gbase + (long long)(d.x0 + ix) * G.NY + d.y0;
        for (int iy = threadIdx.x; iy < d.ny; iy += blockDim.x) {
            if (dir == 0)
                dst[gi + iy] = src[ps + iy];
            else
                dst[ps + iy] = src[gi + iy];
        }
    }
}

void launch_pack(const double* src, double* dst, const nest_patch_desc* P, int npatch, int max_nx,
                 const nest_prob_geom* geoms, const long long* goff, int dir, hipStream_t s) {
    if (npatch <= 0 || max_nx <= 0) return;
    const int nz = (max_nx + PACK_ROWS - 1) / PACK_ROWS;
    hipLaunchKernelGGL(pack_kernel, dim3(npatch, 4, nz), dim3(256), 0, s, src, dst, P, geoms, goff, dir);
}

void launch_adj_y(int lim, const adj_args& a, int grid, int block, hipStream_t s) {
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL(adj_y_kernel<LIM_MC>, grid, block, 0, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL(adj_y_kernel<LIM_NONE>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(adj_y_kernel<LIM_MINMOD>, grid, block, 0, s, a); break;
    }
}
template <bool OUTX>
void launch_adj_x_bc(int lim, const adj_args& a, int grid, int block, hipStream_t s) {
    switch (lim) {
        case LIM_MC: hipLaunchKernelGGL((adj_x_kernel<LIM_MC, OUTX>), grid, block, 0, s, a); break;
        case LIM_NONE: hipLaunchKernelGGL((adj_x_kernel<LIM_NONE, OUTX>), grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL((adj_x_kernel<LIM_MINMOD, OUTX>), grid, block, 0, s, a); break;
    }
}
void launch_adj_x(int lim, const adj_args& a, int grid, int block, hipStream_t s) {
    if (a.outflow_x)
        launch_adj_x_bc<true>(lim, a, grid, block, s);
    else
        launch_adj_x_bc<false>(lim, a, grid, block, s);
}

}  // namespace NEST_NS
