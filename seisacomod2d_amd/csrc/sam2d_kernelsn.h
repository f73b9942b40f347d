// sam2d_kernelsn.h -- the forward step advanced T time steps per pass
// (temporal blocking of depth T = 2..5) for gfx950.
//
// The per-cell arithmetic is that of k_step (sam2d_kernels.hip), so every T
// gives the same bits.  One launch carries the state from step n-1 to step
// n+T-1; the states in between stay in registers, so a cell's state and
// coefficients cross HBM once per T updates (outside the sponge frame 32
// bytes per T cell-updates, plus halos).
//
// A wavefront covers 64 rows: it owns 64 - 6T and carries a 3T-row halo above
// and below, the dependency cone of T steps (each step reaches two rows one
// way and one row the other in each phase).  It marches along x; at column c
// step s (s = 0 .. T-1) runs its p-phase at column c - 3s and its v-phase at
// column c - 3s - 2:
//   p-phase s at cp = c-3s    needs vx(s-1) at cp-2..cp+1, vz, px, pz(s-1) at cp
//   v-phase s at jv = c-3s-2  needs p(s) at jv-1..jv+2, vx, vz(s-1) at jv
// Every intermediate quantity lives in a four-column window indexed by
// compile-time position over a march unrolled by four; a coefficient needed
// by a later step moves from one step's window to the next.
//
// Receivers: the pressure after each intermediate step is kept for the
// receiver rows in a band slice; the state after the last step is stored.
// The next launch gathers all T samples (gather_deferred).
//
// State layout (slot_cells, slot_pz): cell-interleaved (P, vx, vz) of a cell
// together, one 12-byte load and store per lane and column, a contiguous 768
// bytes per wave, which streams faster on MI355X than fields in separate
// arrays.  P is px in the sponge frame and the pressure p = px + pz outside
// it, where both p-profiles are exactly 1 and px, pz are only ever used as
// their sum; the frame's pz is a field of its own, loaded and stored only by
// frame lanes.  Outside the frame the step computes
//   P' = (fma(KX, Dx- vx, P) + source) + fl(KZ * Dz- vz),
// the frame's p = px' + pz' with pz = 0 (all profile factors being 1 there).
//
// Instruction economy:
//  * z-neighbours move between lanes with DPP wavefront shifts;
//  * memory is addressed as a uniform base plus a 32-bit per-lane byte offset,
//    one offset per column shared by every field, and the prefetch slots are
//    emptied through keep() so each next load lands in place;
//  * groups of four columns where the x sponge profiles are exactly 1 and the
//    stencil has its 4th-order coefficients take a fast path without profile
//    loads or coefficient selects (multiplying by exactly 1 is exact, so both
//    paths give the same bits); slow groups, the fast range and slow groups
//    run as three loops in sequence.  Fast groups touch no pz; a wave with
//    rows in the z-frame runs all its groups as slow groups.

#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"
#include "sam2d_round.h"

namespace sam2d {

struct ColN {
    float kx, kz;               // at column c (kz unused when KZ aliases KX)
    float b;                    // buoyancy 1/rho at column c-1
    ftri s;                     // the cells (P, vx, vz) at column c+1, kept whole
                                // until the column consumes it
    float pz;                   // pz at column c+1 (frame lanes; 0 elsewhere)
};

// Arguments of one tile's march.
struct StepNTile {
    int zb, xa, xb;             // first owned row; columns [xa, xb) to store
    int step, in_slot, out_slot, band_mask, band_w, ring_slot;
};

// One tile: the march over columns xa - (3T-2) .. xb + 3T-2 of a wave's
// z-strip.  One 12-byte load per column and row brings P, vx, vz of column
// c+1 (and pz where the march touches the frame); P, pz, vz wait a column (vz
// three) in windows, and the last step stores the cell of column jv at once,
// P and pz kept from its p-phase two columns earlier.  GENERAL: any tile, every column group on the general path (edge
// stencil coefficients, x-profiles, sponge-frame lanes by row and column).
// Otherwise the march must lie in the fast range (fast_tile): all groups take
// the fast path, and only a wave with rows in the z-frame touches pz.
template <int T, int WAVES, int PF, bool KZA, bool RING, bool GENERAL>
__device__ __forceinline__ void stepn_tile(const ShotDev &d, int lane, const StepNTile &a)
{
    constexpr int HALO = 3 * T;
    constexpr int OWN = 64 - 6 * T;
    using gcchar = const SAM2D_GLOBAL char;
    using gchar = SAM2D_GLOBAL char;
    const int nnz = d.nnz, nnx = d.nnx;
    const int zb = a.zb, xa = a.xa, xb = a.xb, step = a.step;
    const int in_slot = a.in_slot, out_slot = a.out_slot, band_mask = a.band_mask;
    const int band_w = a.band_w, ring_slot = a.ring_slot;
    const int r = zb - HALO + lane;
    const bool rown = (lane >= HALO) && (lane < HALO + OWN) && (r < nnz);
    const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
    const unsigned rc4 = rc * 4u;
    const unsigned rc12 = rc * 12u;

    // Inputs, outputs and coefficients never overlap.
    gcchar *__restrict__ si = (gcchar *)slot_cells(d, in_slot);
    gchar *__restrict__ so = (gchar *)slot_cells(d, out_slot);
    // pz lies 12 ncell bytes past a slot's cells (slot_pz): addressed
    // from the cells' base, so it needs no base registers of its own.
    const unsigned pzb = (unsigned)d.ncell * 12u;
    gcchar *__restrict__ KX = (gcchar *)d.KX;
    gcchar *__restrict__ KZ = (gcchar *)d.KZ;
    gcchar *__restrict__ Bf = (gcchar *)d.B;
    const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);
    const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);
    // Ring slots of the T steps (RING only): the pairs, the frame's side
    // values 8 ncell bytes further (ring_pair, ring_side).
    gchar *__restrict__ ring[T];
    #pragma unroll
    for (int t = 0; t < T; t++)
        ring[t] = RING ? (gchar *)ring_pair(d.ring, d.ncell, ring_slot + t) : nullptr;
    const unsigned rside = (unsigned)d.ncell * 8u;

    auto ld = [](gcchar *__restrict__ base, unsigned off) -> float {
        return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
    };
    auto st1 = [](gchar *__restrict__ base, unsigned off, float v) {
        *(SAM2D_GLOBAL float *__restrict__)(base + off) = v;
    };
    // Byte offset of this lane's row in column col (clamped to the grid).
    auto coff = [&](int col) -> unsigned {
        return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 4u + rc4;
    };
    // Byte offset of this lane's cell (P, vx, vz) in column col (clamped).
    auto coff3 = [&](int col) -> unsigned {
        return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 12u + rc12;
    };
    auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };
    // Sponge frame: this lane's (clamped) row in the z-frame, a (clamped)
    // column outside the x-frame; whether any row of the wave is in the
    // z-frame (uniform).
    const bool zfr = (int)rc < d.zin_lo || (int)rc > d.zin_hi;
    const bool wave_zf = max(zb - HALO, 0) < d.zin_lo || min(zb - HALO + 63, nnz - 1) > d.zin_hi;
    auto xin = [&](int col) -> bool {
        const int cc = clampc(col);
        return cc >= d.xin_lo && cc <= d.xin_hi;
    };

    const float daz = prz[rc], skz = prz[nnz + rc];
    const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
    float zm0, zm1, zp0, zp1;
    dminus_coef(r, nnz, zm0, zm1);
    dplus_coef(r, nnz, zp0, zp1);
    const bool src_row = (r == d.src_z);
    // Source of step step+k: injected while step+k <= nts, the increment
    // past the wavelet's end being 0.
    float amp[T];
    int srcx[T];
    #pragma unroll
    for (int k = 0; k < T; k++) {
        amp[k] = (step + k < d.nts) ? gp(d.amp)[step + k] : 0.0f;
        srcx[k] = (step + k <= d.nts) ? d.src_x : -1;
    }

    const bool strip_band = d.rrow_lo <= d.rrow_hi && d.rrow_lo <= zb + OWN - 1 &&
                            d.rrow_hi >= zb;
    // Band slices this launch writes (restrict: they overlap nothing else
    // the march touches, so the descriptor is not read again).
    SAM2D_GLOBAL float *__restrict__ bandp[T];
    #pragma unroll
    for (int t = 0; t < T; t++) bandp[t] = band_slice(d, band_w, t < T - 1 ? t : 0);
    const bool band_lane = rown && r >= d.rrow_lo && r <= d.rrow_hi;
    // Operands of column c.  MODE 0: the general path, any column;
    // 1: the fast path (every column outside the x-frame), a wave with rows
    // in the z-frame; 2: the fast path, a wave without, where no lane touches
    // pz.  pz is 0 outside the frame (never written there but as 0), so
    // MODE 0 and 1 load it for every lane: the waves and columns that take
    // them are in, or next to, the frame.
    auto load_col = [&](auto mode_tag, int c) -> ColN {
        constexpr int MODE = decltype(mode_tag)::value;
        ColN o;
        o.s = ld3(si, coff3(c + 1));
        o.pz = MODE == 2 ? 0.0f : ld(si, pzb + coff(c + 1));
        const unsigned o0 = coff(c);
        o.kx = ld(KX, o0);
        o.kz = KZA ? 0.0f : ld(KZ, o0);
        o.b = ld(Bf, coff(c - 1));
        return o;
    };

    // Windows, [k & 3] relative to the unrolled base:
    //   vxin          vx(n-1) at c-2 .. c+1
    //   P, PX, PZ[s]  p-phase s results at c-3s-3 .. c-3s (and KXw, KZw)
    //   VX, VZ[s]     v-phase s results at c-3s-5 .. c-3s-2 (and BXw, BZw)
    //   pxin, pzin, vzin  px, pz, vz(n-1) at c-2 .. c+1
    float vxin[4], pxin[4], pzin[4], vzin[4];
    float P[T][4], PX[T][4], PZ[T][4], KXw[T][4], KZw[T][4];
    float VX[T][4], VZ[T][4], BXw[T][4], BZw[T][4];
    #pragma unroll
    for (int k = 0; k < 4; k++) {
        vxin[k] = pxin[k] = pzin[k] = vzin[k] = 0.0f;
        #pragma unroll
        for (int t = 0; t < T; t++) {
            P[t][k] = PX[t][k] = PZ[t][k] = KXw[t][k] = KZw[t][k] = 0.0f;
            VX[t][k] = VZ[t][k] = BXw[t][k] = BZw[t][k] = 0.0f;
        }
    }
    // Columns c_first .. c_last: the last step stores jv = c - 3T + 1 in
    // [xa, xb).  Prefetches past c_last re-read column c_last (a cache
    // hit) and the march ends at c_last exactly.
    const int c_first = xa - (3 * T - 2);
    const int c_last = xb + 3 * T - 2;
    #pragma unroll
    for (int k = 0; k < 3; k++) {
        const int col = c_first - 2 + k;
        const ftri sv = ld3(si, coff3(col));
        pxin[k] = sv.x; vxin[k] = sv.y; vzin[k] = sv.z;
        pzin[k] = ld(si, pzb + coff(col));
    }

    // Buoyancy at the v-phase column of the first march column,
    // jv = c_first - 2.
    float b_0 = ld(Bf, coff(c_first - 2));
    const float dtdx = d.dtdx, dtdz = d.dtdz;

    using M0 = std::integral_constant<int, 0>;
    using M1 = std::integral_constant<int, 1>;
    using M2 = std::integral_constant<int, 2>;
    ColN q[PF];
    #pragma unroll
    for (int u = 0; u < PF; u++) q[u] = load_col(M0(), c_first + u);

    // One column of the march.  MODE as for load_col: 0 the general path
    // (sponge-frame lanes by row and column), 1 and 2 the fast path (every
    // column the group computes or loads lies outside the x-frame; MODE 1
    // tells the lanes apart by row, under MODE 2 every row of the wave is
    // outside the z-frame).
    auto column = [&](auto mode_tag, auto u_tag, int c0) {
        constexpr int MODE = decltype(mode_tag)::value;
        constexpr bool FAST = MODE > 0;
        constexpr int u = decltype(u_tag)::value;
        constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
        const int c = c0 + u;
        ColN &cur = q[u % PF];
        pxin[i3] = keep(cur.s.x);
        pzin[i3] = MODE == 2 ? 0.0f : keep(cur.pz);
        vxin[i3] = keep(cur.s.y);
        vzin[i3] = keep(cur.s.z);

        #pragma unroll
        for (int t = 0; t < T; t++) {
            // ---- p-phase of step t at column cp
            const int cp = c - 3 * t;
            {
                float m0 = SAM2D_A0, m1 = SAM2D_A1;
                if (!FAST) dminus_coef(cp, nnx, m0, m1);
                const float *vxw = t == 0 ? vxin : VX[t > 0 ? t - 1 : 0];
                const float dvx = fmaf(m0, vxw[i2] - vxw[i1], m1 * (vxw[i3] - vxw[i0]));
                const float vz_0 = t == 0 ? vzin[i2] : VZ[t > 0 ? t - 1 : 0][i2];
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                float kx, kz, px_old, pz_old;
                if (t == 0) {
                    kx = keep(cur.kx);
                    kz = KZA ? kx : keep(cur.kz);
                    px_old = pxin[i2];
                    pz_old = pzin[i2];
                } else {
                    kx = KXw[t - 1][i0];
                    kz = KZA ? kx : KZw[t - 1][i0];
                    px_old = PX[t - 1][i0];
                    pz_old = PZ[t - 1][i0];
                }
                float pxn;
                if (FAST) {
                    pxn = fmaf(kx, dvx, px_old);
                } else {
                    const int cc = clampc(cp);
                    pxn = fmaf(kx * prx[nnx + cc], dvx, prx[cc] * px_old);
                }
                if (src_row && cp == srcx[t]) pxn += amp[t];
                // MODE 2: every z-profile factor of the wave is exactly 1 and
                // pz is 0, so the general form reduces to these (same bits)
                const float pzn = MODE == 2 ? fmaf(kz, dvz, 0.0f)
                                            : fmaf(kz * skz, dvz, daz * pz_old);
                const float pn = pxn + pzn;
                // outside the frame P carries p and pz stays 0
                const bool inner = MODE == 2 ? true : (MODE == 1 ? !zfr : !zfr && xin(cp));
                P[t][i3] = pn;
                PX[t][i3] = inner ? pn : pxn;         // the last step's are stored
                PZ[t][i3] = inner ? 0.0f : pzn;       // with vx, vz at jv
                if (t < T - 1) {
                    KXw[t][i3] = kx;
                    if (!KZA) KZw[t][i3] = kz;
                    if (((band_mask >> t) & 1) && strip_band && cp >= xa && cp < xb) {
                        if (band_lane) bandp[t][(size_t)r * nnx + cp] = P[t][i3];
                    }
                }
                if (RING && rown && cp >= xa && cp < xb) {
                    // the step's pressure and the dJ/dkappa factors of the
                    // state before it (ring_pair, ring_side)
                    const float skx = FAST ? 1.0f : prx[nnx + clampc(cp)];
                    float ez;
                    const float e = ring_factor(dvx, dtdx * skx, dvz, dtdz * skz, inner, ez);
                    const unsigned cell = (unsigned)cp * (unsigned)nnz + rc;
                    st2(ring[t], cell * 8u, P[t][i3], e);
                    if (MODE != 2 && !inner) st1(ring[t], rside + cell * 4u, ez);
                }
            }
            // ---- v-phase of step t at column jv
            const int jv = cp - 2;
            {
                float q0 = SAM2D_A0, q1 = SAM2D_A1;
                if (!FAST) dplus_coef(jv, nnx, q0, q1);
                const float pm3 = P[t][i0], pm2 = P[t][i1], pm1 = P[t][i2], pn = P[t][i3];
                const float dpx = fmaf(q0, pm1 - pm2, q1 * (pn - pm3));
                const float p_u1 = lane_up1(pm2);
                const float p_d1 = lane_dn1(pm2), p_d2 = lane_dn1(p_d1);
                const float dpz = fmaf(zp0, p_d1 - pm2, zp1 * (p_d2 - p_u1));
                float bx, bz, vx_old, vz_old;
                if (t == 0) {
                    // BX, BZ at jv from the buoyancy (the host's fields, bit
                    // for bit: cpu_stagger's averages and edge rules,
                    // cpu_modelling_kernels.cpp:64-121, times dt/dx and dt/dz
                    // in FP64, rounded without bias)
                    const float b_1 = keep(cur.b);            // b(jv + 1)
                    const bool last_row = r >= nnz - 1;
                    // The lane shifts run before the row test: inside an arm
                    // of ?: they would run with the last row's lanes disabled,
                    // which read as 0 to the row above.
                    const float b_d1 = lane_dn1(b_0);         // b(jv, r + 1)
                    const float bu1 = last_row ? 0.0f : 0.5f * (b_d1 + b_0);
                    float bu2;
                    if (FAST || jv < nnx - 1) {
                        bu2 = 0.5f * (b_1 + b_0);
                    } else {        // last column: bu1 of column nnx - 2
                        const float b_m1 = ld(Bf, coff(jv - 1));
                        const float b_m1d1 = lane_dn1(b_m1);
                        bu2 = last_row ? 0.0f : 0.5f * (b_m1d1 + b_m1);
                    }
                    bx = round_unbiased_prod_pos(dtdx, bu2, cell_draw(RND_BX, jv, r));
                    bz = round_unbiased_prod_pos(dtdz, bu1, cell_draw(RND_BZ, jv, r));
                    b_0 = b_1;
                    vx_old = vxin[i0];
                    vz_old = vzin[i0];
                } else {
                    bx = BXw[t - 1][i0];
                    bz = BZw[t - 1][i0];
                    vx_old = VX[t - 1][i0];
                    vz_old = VZ[t - 1][i0];
                }
                if (t < T - 1) {
                    if (FAST) {
                        VX[t][i3] = fmaf(bx, dpx, vx_old);
                    } else {
                        const int jc = clampc(jv);
                        VX[t][i3] = fmaf(bx * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx_old);
                    }
                    VZ[t][i3] = MODE == 2 ? fmaf(bz, dpz, vz_old)
                                          : fmaf(bz * sbz, dpz, dbz * vz_old);
                    BXw[t][i3] = bx;
                    BZw[t][i3] = bz;
                } else if (rown && jv >= xa && jv < xb) {
                    float vxn;
                    if (FAST) {
                        vxn = fmaf(bx, dpx, vx_old);
                    } else {
                        const int jc = clampc(jv);
                        vxn = fmaf(bx * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx_old);
                    }
                    const float vzn = MODE == 2 ? fmaf(bz, dpz, vz_old)
                                                : fmaf(bz * sbz, dpz, dbz * vz_old);
                    const unsigned cn = (unsigned)jv * (unsigned)nnz;
                    st3(so, cn * 12u + rc12, PX[t][i1], vxn, vzn);
                    if (MODE != 2) st1(so, pzb + cn * 4u + rc4, PZ[t][i1]);
                }
            }
        }
        cur = load_col(mode_tag, min(c + PF, c_last));
    };

    using U0 = std::integral_constant<int, 0>;
    using U1 = std::integral_constant<int, 1>;
    using U2 = std::integral_constant<int, 2>;
    using U3 = std::integral_constant<int, 3>;
    // Groups of four columns from c0 to the end of the march.
    auto march_to_end = [&](auto mode_tag, int c0) {
        for (; c0 <= c_last; c0 += 4) {
            column(mode_tag, U0(), c0);
            if (c0 + 1 > c_last) break;
            column(mode_tag, U1(), c0);
            if (c0 + 2 > c_last) break;
            column(mode_tag, U2(), c0);
            if (c0 + 3 > c_last) break;
            column(mode_tag, U3(), c0);
        }
    };
    int c0 = c_first;
    if (GENERAL) {
        march_to_end(M0(), c0);
    } else if (wave_zf) {
        for (; c0 + 3 <= c_last; c0 += 4) {
            column(M1(), U0(), c0); column(M1(), U1(), c0);
            column(M1(), U2(), c0); column(M1(), U3(), c0);
        }
    } else {
        for (; c0 + 3 <= c_last; c0 += 4) {
            column(M2(), U0(), c0); column(M2(), U1(), c0);
            column(M2(), U2(), c0); column(M2(), U3(), c0);
        }
    }
}

// Whether the march of tile columns [xa, xb) lies in the fast range: whole
// four-column groups whose columns, and the state columns they load (up to
// c0 + 4 + PF), are outside the x-frame and in the 4th-order stencil range.
template <int T, int PF>
__device__ __forceinline__ bool fast_tile(const ShotDev &d, int xa, int xb)
{
    const int c_first = xa - (3 * T - 2), c_last = xb + 3 * T - 2;
    const int fast_lo = max(d.xin_lo + 3 * T - 1, 3 * T);
    const int fast_hi = min(d.xin_hi - 4 - PF, d.nnx - 5);
    return c_first >= fast_lo && c_last - 3 <= fast_hi && ((c_last - c_first + 1) & 3) == 0;
}

// The kernel: T steps of every tile of every staged shot.
// RING: the gradient's recomputation of a checkpoint segment.  For each step
// t the p-phase also stores, for every owned cell of its column, what one
// backward step reads of the forward into ring slot ring_slot + t: the pair
// (p after step+t, E) in one dwordx2 store per lane, and a frame lane's Ez
// (ring_pair, ring_side).  No receivers are recorded.
// Three-step kernels are held to three waves per SIMD (at most 168 VGPRs),
// the occupancy their measured best shapes need.
template <int T, int WAVES, int PF, bool KZA, bool RING>
__global__ __launch_bounds__(WAVES * 64, T == 3 ? 3 : 1)
void k_stepn(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd, int ntiles,
             int step, int in_slot, int out_slot, int band_mask, int band_w, int4 band_s,
             int state_sample, int ngroups, int xorder, int ring_slot)
{
    static_assert(T >= 2 && T <= BAND_SLICES + 1, "steps per pass");
    static_assert(PF == 2 || PF == 4, "PF must divide the unroll of 4");
    constexpr int OWN = 64 - 6 * T;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    if (band_s.x >= 0 || band_s.y >= 0 || band_s.z >= 0 || band_s.w >= 0 || state_sample >= 0)
        gather_deferred(shots, nshots, band_w ^ 1, band_s, in_slot, state_sample,
                        blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile_ordered(v, per_xcd, ngroups, xorder);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const ShotDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int zb = (group * WAVES + wv) * OWN;
        if (zb >= d.nnz) continue;                    // wave-uniform
        StepNTile a;
        tile_columns(d, chunk, a.xa, a.xb);
        a.zb = zb;
        a.step = step; a.in_slot = in_slot; a.out_slot = out_slot;
        a.band_mask = band_mask; a.band_w = band_w; a.ring_slot = ring_slot;
        // Most tiles take the fast form; the sponge frame's own narrow chunk
        // columns (shot_tiles) take the general one.
        if (fast_tile<T, PF>(d, a.xa, a.xb))
            stepn_tile<T, WAVES, PF, KZA, RING, false>(d, lane, a);
        else
            stepn_tile<T, WAVES, PF, KZA, RING, true>(d, lane, a);
    }
}

struct StepNArgs {
    const ShotDev *d;
    int ns, vt, per, nt, step, in, out, bmask, bw;
    int4 bs;
    int ss, ng, xo, ring;
};

template <int T, int W, int PF, bool KZA, bool RING>
static void launchn_t(dim3 g, hipStream_t s, const StepNArgs &a)
{
    hipLaunchKernelGGL((k_stepn<T, W, PF, KZA, RING>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt,
                       a.per, a.nt, a.step, a.in, a.out, a.bmask, a.bw, a.bs, a.ss, a.ng, a.xo,
                       a.ring);
}

template <int T, int W, bool RING>
static void launchn_w(int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a)
{
    if (pf == 2 && kza) launchn_t<T, W, 2, true, RING>(g, s, a);
    else if (pf == 2) launchn_t<T, W, 2, false, RING>(g, s, a);
    else if (kza) launchn_t<T, W, 4, true, RING>(g, s, a);
    else launchn_t<T, W, 4, false, RING>(g, s, a);
}

// Launch of k_stepn<T> (RING: the recomputation form) with the variant's
// waves per block and prefetch depth.  Each depth, and the ring form, is
// instantiated in a translation unit of its own (sam2d_kernelsn_t*.hip) so
// that the builds run in parallel.
template <int T, bool RING>
static hipError_t launchn_waves(int waves, int pf, bool kza, dim3 g, hipStream_t s,
                                const StepNArgs &a)
{
    switch (waves) {
    case 1: launchn_w<T, 1, RING>(pf, kza, g, s, a); break;
    case 2: launchn_w<T, 2, RING>(pf, kza, g, s, a); break;
    case 4: launchn_w<T, 4, RING>(pf, kza, g, s, a); break;
    case 8: launchn_w<T, 8, RING>(pf, kza, g, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_stepn_t2(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a);
hipError_t launch_stepn_t3(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a);
hipError_t launch_stepn_t4(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a);
hipError_t launch_stepn_t5(int waves, int pf, bool kza, dim3 g, hipStream_t s, const StepNArgs &a);
hipError_t launch_stepn_t3_ring(int waves, int pf, bool kza, dim3 g, hipStream_t s,
                                const StepNArgs &a);

}  // namespace sam2d
