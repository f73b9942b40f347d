// sam2d_kernels3.hip -- the forward step advanced three time steps per pass
// (temporal blocking of depth 3) for gfx950.
//
// The same per-cell arithmetic as k_step (sam2d_kernels.hip) and k_step2
// (sam2d_kernels2.hip), so all three give the same bits.  One launch carries
// the state from step n-1 to step n+2; the states after steps n and n+1 stay
// in registers, so a cell's fields and coefficients cross HBM once per three
// updates (44 bytes per three cell-updates, plus halos).
//
// A wavefront owns OWN_Z3 = 46 rows with a nine-row halo above and below (the
// dependency cone of three steps) and marches along x through six pipeline
// stages per column c:
//   A  step-n   p-phase at c      B  step-n   v-phase at c-2
//   C  step-n+1 p-phase at c-3    D  step-n+1 v-phase at c-5
//   E  step-n+2 p-phase at c-6    F  step-n+2 v-phase at c-8
// Every window holds four columns and rotates by compile-time index over a
// march unrolled by four; a coefficient needed three columns later passes
// from one window to the next (KX: A -> C -> E, BX/BZ: B -> D -> F).
//
// Receivers: the pressure after steps n and n+1 is kept for the receiver rows
// in band slices 0 and 1; the state after step n+2 is stored.  The next
// launch gathers all three samples (gather_deferred).
//
// Loads and stores, DPP lane shifts, fast path, prefetch and the three-loop
// march are those of k_step2.

#include <hip/hip_runtime.h>
#include <type_traits>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

struct Col3 {
    float vx_next;              // vx(n-1) at column c+1
    float vz, px, pz, kx, kz;   // at column c (kz unused when KZ aliases KX)
    float bx, bz, vz_j;         // BX, BZ, vz(n-1) at column c-2
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_step3(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd, int ntiles,
             int step, int in_slot, int out_slot, int band_mask, int band_w, int band_s0,
             int band_s1, int state_sample)
{
    static_assert(PF == 2 || PF == 4, "PF must divide the unroll of 4");
    using gcchar = const SAM2D_GLOBAL char;
    using gchar = SAM2D_GLOBAL char;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    if (band_s0 >= 0 || band_s1 >= 0 || state_sample >= 0)
        gather_deferred(shots, nshots, band_w ^ 1, band_s0, band_s1, in_slot, state_sample,
                        blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile(v, per_xcd);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const ShotDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int nnz = d.nnz, nnx = d.nnx;
        const int zb = (group * WAVES + wv) * OWN_Z3;
        if (zb >= nnz) continue;                      // wave-uniform
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);

        const int r = zb - HALO_T3 + lane;
        const bool rown = (lane >= HALO_T3) && (lane < HALO_T3 + OWN_Z3) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
        const unsigned rc4 = rc * 4u;

        gcchar *__restrict__ pxi = (gcchar *)slot_field(d, in_slot, 0);
        gcchar *__restrict__ pzi = (gcchar *)slot_field(d, in_slot, 1);
        gcchar *__restrict__ vxi = (gcchar *)slot_field(d, in_slot, 2);
        gcchar *__restrict__ vzi = (gcchar *)slot_field(d, in_slot, 3);
        gchar *__restrict__ pxo = (gchar *)slot_field(d, out_slot, 0);
        gchar *__restrict__ pzo = (gchar *)slot_field(d, out_slot, 1);
        gchar *__restrict__ vxo = (gchar *)slot_field(d, out_slot, 2);
        gchar *__restrict__ vzo = (gchar *)slot_field(d, out_slot, 3);
        gcchar *__restrict__ KX = (gcchar *)d.KX;
        gcchar *__restrict__ KZ = (gcchar *)d.KZ;
        gcchar *__restrict__ BX = (gcchar *)d.BX;
        gcchar *__restrict__ BZ = (gcchar *)d.BZ;
        const SAM2D_GLOBAL float *__restrict__ prx = gp(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);
        SAM2D_GLOBAL float *__restrict__ band0 = band_slice(d, band_w, 0);
        SAM2D_GLOBAL float *__restrict__ band1 = band_slice(d, band_w, 1);

        auto ld = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        auto stf = [](gchar *__restrict__ base, unsigned off, float val) {
            *(SAM2D_GLOBAL float *__restrict__)(base + off) = val;
        };
        auto coff = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 4u + rc4;
        };
        auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };

        const float daz = prz[rc], skz = prz[nnz + rc];
        const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
        float zm0, zm1, zp0, zp1;
        dminus_coef(r, nnz, zm0, zm1);
        dplus_coef(r, nnz, zp0, zp1);
        const bool src_row = (r == d.src_z);
        // Source increments of the three steps (injection runs while
        // step <= nts; the increment past the wavelet's end is 0).
        const float amp1 = (step < d.nts) ? gp(d.amp)[step] : 0.0f;
        const int src_x1 = (step <= d.nts) ? d.src_x : -1;
        const float amp2 = (step + 1 < d.nts) ? gp(d.amp)[step + 1] : 0.0f;
        const int src_x2 = (step + 1 <= d.nts) ? d.src_x : -1;
        const float amp3 = (step + 2 < d.nts) ? gp(d.amp)[step + 2] : 0.0f;
        const int src_x3 = (step + 2 <= d.nts) ? d.src_x : -1;

        const bool strip_band = d.rrow_lo <= d.rrow_hi && d.rrow_lo <= zb + OWN_Z3 - 1 &&
                                d.rrow_hi >= zb;
        const bool tile_band0 = (band_mask & 1) && strip_band;
        const bool tile_band1 = (band_mask & 2) && strip_band;
        const bool band_lane = rown && r >= d.rrow_lo && r <= d.rrow_hi;
        // Fast groups start at c0 in [fast_lo, fast_hi]: columns c0-8 .. c0+3
        // inside the all-ones x-profile range and the 4th-order stencil range
        // (c0 >= 9, c0 <= nnx-5).
        const int fast_lo = max(d.xin_lo + 8, 9);
        const int fast_hi = min(d.xin_hi - 3, nnx - 5);

        auto load_col = [&](int c) -> Col3 {
            Col3 o;
            const unsigned o1 = coff(c + 1), o0 = coff(c), oj = coff(c - 2);
            o.vx_next = ld(vxi, o1);
            o.vz = ld(vzi, o0); o.px = ld(pxi, o0); o.pz = ld(pzi, o0);
            o.kx = ld(KX, o0);
            o.kz = KZA ? 0.0f : ld(KZ, o0);
            o.bx = ld(BX, oj); o.bz = ld(BZ, oj); o.vz_j = ld(vzi, oj);
            return o;
        };

        // Windows, [k & 3] relative to the unrolled base (the columns held):
        //   vx0 c-2..c+1 | p1 px1 pz1 kx1 kz1 c-3..c | vx1 vz1 bx1 bz1 c-5..c-2
        //   p2 px2 pz2 kx2 kz2 c-6..c-3 | vx2 vz2 bx2 bz2 c-8..c-5 | p3 c-9..c-6
        float vx0[4], p1[4], px1[4], pz1[4], kx1[4], kz1[4];
        float vx1[4], vz1[4], bx1[4], bz1[4];
        float p2[4], px2[4], pz2[4], kx2[4], kz2[4];
        float vx2[4], vz2[4], bx2[4], bz2[4], p3[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            vx0[k] = p1[k] = px1[k] = pz1[k] = kx1[k] = kz1[k] = 0.0f;
            vx1[k] = vz1[k] = bx1[k] = bz1[k] = 0.0f;
            p2[k] = px2[k] = pz2[k] = kx2[k] = kz2[k] = 0.0f;
            vx2[k] = vz2[k] = bx2[k] = bz2[k] = p3[k] = 0.0f;
        }
        const int c_first = xa - 7;
        vx0[0] = ld(vxi, coff(c_first - 2));
        vx0[1] = ld(vxi, coff(c_first - 1));
        vx0[2] = ld(vxi, coff(c_first));

        Col3 q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(c_first + u);

        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            Col3 &cur = q[u % PF];

            // ---- A: step-n p-phase at column c
            vx0[i3] = keep(cur.vx_next);
            {
                float m0 = SAM2D_A0, m1 = SAM2D_A1;
                if (!FAST) dminus_coef(c, nnx, m0, m1);
                const float dvx = fmaf(m0, vx0[i2] - vx0[i1], m1 * (vx0[i3] - vx0[i0]));
                const float vz_0 = cur.vz;
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                const float kx = keep(cur.kx);
                const float kz = KZA ? kx : keep(cur.kz);
                float pxn;
                if (FAST) {
                    pxn = fmaf(kx, dvx, cur.px);
                } else {
                    const int cc = clampc(c);
                    pxn = fmaf(kx * prx[nnx + cc], dvx, prx[cc] * cur.px);
                }
                if (src_row && c == src_x1) pxn += amp1;
                const float pzn = fmaf(kz * skz, dvz, daz * cur.pz);
                px1[i3] = pxn;
                pz1[i3] = pzn;
                p1[i3] = pxn + pzn;
                kx1[i3] = kx;
                if (!KZA) kz1[i3] = kz;
                if (tile_band0 && c >= xa && c < xb) {
                    if (band_lane) band0[(size_t)r * nnx + c] = p1[i3];
                }
            }

            // ---- B: step-n v-phase at column j1 = c-2
            {
                const int j1 = c - 2;
                float q0 = SAM2D_A0, q1 = SAM2D_A1;
                if (!FAST) dplus_coef(j1, nnx, q0, q1);
                const float pm3 = p1[i0], pm2 = p1[i1], pm1 = p1[i2], pn = p1[i3];
                const float dpx = fmaf(q0, pm1 - pm2, q1 * (pn - pm3));
                const float p_u1 = lane_up1(pm2);
                const float p_d1 = lane_dn1(pm2), p_d2 = lane_dn1(p_d1);
                const float dpz = fmaf(zp0, p_d1 - pm2, zp1 * (p_d2 - p_u1));
                const float bx = keep(cur.bx), bz = keep(cur.bz);
                if (FAST) {
                    vx1[i3] = fmaf(bx, dpx, vx0[i0]);
                } else {
                    const int jc = clampc(j1);
                    vx1[i3] = fmaf(bx * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx0[i0]);
                }
                vz1[i3] = fmaf(bz * sbz, dpz, dbz * cur.vz_j);
                bx1[i3] = bx;
                bz1[i3] = bz;
            }

            // ---- C: step-(n+1) p-phase at column c2 = c-3
            {
                const int c2 = c - 3;
                float m0 = SAM2D_A0, m1 = SAM2D_A1;
                if (!FAST) dminus_coef(c2, nnx, m0, m1);
                const float dvx = fmaf(m0, vx1[i2] - vx1[i1], m1 * (vx1[i3] - vx1[i0]));
                const float vz_0 = vz1[i2];
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                const float kx = kx1[i0];
                const float kz = KZA ? kx : kz1[i0];
                float pxn;
                if (FAST) {
                    pxn = fmaf(kx, dvx, px1[i0]);
                } else {
                    const int cc = clampc(c2);
                    pxn = fmaf(kx * prx[nnx + cc], dvx, prx[cc] * px1[i0]);
                }
                if (src_row && c2 == src_x2) pxn += amp2;
                const float pzn = fmaf(kz * skz, dvz, daz * pz1[i0]);
                px2[i3] = pxn;
                pz2[i3] = pzn;
                p2[i3] = pxn + pzn;
                kx2[i3] = kx;
                if (!KZA) kz2[i3] = kz;
                if (tile_band1 && c2 >= xa && c2 < xb) {
                    if (band_lane) band1[(size_t)r * nnx + c2] = p2[i3];
                }
            }

            // ---- D: step-(n+1) v-phase at column j2 = c-5
            {
                const int j2 = c - 5;
                float q0 = SAM2D_A0, q1 = SAM2D_A1;
                if (!FAST) dplus_coef(j2, nnx, q0, q1);
                const float pm3 = p2[i0], pm2 = p2[i1], pm1 = p2[i2], pn = p2[i3];
                const float dpx = fmaf(q0, pm1 - pm2, q1 * (pn - pm3));
                const float p_u1 = lane_up1(pm2);
                const float p_d1 = lane_dn1(pm2), p_d2 = lane_dn1(p_d1);
                const float dpz = fmaf(zp0, p_d1 - pm2, zp1 * (p_d2 - p_u1));
                const float bx = bx1[i0], bz = bz1[i0];
                if (FAST) {
                    vx2[i3] = fmaf(bx, dpx, vx1[i0]);
                } else {
                    const int jc = clampc(j2);
                    vx2[i3] = fmaf(bx * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx1[i0]);
                }
                vz2[i3] = fmaf(bz * sbz, dpz, dbz * vz1[i0]);
                bx2[i3] = bx;
                bz2[i3] = bz;
            }

            // ---- E: step-(n+2) p-phase at column c3 = c-6 (stored)
            {
                const int c3 = c - 6;
                float m0 = SAM2D_A0, m1 = SAM2D_A1;
                if (!FAST) dminus_coef(c3, nnx, m0, m1);
                const float dvx = fmaf(m0, vx2[i2] - vx2[i1], m1 * (vx2[i3] - vx2[i0]));
                const float vz_0 = vz2[i2];
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                const float kx = kx2[i0];
                const float kz = KZA ? kx : kz2[i0];
                float pxn;
                if (FAST) {
                    pxn = fmaf(kx, dvx, px2[i0]);
                } else {
                    const int cc = clampc(c3);
                    pxn = fmaf(kx * prx[nnx + cc], dvx, prx[cc] * px2[i0]);
                }
                if (src_row && c3 == src_x3) pxn += amp3;
                const float pzn = fmaf(kz * skz, dvz, daz * pz2[i0]);
                p3[i3] = pxn + pzn;
                if (rown && c3 >= xa && c3 < xb) {
                    const unsigned id = (unsigned)c3 * (unsigned)nnz * 4u + rc4;
                    stf(pxo, id, pxn);
                    stf(pzo, id, pzn);
                }
            }

            // ---- F: step-(n+2) v-phase at column j3 = c-8 (stored)
            {
                const int j3 = c - 8;
                float q0 = SAM2D_A0, q1 = SAM2D_A1;
                if (!FAST) dplus_coef(j3, nnx, q0, q1);
                const float pm3 = p3[i0], pm2 = p3[i1], pm1 = p3[i2], pn = p3[i3];
                const float dpx = fmaf(q0, pm1 - pm2, q1 * (pn - pm3));
                const float p_u1 = lane_up1(pm2);
                const float p_d1 = lane_dn1(pm2), p_d2 = lane_dn1(p_d1);
                const float dpz = fmaf(zp0, p_d1 - pm2, zp1 * (p_d2 - p_u1));
                if (rown && j3 >= xa && j3 < xb) {
                    float vxn;
                    if (FAST) {
                        vxn = fmaf(bx2[i0], dpx, vx2[i0]);
                    } else {
                        const int jc = clampc(j3);
                        vxn = fmaf(bx2[i0] * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx2[i0]);
                    }
                    const float vzn = fmaf(bz2[i0] * sbz, dpz, dbz * vz2[i0]);
                    const unsigned id = (unsigned)j3 * (unsigned)nnz * 4u + rc4;
                    stf(vxo, id, vxn);
                    stf(vzo, id, vzn);
                }
            }
            cur = load_col(c + PF);
        };

        using T = std::integral_constant<bool, true>;
        using F = std::integral_constant<bool, false>;
        using U0 = std::integral_constant<int, 0>;
        using U1 = std::integral_constant<int, 1>;
        using U2 = std::integral_constant<int, 2>;
        using U3 = std::integral_constant<int, 3>;
        // Slow groups, the fast range, slow groups (see k_step2).
        const int c_last = xb + 7;
        int c0 = c_first;
        for (; c0 <= c_last && c0 < fast_lo; c0 += 4) {
            column(F(), U0(), c0); column(F(), U1(), c0);
            column(F(), U2(), c0); column(F(), U3(), c0);
        }
        const int c_fast = min(fast_hi, c_last);
        for (; c0 <= c_fast; c0 += 4) {
            column(T(), U0(), c0); column(T(), U1(), c0);
            column(T(), U2(), c0); column(T(), U3(), c0);
        }
        for (; c0 <= c_last; c0 += 4) {
            column(F(), U0(), c0); column(F(), U1(), c0);
            column(F(), U2(), c0); column(F(), U3(), c0);
        }
    }
}

struct Step3Args {
    const ShotDev *d;
    int ns, vt, per, nt, step, in, out, bmask, bw, b0, b1, ss;
};

template <int W, int PF, bool KZA>
static void launch3_t(dim3 g, hipStream_t s, const Step3Args &a)
{
    hipLaunchKernelGGL((k_step3<W, PF, KZA>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt, a.per, a.nt,
                       a.step, a.in, a.out, a.bmask, a.bw, a.b0, a.b1, a.ss);
}

template <int W>
static void launch3_w(int pf, bool kza, dim3 g, hipStream_t s, const Step3Args &a)
{
    if (pf == 2 && kza) launch3_t<W, 2, true>(g, s, a);
    else if (pf == 2) launch3_t<W, 2, false>(g, s, a);
    else if (kza) launch3_t<W, 4, true>(g, s, a);
    else launch3_t<W, 4, false>(g, s, a);
}

hipError_t launch_step3(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                        bool kz_alias, int step, int in_slot, int out_slot, int band_mask,
                        int band_w, int band_s0, int band_s1, int state_sample, int max_ng,
                        hipStream_t stream)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const dim3 g(grid);
    if (max_ng <= 0) { band_s0 = band_s1 = state_sample = -1; band_mask = 0; }
    const Step3Args a{d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot,
                      band_mask, band_w, band_s0, band_s1, state_sample};
    switch (waves) {
    case 1: launch3_w<1>(pf, kz_alias, g, stream, a); break;
    case 2: launch3_w<2>(pf, kz_alias, g, stream, a); break;
    case 4: launch3_w<4>(pf, kz_alias, g, stream, a); break;
    case 8: launch3_w<8>(pf, kz_alias, g, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sam2d
