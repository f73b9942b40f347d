// sam2d_kernels2.hip -- the forward step advanced two time steps per pass
// (temporal blocking) for gfx950.
//
// Same arithmetic per cell and step as k_step in sam2d_kernels.hip (so the
// results are bit-identical), but one launch carries the state from step n-1
// to step n+1 and the state after step n never leaves the registers: HBM
// traffic per cell-update halves (44 bytes per two updates, plus halos).
//
// A wavefront owns OWN_Z2 = 52 rows with a six-row halo above and below (the
// dependency cone of the second step) and marches along x through four
// pipeline stages per column c:
//   A  step-n   p-phase at column c      (needs vx at c-2..c+1, vz at c)
//   B  step-n   v-phase at column c-2    (needs p(n) at c-3..c)
//   C  step-n+1 p-phase at column c-3    (needs vx(n) at c-5..c-2)
//   D  step-n+1 v-phase at column c-5    (needs p(n+1) at c-6..c-3)
// Every window holds four columns and rotates by compile-time index over a
// march unrolled by four.
//
// Receivers: the state after step n+1 is stored, so its receivers are
// gathered from it by the next launch; the intermediate pressure after step
// n is kept only for the rows that hold receivers (a band buffer, two
// parities), from which the next launch gathers as well.  No tile waits on a
// receiver lookup.
//
// Instruction economy (the kernel is issue- and latency-bound once its HBM
// traffic is halved):
//  * z-neighbours move between lanes with DPP wavefront shifts (one VALU
//    operation, often folded into the consumer) instead of LDS permutes;
//  * loads and stores address memory as a uniform base plus a 32-bit per-lane
//    byte offset, one offset per column shared by every field;
//  * each group of four columns takes a fast path when all of them lie where
//    the x sponge profiles are exactly 1 and the stencil has its 4th-order
//    coefficients: no profile loads, no coefficient selects.  Multiplying by
//    a profile value of exactly 1 is exact, so both paths give the same bits.

#include <hip/hip_runtime.h>
#include <type_traits>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

__global__ void k_gather(const ShotDev *__restrict__ shots, int nshots, int band_r, int band_s0,
                         int band_s1, int state_slot, int state_sample)
{
    gather_deferred(shots, nshots, band_r, band_s0, band_s1, state_slot, state_sample,
                    blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Operands of one column that the march prefetches (one row per lane).
struct Col2 {
    float vx_next;              // vx(n-1) at column c+1
    float vz, px, pz, kx, kz;   // at column c (kz unused when KZ aliases KX)
    float bx, bz, vz_j;         // BX, BZ, vz(n-1) at column c-2
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_step2(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd, int ntiles,
             int step, int in_slot, int out_slot, int band_store, int band_w, int band_sample,
             int state_sample)
{
    static_assert(PF == 2 || PF == 4, "PF must divide the unroll of 4");
    using gcchar = const SAM2D_GLOBAL char;
    using gchar = SAM2D_GLOBAL char;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    if (band_sample >= 0 || state_sample >= 0)
        gather_deferred(shots, nshots, band_w ^ 1, band_sample, -1, in_slot, state_sample,
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
        const int zb = (group * WAVES + wv) * OWN_Z2;
        if (zb >= nnz) continue;                      // wave-uniform
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);

        const int r = zb - HALO_T2 + lane;
        const bool rown = (lane >= HALO_T2) && (lane < HALO_T2 + OWN_Z2) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
        const unsigned rc4 = rc * 4u;

        // Inputs, outputs and coefficients never overlap (restrict lets the
        // prefetch of later columns move ahead of this column's stores).
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
        SAM2D_GLOBAL float *__restrict__ band = band_slice(d, band_w, 0);

        auto ld = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        auto stf = [](gchar *__restrict__ base, unsigned off, float val) {
            *(SAM2D_GLOBAL float *__restrict__)(base + off) = val;
        };
        // Byte offset of this lane's row in column col (clamped to the grid).
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
        const float amp1 = (step < d.nts) ? gp(d.amp)[step] : 0.0f;
        const int src_x1 = (step <= d.nts) ? d.src_x : -1;
        const float amp2 = (step + 1 < d.nts) ? gp(d.amp)[step + 1] : 0.0f;
        const int src_x2 = (step + 1 <= d.nts) ? d.src_x : -1;

        // Does this strip own receiver rows, and does this launch keep the
        // intermediate pressure for them?  (uniform)
        const bool tile_band = band_store && d.rrow_lo <= d.rrow_hi &&
                               d.rrow_lo <= zb + OWN_Z2 - 1 && d.rrow_hi >= zb;
        const bool band_lane = rown && r >= d.rrow_lo && r <= d.rrow_hi;
        // Fast groups of four columns start at c0 in [fast_lo, fast_hi]:
        // columns c0-5 .. c0+3 inside the all-ones x-profile range and the
        // 4th-order stencil range (c0 >= 6, c0 <= nnx-5).
        const int fast_lo = max(d.xin_lo + 5, 6);
        const int fast_hi = min(d.xin_hi - 3, nnx - 5);

        auto load_col = [&](int c) -> Col2 {
            Col2 o;
            const unsigned o1 = coff(c + 1), o0 = coff(c), oj = coff(c - 2);
            o.vx_next = ld(vxi, o1);
            o.vz = ld(vzi, o0); o.px = ld(pxi, o0); o.pz = ld(pzi, o0);
            o.kx = ld(KX, o0);
            o.kz = KZA ? 0.0f : ld(KZ, o0);
            o.bx = ld(BX, oj); o.bz = ld(BZ, oj); o.vz_j = ld(vzi, oj);
            return o;
        };
        // Windows, [k & 3] relative to the unrolled base: vx0 holds columns
        // c-2..c+1; p1, px1, pz1, kx, kz hold c-3..c; vx1, vz1, bx, bz hold
        // c-5..c-2; p2 holds c-6..c-3.
        float vx0[4], p1[4], px1[4], pz1[4], kxw[4], kzw[4];
        float vx1[4], vz1[4], bxw[4], bzw[4], p2[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            vx0[k] = p1[k] = px1[k] = pz1[k] = kxw[k] = kzw[k] = 0.0f;
            vx1[k] = vz1[k] = bxw[k] = bzw[k] = p2[k] = 0.0f;
        }
        const int c_first = xa - 4;
        vx0[0] = ld(vxi, coff(c_first - 2));
        vx0[1] = ld(vxi, coff(c_first - 1));
        vx0[2] = ld(vxi, coff(c_first));

        Col2 q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(c_first + u);

        // One column of the four-stage pipeline; FAST as described above.
        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            Col2 &cur = q[u % PF];
            const float kzc = KZA ? cur.kx : cur.kz;

            // ---- A: step-n p-phase at column c
            vx0[i3] = keep(cur.vx_next);
            {
                float xm0 = SAM2D_A0, xm1 = SAM2D_A1;
                if (!FAST) dminus_coef(c, nnx, xm0, xm1);
                const float dvx = fmaf(xm0, vx0[i2] - vx0[i1], xm1 * (vx0[i3] - vx0[i0]));
                const float vz_0 = cur.vz;
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                float pxn;
                if (FAST) {
                    pxn = fmaf(cur.kx, dvx, cur.px);
                } else {
                    const int cc = clampc(c);
                    pxn = fmaf(cur.kx * prx[nnx + cc], dvx, prx[cc] * cur.px);
                }
                if (src_row && c == src_x1) pxn += amp1;
                const float pzn = fmaf(kzc * skz, dvz, daz * cur.pz);
                px1[i3] = pxn;
                pz1[i3] = pzn;
                p1[i3] = pxn + pzn;
                kxw[i3] = keep(cur.kx);
                if (!KZA) kzw[i3] = keep(cur.kz);
            }
            if (tile_band && c >= xa && c < xb) {
                if (band_lane) band[(size_t)r * nnx + c] = p1[i3];
            }

            // ---- B: step-n v-phase at column j1 = c-2
            {
                const int j1 = c - 2;
                float xp0 = SAM2D_A0, xp1 = SAM2D_A1;
                if (!FAST) dplus_coef(j1, nnx, xp0, xp1);
                const float pm3 = p1[i0], pm2 = p1[i1], pm1 = p1[i2], pn = p1[i3];
                const float dpx = fmaf(xp0, pm1 - pm2, xp1 * (pn - pm3));
                const float q_u1 = lane_up1(pm2);
                const float q_d1 = lane_dn1(pm2), q_d2 = lane_dn1(q_d1);
                const float dpz = fmaf(zp0, q_d1 - pm2, zp1 * (q_d2 - q_u1));
                if (FAST) {
                    vx1[i3] = fmaf(cur.bx, dpx, vx0[i0]);
                } else {
                    const int jc = clampc(j1);
                    vx1[i3] = fmaf(cur.bx * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx0[i0]);
                }
                vz1[i3] = fmaf(cur.bz * sbz, dpz, dbz * cur.vz_j);
                bxw[i3] = keep(cur.bx);
                bzw[i3] = keep(cur.bz);
            }

            // ---- C: step-(n+1) p-phase at column c2 = c-3
            {
                const int c2 = c - 3;
                float ym0 = SAM2D_A0, ym1 = SAM2D_A1;
                if (!FAST) dminus_coef(c2, nnx, ym0, ym1);
                // vx(n) at c2-2..c2+1 = c-5..c-2 -> vx1[i0..i3]
                const float dvx = fmaf(ym0, vx1[i2] - vx1[i1], ym1 * (vx1[i3] - vx1[i0]));
                const float vz_0 = vz1[i2];   // column c-3
                const float vz_u1 = lane_up1(vz_0), vz_u2 = lane_up1(vz_u1);
                const float vz_d1 = lane_dn1(vz_0);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                // px(n), pz(n), KX, KZ at column c-3 -> index i0 of the c-3..c windows
                float pxn;
                if (FAST) {
                    pxn = fmaf(kxw[i0], dvx, px1[i0]);
                } else {
                    const int cc = clampc(c2);
                    pxn = fmaf(kxw[i0] * prx[nnx + cc], dvx, prx[cc] * px1[i0]);
                }
                if (src_row && c2 == src_x2) pxn += amp2;
                const float pzn = fmaf((KZA ? kxw[i0] : kzw[i0]) * skz, dvz, daz * pz1[i0]);
                const float pn = pxn + pzn;
                p2[i3] = pn;
                if (rown && c2 >= xa && c2 < xb) {
                    const unsigned id = (unsigned)c2 * (unsigned)nnz * 4u + rc4;
                    stf(pxo, id, pxn);
                    stf(pzo, id, pzn);
                }
            }

            // ---- D: step-(n+1) v-phase at column j2 = c-5
            {
                const int j2 = c - 5;
                float xp0 = SAM2D_A0, xp1 = SAM2D_A1;
                if (!FAST) dplus_coef(j2, nnx, xp0, xp1);
                const float pm3 = p2[i0], pm2 = p2[i1], pm1 = p2[i2], pn = p2[i3];
                const float dpx = fmaf(xp0, pm1 - pm2, xp1 * (pn - pm3));
                const float q_u1 = lane_up1(pm2);
                const float q_d1 = lane_dn1(pm2), q_d2 = lane_dn1(q_d1);
                const float dpz = fmaf(zp0, q_d1 - pm2, zp1 * (q_d2 - q_u1));
                if (rown && j2 >= xa && j2 < xb) {
                    // vx(n), vz(n), BX, BZ at column c-5 -> index i0 of the c-5..c-2 windows
                    float vxn;
                    if (FAST) {
                        vxn = fmaf(bxw[i0], dpx, vx1[i0]);
                    } else {
                        const int jc = clampc(j2);
                        vxn = fmaf(bxw[i0] * prx[3 * nnx + jc], dpx, prx[2 * nnx + jc] * vx1[i0]);
                    }
                    const float vzn = fmaf(bzw[i0] * sbz, dpz, dbz * vz1[i0]);
                    const unsigned id = (unsigned)j2 * (unsigned)nnz * 4u + rc4;
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
        // Three loops in sequence: slow groups before the fast range, the
        // fast range, slow groups after it.  A loop of one kind only lets the
        // register allocator keep every prefetched operand in place across
        // the back edge, so no wait on a load in flight is forced there.
        const int c_last = xb + 4;
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

struct Step2Args {
    const ShotDev *d;
    int ns, vt, per, nt, step, in, out, bstore, bw, bs, ss;
};

template <int W, int PF, bool KZA>
static void launch2_t(dim3 g, hipStream_t s, const Step2Args &a)
{
    hipLaunchKernelGGL((k_step2<W, PF, KZA>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt, a.per, a.nt,
                       a.step, a.in, a.out, a.bstore, a.bw, a.bs, a.ss);
}

template <int W>
static void launch2_w(int pf, bool kza, dim3 g, hipStream_t s, const Step2Args &a)
{
    if (pf == 2 && kza) launch2_t<W, 2, true>(g, s, a);
    else if (pf == 2) launch2_t<W, 2, false>(g, s, a);
    else if (kza) launch2_t<W, 4, true>(g, s, a);
    else launch2_t<W, 4, false>(g, s, a);
}

hipError_t launch_step2(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                        bool kz_alias, int step, int in_slot, int out_slot, bool band_store,
                        int band_w, int band_sample, int state_sample, int max_ng,
                        hipStream_t stream)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const dim3 g(grid);
    if (max_ng <= 0) { band_sample = -1; state_sample = -1; band_store = false; }
    const Step2Args a{d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot,
                      band_store ? 1 : 0, band_w, band_sample, state_sample};
    switch (waves) {
    case 1: launch2_w<1>(pf, kz_alias, g, stream, a); break;
    case 2: launch2_w<2>(pf, kz_alias, g, stream, a); break;
    case 4: launch2_w<4>(pf, kz_alias, g, stream, a); break;
    case 8: launch2_w<8>(pf, kz_alias, g, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gather(const ShotDev *d_shots, int nshots, int max_ng, int band_r, int band_s0,
                         int band_s1, int state_slot, int state_sample, hipStream_t stream)
{
    if (max_ng <= 0 || (band_s0 < 0 && band_s1 < 0 && state_sample < 0)) return hipSuccess;
    const int blocks = (max_ng + 255) / 256;
    hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, stream, d_shots, nshots, band_r,
                       band_s0, band_s1, state_slot, state_sample);
    return hipGetLastError();
}

}  // namespace sam2d
