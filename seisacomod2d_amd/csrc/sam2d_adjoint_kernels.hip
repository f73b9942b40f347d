// sam2d_adjoint_kernels.hip -- gfx950 kernels of the backward (adjoint) steps
// of the staggered-grid scheme in sam2d_kernels.hip, with the FWI gradient
// accumulation fused in: k_adjoint (one step per pass), k_adjoint2 (two
// steps per pass, the same bits), the residual band builder k_band and the
// fallback injection k_inject.
//
// The reference has no adjoint (job_t.adj is declared and unused,
// src/src_cpu/modelling.h:51).  This is the exact transpose of the FP32
// forward step, run backward in time.  With lambda the adjoint of the state
// after step n (L_px, L_pz, L_vx, L_vz) and r_n the residual injected where
// step n was recorded:
//
//   L_px += r_n, L_pz += r_n at the receiver cells               (record)
//   L_p'  = Dx+^T (BXe L_vx) + Dz+^T (BZe L_vz)                   (v-update)
//   L_px' = L_px + L_p',   L_pz' = L_pz + L_p'                   (p = px + pz)
//   out:  L_px = dax L_px',  L_pz = daz L_pz'                     (p sponge)
//         L_vx = dbx L_vx + Dx-^T (KXe L_px')                     (v sponge, p-update)
//         L_vz = dbz L_vz + Dz-^T (KZe L_pz')
//   dJ/dkappa += L_px' (Dx- vx[n-1]) dtdx skx + L_pz' (Dz- vz[n-1]) dtdz skz
//                (+ L_px'[src] fd_dt w[n]/(dx dz) at the source)
//   dJ/dbu2   += L_vx (Dx+ p[n]) dtdx sbx,   dJ/dbu1 += L_vz (Dz+ p[n]) dtdz sbz
//
// where KXe = KX*skx etc. are the forward's effective coefficients.  The
// transposes are written as gathers: Dx+^T g at column c is
// c0(c-1) g[c-1] - c0(c) g[c] + c1(c-2) g[c-2] - c1(c+1) g[c+1] with the
// forward's per-index coefficients (zero outside the grid), so the edge rules
// transpose exactly.
//
// The tiling, the halo recomputation, the x-march with mod-4 register windows
// and the clamped loads follow the forward kernel: phase A (the L_p' part)
// runs at column c, phase B (the velocity part) at column j = c - 2.  As in
// the multi-step forward kernel, z-neighbours move by DPP lane shifts, memory
// is addressed as a uniform base plus a 32-bit per-lane byte offset, the
// prefetch slots are emptied through keep(), and groups of four columns in
// the all-ones x-profile range run a fast path without profile loads.  The
// residuals come from a dense band over the receiver rows, prefetched with the
// other operands of a column, so no tile waits on a receiver lookup; shots
// whose band would be too large are injected by k_inject beforehand.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

// Every product and sum is rounded as written, so the one-step and two-step
// adjoints, which form the same terms in different places, give the same bits.
#pragma clang fp contract(off)

// Residual injection before backward step n, for shots without a residual
// band: for each distinct receiver cell,
// the residuals of its receivers (sample rec_k, receiver order) are summed and
// added to L_px and L_pz of the adjoint state slot `lam` (the transpose of
// recording p = px + pz there, cpu_record, cpu_modelling_kernels.cpp:24-31).
__global__ void k_inject(const AdjDev *__restrict__ shots, int nshots, int lam, int rec_k)
{
    const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const int gstride = gridDim.x * blockDim.x;
    for (int s = 0; s < nshots; ++s) {
        const AdjDev &d = shots[s];
        if (d.rband) continue;                        // injected inside k_adjoint
        // cell-interleaved adjoint state: L_px, L_pz of cell c at 4c, 4c+1
        SAM2D_GLOBAL float *L = gp(d.lam + 4 * (size_t)lam * d.ncell);
        for (int u = gtid; u < d.ncells; u += gstride) {
            float sum = 0.0f;
            for (int e = d.rptr[u]; e < d.rptr[u + 1]; e++)
                sum += gp(d.res)[(size_t)d.rid[e] * d.real_nt + rec_k];
            const size_t c = 4 * (size_t)d.rcell[u];
            L[c] += sum;
            L[c + 1] += sum;
        }
    }
}

// One element of the residual band per thread, consecutive threads along a
// band row (coalesced stores).
__global__ void k_band(const int *__restrict__ bmap, const int *__restrict__ rptr,
                       const int *__restrict__ rid, const float *__restrict__ res, int real_nt,
                       int band_cells, float *__restrict__ band)
{
    const size_t total = (size_t)real_nt * band_cells;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int k = (int)(i / band_cells);
        const int u = bmap[i - (size_t)k * band_cells];
        float v = -0.0f;
        if (u >= 0)
            for (int e = rptr[u]; e < rptr[u + 1]; e++)
                v += res[(size_t)rid[e] * real_nt + k];
        band[i] = v;
    }
}

hipError_t launch_band(const int *bmap, const int *rptr, const int *rid, const float *res,
                       int real_nt, int band_cells, float *band, hipStream_t stream)
{
    const size_t total = (size_t)real_nt * band_cells;
    if (total == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_band, dim3((unsigned)blocks), dim3(256), 0, stream, bmap, rptr, rid, res,
                       real_nt, band_cells, band);
    return hipGetLastError();
}

// The adjoint state is cell-interleaved (L_px, L_pz, L_vx, L_vz of cell i at
// 4i .. 4i+3): a column's state arrives as one 16-byte load per lane at
// column c+1 (kept whole in the prefetch slot) and leaves as one 16-byte
// store at the phase-B column, L_px and L_pz kept from phase A.
struct AdjCol {
    fquad lam;                  // column c+1: adjoint state
    float bx_n, fvx_n;          //             BX, fwd vx[n-1]
    float bz;                   // column c:   BZ
    float fvz, fpx, fpz;        //             fwd vz[n-1], px[n], pz[n]
    float kx, kz, gk;           //             KX, KZ (unused when aliased), dJ/dkappa
    float res;                  //             residual of this step (band rows)
    float gbx_j, gbz_j;         // column j:   dJ/dbu2, dJ/dbu1
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_adjoint(const AdjDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
               int ntiles, int step, int lam_in, int lam_out, int fwd_prev, int fwd_cur,
               int rec_k, int psum, int ngroups, int xorder)
{
    static_assert(PF == 2 || PF == 4, "PF must divide the unroll of 4");
    using gcchar = const SAM2D_GLOBAL char;
    using gchar = SAM2D_GLOBAL char;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile_ordered(v, per_xcd, ngroups, xorder);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const AdjDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int nnz = d.nnz, nnx = d.nnx;
        const int zb = (group * WAVES + wv) * OWN_Z;
        if (zb >= nnz) continue;                      // wave-uniform
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);
        // Columns up to c_last: the stores need j = c - 2 in [xa, xb).
        const int c_last = xb + 1;

        const int r = zb - HALO_T + lane;
        const bool rown = (lane >= HALO_T) && (lane < HALO_T + OWN_Z) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
        const unsigned rc4 = rc * 4u;
        const size_t N = d.ncell;

        // Adjoint state, forward ring slots, coefficients and accumulators
        // never overlap.
        gcchar *__restrict__ Lin = (gcchar *)(d.lam + 4 * (size_t)lam_in * N);
        gchar *__restrict__ Lout = (gchar *)(d.lam + 4 * (size_t)lam_out * N);
        // psum & 1: slot fwd_cur holds p in its px field and no pz; the pz
        // loads then read the px field again (a cache hit) and are not used.
        const bool ps_cur = psum & 1;
        gcchar *__restrict__ Fpx = (gcchar *)(d.fwd + (4 * (size_t)fwd_cur + 0) * N);
        gcchar *__restrict__ Fpz = (gcchar *)(d.fwd + (4 * (size_t)fwd_cur + (ps_cur ? 0 : 1)) * N);
        gcchar *__restrict__ Fvx = (gcchar *)(d.fwd + (4 * (size_t)fwd_prev + 2) * N);
        gcchar *__restrict__ Fvz = (gcchar *)(d.fwd + (4 * (size_t)fwd_prev + 3) * N);
        gcchar *__restrict__ KX = (gcchar *)d.KX;
        gcchar *__restrict__ KZ = (gcchar *)d.KZ;
        gcchar *__restrict__ BX = (gcchar *)d.BX;
        gcchar *__restrict__ BZ = (gcchar *)d.BZ;
        // The accumulators are read at a column and written back at the same
        // cell by the same lane only.
        gchar *__restrict__ GK = (gchar *)d.gk;
        gchar *__restrict__ GBX = (gchar *)d.gbx;
        gchar *__restrict__ GBZ = (gchar *)d.gbz;
        const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);

        auto ld = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        // Adjoint state: byte offset 4x that of one field.
        auto ld4 = [](gcchar *__restrict__ base, unsigned off) -> fquad {
            return *(const SAM2D_GLOBAL fquad *__restrict__)(base + off);
        };
        auto st4 = [](gchar *__restrict__ base, unsigned off, float a, float b, float c, float e) {
            fquad v;
            v.x = a; v.y = b; v.z = c; v.w = e;
            *(SAM2D_GLOBAL fquad *__restrict__)(base + off) = v;
        };
        auto ldw = [](gchar *__restrict__ base, unsigned off) -> float {
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
        const float dtdx = d.dtdx, dtdz = d.dtdz;
        const bool src_lane = (r == d.src_z);
        const float srcg = (step < d.nts) ? gp(d.srcg)[step] : 0.0f;
        const int src_x = (step < d.nts) ? d.src_x : -1;
        // Fast groups start at c0 in [fast_lo, fast_hi]: columns c0-2 .. c0+4
        // in the all-ones x-profile range, 4th-order stencils at c0-2 .. c0+4.
        const int fast_lo = max(d.xin_lo + 2, 3);
        const int fast_hi = min(d.xin_hi - 4, nnx - 7);
        // Residual of this step at the receiver rows, prefetched with the
        // column (x + -0 = x, so cells without a receiver are unchanged).  Halo
        // rows take it too: phase B of an own row reads phase A of its halo.
        const bool tile_band = rec_k >= 0 && d.rband != nullptr &&
                               d.rrow_lo <= zb - HALO_T + 63 && d.rrow_hi >= zb - HALO_T;
        const bool band_lane = r >= d.rrow_lo && r <= d.rrow_hi;
        const SAM2D_GLOBAL float *__restrict__ rb =
            tile_band ? gp(d.rband) + ((ptrdiff_t)rec_k * d.band_h - d.rrow_lo) * (ptrdiff_t)nnx
                      : nullptr;

        auto load_col = [&](int c) -> AdjCol {
            AdjCol o;
            const unsigned o1 = coff(c + 1), o0 = coff(c), oj = coff(c - 2);
            o.lam = ld4(Lin, 4 * o1); o.bx_n = ld(BX, o1); o.fvx_n = ld(Fvx, o1);
            o.bz = ld(BZ, o0);
            o.fvz = ld(Fvz, o0); o.fpx = ld(Fpx, o0); o.fpz = ld(Fpz, o0);
            o.kx = ld(KX, o0);
            o.kz = KZA ? 0.0f : ld(KZ, o0);
            o.gk = ldw(GK, o0);
            o.gbx_j = ldw(GBX, oj); o.gbz_j = ldw(GBZ, oj);
            o.res = -0.0f;
            if (tile_band) {
                if (band_lane) o.res = rb[(size_t)r * nnx + clampc(c)];
            }
            return o;
        };
        // Per-column products of the transposed D+ (x): c0(i) gx(i), c1(i) gx(i)
        // with gx = BXe L_vx at column i.
        auto gx_terms = [&](int i, float lvx, float bx, float &a, float &b) {
            float p0, p1;
            dplus_coef(i, nnx, p0, p1);
            const float g = (bx * prx[3 * nnx + clampc(i)]) * lvx;
            a = p0 * g;
            b = p1 * g;
        };

        // vx-style windows: [k & 3] = column c-2+k (LVX, LPX, LPZ, LVZ: the
        // input adjoint state).  p-style: [k & 3] = column c-3+k (OPX, OPZ:
        // the new L_px, L_pz, stored with phase B).
        float LVX[4], LPX[4], LPZ[4], LVZ[4], FVX[4], AXP[4], BXP[4];
        float AM[4], BM[4], AMZ[4], BMZ[4], PN[4], OPX[4], OPZ[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            LVX[k] = LPX[k] = LPZ[k] = LVZ[k] = FVX[k] = AXP[k] = BXP[k] = 0.0f;
            AM[k] = BM[k] = AMZ[k] = BMZ[k] = PN[k] = OPX[k] = OPZ[k] = 0.0f;
        }
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const int col = xa - 3 + k;
            const unsigned o = coff(col);
            const fquad lv = ld4(Lin, 4 * o);
            LPX[k] = lv.x; LPZ[k] = lv.y; LVX[k] = lv.z; LVZ[k] = lv.w;
            FVX[k] = ld(Fvx, o);
            gx_terms(col, LVX[k], ld(BX, o), AXP[k], BXP[k]);
        }

        AdjCol q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(xa - 1 + u);

        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            AdjCol &cur = q[u % PF];

            // ---- phase A at column c
            LPX[i3] = keep(cur.lam.x);
            LPZ[i3] = keep(cur.lam.y);
            LVZ[i3] = keep(cur.lam.w);
            const float lvx_n = keep(cur.lam.z);
            LVX[i3] = lvx_n;
            FVX[i3] = keep(cur.fvx_n);
            if (FAST) {
                const float g = cur.bx_n * lvx_n;
                AXP[i3] = SAM2D_A0 * g;
                BXP[i3] = SAM2D_A1 * g;
            } else {
                gx_terms(c + 1, lvx_n, cur.bx_n, AXP[i3], BXP[i3]);
            }
            const float ax = AXP[i1] - AXP[i2] + (BXP[i0] - BXP[i3]);
            const float gz = (cur.bz * sbz) * LVZ[i2];
            const float az = zp0 * gz, bzv = zp1 * gz;
            const float bz_u1 = lane_up1(bzv);
            const float azz = (lane_up1(az) - az) + (lane_up1(bz_u1) - lane_dn1(bzv));
            const float lp = ax + azz;
            const float lpx1 = (LPX[i2] + cur.res) + lp;
            const float lpz1 = (LPZ[i2] + cur.res) + lp;
            float xm0 = SAM2D_A0, xm1 = SAM2D_A1;
            if (!FAST) dminus_coef(c, nnx, xm0, xm1);
            // dJ/dkappa from the p-update needs Dx- vx[n-1] and Dz- vz[n-1];
            // the lane shifts run with every lane active.
            const float f0 = cur.fvz;
            const float f_u1 = lane_up1(f0), f_u2 = lane_up1(f_u1), f_d1 = lane_dn1(f0);
            float skx = 1.0f, dax = 1.0f;
            if (!FAST) {
                const int cc = clampc(c);
                dax = prx[cc];
                skx = prx[nnx + cc];
            }
            if (rown && c >= xa && c < xb) {
                const float dvx = fmaf(xm0, FVX[i2] - FVX[i1], xm1 * (FVX[i3] - FVX[i0]));
                const float dvz = fmaf(zm0, f0 - f_u1, zm1 * (f_d1 - f_u2));
                float g = cur.gk + lpx1 * dvx * (dtdx * skx) + lpz1 * dvz * (dtdz * skz);
                if (src_lane && c == src_x) g += lpx1 * srcg;
                const unsigned id = (unsigned)c * (unsigned)nnz * 4u + rc4;
                stf(GK, id, g);
            }
            OPX[i3] = dax * lpx1;
            OPZ[i3] = daz * lpz1;
            const float kz = KZA ? cur.kx : cur.kz;
            const float hx = (cur.kx * skx) * lpx1;
            const float hz = (kz * skz) * lpz1;
            AM[i3] = xm0 * hx;
            BM[i3] = xm1 * hx;
            AMZ[i3] = zm0 * hz;
            BMZ[i3] = zm1 * hz;
            PN[i3] = ps_cur ? cur.fpx : cur.fpx + cur.fpz;

            // ---- phase B at column j = c - 2
            const int j = c - 2;
            const float tx = AM[i1] - AM[i2] + (BM[i0] - BM[i3]);
            const float amz = AMZ[i1], bmz = BMZ[i1];
            const float bm_d1 = lane_dn1(bmz);
            const float tz = (amz - lane_dn1(amz)) + (lane_up1(bmz) - lane_dn1(bm_d1));
            const float lvx_j = LVX[i0];
            const float pj = PN[i1];
            const float q_u1 = lane_up1(pj), q_d1 = lane_dn1(pj), q_d2 = lane_dn1(q_d1);
            if (rown && j >= xa && j < xb) {
                float xp0 = SAM2D_A0, xp1 = SAM2D_A1, dbx = 1.0f, sbx = 1.0f;
                if (!FAST) {
                    dplus_coef(j, nnx, xp0, xp1);
                    const int jc = clampc(j);
                    dbx = prx[2 * nnx + jc];
                    sbx = prx[3 * nnx + jc];
                }
                const float dpx = fmaf(xp0, PN[i2] - pj, xp1 * (PN[i3] - PN[i0]));
                const float dpz = fmaf(zp0, q_d1 - pj, zp1 * (q_d2 - q_u1));
                const unsigned id = (unsigned)j * (unsigned)nnz * 4u + rc4;
                const float lvz_j = LVZ[i0];
                st4(Lout, 4 * id, OPX[i1], OPZ[i1], fmaf(dbx, lvx_j, tx), fmaf(dbz, lvz_j, tz));
                stf(GBX, id, cur.gbx_j + lvx_j * dpx * (dtdx * sbx));
                stf(GBZ, id, cur.gbz_j + lvz_j * dpz * (dtdz * sbz));
            }
            cur = load_col(min(c + PF, c_last));   // past c_last: a cache hit
        };

        using TT = std::integral_constant<bool, true>;
        using FF = std::integral_constant<bool, false>;
        using U0 = std::integral_constant<int, 0>;
        using U1 = std::integral_constant<int, 1>;
        using U2 = std::integral_constant<int, 2>;
        using U3 = std::integral_constant<int, 3>;
        // Slow groups, the fast range, slow groups; the march ends at c_last.
        int c0 = xa - 1;
        for (; c0 + 3 <= c_last && c0 < fast_lo; c0 += 4) {
            column(FF(), U0(), c0); column(FF(), U1(), c0);
            column(FF(), U2(), c0); column(FF(), U3(), c0);
        }
        const int c_fast = min(fast_hi, c_last - 3);
        for (; c0 <= c_fast; c0 += 4) {
            column(TT(), U0(), c0); column(TT(), U1(), c0);
            column(TT(), U2(), c0); column(TT(), U3(), c0);
        }
        for (; c0 <= c_last; c0 += 4) {
            column(FF(), U0(), c0);
            if (c0 + 1 > c_last) break;
            column(FF(), U1(), c0);
            if (c0 + 2 > c_last) break;
            column(FF(), U2(), c0);
            if (c0 + 3 > c_last) break;
            column(FF(), U3(), c0);
        }
    }
}

// Two backward steps, n and n-1, per pass: the adjoint state after step n's
// adjoint never leaves the registers, and dJ/dkappa, dJ/dbu2, dJ/dbu1 are
// read and written once for both steps.  A wavefront owns OWN_Z2 rows with a
// HALO_T2-row halo on each side (the cone of two adjoint steps is that of two
// forward steps), and at column c runs
//   A1  step n   phase A at c        B1  step n   phase B at c-2
//   A2  step n-1 phase A at c-3      B2  step n-1 phase B at c-5
// with the same arithmetic per cell as k_adjoint, so both give the same bits.
// The x-transposes are formed at use from windows of the per-column factors
// (G = BXe L_vx, H = KXe L_px', HZ = KZe L_pz') and per-index coefficients.
// Forward states: vx, vz after step n-2 in ring slot fa, after step n-1 in fb;
// px, pz after step n-1 in fb, after step n in fc.
struct AdjCol2 {
    fquad lam;                  // column c+1: adjoint state (cell-interleaved)
    float bx_n, fvx_n;          //             BX, fwd vx[n-1]
    float bz;                   // column c:   BZ
    float fvz, fpx, fpz;        //             fwd vz[n-1], px[n], pz[n]
    float kx, kz, gk, res1;     //             KX, KZ, dJ/dkappa, residual of step n
    float gbx_j, gbz_j;         // column c-2: dJ/dbu2, dJ/dbu1
    float fvx2;                 //             fwd vx[n-2]
    float fvz2, fpx2, fpz2;     // column c-3: fwd vz[n-2], px[n-1], pz[n-1]
    float res2;                 //             residual of step n-1
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_adjoint2(const AdjDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
                int ntiles, int step, int lam_in, int lam_out, int fa, int fb, int fc,
                int rec1, int rec2, int psum, int ngroups, int xorder)
{
    static_assert(PF == 2 || PF == 4, "PF must divide the unroll of 4");
    using gcchar = const SAM2D_GLOBAL char;
    using gchar = SAM2D_GLOBAL char;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile_ordered(v, per_xcd, ngroups, xorder);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const AdjDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int nnz = d.nnz, nnx = d.nnx;
        const int zb = (group * WAVES + wv) * OWN_Z2;
        if (zb >= nnz) continue;                      // wave-uniform
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);
        // Columns up to c_last: the B-phase stores j = c - 5 in [xa, xb).
        const int c_last = xb + 4;

        const int r = zb - HALO_T2 + lane;
        const bool rown = (lane >= HALO_T2) && (lane < HALO_T2 + OWN_Z2) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
        const unsigned rc4 = rc * 4u;
        const size_t N = d.ncell;

        gcchar *__restrict__ Lin = (gcchar *)(d.lam + 4 * (size_t)lam_in * N);
        gchar *__restrict__ Lout = (gchar *)(d.lam + 4 * (size_t)lam_out * N);
        // psum bit 0 (slot fc) and bit 1 (slot fb): the slot holds p in its px
        // field and no pz; the pz loads then read the px field again.
        const bool ps_c = psum & 1, ps_b = (psum >> 1) & 1;
        gcchar *__restrict__ Fpx1 = (gcchar *)(d.fwd + (4 * (size_t)fc + 0) * N);
        gcchar *__restrict__ Fpz1 = (gcchar *)(d.fwd + (4 * (size_t)fc + (ps_c ? 0 : 1)) * N);
        gcchar *__restrict__ Fvx1 = (gcchar *)(d.fwd + (4 * (size_t)fb + 2) * N);
        gcchar *__restrict__ Fvz1 = (gcchar *)(d.fwd + (4 * (size_t)fb + 3) * N);
        gcchar *__restrict__ Fpx2 = (gcchar *)(d.fwd + (4 * (size_t)fb + 0) * N);
        gcchar *__restrict__ Fpz2 = (gcchar *)(d.fwd + (4 * (size_t)fb + (ps_b ? 0 : 1)) * N);
        gcchar *__restrict__ Fvx2 = (gcchar *)(d.fwd + (4 * (size_t)fa + 2) * N);
        gcchar *__restrict__ Fvz2 = (gcchar *)(d.fwd + (4 * (size_t)fa + 3) * N);
        gcchar *__restrict__ KX = (gcchar *)d.KX;
        gcchar *__restrict__ KZ = (gcchar *)d.KZ;
        gcchar *__restrict__ BX = (gcchar *)d.BX;
        gcchar *__restrict__ BZ = (gcchar *)d.BZ;
        gchar *__restrict__ GK = (gchar *)d.gk;
        gchar *__restrict__ GBX = (gchar *)d.gbx;
        gchar *__restrict__ GBZ = (gchar *)d.gbz;
        const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);

        auto ld = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        // Adjoint state: byte offset 4x that of one field.
        auto ld4 = [](gcchar *__restrict__ base, unsigned off) -> fquad {
            return *(const SAM2D_GLOBAL fquad *__restrict__)(base + off);
        };
        auto st4 = [](gchar *__restrict__ base, unsigned off, float a, float b, float c, float e) {
            fquad v;
            v.x = a; v.y = b; v.z = c; v.w = e;
            *(SAM2D_GLOBAL fquad *__restrict__)(base + off) = v;
        };
        auto ldw = [](gchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        auto stf = [](gchar *__restrict__ base, unsigned off, float val) {
            *(SAM2D_GLOBAL float *__restrict__)(base + off) = val;
        };
        auto coff = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 4u + rc4;
        };
        auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };
        auto p0_at = [&](int i) { float a, b; dplus_coef(i, nnx, a, b); return a; };
        auto p1_at = [&](int i) { float a, b; dplus_coef(i, nnx, a, b); return b; };
        auto m0_at = [&](int i) { float a, b; dminus_coef(i, nnx, a, b); return a; };
        auto m1_at = [&](int i) { float a, b; dminus_coef(i, nnx, a, b); return b; };

        const float daz = prz[rc], skz = prz[nnz + rc];
        const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
        float zm0, zm1, zp0, zp1;
        dminus_coef(r, nnz, zm0, zm1);
        dplus_coef(r, nnz, zp0, zp1);
        const float dtdx = d.dtdx, dtdz = d.dtdz;
        const bool src_lane = (r == d.src_z);
        const float srcg1 = (step < d.nts) ? gp(d.srcg)[step] : 0.0f;
        const int src_x1 = (step < d.nts) ? d.src_x : -1;
        const float srcg2 = (step - 1 < d.nts) ? gp(d.srcg)[step - 1] : 0.0f;
        const int src_x2 = (step - 1 < d.nts) ? d.src_x : -1;
        // Fast groups start at c0 in [fast_lo, fast_hi]: profiles read at
        // columns c0-5 .. c0+4 are all ones, D+ is 4th order on c0-5 .. c0+4
        // and D- on c0-6 .. c0+3.
        const int fast_lo = max(d.xin_lo + 6, 8);
        const int fast_hi = min(d.xin_hi - 4, nnx - 7);
        const bool strip_band = d.rband != nullptr && d.rrow_lo <= zb - HALO_T2 + 63 &&
                                d.rrow_hi >= zb - HALO_T2;
        const bool band1 = rec1 >= 0 && strip_band, band2 = rec2 >= 0 && strip_band;
        const bool band_lane = r >= d.rrow_lo && r <= d.rrow_hi;
        const SAM2D_GLOBAL float *__restrict__ rb1 =
            band1 ? gp(d.rband) + ((ptrdiff_t)rec1 * d.band_h - d.rrow_lo) * (ptrdiff_t)nnx
                  : nullptr;
        const SAM2D_GLOBAL float *__restrict__ rb2 =
            band2 ? gp(d.rband) + ((ptrdiff_t)rec2 * d.band_h - d.rrow_lo) * (ptrdiff_t)nnx
                  : nullptr;

        auto load_col = [&](int c) -> AdjCol2 {
            AdjCol2 o;
            const unsigned o1 = coff(c + 1), o0 = coff(c), oj = coff(c - 2), o3 = coff(c - 3);
            o.lam = ld4(Lin, 4 * o1); o.bx_n = ld(BX, o1); o.fvx_n = ld(Fvx1, o1);
            o.bz = ld(BZ, o0);
            o.fvz = ld(Fvz1, o0); o.fpx = ld(Fpx1, o0); o.fpz = ld(Fpz1, o0);
            o.kx = ld(KX, o0);
            o.kz = KZA ? 0.0f : ld(KZ, o0);
            o.gk = ldw(GK, o0);
            o.gbx_j = ldw(GBX, oj); o.gbz_j = ldw(GBZ, oj);
            o.fvx2 = ld(Fvx2, oj);
            o.fvz2 = ld(Fvz2, o3); o.fpx2 = ld(Fpx2, o3); o.fpz2 = ld(Fpz2, o3);
            o.res1 = -0.0f;
            o.res2 = -0.0f;
            if (band1) {
                if (band_lane) o.res1 = rb1[(size_t)r * nnx + clampc(c)];
            }
            if (band2) {
                if (band_lane) o.res2 = rb2[(size_t)r * nnx + clampc(c - 3)];
            }
            return o;
        };

        // Windows, [k & 3] relative to the unrolled base (columns held):
        //   G1 LVX0 FVX1 BXW                      c-2 .. c+1
        //   LVZ0 BZW KXW KZW LPX1 LPZ1 GKW H1 HZ1 PN1   c-3 .. c
        //   G2 LVX1 LVZ1 GBXW GBZW FVX2           c-5 .. c-2
        //   H2 HZ2 PN2                            c-6 .. c-3
        //   LPXn LPZn LVZn (input L_px, L_pz, L_vz)  c-2 .. c+1
        //   OPA OPZA (new L_px, L_pz of step n-1)   c-6 .. c-3, stored with B2
        float G1[4], LVX0[4], FVX1[4], BXW[4], LPXn[4], LPZn[4], LVZn[4];
        float LVZ0[4], BZW[4], KXW[4], KZW[4], LPX1[4], LPZ1[4], GKW[4], H1[4], HZ1[4], PN1[4];
        float G2[4], LVX1[4], LVZ1[4], GBXW[4], GBZW[4], FVX2[4];
        float H2[4], HZ2[4], PN2[4], OPA[4], OPZA[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            G1[k] = LVX0[k] = FVX1[k] = BXW[k] = LPXn[k] = LPZn[k] = LVZn[k] = 0.0f;
            LVZ0[k] = BZW[k] = KXW[k] = KZW[k] = LPX1[k] = LPZ1[k] = GKW[k] = 0.0f;
            H1[k] = HZ1[k] = PN1[k] = 0.0f;
            G2[k] = LVX1[k] = LVZ1[k] = GBXW[k] = GBZW[k] = FVX2[k] = 0.0f;
            H2[k] = HZ2[k] = PN2[k] = OPA[k] = OPZA[k] = 0.0f;
        }
        const int c_first = xa - 4;
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const int col = c_first - 2 + k;
            const unsigned o = coff(col);
            const fquad lv = ld4(Lin, 4 * o);
            LPXn[k] = lv.x; LPZn[k] = lv.y; LVX0[k] = lv.z; LVZn[k] = lv.w;
            FVX1[k] = ld(Fvx1, o);
            BXW[k] = ld(BX, o);
            G1[k] = (BXW[k] * prx[3 * nnx + clampc(col)]) * LVX0[k];
        }

        AdjCol2 q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(c_first + u);

        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            AdjCol2 &cur = q[u % PF];

            // ---- A1: step n, phase A at column c
            {
                LPXn[i3] = keep(cur.lam.x);
                LPZn[i3] = keep(cur.lam.y);
                LVZn[i3] = keep(cur.lam.w);
                const float lvx_n = keep(cur.lam.z);
                LVX0[i3] = lvx_n;
                FVX1[i3] = keep(cur.fvx_n);
                const float bxn = keep(cur.bx_n);
                BXW[i3] = bxn;
                G1[i3] = FAST ? bxn * lvx_n : (bxn * prx[3 * nnx + clampc(c + 1)]) * lvx_n;
                float ax;
                if (FAST) {
                    ax = SAM2D_A0 * G1[i1] - SAM2D_A0 * G1[i2] +
                         (SAM2D_A1 * G1[i0] - SAM2D_A1 * G1[i3]);
                } else {
                    ax = p0_at(c - 1) * G1[i1] - p0_at(c) * G1[i2] +
                         (p1_at(c - 2) * G1[i0] - p1_at(c + 1) * G1[i3]);
                }
                const float lvz = LVZn[i2];
                LVZ0[i3] = lvz;
                const float bz = keep(cur.bz);
                BZW[i3] = bz;
                const float gz = (bz * sbz) * lvz;
                const float az = zp0 * gz, bzv = zp1 * gz;
                const float bz_u1 = lane_up1(bzv);
                const float azz = (lane_up1(az) - az) + (lane_up1(bz_u1) - lane_dn1(bzv));
                const float lp = ax + azz;
                const float lpx1 = (LPXn[i2] + cur.res1) + lp;
                const float lpz1 = (LPZn[i2] + cur.res1) + lp;
                float xm0 = SAM2D_A0, xm1 = SAM2D_A1, skx = 1.0f, dax = 1.0f;
                if (!FAST) {
                    dminus_coef(c, nnx, xm0, xm1);
                    const int cc = clampc(c);
                    dax = prx[cc];
                    skx = prx[nnx + cc];
                }
                const float f0 = cur.fvz;
                const float f_u1 = lane_up1(f0), f_u2 = lane_up1(f_u1), f_d1 = lane_dn1(f0);
                const float dvx = fmaf(xm0, FVX1[i2] - FVX1[i1], xm1 * (FVX1[i3] - FVX1[i0]));
                const float dvz = fmaf(zm0, f0 - f_u1, zm1 * (f_d1 - f_u2));
                float g = cur.gk + lpx1 * dvx * (dtdx * skx) + lpz1 * dvz * (dtdz * skz);
                if (src_lane && c == src_x1) g += lpx1 * srcg1;
                GKW[i3] = g;
                LPX1[i3] = dax * lpx1;
                LPZ1[i3] = daz * lpz1;
                const float kx = keep(cur.kx);
                const float kz = KZA ? kx : keep(cur.kz);
                KXW[i3] = kx;
                if (!KZA) KZW[i3] = kz;
                H1[i3] = (kx * skx) * lpx1;
                HZ1[i3] = (kz * skz) * lpz1;
                PN1[i3] = ps_c ? cur.fpx : cur.fpx + cur.fpz;
            }

            // ---- B1: step n, phase B at column j1 = c-2
            {
                const int j1 = c - 2;
                float tx;
                if (FAST) {
                    tx = SAM2D_A0 * H1[i1] - SAM2D_A0 * H1[i2] +
                         (SAM2D_A1 * H1[i0] - SAM2D_A1 * H1[i3]);
                } else {
                    tx = m0_at(j1) * H1[i1] - m0_at(j1 + 1) * H1[i2] +
                         (m1_at(j1 - 1) * H1[i0] - m1_at(j1 + 2) * H1[i3]);
                }
                const float amz = zm0 * HZ1[i1], bmz = zm1 * HZ1[i1];
                const float bm_d1 = lane_dn1(bmz);
                const float tz = (amz - lane_dn1(amz)) + (lane_up1(bmz) - lane_dn1(bm_d1));
                const float lvx_j = LVX0[i0], lvz_j = LVZ0[i1];
                const float pj = PN1[i1];
                const float q_u1 = lane_up1(pj), q_d1 = lane_dn1(pj), q_d2 = lane_dn1(q_d1);
                float xp0 = SAM2D_A0, xp1 = SAM2D_A1, dbx = 1.0f, sbx = 1.0f;
                if (!FAST) {
                    dplus_coef(j1, nnx, xp0, xp1);
                    const int jc = clampc(j1);
                    dbx = prx[2 * nnx + jc];
                    sbx = prx[3 * nnx + jc];
                }
                const float dpx = fmaf(xp0, PN1[i2] - pj, xp1 * (PN1[i3] - PN1[i0]));
                const float dpz = fmaf(zp0, q_d1 - pj, zp1 * (q_d2 - q_u1));
                const float ovx = fmaf(dbx, lvx_j, tx);
                LVX1[i3] = ovx;
                LVZ1[i3] = fmaf(dbz, lvz_j, tz);
                GBXW[i3] = keep(cur.gbx_j) + lvx_j * dpx * (dtdx * sbx);
                GBZW[i3] = keep(cur.gbz_j) + lvz_j * dpz * (dtdz * sbz);
                G2[i3] = (BXW[i0] * sbx) * ovx;
                FVX2[i3] = keep(cur.fvx2);
            }

            // ---- A2: step n-1, phase A at column c2 = c-3 (results stored)
            {
                const int c2 = c - 3;
                float ax;
                if (FAST) {
                    ax = SAM2D_A0 * G2[i1] - SAM2D_A0 * G2[i2] +
                         (SAM2D_A1 * G2[i0] - SAM2D_A1 * G2[i3]);
                } else {
                    ax = p0_at(c2 - 1) * G2[i1] - p0_at(c2) * G2[i2] +
                         (p1_at(c2 - 2) * G2[i0] - p1_at(c2 + 1) * G2[i3]);
                }
                const float gz = (BZW[i0] * sbz) * LVZ1[i2];
                const float az = zp0 * gz, bzv = zp1 * gz;
                const float bz_u1 = lane_up1(bzv);
                const float azz = (lane_up1(az) - az) + (lane_up1(bz_u1) - lane_dn1(bzv));
                const float lp = ax + azz;
                const float lpx2 = (LPX1[i0] + cur.res2) + lp;
                const float lpz2 = (LPZ1[i0] + cur.res2) + lp;
                float xm0 = SAM2D_A0, xm1 = SAM2D_A1, skx = 1.0f, dax = 1.0f;
                if (!FAST) {
                    dminus_coef(c2, nnx, xm0, xm1);
                    const int cc = clampc(c2);
                    dax = prx[cc];
                    skx = prx[nnx + cc];
                }
                const float f0 = cur.fvz2;
                const float f_u1 = lane_up1(f0), f_u2 = lane_up1(f_u1), f_d1 = lane_dn1(f0);
                if (rown && c2 >= xa && c2 < xb) {
                    const float dvx = fmaf(xm0, FVX2[i2] - FVX2[i1], xm1 * (FVX2[i3] - FVX2[i0]));
                    const float dvz = fmaf(zm0, f0 - f_u1, zm1 * (f_d1 - f_u2));
                    float g = GKW[i0] + lpx2 * dvx * (dtdx * skx) + lpz2 * dvz * (dtdz * skz);
                    if (src_lane && c2 == src_x2) g += lpx2 * srcg2;
                    const unsigned id = (unsigned)c2 * (unsigned)nnz * 4u + rc4;
                    stf(GK, id, g);
                }
                OPA[i3] = dax * lpx2;
                OPZA[i3] = daz * lpz2;
                const float kz = KZA ? KXW[i0] : KZW[i0];
                H2[i3] = (KXW[i0] * skx) * lpx2;
                HZ2[i3] = (kz * skz) * lpz2;
                PN2[i3] = ps_b ? cur.fpx2 : cur.fpx2 + cur.fpz2;
            }

            // ---- B2: step n-1, phase B at column j2 = c-5 (results stored)
            {
                const int j2 = c - 5;
                float tx;
                if (FAST) {
                    tx = SAM2D_A0 * H2[i1] - SAM2D_A0 * H2[i2] +
                         (SAM2D_A1 * H2[i0] - SAM2D_A1 * H2[i3]);
                } else {
                    tx = m0_at(j2) * H2[i1] - m0_at(j2 + 1) * H2[i2] +
                         (m1_at(j2 - 1) * H2[i0] - m1_at(j2 + 2) * H2[i3]);
                }
                const float amz = zm0 * HZ2[i1], bmz = zm1 * HZ2[i1];
                const float bm_d1 = lane_dn1(bmz);
                const float tz = (amz - lane_dn1(amz)) + (lane_up1(bmz) - lane_dn1(bm_d1));
                const float lvx_j = LVX1[i0], lvz_j = LVZ1[i0];
                const float pj = PN2[i1];
                const float q_u1 = lane_up1(pj), q_d1 = lane_dn1(pj), q_d2 = lane_dn1(q_d1);
                if (rown && j2 >= xa && j2 < xb) {
                    float xp0 = SAM2D_A0, xp1 = SAM2D_A1, dbx = 1.0f, sbx = 1.0f;
                    if (!FAST) {
                        dplus_coef(j2, nnx, xp0, xp1);
                        const int jc = clampc(j2);
                        dbx = prx[2 * nnx + jc];
                        sbx = prx[3 * nnx + jc];
                    }
                    const float dpx = fmaf(xp0, PN2[i2] - pj, xp1 * (PN2[i3] - PN2[i0]));
                    const float dpz = fmaf(zp0, q_d1 - pj, zp1 * (q_d2 - q_u1));
                    const unsigned id = (unsigned)j2 * (unsigned)nnz * 4u + rc4;
                    st4(Lout, 4 * id, OPA[i1], OPZA[i1], fmaf(dbx, lvx_j, tx), fmaf(dbz, lvz_j, tz));
                    stf(GBX, id, GBXW[i0] + lvx_j * dpx * (dtdx * sbx));
                    stf(GBZ, id, GBZW[i0] + lvz_j * dpz * (dtdz * sbz));
                }
            }
            cur = load_col(min(c + PF, c_last));   // past c_last: a cache hit
        };

        using TT = std::integral_constant<bool, true>;
        using FF = std::integral_constant<bool, false>;
        using U0 = std::integral_constant<int, 0>;
        using U1 = std::integral_constant<int, 1>;
        using U2 = std::integral_constant<int, 2>;
        using U3 = std::integral_constant<int, 3>;
        int c0 = c_first;
        for (; c0 + 3 <= c_last && c0 < fast_lo; c0 += 4) {
            column(FF(), U0(), c0); column(FF(), U1(), c0);
            column(FF(), U2(), c0); column(FF(), U3(), c0);
        }
        const int c_fast = min(fast_hi, c_last - 3);
        for (; c0 <= c_fast; c0 += 4) {
            column(TT(), U0(), c0); column(TT(), U1(), c0);
            column(TT(), U2(), c0); column(TT(), U3(), c0);
        }
        for (; c0 <= c_last; c0 += 4) {
            column(FF(), U0(), c0);
            if (c0 + 1 > c_last) break;
            column(FF(), U1(), c0);
            if (c0 + 2 > c_last) break;
            column(FF(), U2(), c0);
            if (c0 + 3 > c_last) break;
            column(FF(), U3(), c0);
        }
    }
}

template <int W, int PF, bool KZA>
static void launch_adj_t(dim3 g, hipStream_t s, const AdjDev *d, int ns, int vt, int per, int nt,
                         int step, int li, int lo, int fp, int fc, int rk, int ps, TileOrder o)
{
    hipLaunchKernelGGL((k_adjoint<W, PF, KZA>), g, dim3(64 * W), 0, s, d, ns, vt, per, nt, step, li,
                       lo, fp, fc, rk, ps, o.ngroups, o.xorder);
}

template <int W>
static void launch_adj_w(int pf, bool kza, dim3 g, hipStream_t s, const AdjDev *d, int ns, int vt,
                         int per, int nt, int step, int li, int lo, int fp, int fc, int rk,
                         int ps, TileOrder o)
{
    if (pf == 2 && kza) launch_adj_t<W, 2, true>(g, s, d, ns, vt, per, nt, step, li, lo, fp, fc, rk, ps, o);
    else if (pf == 2) launch_adj_t<W, 2, false>(g, s, d, ns, vt, per, nt, step, li, lo, fp, fc, rk, ps, o);
    else if (kza) launch_adj_t<W, 4, true>(g, s, d, ns, vt, per, nt, step, li, lo, fp, fc, rk, ps, o);
    else launch_adj_t<W, 4, false>(g, s, d, ns, vt, per, nt, step, li, lo, fp, fc, rk, ps, o);
}

hipError_t launch_adjoint_step(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                               bool kz_alias, int step, int lam_in, int lam_out, int fwd_prev,
                               int fwd_cur, int rec_k, int psum, TileOrder order,
                               hipStream_t stream)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int per = per_xcd_for(ntiles, order);
    const int vtotal = per * 8;
    const dim3 g(vtotal);
    switch (waves) {
    case 1: launch_adj_w<1>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k, psum, order); break;
    case 2: launch_adj_w<2>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k, psum, order); break;
    case 4: launch_adj_w<4>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k, psum, order); break;
    case 8: launch_adj_w<8>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k, psum, order); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int W, int PF, bool KZA>
static void launch_adj2_t(dim3 g, hipStream_t s, const AdjDev *d, int ns, int vt, int per,
                          int nt, int step, int li, int lo, int fa, int fb, int fc, int r1, int r2,
                          int ps, TileOrder o)
{
    hipLaunchKernelGGL((k_adjoint2<W, PF, KZA>), g, dim3(64 * W), 0, s, d, ns, vt, per, nt, step,
                       li, lo, fa, fb, fc, r1, r2, ps, o.ngroups, o.xorder);
}

template <int W>
static void launch_adj2_w(int pf, bool kza, dim3 g, hipStream_t s, const AdjDev *d, int ns,
                          int vt, int per, int nt, int step, int li, int lo, int fa, int fb, int fc,
                          int r1, int r2, int ps, TileOrder o)
{
    if (pf == 2 && kza) launch_adj2_t<W, 2, true>(g, s, d, ns, vt, per, nt, step, li, lo, fa, fb, fc, r1, r2, ps, o);
    else if (pf == 2) launch_adj2_t<W, 2, false>(g, s, d, ns, vt, per, nt, step, li, lo, fa, fb, fc, r1, r2, ps, o);
    else if (kza) launch_adj2_t<W, 4, true>(g, s, d, ns, vt, per, nt, step, li, lo, fa, fb, fc, r1, r2, ps, o);
    else launch_adj2_t<W, 4, false>(g, s, d, ns, vt, per, nt, step, li, lo, fa, fb, fc, r1, r2, ps, o);
}

hipError_t launch_adjoint_step2(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                                bool kz_alias, int step, int lam_in, int lam_out, int fa, int fb,
                                int fc, int rec1, int rec2, int psum, TileOrder order,
                                hipStream_t stream)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int per = per_xcd_for(ntiles, order);
    const int vtotal = per * 8;
    const dim3 g(vtotal);
    switch (waves) {
    case 1: launch_adj2_w<1>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fa, fb, fc, rec1, rec2, psum, order); break;
    case 2: launch_adj2_w<2>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fa, fb, fc, rec1, rec2, psum, order); break;
    case 4: launch_adj2_w<4>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fa, fb, fc, rec1, rec2, psum, order); break;
    case 8: launch_adj2_w<8>(pf, kz_alias, g, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fa, fb, fc, rec1, rec2, psum, order); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_inject(const AdjDev *d_adj, int nshots, int max_cells, int lam, int rec_k,
                         hipStream_t stream)
{
    if (max_cells <= 0 || rec_k < 0) return hipSuccess;
    const int blocks = (max_cells + 255) / 256;
    hipLaunchKernelGGL(k_inject, dim3(blocks), dim3(256), 0, stream, d_adj, nshots, lam, rec_k);
    return hipGetLastError();
}

}  // namespace sam2d
