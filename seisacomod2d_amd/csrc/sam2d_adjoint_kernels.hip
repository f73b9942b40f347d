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
//   dJ/dkappa += L_px' Ex + L_pz' Ez, Ex = (Dx- vx[n-1]) dtdx skx,
//                Ez = (Dz- vz[n-1]) dtdz skz (+ L_px'[src] fd_dt w[n]/(dx dz)
//                at the source)
//   dJ/dbu2   += L_vx (Dx+ p[n]) dtdx sbx,   dJ/dbu1 += L_vz (Dz+ p[n]) dtdz sbz
//
// where KXe = KX*skx etc. are the forward's effective coefficients.  The
// forward's recomputation stores p[n] and the factors of dJ/dkappa per cell
// (ring_pair, ring_side): outside the sponge frame L_px' = L_pz' and one
// factor E = Ex + Ez suffices; in the frame E = Ex and Ez apart.  The
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

// The sponge frame of a domain: cells where an x- or z-profile differs from
// 1.  Outside it the p-sponge factors are exactly 1, so L_px and L_pz take the
// same values (they start at zero and receive the same increments), and only
// L_px is stored.
__device__ __forceinline__ bool in_frame(const AdjDev &d, int col, int row)
{
    return col < d.xin_lo || col > d.xin_hi || row < d.zin_lo || row > d.zin_hi;
}

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
        // L_px of cell c at 3c; L_pz in the sponge frame only
        SAM2D_GLOBAL float *L = gp(d.lam + 3 * (size_t)lam * d.ncell);
        SAM2D_GLOBAL float *Z = gp(d.lpz + (size_t)lam * d.ncell);
        for (int u = gtid; u < d.ncells; u += gstride) {
            float sum = 0.0f;
            for (int e = d.rptr[u]; e < d.rptr[u + 1]; e++)
                sum += gp(d.res)[(size_t)d.rid[e] * d.real_nt + rec_k];
            const int c = d.rcell[u];
            const int col = c / d.nnz, row = c - col * d.nnz;
            L[3 * (size_t)c] += sum;
            if (in_frame(d, col, row)) Z[c] += sum;
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

// Coefficient cells of the adjoint: KX, BX, BZ (KZ aliases KX) or KX, BX,
// BZ, KZ, one dwordx3 or dwordx4 load per lane and column.
__global__ void k_pack_coef(const float *__restrict__ kx, const float *__restrict__ kz,
                            const float *__restrict__ bx, const float *__restrict__ bz, size_t n,
                            float *__restrict__ out)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const int w = kz ? 4 : 3;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float *o = out + w * i;
        o[0] = kx[i];
        o[1] = bx[i];
        o[2] = bz[i];
        if (kz) o[3] = kz[i];
    }
}

hipError_t launch_pack_coef(const float *kx, const float *kz, const float *bx, const float *bz,
                            size_t n, float *out, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_pack_coef, dim3((unsigned)blocks), dim3(256), 0, stream, kx, kz, bx, bz,
                       n, out);
    return hipGetLastError();
}

// The adjoint state is cell-interleaved (L_px, L_vx, L_vz of cell i at
// 3i .. 3i+2, L_pz apart and only in the sponge frame): a column's state
// arrives as one 12-byte load per lane at column c+1 (kept whole in the
// prefetch slot) and leaves as one 12-byte store at the phase-B column, L_px
// kept from phase A; frame lanes add a 4-byte load and store.  The forward
// ring pair of step n (p[n], E) and the coefficient cell (KX, BX, BZ[, KZ])
// arrive at column c+1 as well (a frame lane adds Ez), their column-c parts
// waiting a column in windows; the accumulator cell (dJ/dkappa, dJ/dbu2, dJ/dbu1)
// arrives at column c and leaves with the adjoint state at the phase-B
// column.
template <bool KZA>
using CoefCell = typename std::conditional<KZA, ftri, fquad>::type;

template <bool KZA>
struct AdjCol {
    ftri lam;                   // column c+1: adjoint L_px, L_vx, L_vz
    float lpz;                  //             L_pz (sponge frame)
    fpair fw;                   //             p[n], E (ring_pair)
    float ez;                   //             Ez (sponge frame, ring_side)
    CoefCell<KZA> cf;           //             KX, BX, BZ[, KZ]
    ftri acc;                   // column c:   dJ/dkappa, dJ/dbu2, dJ/dbu1
    float res;                  //             residual of this step
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_adjoint(const AdjDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
               int ntiles, int step, int lam_in, int lam_out, int fslot, int rec_k, int ngroups,
               int xorder)
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

        // Adjoint state, forward ring, coefficients and accumulators never
        // overlap.  The accumulators are read at a cell and written back there
        // by the same lane only.
        gcchar *__restrict__ Lin = (gcchar *)(d.lam + 3 * (size_t)lam_in * N);
        gchar *__restrict__ Lout = (gchar *)(d.lam + 3 * (size_t)lam_out * N);
        gcchar *__restrict__ Zin = (gcchar *)(d.lpz + (size_t)lam_in * N);
        gchar *__restrict__ Zout = (gchar *)(d.lpz + (size_t)lam_out * N);
        gcchar *__restrict__ FW = (gcchar *)(d.ring + 3 * (size_t)fslot * N);
        gcchar *__restrict__ FWZ = (gcchar *)(d.ring + 3 * (size_t)fslot * N + 2 * N);
        gcchar *__restrict__ CF = (gcchar *)d.coef;
        gchar *__restrict__ ACC = (gchar *)d.acc;
        const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);

        auto ld4 = [](gcchar *__restrict__ base, unsigned off) -> fquad {
            return *(const SAM2D_GLOBAL fquad *__restrict__)(base + off);
        };
        auto ldcf = [&](unsigned off) -> CoefCell<KZA> {
            if constexpr (KZA) return ld3(CF, 3 * off);
            else return ld4(CF, 4 * off);
        };
        // Byte offset of this lane's row in column col of a one-float field
        // (clamped to the grid); a cell of w floats is at w times it.
        auto coff = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 4u + rc4;
        };
        auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };
        // Sponge-frame lanes of a column (clamped as the loads are).
        const bool row_frame = r < d.zin_lo || r > d.zin_hi;
        auto frame = [&](int col) -> bool {
            const int cc = clampc(col);
            return row_frame || cc < d.xin_lo || cc > d.xin_hi;
        };
        auto ldf = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        auto stz = [](gchar *__restrict__ base, unsigned off, float val) {
            *(SAM2D_GLOBAL float *__restrict__)(base + off) = val;
        };

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

        auto load_col = [&](int c) -> AdjCol<KZA> {
            AdjCol<KZA> o;
            const unsigned o1 = coff(c + 1), o0 = coff(c);
            o.lam = ld3(Lin, 3 * o1);
            o.lpz = 0.0f;
            if (frame(c + 1)) o.lpz = ldf(Zin, o1);
            o.fw = ld2(FW, 2 * o1);
            o.ez = 0.0f;
            if (frame(c + 1)) o.ez = ldf(FWZ, o1);
            o.cf = ldcf(o1);
            o.acc = ld3(ACC, 3 * o0);
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
        // input adjoint state; FP, FE, FEZ: the forward ring; KXn, KZn,
        // BZn: coefficients; GBX, GBZ: accumulators as loaded; GK: dJ/dkappa
        // after phase A).  p-style: [k & 3] = column c-3+k (OPX, OPZ: the new
        // L_px, L_pz, stored with phase B).
        float LVX[4], LPX[4], LPZ[4], LVZ[4], FE[4], FEZ[4], FP[4], AXP[4], BXP[4];
        float KXn[4], KZn[4], BZn[4], GBX[4], GBZ[4], GK[4];
        float AM[4], BM[4], AMZ[4], BMZ[4], PN[4], OPX[4], OPZ[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            LVX[k] = LPX[k] = LPZ[k] = LVZ[k] = FE[k] = FEZ[k] = FP[k] = AXP[k] = BXP[k] = 0.0f;
            KXn[k] = KZn[k] = BZn[k] = GBX[k] = GBZ[k] = GK[k] = 0.0f;
            AM[k] = BM[k] = AMZ[k] = BMZ[k] = PN[k] = OPX[k] = OPZ[k] = 0.0f;
        }
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const int col = xa - 3 + k;
            const unsigned o = coff(col);
            const ftri lv = ld3(Lin, 3 * o);
            LPX[k] = lv.x; LVX[k] = lv.y; LVZ[k] = lv.z;
            LPZ[k] = frame(col) ? ldf(Zin, o) : lv.x;
            const fpair fw = ld2(FW, 2 * o);
            FP[k] = fw.x; FE[k] = fw.y;
            FEZ[k] = frame(col) ? ldf(FWZ, o) : 0.0f;
            const CoefCell<KZA> cf = ldcf(o);
            KXn[k] = cf.x; BZn[k] = cf.z;
            if constexpr (!KZA) KZn[k] = cf.w;
            gx_terms(col, LVX[k], cf.y, AXP[k], BXP[k]);
        }

        AdjCol<KZA> q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(xa - 1 + u);

        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            AdjCol<KZA> &cur = q[u % PF];

            // ---- phase A at column c
            LPX[i3] = keep(cur.lam.x);
            LPZ[i3] = frame(c + 1) ? keep(cur.lpz) : LPX[i3];
            LVZ[i3] = keep(cur.lam.z);
            const float lvx_n = keep(cur.lam.y);
            LVX[i3] = lvx_n;
            FP[i3] = keep(cur.fw.x);
            FE[i3] = keep(cur.fw.y);
            FEZ[i3] = frame(c + 1) ? keep(cur.ez) : 0.0f;
            const float bx_n = keep(cur.cf.y);
            KXn[i3] = keep(cur.cf.x);
            BZn[i3] = keep(cur.cf.z);
            if constexpr (!KZA) KZn[i3] = keep(cur.cf.w);
            GBX[i2] = keep(cur.acc.y);
            GBZ[i2] = keep(cur.acc.z);
            if (FAST) {
                const float g = bx_n * lvx_n;
                AXP[i3] = SAM2D_A0 * g;
                BXP[i3] = SAM2D_A1 * g;
            } else {
                gx_terms(c + 1, lvx_n, bx_n, AXP[i3], BXP[i3]);
            }
            const float ax = AXP[i1] - AXP[i2] + (BXP[i0] - BXP[i3]);
            const float gz = (BZn[i2] * sbz) * LVZ[i2];
            const float az = zp0 * gz, bzv = zp1 * gz;
            const float bz_u1 = lane_up1(bzv);
            const float azz = (lane_up1(az) - az) + (lane_up1(bz_u1) - lane_dn1(bzv));
            const float lp = ax + azz;
            const float lpx1 = (LPX[i2] + cur.res) + lp;
            const float lpz1 = (LPZ[i2] + cur.res) + lp;
            float xm0 = SAM2D_A0, xm1 = SAM2D_A1;
            if (!FAST) dminus_coef(c, nnx, xm0, xm1);
            float skx = 1.0f, dax = 1.0f;
            if (!FAST) {
                const int cc = clampc(c);
                dax = prx[cc];
                skx = prx[nnx + cc];
            }
            {
                // dJ/dkappa of the p-update, from the forward's factors
                const float gk = cur.acc.x + lpx1 * FE[i2];
                float g = frame(c) ? gk + lpz1 * FEZ[i2] : gk;
                if (src_lane && c == src_x) g += lpx1 * srcg;
                GK[i2] = g;
            }
            OPX[i3] = dax * lpx1;
            OPZ[i3] = daz * lpz1;
            const float kx = KXn[i2];
            const float kz = KZA ? kx : KZn[i2];
            const float hx = (kx * skx) * lpx1;
            const float hz = (kz * skz) * lpz1;
            AM[i3] = xm0 * hx;
            BM[i3] = xm1 * hx;
            AMZ[i3] = zm0 * hz;
            BMZ[i3] = zm1 * hz;
            PN[i3] = FP[i2];

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
                st3(Lout, 3 * id, OPX[i1], fmaf(dbx, lvx_j, tx), fmaf(dbz, lvz_j, tz));
                if (frame(j)) stz(Zout, id, OPZ[i1]);
                st3(ACC, 3 * id, GK[i0], GBX[i0] + lvx_j * dpx * (dtdx * sbx),
                    GBZ[i0] + lvz_j * dpz * (dtdz * sbz));
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
// Forward ring: step n's pairs (p[n], E) in ring slot fslot, loaded at column
// c+1; step n-1's in fslot - 1, loaded at column c-2 (frame lanes add Ez).
template <bool KZA>
struct AdjCol2 {
    ftri lam;                   // column c+1: adjoint L_px, L_vx, L_vz
    float lpz;                  //             L_pz (sponge frame)
    fpair fw1;                  //             forward ring pair of step n
    float ez1;                  //             its Ez (sponge frame)
    CoefCell<KZA> cf;           //             KX, BX, BZ[, KZ]
    ftri acc;                   // column c:   dJ/dkappa, dJ/dbu2, dJ/dbu1
    float res1;                 //             residual of step n
    fpair fw2;                  // column c-2: forward ring pair of step n-1
    float ez2;                  //             its Ez (sponge frame)
    float res2;                 // column c-3: residual of step n-1
};

template <int WAVES, int PF, bool KZA>
__global__ __launch_bounds__(WAVES * 64)
void k_adjoint2(const AdjDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
                int ntiles, int step, int lam_in, int lam_out, int fslot, int rec1, int rec2,
                int ngroups, int xorder)
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

        gcchar *__restrict__ Lin = (gcchar *)(d.lam + 3 * (size_t)lam_in * N);
        gchar *__restrict__ Lout = (gchar *)(d.lam + 3 * (size_t)lam_out * N);
        gcchar *__restrict__ Zin = (gcchar *)(d.lpz + (size_t)lam_in * N);
        gchar *__restrict__ Zout = (gchar *)(d.lpz + (size_t)lam_out * N);
        gcchar *__restrict__ FW1 = (gcchar *)(d.ring + 3 * (size_t)fslot * N);
        gcchar *__restrict__ FW2 = (gcchar *)(d.ring + 3 * (size_t)(fslot - 1) * N);
        gcchar *__restrict__ FWZ1 = FW1 + 8 * N;      // ring_side of each slot
        gcchar *__restrict__ FWZ2 = FW2 + 8 * N;
        gcchar *__restrict__ CF = (gcchar *)d.coef;
        gchar *__restrict__ ACC = (gchar *)d.acc;
        const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);

        // Coefficient cells of four floats: byte offset 4x that of one field.
        auto ld4 = [](gcchar *__restrict__ base, unsigned off) -> fquad {
            return *(const SAM2D_GLOBAL fquad *__restrict__)(base + off);
        };
        auto ldcf = [&](unsigned off) -> CoefCell<KZA> {
            if constexpr (KZA) return ld3(CF, 3 * off);
            else return ld4(CF, 4 * off);
        };
        auto coff = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz * 4u + rc4;
        };
        auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };
        // Sponge-frame lanes of a column (clamped as the loads are).
        const bool row_frame = r < d.zin_lo || r > d.zin_hi;
        auto frame = [&](int col) -> bool {
            const int cc = clampc(col);
            return row_frame || cc < d.xin_lo || cc > d.xin_hi;
        };
        auto ldf = [](gcchar *__restrict__ base, unsigned off) -> float {
            return *(const SAM2D_GLOBAL float *__restrict__)(base + off);
        };
        auto stz = [](gchar *__restrict__ base, unsigned off, float val) {
            *(SAM2D_GLOBAL float *__restrict__)(base + off) = val;
        };
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

        auto load_col = [&](int c) -> AdjCol2<KZA> {
            AdjCol2<KZA> o;
            const unsigned o1 = coff(c + 1), o0 = coff(c), oj = coff(c - 2);
            o.lam = ld3(Lin, 3 * o1);
            o.lpz = 0.0f;
            if (frame(c + 1)) o.lpz = ldf(Zin, o1);
            o.fw1 = ld2(FW1, 2 * o1);
            o.ez1 = 0.0f;
            if (frame(c + 1)) o.ez1 = ldf(FWZ1, o1);
            o.cf = ldcf(o1);
            o.acc = ld3(ACC, 3 * o0);
            o.fw2 = ld2(FW2, 2 * oj);
            o.ez2 = 0.0f;
            if (frame(c - 2)) o.ez2 = ldf(FWZ2, oj);
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
        //   G1 LVX0 BXW                           c-2 .. c+1
        //   FP1n FE1 FEZ1 KXn KZn BZn (as loaded) c-2 .. c+1
        //   LVZ0 BZW KXW KZW LPX1 LPZ1 GKW H1 HZ1 PN1   c-3 .. c
        //   GBXn GBZn (accumulators as loaded)    c-3 .. c
        //   G2 LVX1 LVZ1 GBXW GBZW                c-5 .. c-2
        //   FP2n FE2 FEZ2 (as loaded)             c-5 .. c-2
        //   H2 HZ2 PN2 GK2                        c-6 .. c-3
        //   LPXn LPZn LVZn (input L_px, L_pz, L_vz)  c-2 .. c+1
        //   OPA OPZA (new L_px, L_pz of step n-1)   c-6 .. c-3, stored with B2
        float G1[4], LVX0[4], FE1[4], BXW[4], LPXn[4], LPZn[4], LVZn[4];
        float FP1n[4], FEZ1[4], KXn[4], KZn[4], BZn[4], GBXn[4], GBZn[4], FP2n[4], FEZ2[4];
        float LVZ0[4], BZW[4], KXW[4], KZW[4], LPX1[4], LPZ1[4], GKW[4], H1[4], HZ1[4], PN1[4];
        float G2[4], LVX1[4], LVZ1[4], GBXW[4], GBZW[4], FE2[4];
        float H2[4], HZ2[4], PN2[4], GK2[4], OPA[4], OPZA[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            G1[k] = LVX0[k] = FE1[k] = BXW[k] = LPXn[k] = LPZn[k] = LVZn[k] = 0.0f;
            FP1n[k] = FEZ1[k] = KXn[k] = KZn[k] = BZn[k] = GBXn[k] = GBZn[k] = 0.0f;
            FP2n[k] = FEZ2[k] = 0.0f;
            LVZ0[k] = BZW[k] = KXW[k] = KZW[k] = LPX1[k] = LPZ1[k] = GKW[k] = 0.0f;
            H1[k] = HZ1[k] = PN1[k] = 0.0f;
            G2[k] = LVX1[k] = LVZ1[k] = GBXW[k] = GBZW[k] = FE2[k] = 0.0f;
            H2[k] = HZ2[k] = PN2[k] = GK2[k] = OPA[k] = OPZA[k] = 0.0f;
        }
        const int c_first = xa - 4;
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const int col = c_first - 2 + k;
            const unsigned o = coff(col);
            const ftri lv = ld3(Lin, 3 * o);
            LPXn[k] = lv.x; LVX0[k] = lv.y; LVZn[k] = lv.z;
            LPZn[k] = frame(col) ? ldf(Zin, o) : lv.x;
            const fpair fw = ld2(FW1, 2 * o);
            FP1n[k] = fw.x; FE1[k] = fw.y;
            FEZ1[k] = frame(col) ? ldf(FWZ1, o) : 0.0f;
            const CoefCell<KZA> cf = ldcf(o);
            KXn[k] = cf.x; BXW[k] = cf.y; BZn[k] = cf.z;
            if constexpr (!KZA) KZn[k] = cf.w;
            G1[k] = (BXW[k] * prx[3 * nnx + clampc(col)]) * LVX0[k];
        }

        AdjCol2<KZA> q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(c_first + u);

        auto column = [&](auto fast_tag, auto u_tag, int c0) {
            constexpr bool FAST = decltype(fast_tag)::value;
            constexpr int u = decltype(u_tag)::value;
            constexpr int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;
            const int c = c0 + u;
            AdjCol2<KZA> &cur = q[u % PF];

            // ---- A1: step n, phase A at column c
            {
                LPXn[i3] = keep(cur.lam.x);
                LPZn[i3] = frame(c + 1) ? keep(cur.lpz) : LPXn[i3];
                LVZn[i3] = keep(cur.lam.z);
                const float lvx_n = keep(cur.lam.y);
                LVX0[i3] = lvx_n;
                FP1n[i3] = keep(cur.fw1.x);
                FE1[i3] = keep(cur.fw1.y);
                FEZ1[i3] = frame(c + 1) ? keep(cur.ez1) : 0.0f;
                const float bxn = keep(cur.cf.y);
                BXW[i3] = bxn;
                KXn[i3] = keep(cur.cf.x);
                BZn[i3] = keep(cur.cf.z);
                if constexpr (!KZA) KZn[i3] = keep(cur.cf.w);
                GBXn[i3] = keep(cur.acc.y);
                GBZn[i3] = keep(cur.acc.z);
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
                const float bz = BZn[i2];
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
                const float gk = cur.acc.x + lpx1 * FE1[i2];
                float g = frame(c) ? gk + lpz1 * FEZ1[i2] : gk;
                if (src_lane && c == src_x1) g += lpx1 * srcg1;
                GKW[i3] = g;
                LPX1[i3] = dax * lpx1;
                LPZ1[i3] = daz * lpz1;
                const float kx = KXn[i2];
                const float kz = KZA ? kx : KZn[i2];
                KXW[i3] = kx;
                if (!KZA) KZW[i3] = kz;
                H1[i3] = (kx * skx) * lpx1;
                HZ1[i3] = (kz * skz) * lpz1;
                PN1[i3] = FP1n[i2];
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
                GBXW[i3] = GBXn[i1] + lvx_j * dpx * (dtdx * sbx);
                GBZW[i3] = GBZn[i1] + lvz_j * dpz * (dtdz * sbz);
                G2[i3] = (BXW[i0] * sbx) * ovx;
                FP2n[i3] = keep(cur.fw2.x);
                FE2[i3] = keep(cur.fw2.y);
                FEZ2[i3] = frame(c - 2) ? keep(cur.ez2) : 0.0f;
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
                {
                    const float gk = GKW[i0] + lpx2 * FE2[i2];
                    float g = frame(c2) ? gk + lpz2 * FEZ2[i2] : gk;
                    if (src_lane && c2 == src_x2) g += lpx2 * srcg2;
                    GK2[i3] = g;                  // stored with phase B2
                }
                OPA[i3] = dax * lpx2;
                OPZA[i3] = daz * lpz2;
                const float kz = KZA ? KXW[i0] : KZW[i0];
                H2[i3] = (KXW[i0] * skx) * lpx2;
                HZ2[i3] = (kz * skz) * lpz2;
                PN2[i3] = FP2n[i2];
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
                    st3(Lout, 3 * id, OPA[i1], fmaf(dbx, lvx_j, tx), fmaf(dbz, lvz_j, tz));
                    if (frame(j2)) stz(Zout, id, OPZA[i1]);
                    st3(ACC, 3 * id, GK2[i1], GBXW[i0] + lvx_j * dpx * (dtdx * sbx),
                        GBZW[i0] + lvz_j * dpz * (dtdz * sbz));
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

struct AdjArgs {
    const AdjDev *d;
    int ns, vt, per, nt, step, li, lo, fslot, r1, r2;
    TileOrder o;
};

template <int W, int PF, bool KZA>
static void launch_adj_t(dim3 g, hipStream_t s, const AdjArgs &a, bool two)
{
    if (two)
        hipLaunchKernelGGL((k_adjoint2<W, PF, KZA>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt, a.per,
                           a.nt, a.step, a.li, a.lo, a.fslot, a.r1, a.r2, a.o.ngroups, a.o.xorder);
    else
        hipLaunchKernelGGL((k_adjoint<W, PF, KZA>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt, a.per,
                           a.nt, a.step, a.li, a.lo, a.fslot, a.r1, a.o.ngroups, a.o.xorder);
}

template <int W>
static void launch_adj_w(int pf, bool kza, dim3 g, hipStream_t s, const AdjArgs &a, bool two)
{
    if (pf == 2 && kza) launch_adj_t<W, 2, true>(g, s, a, two);
    else if (pf == 2) launch_adj_t<W, 2, false>(g, s, a, two);
    else if (kza) launch_adj_t<W, 4, true>(g, s, a, two);
    else launch_adj_t<W, 4, false>(g, s, a, two);
}

static hipError_t launch_adj(int variant, bool kz_alias, bool two, hipStream_t stream, AdjArgs a)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    a.per = per_xcd_for(a.nt, a.o);
    a.vt = a.per * 8;
    const dim3 g(a.vt);
    switch (waves) {
    case 1: launch_adj_w<1>(pf, kz_alias, g, stream, a, two); break;
    case 2: launch_adj_w<2>(pf, kz_alias, g, stream, a, two); break;
    case 4: launch_adj_w<4>(pf, kz_alias, g, stream, a, two); break;
    case 8: launch_adj_w<8>(pf, kz_alias, g, stream, a, two); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_step(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                               bool kz_alias, int step, int lam_in, int lam_out, int fslot,
                               int rec_k, TileOrder order, hipStream_t stream)
{
    return launch_adj(variant, kz_alias, false, stream,
                      AdjArgs{d_adj, nshots, 0, 0, ntiles, step, lam_in, lam_out, fslot, rec_k, -1,
                              order});
}

hipError_t launch_adjoint_step2(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                                bool kz_alias, int step, int lam_in, int lam_out, int fslot,
                                int rec1, int rec2, TileOrder order, hipStream_t stream)
{
    return launch_adj(variant, kz_alias, true, stream,
                      AdjArgs{d_adj, nshots, 0, 0, ntiles, step, lam_in, lam_out, fslot, rec1, rec2,
                              order});
}

hipError_t launch_inject(const AdjDev *d_adj, int nshots, int max_cells, int lam, int rec_k,
                         hipStream_t stream)
{
    if (max_cells <= 0 || rec_k < 0) return hipSuccess;
    const int blocks = (max_cells + 255) / 256;
    hipLaunchKernelGGL(k_inject, dim3(blocks), dim3(256), 0, stream, d_adj, nshots, lam, rec_k);
    return hipGetLastError();
}

// dJ/dvp and dJ/drho of one shot's padded domain added into the model's FP64
// gradients, one thread per model cell of the window [wx0, wx0 + w): the
// transpose of expand (cpumodelling.cpp:39-60, sponge cells copy the nearest
// model cell), of cpu_stagger (cpu_modelling_kernels.cpp:64-121: bu1 is the
// mean of b at rows iz, iz+1 except the last row, bu2 that of b at columns
// ix, ix+1, its last column a copy of bu1's column nnx-2) and of kappa =
// rho vp^2, b = 1/rho.  A model cell takes the padded cells that copy it in
// ascending (ix, iz) order and each term is rounded as written, so the sums
// are those of the same loop run on the host, and shots are added one launch
// after another in a fixed order.
__global__ void k_chain(const float *__restrict__ acc, int nnz, int nnx, int npml, int wx0, int w,
                        int nz, const float *__restrict__ vp, const float *__restrict__ rho,
                        double *__restrict__ g_vp, double *__restrict__ g_rho)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w * nz) return;
    const int mx = t / nz, mz = t - mx * nz;
    const size_t col = (size_t)(wx0 + mx) * nz + mz;
    const int ixl = mx == 0 ? 0 : npml + mx, ixh = mx == w - 1 ? nnx - 1 : npml + mx;
    const int izl = mz == 0 ? 0 : npml + mz, izh = mz == nz - 1 ? nnz - 1 : npml + mz;
    const double v = vp[col], r = rho[col];
    const bool pos = rho[col] > 0.0f;
    double gv = g_vp[col], gr = g_rho[col];
    // b1 = dJ/dbu1 in FP64, with bu2's last column folded into column nnx-2
    auto b1 = [&](int ix, size_t id) -> double {
        double x = acc[3 * id + 2];
        if (ix == nnx - 2) x += (double)acc[3 * (id + nnz) + 1];
        return x;
    };
    auto b2 = [&](size_t id) -> double { return acc[3 * id + 1]; };
    for (int ix = ixl; ix <= ixh; ix++)
        for (int iz = izl; iz <= izh; iz++) {
            const size_t id = (size_t)ix * nnz + iz;
            const double gk = acc[3 * id];
            double gb = 0.0;
            if (iz >= 1) gb += 0.5 * b1(ix, id - 1);
            if (iz < nnz - 1) gb += 0.5 * b1(ix, id);
            if (ix >= 1) gb += 0.5 * b2(id - nnz);
            if (ix < nnx - 1) gb += 0.5 * b2(id);
            gv += 2.0 * v * r * gk;
            double g = v * v * gk;
            if (pos) g -= gb / (r * r);
            gr += g;
        }
    g_vp[col] = gv;
    g_rho[col] = gr;
}

hipError_t launch_chain(const float *acc, int nnz, int nnx, int npml, int wx0, int w, int nz,
                        const float *vp, const float *rho, double *g_vp, double *g_rho,
                        hipStream_t stream)
{
    const size_t n = (size_t)w * nz;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_chain, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, acc, nnz,
                       nnx, npml, wx0, w, nz, vp, rho, g_vp, g_rho);
    return hipGetLastError();
}

}  // namespace sam2d
