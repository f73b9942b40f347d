// sam2d_adjoint_kernels.hip -- gfx950 kernel of one backward (adjoint) step
// of the staggered-grid scheme in sam2d_kernels.hip, with the FWI gradient
// accumulation fused in.
//
// The reference has no adjoint (job_t.adj is declared and unused,
// src/src_cpu/modelling.h:51).  This is the exact transpose of the FP32
// forward step, run backward in time.  With lambda the adjoint of the state
// after step n (L_px, L_pz, L_vx, L_vz) and r_n the residual injected where
// step n was recorded:
//
//   L_p'  = Dx+^T (BXe L_vx) + Dz+^T (BZe L_vz) + r_n           (v-update, record)
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
// runs at column c, phase B (the velocity part) at column j = c - 2.

#include <hip/hip_runtime.h>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

struct AdjCol {
    float lvx_n, bx_n, sbx_n, fvx_n;   // column c+1: L_vx, BX, damp2b1, fwd vx[n-1]
    float lvz, bz;                     // column c:   L_vz, BZ
    float lpx, lpz;                    //             L_px, L_pz
    float fvz, fpx, fpz;               //             fwd vz[n-1], px[n], pz[n]
    float kx, kz, dax, skx, gk;        //             KX, KZ, damp2a, damp2a1, dJ/dkappa
    float lvz_j, gbx_j, gbz_j;         // column j:   L_vz, dJ/dbu2, dJ/dbu1
    float dbx_j, sbx_j;                //             damp2b, damp2b1
};

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64)
void k_adjoint(const AdjDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
               int ntiles, int step, int lam_in, int lam_out, int fwd_prev, int fwd_cur,
               int rec_k)
{
    constexpr int PF = 2;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile(v, per_xcd);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const AdjDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int nnz = d.nnz, nnx = d.nnx;
        const int zb = (group * WAVES + wv) * OWN_Z;
        if (zb >= nnz) continue;
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);

        const int r = zb - HALO_T + lane;
        const bool rown = (lane >= HALO_T) && (lane < HALO_T + OWN_Z) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);
        const size_t N = d.ncell;

        const SAM2D_GLOBAL float *__restrict__ Lpx = gp(d.lam + (4 * (size_t)lam_in + 0) * N);
        const SAM2D_GLOBAL float *__restrict__ Lpz = gp(d.lam + (4 * (size_t)lam_in + 1) * N);
        const SAM2D_GLOBAL float *__restrict__ Lvx = gp(d.lam + (4 * (size_t)lam_in + 2) * N);
        const SAM2D_GLOBAL float *__restrict__ Lvz = gp(d.lam + (4 * (size_t)lam_in + 3) * N);
        SAM2D_GLOBAL float *__restrict__ Opx = gp(d.lam + (4 * (size_t)lam_out + 0) * N);
        SAM2D_GLOBAL float *__restrict__ Opz = gp(d.lam + (4 * (size_t)lam_out + 1) * N);
        SAM2D_GLOBAL float *__restrict__ Ovx = gp(d.lam + (4 * (size_t)lam_out + 2) * N);
        SAM2D_GLOBAL float *__restrict__ Ovz = gp(d.lam + (4 * (size_t)lam_out + 3) * N);
        const SAM2D_GLOBAL float *__restrict__ Fpx = gp(d.fwd + (4 * (size_t)fwd_cur + 0) * N);
        const SAM2D_GLOBAL float *__restrict__ Fpz = gp(d.fwd + (4 * (size_t)fwd_cur + 1) * N);
        const SAM2D_GLOBAL float *__restrict__ Fvx = gp(d.fwd + (4 * (size_t)fwd_prev + 2) * N);
        const SAM2D_GLOBAL float *__restrict__ Fvz = gp(d.fwd + (4 * (size_t)fwd_prev + 3) * N);
        const SAM2D_GLOBAL float *__restrict__ KX = gp(d.KX);
        const SAM2D_GLOBAL float *__restrict__ KZ = gp(d.KZ);
        const SAM2D_GLOBAL float *__restrict__ BX = gp(d.BX);
        const SAM2D_GLOBAL float *__restrict__ BZ = gp(d.BZ);
        const SAM2D_GLOBAL float *__restrict__ prx = gp(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);
        SAM2D_GLOBAL float *__restrict__ GK = gp(d.gk);
        SAM2D_GLOBAL float *__restrict__ GBX = gp(d.gbx);
        SAM2D_GLOBAL float *__restrict__ GBZ = gp(d.gbz);

        const float daz = prz[rc], skz = prz[nnz + rc];
        const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
        float zm0, zm1, zp0, zp1;
        dminus_coef(r, nnz, zm0, zm1);
        dplus_coef(r, nnz, zp0, zp1);
        const float dtdx = d.dtdx, dtdz = d.dtdz;
        const bool src_lane = (r == d.src_z);
        const float srcg = (step < d.nts) ? gp(d.srcg)[step] : 0.0f;
        const int src_x = (step < d.nts) ? d.src_x : -1;
        const SAM2D_GLOBAL int *colptr = gp(d.colptr);
        const SAM2D_GLOBAL int *rrow = gp(d.rrow);
        const SAM2D_GLOBAL int *rid = gp(d.rid);
        const SAM2D_GLOBAL float *res = gp(d.res);

        auto colbase = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz;
        };
        auto load_col = [&](int c) -> AdjCol {
            AdjCol o;
            const unsigned b1 = colbase(c + 1), b0 = colbase(c), bj = colbase(c - 2);
            const int c1 = min(max(c + 1, 0), nnx - 1), cc = min(max(c, 0), nnx - 1);
            const int jc = min(max(c - 2, 0), nnx - 1);
            o.lvx_n = Lvx[b1 + rc]; o.bx_n = BX[b1 + rc]; o.sbx_n = prx[3 * nnx + c1];
            o.fvx_n = Fvx[b1 + rc];
            o.lvz = Lvz[b0 + rc]; o.bz = BZ[b0 + rc];
            o.lpx = Lpx[b0 + rc]; o.lpz = Lpz[b0 + rc];
            o.fvz = Fvz[b0 + rc]; o.fpx = Fpx[b0 + rc]; o.fpz = Fpz[b0 + rc];
            o.kx = KX[b0 + rc]; o.kz = KZ[b0 + rc];
            o.dax = prx[cc]; o.skx = prx[nnx + cc]; o.gk = GK[b0 + rc];
            o.lvz_j = Lvz[bj + rc]; o.gbx_j = GBX[bj + rc]; o.gbz_j = GBZ[bj + rc];
            o.dbx_j = prx[2 * nnx + jc]; o.sbx_j = prx[3 * nnx + jc];
            return o;
        };
        // Per-column products of the transposed D+ (x): c0(i) gx(i), c1(i) gx(i)
        // with gx = BXe L_vx at column i.
        auto gx_terms = [&](int i, float lvx, float bx, float sbx, float &a, float &b) {
            float p0, p1;
            dplus_coef(i, nnx, p0, p1);
            const float g = (bx * sbx) * lvx;
            a = p0 * g;
            b = p1 * g;
        };

        // vx-style windows: [k & 3] = column c-2+k.  p-style: [k & 3] = column c-3+k.
        float LVX[4], FVX[4], AXP[4], BXP[4];
        float AM[4], BM[4], AMZ[4], BMZ[4], PN[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            LVX[k] = FVX[k] = AXP[k] = BXP[k] = 0.0f;
            AM[k] = BM[k] = AMZ[k] = BMZ[k] = PN[k] = 0.0f;
        }
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const int col = xa - 3 + k;
            const unsigned b = colbase(col);
            const int cl = min(max(col, 0), nnx - 1);
            LVX[k] = Lvx[b + rc];
            FVX[k] = Fvx[b + rc];
            gx_terms(col, LVX[k], BX[b + rc], prx[3 * nnx + cl], AXP[k], BXP[k]);
        }

        AdjCol q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(xa - 1 + u);

        for (int c0 = xa - 1; c0 <= xb + 1; c0 += 4) {
            #pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u;
                AdjCol &cur = q[u % PF];
                // ---- phase A at column c
                LVX[(u + 3) & 3] = cur.lvx_n;
                FVX[(u + 3) & 3] = cur.fvx_n;
                gx_terms(c + 1, cur.lvx_n, cur.bx_n, cur.sbx_n, AXP[(u + 3) & 3], BXP[(u + 3) & 3]);
                const float ax = AXP[(u + 1) & 3] - AXP[(u + 2) & 3]
                               + (BXP[u & 3] - BXP[(u + 3) & 3]);
                const float gz = (cur.bz * sbz) * cur.lvz;
                const float az = zp0 * gz, bzv = zp1 * gz;
                const float azz = (shfl_up(az, 1) - az) + (shfl_up(bzv, 2) - shfl_dn(bzv, 1));
                float lp = ax + azz;
                if (rec_k >= 0 && c >= 0 && c < nnx) {
                    const int e1 = colptr[c + 1];
                    for (int e = colptr[c]; e < e1; e++) {
                        const int row = rrow[e];
                        const float val = res[(size_t)rid[e] * d.real_nt + rec_k];
                        if (r == row) lp += val;
                    }
                }
                const float lpx1 = cur.lpx + lp;
                const float lpz1 = cur.lpz + lp;
                float xm0, xm1;
                dminus_coef(c, nnx, xm0, xm1);
                // dJ/dkappa from the p-update needs Dx- vx[n-1] and Dz- vz[n-1];
                // the lane shuffles run with every lane active (a shuffle only
                // sees the values of active lanes).
                const float f0 = cur.fvz;
                const float f_u1 = shfl_up(f0, 1), f_u2 = shfl_up(f0, 2), f_d1 = shfl_dn(f0, 1);
                if (rown && c >= xa && c < xb) {
                    const float dvx = fmaf(xm0, FVX[(u + 2) & 3] - FVX[(u + 1) & 3],
                                           xm1 * (FVX[(u + 3) & 3] - FVX[u & 3]));
                    const float dvz = fmaf(zm0, f0 - f_u1, zm1 * (f_d1 - f_u2));
                    float g = cur.gk + lpx1 * dvx * (dtdx * cur.skx) + lpz1 * dvz * (dtdz * skz);
                    if (src_lane && c == src_x) g += lpx1 * srcg;
                    const unsigned id = (unsigned)c * (unsigned)nnz + (unsigned)r;
                    GK[id] = g;
                    Opx[id] = cur.dax * lpx1;
                    Opz[id] = daz * lpz1;
                }
                const float hx = (cur.kx * cur.skx) * lpx1;
                const float hz = (cur.kz * skz) * lpz1;
                AM[(u + 3) & 3] = xm0 * hx;
                BM[(u + 3) & 3] = xm1 * hx;
                AMZ[(u + 3) & 3] = zm0 * hz;
                BMZ[(u + 3) & 3] = zm1 * hz;
                PN[(u + 3) & 3] = cur.fpx + cur.fpz;

                // ---- phase B at column j = c - 2
                const int j = c - 2;
                if (j >= xa && j < xb) {
                    const float tx = AM[(u + 1) & 3] - AM[(u + 2) & 3]
                                   + (BM[u & 3] - BM[(u + 3) & 3]);
                    const float amz = AMZ[(u + 1) & 3], bmz = BMZ[(u + 1) & 3];
                    const float tz = (amz - shfl_dn(amz, 1)) + (shfl_up(bmz, 1) - shfl_dn(bmz, 2));
                    const float lvx_j = LVX[u & 3];
                    const float pj = PN[(u + 1) & 3];
                    const float q_u1 = shfl_up(pj, 1), q_d1 = shfl_dn(pj, 1), q_d2 = shfl_dn(pj, 2);
                    if (rown) {
                        float xp0, xp1;
                        dplus_coef(j, nnx, xp0, xp1);
                        const float dpx = fmaf(xp0, PN[(u + 2) & 3] - pj, xp1 * (PN[(u + 3) & 3] - PN[u & 3]));
                        const float dpz = fmaf(zp0, q_d1 - pj, zp1 * (q_d2 - q_u1));
                        const unsigned id = (unsigned)j * (unsigned)nnz + (unsigned)r;
                        Ovx[id] = fmaf(cur.dbx_j, lvx_j, tx);
                        Ovz[id] = fmaf(dbz, cur.lvz_j, tz);
                        GBX[id] = cur.gbx_j + lvx_j * dpx * (dtdx * cur.sbx_j);
                        GBZ[id] = cur.gbz_j + cur.lvz_j * dpz * (dtdz * sbz);
                    }
                }
                cur = load_col(c + PF);
            }
        }
    }
}

hipError_t launch_adjoint_step(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                               int step, int lam_in, int lam_out, int fwd_prev, int fwd_cur,
                               int rec_k, hipStream_t stream)
{
    int waves, chunk;
    variant_shape(variant, &waves, &chunk);
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    const dim3 g(vtotal);
    switch (waves) {
    case 1: hipLaunchKernelGGL(k_adjoint<1>, g, dim3(64), 0, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k); break;
    case 2: hipLaunchKernelGGL(k_adjoint<2>, g, dim3(128), 0, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k); break;
    case 4: hipLaunchKernelGGL(k_adjoint<4>, g, dim3(256), 0, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k); break;
    case 8: hipLaunchKernelGGL(k_adjoint<8>, g, dim3(512), 0, stream, d_adj, nshots, vtotal, per, ntiles, step, lam_in, lam_out, fwd_prev, fwd_cur, rec_k); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sam2d
