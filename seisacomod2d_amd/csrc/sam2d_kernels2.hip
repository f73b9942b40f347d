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
// march unrolled by four.  Receivers are recorded in place after each of the
// two steps from per-column receiver lists, since the intermediate state is
// never stored.

#include <hip/hip_runtime.h>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

struct ColOps2 {
    float vx_next;              // vx(n-1) at column c+1
    float vz, px, pz, kx, kz;   // at column c
    float dax, skx;             // x-profiles at column c
    float bx, bz, vz_j;         // BX, BZ, vz(n-1) at column c-2
    float dbx, sbx;             // x-profiles at column c-2
    float dax2, skx2;           // x-profiles at column c-3 (second step's p-phase)
    float dbx2, sbx2;           // x-profiles at column c-5 (second step's v-phase)
};

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64)
void k_step2(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd, int ntiles,
             int step, int in_slot, int out_slot, int rec1, int rec2)
{
    constexpr int PF = 2;
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;

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
        if (zb >= nnz) continue;
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);

        const int r = zb - HALO_T2 + lane;
        const bool rown = (lane >= HALO_T2) && (lane < HALO_T2 + OWN_Z2) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);

        const SAM2D_GLOBAL float *__restrict__ pxi = gp(slot_field(d, in_slot, 0));
        const SAM2D_GLOBAL float *__restrict__ pzi = gp(slot_field(d, in_slot, 1));
        const SAM2D_GLOBAL float *__restrict__ vxi = gp(slot_field(d, in_slot, 2));
        const SAM2D_GLOBAL float *__restrict__ vzi = gp(slot_field(d, in_slot, 3));
        SAM2D_GLOBAL float *__restrict__ pxo = gp(slot_field(d, out_slot, 0));
        SAM2D_GLOBAL float *__restrict__ pzo = gp(slot_field(d, out_slot, 1));
        SAM2D_GLOBAL float *__restrict__ vxo = gp(slot_field(d, out_slot, 2));
        SAM2D_GLOBAL float *__restrict__ vzo = gp(slot_field(d, out_slot, 3));
        const SAM2D_GLOBAL float *__restrict__ KX = gp(d.KX);
        const SAM2D_GLOBAL float *__restrict__ KZ = gp(d.KZ);
        const SAM2D_GLOBAL float *__restrict__ BX = gp(d.BX);
        const SAM2D_GLOBAL float *__restrict__ BZ = gp(d.BZ);
        const SAM2D_GLOBAL float *__restrict__ prx = gp(d.prof_x);
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);
        const SAM2D_GLOBAL int *__restrict__ rcolptr = gp(d.rcolptr);
        const SAM2D_GLOBAL int *__restrict__ rrow = gp(d.rrow);
        const SAM2D_GLOBAL int *__restrict__ rid = gp(d.rid);
        SAM2D_GLOBAL float *__restrict__ seis = gp(d.seis);

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

        auto colbase = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz;
        };
        auto clampc = [&](int col) -> int { return min(max(col, 0), nnx - 1); };
        auto load_col = [&](int c) -> ColOps2 {
            ColOps2 o;
            const unsigned b1 = colbase(c + 1), b0 = colbase(c), bj = colbase(c - 2);
            o.vx_next = vxi[b1 + rc];
            o.vz = vzi[b0 + rc]; o.px = pxi[b0 + rc]; o.pz = pzi[b0 + rc];
            o.kx = KX[b0 + rc]; o.kz = KZ[b0 + rc];
            o.bx = BX[bj + rc]; o.bz = BZ[bj + rc]; o.vz_j = vzi[bj + rc];
            const int cc = clampc(c), jc = clampc(c - 2);
            o.dax = prx[cc]; o.skx = prx[nnx + cc];
            o.dbx = prx[2 * nnx + jc]; o.sbx = prx[3 * nnx + jc];
            const int c3 = clampc(c - 3), c5 = clampc(c - 5);
            o.dax2 = prx[c3]; o.skx2 = prx[nnx + c3];
            o.dbx2 = prx[2 * nnx + c5]; o.sbx2 = prx[3 * nnx + c5];
            return o;
        };
        // In-place recording of the pressure p at column col into sample k
        // (owned rows and columns only; the gather of cpu_record,
        // cpu_modelling_kernels.cpp:24-31).
        auto record = [&](int col, int k, float p) {
            if (k < 0 || col < xa || col >= xb) return;
            const int e1 = rcolptr[col + 1];
            for (int e = rcolptr[col]; e < e1; e++)
                if (rown && r == rrow[e]) seis[(size_t)rid[e] * d.real_nt + k] = p;
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
        vx0[0] = vxi[colbase(c_first - 2) + rc];
        vx0[1] = vxi[colbase(c_first - 1) + rc];
        vx0[2] = vxi[colbase(c_first) + rc];

        ColOps2 q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(c_first + u);

        for (int c0 = c_first; c0 <= xb + 4; c0 += 4) {
            #pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u;
                ColOps2 &cur = q[u % PF];
                const int i0 = u & 3, i1 = (u + 1) & 3, i2 = (u + 2) & 3, i3 = (u + 3) & 3;

                // ---- A: step-n p-phase at column c
                vx0[i3] = cur.vx_next;
                float xm0, xm1;
                dminus_coef(c, nnx, xm0, xm1);
                {
                    const float dvx = fmaf(xm0, vx0[i2] - vx0[i1], xm1 * (vx0[i3] - vx0[i0]));
                    const float vz_0 = cur.vz;
                    const float vz_u2 = shfl_up(vz_0, 2), vz_u1 = shfl_up(vz_0, 1);
                    const float vz_d1 = shfl_dn(vz_0, 1);
                    const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                    float pxn = fmaf(cur.kx * cur.skx, dvx, cur.dax * cur.px);
                    const float pzn = fmaf(cur.kz * skz, dvz, daz * cur.pz);
                    if (src_row && c == src_x1) pxn += amp1;
                    px1[i3] = pxn;
                    pz1[i3] = pzn;
                    p1[i3] = pxn + pzn;
                    kxw[i3] = cur.kx;
                    kzw[i3] = cur.kz;
                }
                record(c, rec1, p1[i3]);

                // ---- B: step-n v-phase at column j1 = c-2
                {
                    const int j1 = c - 2;
                    float xp0, xp1;
                    dplus_coef(j1, nnx, xp0, xp1);
                    const float pm3 = p1[i0], pm2 = p1[i1], pm1 = p1[i2], pn = p1[i3];
                    const float dpx = fmaf(xp0, pm1 - pm2, xp1 * (pn - pm3));
                    const float q_u1 = shfl_up(pm2, 1);
                    const float q_d1 = shfl_dn(pm2, 1), q_d2 = shfl_dn(pm2, 2);
                    const float dpz = fmaf(zp0, q_d1 - pm2, zp1 * (q_d2 - q_u1));
                    vx1[i3] = fmaf(cur.bx * cur.sbx, dpx, cur.dbx * vx0[i0]);
                    vz1[i3] = fmaf(cur.bz * sbz, dpz, dbz * cur.vz_j);
                    bxw[i3] = cur.bx;
                    bzw[i3] = cur.bz;
                }

                // ---- C: step-(n+1) p-phase at column c2 = c-3
                {
                    const int c2 = c - 3;
                    const float dax2 = cur.dax2, skx2 = cur.skx2;
                    float ym0, ym1;
                    dminus_coef(c2, nnx, ym0, ym1);
                    // vx(n) at c2-2..c2+1 = c-5..c-2 -> vx1[i0..i3]
                    const float dvx = fmaf(ym0, vx1[i2] - vx1[i1], ym1 * (vx1[i3] - vx1[i0]));
                    const float vz_0 = vz1[i2];   // column c-3
                    const float vz_u2 = shfl_up(vz_0, 2), vz_u1 = shfl_up(vz_0, 1);
                    const float vz_d1 = shfl_dn(vz_0, 1);
                    const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                    // px(n), pz(n), KX, KZ at column c-3 -> index i0 of the c-3..c windows
                    float pxn = fmaf(kxw[i0] * skx2, dvx, dax2 * px1[i0]);
                    const float pzn = fmaf(kzw[i0] * skz, dvz, daz * pz1[i0]);
                    if (src_row && c2 == src_x2) pxn += amp2;
                    const float pn = pxn + pzn;
                    p2[i3] = pn;
                    if (rown && c2 >= xa && c2 < xb) {
                        const unsigned id = (unsigned)c2 * (unsigned)nnz + (unsigned)r;
                        pxo[id] = pxn;
                        pzo[id] = pzn;
                    }
                    record(c2, rec2, pn);
                }

                // ---- D: step-(n+1) v-phase at column j2 = c-5
                {
                    const int j2 = c - 5;
                    const float dbx2 = cur.dbx2, sbx2 = cur.sbx2;
                    float xp0, xp1;
                    dplus_coef(j2, nnx, xp0, xp1);
                    const float pm3 = p2[i0], pm2 = p2[i1], pm1 = p2[i2], pn = p2[i3];
                    const float dpx = fmaf(xp0, pm1 - pm2, xp1 * (pn - pm3));
                    const float q_u1 = shfl_up(pm2, 1);
                    const float q_d1 = shfl_dn(pm2, 1), q_d2 = shfl_dn(pm2, 2);
                    const float dpz = fmaf(zp0, q_d1 - pm2, zp1 * (q_d2 - q_u1));
                    if (rown && j2 >= xa && j2 < xb) {
                        // vx(n), vz(n), BX, BZ at column c-5 -> index i0 of the c-5..c-2 windows
                        const float vxn = fmaf(bxw[i0] * sbx2, dpx, dbx2 * vx1[i0]);
                        const float vzn = fmaf(bzw[i0] * sbz, dpz, dbz * vz1[i0]);
                        const unsigned id = (unsigned)j2 * (unsigned)nnz + (unsigned)r;
                        vxo[id] = vxn;
                        vzo[id] = vzn;
                    }
                }
                cur = load_col(c + PF);
            }
        }
    }
}

hipError_t launch_step2(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                        int step, int in_slot, int out_slot, int rec1, int rec2, hipStream_t stream)
{
    int waves, chunk;
    variant_shape(variant, &waves, &chunk);
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const dim3 g(grid);
    switch (waves) {
    case 1: hipLaunchKernelGGL(k_step2<1>, g, dim3(64), 0, stream, d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot, rec1, rec2); break;
    case 2: hipLaunchKernelGGL(k_step2<2>, g, dim3(128), 0, stream, d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot, rec1, rec2); break;
    case 4: hipLaunchKernelGGL(k_step2<4>, g, dim3(256), 0, stream, d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot, rec1, rec2); break;
    case 8: hipLaunchKernelGGL(k_step2<8>, g, dim3(512), 0, stream, d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot, rec1, rec2); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace sam2d
