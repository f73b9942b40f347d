/*
 * sam2d_fp32.c -- the product's FP32 arithmetic of one time step, restated on
 * the CPU.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/ load it (through oracle/pyoracle.py);
 * the product library never links, loads or calls it.
 *
 * The FP64 oracle (sam2d_oracle.c) restates the reference; this file restates
 * what the HIP kernels compute from the FP32 fields the library builds
 * (sam2d_coefficients: KX, KZ, BX, BZ and the sponge profiles), operation by
 * operation and in the same order, so that the GPU's seismograms can be
 * checked bit for bit on the CPU and the FP32 error budget measured against
 * the FP64 oracle.  The step is the reference's FDOperator_4
 * (src/src_cpu/cpumodelling.cpp:115-156) with its kernels
 * (src/src_cpu/cpu_modelling_kernels.cpp:201-368):
 *
 *   Dx- vx  = fma(c0, vx[i] - vx[i-1], c1 * (vx[i+1] - vx[i-2]))   (c0, c1 by
 *             the edge rules of fdoperator420/410; likewise Dz- vz)
 *   px'     = fma(KX * sx1[ix], Dx- vx, sx0[ix] * px)  (+ the source increment)
 *   pz'     = fma(KZ * sz1[iz], Dz- vz, sz0[iz] * pz)
 *   p       = px' + pz'
 *
 * Outside the sponge frame (columns and rows where all four profiles of the
 * axis are exactly 1, the middle run of such indices, as the library's
 * ones_range finds it) the kernels keep p itself in the px slot and pz = 0:
 * there px' = fma(KX, Dx- vx, p) (+ source), pz' = fl(KZ * Dz- vz) and the
 * stored p = px' + pz'; the same formulas with pz = 0.
 *   Dx+ p   = fma(c0, p[i+1] - p[i], c1 * (p[i+2] - p[i-1]))   (fdoperator421/411)
 *   vx'     = fma(BX * bx1[ix], Dx+ p, bx0[ix] * vx),  vz' likewise
 *
 * with a0 = 1.125f, a1 = (float)(-1/24), every operation in FP32 and fma
 * fused (fmaf), nothing else contracted (-ffp-contract=off).  Receivers
 * sample p after step kt when (kt+1) % dt_factor == 0 (cpu_record,
 * cpu_modelling_kernels.cpp:24-31).
 *
 * Pinned by: the GPU kernels' own output (tests/test_gpu_fp32_model.py, bit
 * for bit), and the FP64 oracle within the FP32 error budget
 * (tests/test_oracle_fp32.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define F32_A0 1.125f
#define F32_A1 (-1.0f / 24.0f)

static inline void f32_dminus(int i, int n, float *c0, float *c1)
{
    if (i >= 2 && i <= n - 2) { *c0 = F32_A0; *c1 = F32_A1; }
    else if (i == 1 || i == n - 1) { *c0 = 1.0f; *c1 = 0.0f; }
    else { *c0 = 0.0f; *c1 = 0.0f; }
}
static inline void f32_dplus(int i, int n, float *c0, float *c1)
{
    if (i >= 1 && i <= n - 3) { *c0 = F32_A0; *c1 = F32_A1; }
    else if (i == 0 || i == n - 2) { *c0 = 1.0f; *c1 = 0.0f; }
    else { *c0 = 0.0f; *c1 = 0.0f; }
}
/* Neighbour index clamped to the grid: out-of-grid neighbours meet a zero
 * coefficient, as in the kernels' clamped loads. */
static inline long f32_cl(long i, long n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }

/* The middle run of indices where all four profiles [4][n] are exactly 1
 * (lo > hi: none). */
static void f32_ones_range(const float *prof, int n, int *lo, int *hi)
{
    const int mid = n / 2;
    int k, a, b, ok = 1;
    for (k = 0; k < 4; k++) ok &= prof[(long)k * n + mid] == 1.0f;
    if (!ok) { *lo = 1; *hi = 0; return; }
    a = b = mid;
    for (;;) {
        if (a == 0) break;
        ok = 1;
        for (k = 0; k < 4; k++) ok &= prof[(long)k * n + a - 1] == 1.0f;
        if (!ok) break;
        a--;
    }
    for (;;) {
        if (b == n - 1) break;
        ok = 1;
        for (k = 0; k < 4; k++) ok &= prof[(long)k * n + b + 1] == 1.0f;
        if (!ok) break;
        b++;
    }
    *lo = a; *hi = b;
}

/* fd_nt steps of one shot from the state (px, pz, vx, vz), each nnx*nnz
 * floats, x-major, updated in place, in the kernels' representation: outside
 * the sponge frame px holds p and pz is 0.  prof_x = [damp2a, damp2a1, damp2b,
 * damp2b1][nnx], prof_z = [damp1a, damp1a1, damp1b, damp1b1][nnz]; amp[nts]
 * source increments at cell src (injected while step <= nts, the increment
 * past the wavelet's end being 0); seis[ng][real_nt] samples p at the cells
 * gcell[ng]. */
int f32_forward(int nnx, int nnz, const float *KX, const float *KZ, const float *BX,
                const float *BZ, const float *prof_x, const float *prof_z, int fd_nt, int nts,
                const float *amp, long src, int dt_factor, int real_nt, int ng, const long *gcell,
                float *seis, float *px, float *pz, float *vx, float *vz)
{
    const long N = (long)nnx * nnz;
    float *p = (float *)calloc(N, sizeof(float));
    if (!p) return -30;
    const float *dax = prof_x, *skx = prof_x + nnx, *dbx = prof_x + 2 * nnx, *sbx = prof_x + 3 * nnx;
    const float *daz = prof_z, *skz = prof_z + nnz, *dbz = prof_z + 2 * nnz, *sbz = prof_z + 3 * nnz;
    int xlo, xhi, zlo, zhi;
    f32_ones_range(prof_x, nnx, &xlo, &xhi);
    f32_ones_range(prof_z, nnz, &zlo, &zhi);
    for (int it = 0; it < fd_nt; it++) {
        /* p-phase: Dx- vx, Dz- vz, cpu_pml_p, source, p = px + pz */
        #pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nnx; ix++) {
            float m0, m1;
            f32_dminus(ix, nnx, &m0, &m1);
            const float *xm2 = vx + f32_cl(ix - 2, nnx) * nnz, *xm1 = vx + f32_cl(ix - 1, nnx) * nnz;
            const float *x0 = vx + (long)ix * nnz, *xp1 = vx + f32_cl(ix + 1, nnx) * nnz;
            const float *vzc = vz + (long)ix * nnz;
            for (int iz = 0; iz < nnz; iz++) {
                const long id = (long)ix * nnz + iz;
                float z0, z1;
                f32_dminus(iz, nnz, &z0, &z1);
                const float dvx = fmaf(m0, x0[iz] - xm1[iz], m1 * (xp1[iz] - xm2[iz]));
                const float vz_0 = vzc[iz], vz_u1 = vzc[f32_cl(iz - 1, nnz)];
                const float vz_u2 = vzc[f32_cl(iz - 2, nnz)], vz_d1 = vzc[f32_cl(iz + 1, nnz)];
                const float dvz = fmaf(z0, vz_0 - vz_u1, z1 * (vz_d1 - vz_u2));
                float pxn = fmaf(KX[id] * skx[ix], dvx, dax[ix] * px[id]);
                const float pzn = fmaf(KZ[id] * skz[iz], dvz, daz[iz] * pz[id]);
                if (id == src && it <= nts) pxn += it < nts ? amp[it] : 0.0f;
                const float pn = pxn + pzn;
                if (ix >= xlo && ix <= xhi && iz >= zlo && iz <= zhi) {
                    px[id] = pn;            /* outside the frame: p, and pz stays 0 */
                } else {
                    px[id] = pxn;
                    pz[id] = pzn;
                }
                p[id] = pn;
            }
        }
        /* v-phase: Dx+ p, Dz+ p, cpu_pml_v */
        #pragma omp parallel for schedule(static)
        for (int ix = 0; ix < nnx; ix++) {
            float q0, q1;
            f32_dplus(ix, nnx, &q0, &q1);
            const float *pm1 = p + f32_cl(ix - 1, nnx) * nnz, *p0 = p + (long)ix * nnz;
            const float *pp1 = p + f32_cl(ix + 1, nnx) * nnz, *pp2 = p + f32_cl(ix + 2, nnx) * nnz;
            for (int iz = 0; iz < nnz; iz++) {
                const long id = (long)ix * nnz + iz;
                float z0, z1;
                f32_dplus(iz, nnz, &z0, &z1);
                const float dpx = fmaf(q0, pp1[iz] - p0[iz], q1 * (pp2[iz] - pm1[iz]));
                const float qu1 = p0[f32_cl(iz - 1, nnz)], qd1 = p0[f32_cl(iz + 1, nnz)];
                const float qd2 = p0[f32_cl(iz + 2, nnz)];
                const float dpz = fmaf(z0, qd1 - p0[iz], z1 * (qd2 - qu1));
                vx[id] = fmaf(BX[id] * sbx[ix], dpx, dbx[ix] * vx[id]);
                vz[id] = fmaf(BZ[id] * sbz[iz], dpz, dbz[iz] * vz[id]);
            }
        }
        if ((it + 1) % dt_factor == 0) {
            const int k = (it + 1) / dt_factor - 1;
            if (k < real_nt)
                for (int g = 0; g < ng; g++) seis[(long)g * real_nt + k] = p[gcell[g]];
        }
    }
    free(p);
    return 0;
}
