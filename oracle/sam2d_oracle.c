/*
 * sam2d_oracle.c -- FP64 CPU restatement of SeisAcoMod2D's forward-modelling
 * hot path (2-D acoustic, staggered grid, 4th order in space, 2nd order in
 * time, split-field exponential sponge, Ricker source, pressure receivers).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product library never links, loads or calls it.
 *
 * Every function below restates one piece of the reference
 * (src/src_cpu of richaras/SeisAcoMod2D) and cites the file:line it follows.
 * The arithmetic keeps the reference's types and operation order (FP64 state,
 * FP32 sponge profiles and buoyancy, float-evaluated time parameters) so that
 * a run matches the compiled reference bit for bit; that is checked against
 * golden seismograms written by the reference executable itself.
 *
 * Pinning (what checks each function, so a plausible slip fails one):
 *   orc_time_params  reference log lines (fd_nt, real_nt, DT factor, nts)
 *   orc_expand       brute-force replication rule on tiny arrays
 *   orc_stagger      closed-form averages + last-row/column rule
 *   orc_ricker       closed form (1-2a^2)exp(-a^2)
 *   orc_pml_a/_b     closed form exp(-(fac*x)^2) on both grids
 *   orc_dminus/dplus exactness on cubics, 2nd-order edges, transposes
 *   orc_forward      golden seismograms from the compiled reference
 *   orc_gradient     finite differences and the dot-product test
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define ORC_PI 3.141592654            /* cpu_modelling_kernels.cpp:7 */
#define ORC_A0 1.125f                 /* cpu_modelling_kernels.cpp:9 */
#define ORC_A1 (-1.0f / 24.0f)        /* cpu_modelling_kernels.cpp:10, float */

/* ------------------------------------------------------------------ A1 */
/* Time / grid parameters, cpumodelling.cpp:168-195.  All quotients are
 * evaluated in float exactly as the reference does; sqrt(6.0f) resolves to
 * the float overload there (<cmath> + using namespace std). */
void orc_time_params(float rec_len, float fd_dt, float real_dt, float fm,
                     float dx, float dz, int *fd_nt, int *real_nt,
                     int *dt_factor, int *nts, double *dtdx, double *dtdz)
{
    float _dx = 1.0f / dx, _dz = 1.0f / dz;
    *fd_nt = (int)ceilf(rec_len / fd_dt);
    *real_nt = (int)ceilf(rec_len / real_dt);
    *dtdx = (double)(fd_dt * _dx);
    *dtdz = (double)(fd_dt * _dz);
    *dt_factor = (int)(real_dt / fd_dt);
    double td = (double)sqrtf(6.0f) / (ORC_PI * fm);
    int n = (int)(5.0f * td / fd_dt) + 1;
    if (n > *fd_nt) n = *fd_nt;
    *nts = n;
}

/* ------------------------------------------------------------------ A2 */
/* expand(), cpumodelling.cpp:39-60: embed a[nx1][nz1] (z fastest) at offset
 * npml, then replicate edge values: top/bottom rows first, then the
 * left/right columns (which therefore also fill the corners). */
void orc_expand(const float *a, int nx1, int nz1, int npml, float *b)
{
    int nnz = nz1 + 2 * npml, nnx = nx1 + 2 * npml;
    for (int ix = 0; ix < nx1; ix++)
        for (int iz = 0; iz < nz1; iz++)
            b[(npml + ix) * nnz + npml + iz] = a[ix * nz1 + iz];
    for (int ix = 0; ix < nnx; ix++) {
        for (int iz = 0; iz < npml; iz++) b[ix * nnz + iz] = b[ix * nnz + npml];
        for (int iz = nz1 + npml; iz < nnz; iz++)
            b[ix * nnz + iz] = b[ix * nnz + npml + nz1 - 1];
    }
    for (int iz = 0; iz < nnz; iz++) {
        for (int ix = 0; ix < npml; ix++) b[ix * nnz + iz] = b[npml * nnz + iz];
        for (int ix = npml + nx1; ix < nnx; ix++)
            b[ix * nnz + iz] = b[(npml + nx1 - 1) * nnz + iz];
    }
}

/* ------------------------------------------------------------------ A3 */
/* cpu_stagger(), cpu_modelling_kernels.cpp:64-121.  kappa = vp^2 rho in
 * FP64; bue = 1/rho in float (1e10 where rho <= 0); bu1 (z-staggered) and
 * bu2 (x-staggered) are float averages.  The last z-row of bu1 copies bu2
 * before bu2 is computed, i.e. 0 (bu2 is zeroed by the caller,
 * cpumodelling.cpp:224-229); the last x-column of bu2 copies bu1's
 * column nnx-2. */
void orc_stagger(const float *vpe, const float *rhoe, int nnz, int nnx,
                 double *kappa, float *bue, float *bu1, float *bu2)
{
    long N = (long)nnz * nnx;
    for (long i = 0; i < N; i++) { bu1[i] = 0.0f; bu2[i] = 0.0f; }
    for (long id = 0; id < N; id++) {
        double vp2 = (double)vpe[id] * (double)vpe[id];
        kappa[id] = vp2 * (double)rhoe[id];
        bue[id] = (rhoe[id] > 0.0f) ? 1.0f / rhoe[id] : 10000000000.0f;
    }
    for (int i2 = 0; i2 < nnx; i2++) {
        for (int i1 = 0; i1 < nnz - 1; i1++) {
            long id = i1 + (long)i2 * nnz;
            bu1[id] = 0.5f * (bue[id + 1] + bue[id]);
        }
        bu1[(long)i2 * nnz + nnz - 1] = bu2[(long)i2 * nnz + nnz - 2];
    }
    for (int i2 = 0; i2 < nnx - 1; i2++)
        for (int i1 = 0; i1 < nnz; i1++) {
            long id = i1 + (long)i2 * nnz;
            bu2[id] = 0.5f * (bue[id + nnz] + bue[id]);
        }
    for (int i1 = 0; i1 < nnz; i1++)
        bu2[(long)(nnx - 1) * nnz + i1] = bu1[(long)(nnx - 2) * nnz + i1];
}

/* ------------------------------------------------------------------ A4 */
/* cpu_ricker(), cpu_modelling_kernels.cpp:33-62, iss == 2.  Note that the
 * sample time is the float product i*dt. */
void orc_ricker(float fm, float dt, int nts, double *wlt)
{
    double t0 = 1.5 * sqrt(6.0) / (ORC_PI * (double)fm);
    for (int i = 0; i < nts; i++) {
        double t = (double)(i * dt);
        double a = ORC_PI * (double)fm * (t - t0);
        double a2 = a * a;
        wlt[i] = (1.0f - 2.0f * a2) * exp(-a2);
    }
}

/* ------------------------------------------------------------------ A5 */
/* cpu_pml_coefficient_a(), cpu_modelling_kernels.cpp:123-153: profiles on
 * the pressure (integer) grid.  damp = expf(-(fac x)^2) (float overload),
 * damp1 = (float)exp(-0.5 (fac x)^2) (double, from the 0.5 literal). */
void orc_pml_a(float fac, int npml, int n, float *damp, float *damp1)
{
    float *spg = (float *)calloc(npml + 1, sizeof(float));
    float *spg1 = (float *)calloc(npml + 1, sizeof(float));
    for (int i = 0; i < npml; i++) {
        float x = (float)(npml - (i + 1));
        float temp = fac * x;
        spg[i] = expf(-(temp * temp));
        spg1[i] = (float)exp(-0.5 * (temp * temp));
    }
    for (int i = 0; i < npml; i++) {
        damp[i] = spg[i];
        damp1[i] = spg1[i];
        damp[npml + n + i] = spg[npml - i - 1];
        damp1[npml + n + i] = spg1[npml - i - 1];
    }
    for (int i = 0; i < n; i++) { damp[i + npml] = 1.0f; damp1[i + npml] = 1.0f; }
    free(spg); free(spg1);
}

/* cpu_pml_coefficient_b(), cpu_modelling_kernels.cpp:155-199: profiles on
 * the velocity (half-integer) grid.  Left: i < npml-1 at x = npml-1.5-i;
 * ones on [npml-1, npml+n]; right: i >= npml+n at x = i-npml-n+0.5, which
 * overwrites the one at npml+n. */
void orc_pml_b(float fac, int npml, int n, float *damp, float *damp1)
{
    float *spg = (float *)calloc(npml + 1, sizeof(float));
    float *spg1 = (float *)calloc(npml + 1, sizeof(float));
    for (int i = 0; i < npml - 1; i++) {
        float x = (float)(npml - 1 - (i + 1)) + 0.5f;
        float temp = fac * x;
        spg[i] = expf(-(temp * temp));
        spg1[i] = (float)exp(-0.5 * (temp * temp));
    }
    for (int i = 0; i < npml - 1; i++) { damp[i] = spg[i]; damp1[i] = spg1[i]; }
    for (int i = 0; i < n + 1; i++) { damp[i + npml - 1] = 1.0f; damp1[i + npml - 1] = 1.0f; }
    for (int i = 0; i < npml; i++) {
        float x = (float)((i + 1) - 1) + 0.5f;
        float temp = fac * x;
        spg[i] = expf(-(temp * temp));
        spg1[i] = (float)exp(-0.5 * (temp * temp));
    }
    for (int i = 0; i < npml; i++) { damp[npml + n + i] = spg[i]; damp1[npml + n + i] = spg1[i]; }
    free(spg); free(spg1);
}

/* ------------------------------------------------------------ A9 / A13 */
/* The four derivative operators on an n-point line with stride s
 * (s = nnz for x, 1 for z).  fdoperator420/410
 * (cpu_modelling_kernels.cpp:201-258): backward-staggered D-, 4th order on
 * [2, n-2], 2nd order at 1 and n-1, 0 at 0.  fdoperator421/411 (:260-316):
 * forward-staggered D+, 4th order on [1, n-3], 2nd order at 0 and n-2,
 * 0 at n-1.  Operation order follows the reference: a0*(d1) + a1*(d2). */
static inline double dminus_at(const double *f, int i, int n, long s)
{
    if (i >= 2 && i <= n - 2)
        return ORC_A0 * (f[0] - f[-s]) + ORC_A1 * (f[s] - f[-2 * s]);
    if (i == 1 || i == n - 1) return f[0] - f[-s];
    return 0.0;
}
static inline double dplus_at(const double *f, int i, int n, long s)
{
    if (i >= 1 && i <= n - 3)
        return ORC_A0 * (f[s] - f[0]) + ORC_A1 * (f[2 * s] - f[-s]);
    if (i == 0 || i == n - 2) return f[s] - f[0];
    return 0.0;
}

/* Line versions exported for the operator pinning tests. */
void orc_dminus(const double *f, int n, double *out)
{
    for (int i = 0; i < n; i++) out[i] = dminus_at(f + i, i, n, 1);
}
void orc_dplus(const double *f, int n, double *out)
{
    for (int i = 0; i < n; i++) out[i] = dplus_at(f + i, i, n, 1);
}

/* --------------------------------------------------------- model setup */
typedef struct {
    int nx1, nz1, npml, nnx, nnz;
    long N;
    float dx, dz, fd_dt;
    int fd_nt, real_nt, dt_factor, nts;
    double dtdx, dtdz;
    double *kappa, *kappaw1, *kappaw2, *buw1, *buw2;
    float *damp1a, *damp2a, *damp1b, *damp2b;
    float *damp1a1, *damp2a1, *damp1b1, *damp2b1;
    double *wlt;
} orc_model;

static void orc_model_free(orc_model *m)
{
    free(m->kappa); free(m->kappaw1); free(m->kappaw2); free(m->buw1); free(m->buw2);
    free(m->damp1a); free(m->damp2a); free(m->damp1b); free(m->damp2b);
    free(m->damp1a1); free(m->damp2a1); free(m->damp1b1); free(m->damp2b1);
    free(m->wlt);
}

static int orc_f64_buoyancy = 0;

/* modelling_module() setup, cpumodelling.cpp:168-281: expand vp/rho, stagger,
 * Ricker, sponge profiles (fac = 0.05f, :254) and the four scaled
 * coefficient fields (:270-281). */
static void orc_model_init(orc_model *m, int nx1, int nz1, float dx, float dz,
                           const float *vp, const float *rho, float rec_len,
                           float fd_dt, float real_dt, float fm, int npml)
{
    memset(m, 0, sizeof(*m));
    m->nx1 = nx1; m->nz1 = nz1; m->npml = npml; m->dx = dx; m->dz = dz; m->fd_dt = fd_dt;
    orc_time_params(rec_len, fd_dt, real_dt, fm, dx, dz, &m->fd_nt, &m->real_nt,
                    &m->dt_factor, &m->nts, &m->dtdx, &m->dtdz);
    int nnz = m->nnz = 2 * npml + nz1;
    int nnx = m->nnx = 2 * npml + nx1;
    long N = m->N = (long)nnz * nnx;
    float *vpe = (float *)calloc(N, sizeof(float));
    float *rhoe = (float *)calloc(N, sizeof(float));
    orc_expand(vp, nx1, nz1, npml, vpe);
    orc_expand(rho, nx1, nz1, npml, rhoe);
    float *bue = (float *)calloc(N, sizeof(float));
    float *bu1 = (float *)calloc(N, sizeof(float));
    float *bu2 = (float *)calloc(N, sizeof(float));
    m->kappa = (double *)calloc(N, sizeof(double));
    orc_stagger(vpe, rhoe, nnz, nnx, m->kappa, bue, bu1, bu2);
    m->wlt = (double *)calloc(m->nts + 1, sizeof(double)); /* [nts] stays 0 */
    orc_ricker(fm, fd_dt, m->nts, m->wlt);
    m->damp2a = (float *)calloc(nnx, 4); m->damp2a1 = (float *)calloc(nnx, 4);
    m->damp2b = (float *)calloc(nnx, 4); m->damp2b1 = (float *)calloc(nnx, 4);
    m->damp1a = (float *)calloc(nnz, 4); m->damp1a1 = (float *)calloc(nnz, 4);
    m->damp1b = (float *)calloc(nnz, 4); m->damp1b1 = (float *)calloc(nnz, 4);
    const float fac = 0.05f;
    orc_pml_a(fac, npml, nx1, m->damp2a, m->damp2a1);
    orc_pml_b(fac, npml, nx1, m->damp2b, m->damp2b1);
    orc_pml_a(fac, npml, nz1, m->damp1a, m->damp1a1);
    orc_pml_b(fac, npml, nz1, m->damp1b, m->damp1b1);
    m->kappaw1 = (double *)calloc(N, 8); m->kappaw2 = (double *)calloc(N, 8);
    m->buw1 = (double *)calloc(N, 8); m->buw2 = (double *)calloc(N, 8);
    double *b1 = NULL, *b2 = NULL;
    if (orc_f64_buoyancy) {
        /* test-only: the same staggering rules with b, bu1, bu2 in FP64 */
        double *b = (double *)calloc(N, 8);
        b1 = (double *)calloc(N, 8); b2 = (double *)calloc(N, 8);
        for (long i = 0; i < N; i++) b[i] = rhoe[i] > 0.0f ? 1.0 / (double)rhoe[i] : 1e10;
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz + 1 < nnz; iz++) {
                long id = iz + (long)ix * nnz;
                b1[id] = 0.5 * (b[id + 1] + b[id]);
            }
        for (int ix = 0; ix + 1 < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                long id = iz + (long)ix * nnz;
                b2[id] = 0.5 * (b[id + nnz] + b[id]);
            }
        for (int iz = 0; iz < nnz; iz++)
            b2[(long)(nnx - 1) * nnz + iz] = b1[(long)(nnx - 2) * nnz + iz];
        free(b);
    }
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            long id = iz + (long)ix * nnz;
            m->kappaw2[id] = m->dtdx * m->kappa[id] * m->damp2a1[ix];
            m->kappaw1[id] = m->dtdz * m->kappa[id] * m->damp1a1[iz];
            m->buw2[id] = m->dtdx * (b2 ? b2[id] : bu2[id]) * m->damp2b1[ix];
            m->buw1[id] = m->dtdz * (b1 ? b1[id] : bu1[id]) * m->damp1b1[iz];
        }
    free(b1); free(b2);
    free(vpe); free(rhoe); free(bue); free(bu1); free(bu2);
}

/* Test-only switch: evaluate the buoyancy chain (1/rho and the staggered
 * averages) in FP64 instead of the reference's FP32, so that finite
 * differences in rho are not swamped by float rounding of b.  The adjoint
 * is the same for both. */
void orc_set_f64_buoyancy(int on) { orc_f64_buoyancy = on; }

/* Grid index of a coordinate in metres, cpumodelling.cpp:288-289, 333-334
 * with cpu_set_sg() (cpu_modelling_kernels.cpp:15-22): float multiply by the
 * float reciprocal, truncate, then offset by npml. */
static long orc_cell(const orc_model *m, float x, float z)
{
    float _dx = 1.0f / m->dx, _dz = 1.0f / m->dz;
    int ix = (int)(x * _dx), iz = (int)(z * _dz);
    return (long)m->nnz * (ix + m->npml) + (iz + m->npml);
}

/* ------------------------------------------------------------------ A8 */
/* One time step, FDOperator_4() (cpumodelling.cpp:115-156) with the kernels
 * it calls.  Optionally stores, for the adjoint, the derivative fields the
 * gradient needs (dvx = D-x vx_old, dvz = D-z vz_old, dpx = D+x p, dpz =
 * D+z p). */
typedef struct { double *p, *px, *pz, *vx, *vz, *der1, *der2; } orc_state;

static void orc_step(const orc_model *m, orc_state *s, int it, long isrc,
                     double src_scale)
{
    int nnz = m->nnz, nnx = m->nnx;
    long N = m->N;
    double *vx = s->vx, *vz = s->vz, *px = s->px, *pz = s->pz, *p = s->p;
    double *der1 = s->der1, *der2 = s->der2;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) { der1[i] = 0.0; der2[i] = 0.0; }
    /* fdoperator420 / fdoperator410 */
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            long id = iz + (long)ix * nnz;
            der2[id] = dminus_at(vx + id, ix, nnx, nnz);
            der1[id] = dminus_at(vz + id, iz, nnz, 1);
        }
    /* cpu_pml_p */
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            long id = iz + (long)ix * nnz;
            px[id] = m->damp2a[ix] * px[id] + m->kappaw2[id] * der2[id];
            pz[id] = m->damp1a[iz] * pz[id] + m->kappaw1[id] * der1[id];
        }
    /* source, cpumodelling.cpp:136-140; w[nts] (read past the end by the
     * reference) is taken as 0 */
    if (it <= m->nts) {
        double w = (it < m->nts) ? m->wlt[it] : 0.0;
        px[isrc] = px[isrc] + m->fd_dt * w * m->kappa[isrc] * src_scale / (m->dx * m->dz);
    }
    /* compute_p */
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) p[i] = px[i] + pz[i];
    /* fdoperator421 / fdoperator411 */
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            long id = iz + (long)ix * nnz;
            der2[id] = dplus_at(p + id, ix, nnx, nnz);
            der1[id] = dplus_at(p + id, iz, nnz, 1);
        }
    /* cpu_pml_v */
    #pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            long id = iz + (long)ix * nnz;
            vx[id] = m->damp2b[ix] * vx[id] + m->buw2[id] * der2[id];
            vz[id] = m->damp1b[iz] * vz[id] + m->buw1[id] * der1[id];
        }
}

static void orc_state_alloc(orc_state *s, long N)
{
    s->p = (double *)calloc(N, 8); s->px = (double *)calloc(N, 8);
    s->pz = (double *)calloc(N, 8); s->vx = (double *)calloc(N, 8);
    s->vz = (double *)calloc(N, 8); s->der1 = (double *)calloc(N, 8);
    s->der2 = (double *)calloc(N, 8);
}
static void orc_state_free(orc_state *s)
{
    free(s->p); free(s->px); free(s->pz); free(s->vx); free(s->vz);
    free(s->der1); free(s->der2);
}

/* Run parameters shared by the entry points below. */
typedef struct {
    int nx, nz;
    float dx, dz;
    float rec_len, fd_dt, real_dt, fm;
    int npml;
    int reset_quirk;     /* 1: zero only the first nnz cells between shots
                            (cpumodelling.cpp:343-348); 0: fresh state */
} orc_params;

/* Shape query: fd_nt, real_nt, dt_factor, nts, nnx, nnz. */
void orc_shape(const orc_params *P, int *out6)
{
    double dtdx, dtdz;
    orc_time_params(P->rec_len, P->fd_dt, P->real_dt, P->fm, P->dx, P->dz,
                    &out6[0], &out6[1], &out6[2], &out6[3], &dtdx, &dtdz);
    out6[4] = P->nx + 2 * P->npml;
    out6[5] = P->nz + 2 * P->npml;
}

/* Expose the scaled coefficient fields and profiles (for the host-setup
 * parity tests of the product): out arrays sized N (kappa, kappaw1, kappaw2,
 * buw1, buw2) and nnx/nnz (damp2a, damp2a1, damp2b, damp2b1 / damp1a, damp1a1,
 * damp1b, damp1b1), wlt sized nts. */
void orc_setup(const orc_params *P, const float *vp, const float *rho,
               double *kappa, double *kappaw1, double *kappaw2, double *buw1,
               double *buw2, float *px8, float *pz8, double *wlt)
{
    orc_model m;
    orc_model_init(&m, P->nx, P->nz, P->dx, P->dz, vp, rho, P->rec_len,
                   P->fd_dt, P->real_dt, P->fm, P->npml);
    size_t N = (size_t)m.N;
    memcpy(kappa, m.kappa, N * 8); memcpy(kappaw1, m.kappaw1, N * 8);
    memcpy(kappaw2, m.kappaw2, N * 8); memcpy(buw1, m.buw1, N * 8);
    memcpy(buw2, m.buw2, N * 8);
    size_t nx = m.nnx, nz = m.nnz;
    memcpy(px8 + 0 * nx, m.damp2a, nx * 4); memcpy(px8 + 1 * nx, m.damp2a1, nx * 4);
    memcpy(px8 + 2 * nx, m.damp2b, nx * 4); memcpy(px8 + 3 * nx, m.damp2b1, nx * 4);
    memcpy(pz8 + 0 * nz, m.damp1a, nz * 4); memcpy(pz8 + 1 * nz, m.damp1a1, nz * 4);
    memcpy(pz8 + 2 * nz, m.damp1b, nz * 4); memcpy(pz8 + 3 * nz, m.damp1b1, nz * 4);
    memcpy(wlt, m.wlt, (size_t)m.nts * 8);
    orc_model_free(&m);
}

/* Test-only: change what the reset between shots clears in reset_quirk
 * mode, so the tests can plant a slip and see the reference's two-shot
 * output reject it.  cells < 0 restores the reference's nnz cells; fields
 * is a mask of p, px, pz, vx, vz (bits 0..4; 31 = all, the reference). */
static long orc_reset_cells = -1;
static int orc_reset_fields = 31;
void orc_set_reset_plant(long cells, int fields)
{
    orc_reset_cells = cells;
    orc_reset_fields = fields;
}

/* ------------------------------------------------------------ forward */
/* modelling_module() shot and time loops, cpumodelling.cpp:326-380, plus
 * cpu_record (:24-31, sampling p after step kt when (kt+1)%dt_factor==0) and
 * cpu_transpose (:370-381).  src_xz[2*is] = (x, z) in metres; receivers are
 * concatenated per shot in rec_xz.  seis receives, per shot, ng traces of
 * real_nt float samples (trace-major, as the reference's output file).
 * src_amp scales the injected wavelet (1.0 = reference); it exists for the
 * linearity invariant. Returns 0 or a negative status. */
int orc_forward(const orc_params *P, const float *vp, const float *rho,
                int ns, const float *src_xz, const int *nrec,
                const float *rec_xz, double src_amp, float *seis)
{
    orc_model m;
    orc_model_init(&m, P->nx, P->nz, P->dx, P->dz, vp, rho, P->rec_len,
                   P->fd_dt, P->real_dt, P->fm, P->npml);
    orc_state s;
    orc_state_alloc(&s, m.N);
    long rec_off = 0, seis_off = 0;
    for (int is = 0; is < ns; is++) {
        int ng = nrec[is];
        long isrc = orc_cell(&m, src_xz[2 * is], src_xz[2 * is + 1]);
        long *gxz = (long *)malloc(sizeof(long) * (ng > 0 ? ng : 1));
        for (int g = 0; g < ng; g++)
            gxz[g] = orc_cell(&m, rec_xz[2 * (rec_off + g)], rec_xz[2 * (rec_off + g) + 1]);
        long clear = P->reset_quirk ? (orc_reset_cells >= 0 ? orc_reset_cells : m.nnz) : m.N;
        int fields = P->reset_quirk ? orc_reset_fields : 31;
        if (is == 0) { clear = m.N; fields = 31; }
        if (clear > m.N) clear = m.N;
        for (long i = 0; i < clear; i++) {
            if (fields & 1) s.p[i] = 0.0;
            if (fields & 2) s.px[i] = 0.0;
            if (fields & 4) s.pz[i] = 0.0;
            if (fields & 8) s.vx[i] = 0.0;
            if (fields & 16) s.vz[i] = 0.0;
        }
        float *out = seis + seis_off;
        memset(out, 0, sizeof(float) * (size_t)ng * m.real_nt);
        for (int kt = 0; kt < m.fd_nt; kt++) {
            orc_step(&m, &s, kt, isrc, src_amp);
            if ((kt + 1) % m.dt_factor == 0) {
                int indx = (kt + 1) / m.dt_factor - 1;
                if (indx < m.real_nt)
                    for (int g = 0; g < ng; g++)
                        out[(long)g * m.real_nt + indx] = (float)s.p[gxz[g]];
            }
        }
        free(gxz);
        rec_off += ng;
        seis_off += (long)ng * m.real_nt;
    }
    orc_state_free(&s);
    orc_model_free(&m);
    return 0;
}

/* CPU baseline timing: builds the model of P, then runs nsteps time steps of
 * the reference-structured FP64 loop (9 sweeps per step, static OpenMP) with a
 * source at the domain centre, and returns the seconds spent in the steps
 * alone (setup excluded).  Used by bench.py's cpu_baseline leg. */
#include <omp.h>
double orc_time_steps(const orc_params *P, const float *vp, const float *rho, int nsteps,
                      int *threads)
{
    orc_model m;
    orc_model_init(&m, P->nx, P->nz, P->dx, P->dz, vp, rho, P->rec_len, P->fd_dt,
                   P->real_dt, P->fm, P->npml);
    orc_state s;
    orc_state_alloc(&s, m.N);
    long isrc = (long)m.nnz * (m.nnx / 2) + m.nnz / 2;
    *threads = omp_get_max_threads();
    double t0 = omp_get_wtime();
    for (int kt = 0; kt < nsteps; kt++) orc_step(&m, &s, kt, isrc, 1.0);
    double t = omp_get_wtime() - t0;
    orc_state_free(&s);
    orc_model_free(&m);
    return t;
}

/* =================================================================== FWI */
/* The reference has no adjoint (job_t.adj is declared and never used,
 * src/src_cpu/modelling.h:51).  What follows is the exact discrete adjoint of
 * the FP64 forward above (orc_step), derived operator by operator, for a
 * linear functional of the recorded pressure:
 *
 *   J(vp, rho) = sum over shots, receivers g, samples k of adj[g][k] * p_k[G_g]
 *
 * where p_k is the FP64 pressure the forward samples (before its cast to
 * float).  For a misfit L(d) pass adj = dL/dd.  The forward states are kept
 * in full (small grids only).  It is pinned by central finite differences of
 * orc_functional and by the transposition identities in the tests. */

/* Transposes written as scatters of the forward stencils (so they share no
 * code with any gather form): out[k] += w(i, k) * g[i] for every tap of
 * D-(i) / D+(i) on a line of n points with stride s. */
static void dminus_T_scatter(const double *g, double *out, int n, long s, long base)
{
    for (int i = 0; i < n; i++) {
        const double gi = g[base + i * s];
        if (gi == 0.0) continue;
        if (i >= 2 && i <= n - 2) {
            out[base + i * s] += ORC_A0 * gi;
            out[base + (i - 1) * s] -= ORC_A0 * gi;
            out[base + (i + 1) * s] += ORC_A1 * gi;
            out[base + (i - 2) * s] -= ORC_A1 * gi;
        } else if (i == 1 || i == n - 1) {
            out[base + i * s] += gi;
            out[base + (i - 1) * s] -= gi;
        }
    }
}
static void dplus_T_scatter(const double *g, double *out, int n, long s, long base)
{
    for (int i = 0; i < n; i++) {
        const double gi = g[base + i * s];
        if (gi == 0.0) continue;
        if (i >= 1 && i <= n - 3) {
            out[base + (i + 1) * s] += ORC_A0 * gi;
            out[base + i * s] -= ORC_A0 * gi;
            out[base + (i + 2) * s] += ORC_A1 * gi;
            out[base + (i - 1) * s] -= ORC_A1 * gi;
        } else if (i == 0 || i == n - 2) {
            out[base + (i + 1) * s] += gi;
            out[base + i * s] -= gi;
        }
    }
}

/* Line operators exported for the transposition tests. */
void orc_dminus_T(const double *g, int n, double *out)
{
    memset(out, 0, sizeof(double) * n);
    dminus_T_scatter(g, out, n, 1, 0);
}
void orc_dplus_T(const double *g, int n, double *out)
{
    memset(out, 0, sizeof(double) * n);
    dplus_T_scatter(g, out, n, 1, 0);
}

/* J for one model (double accumulation of the FP64 samples). */
double orc_functional(const orc_params *P, const float *vp, const float *rho, int ns,
                      const float *src_xz, const int *nrec, const float *rec_xz,
                      const float *adj)
{
    orc_model m;
    orc_model_init(&m, P->nx, P->nz, P->dx, P->dz, vp, rho, P->rec_len, P->fd_dt,
                   P->real_dt, P->fm, P->npml);
    orc_state s;
    orc_state_alloc(&s, m.N);
    double J = 0.0;
    long rec_off = 0, a_off = 0;
    for (int is = 0; is < ns; is++) {
        const int ng = nrec[is];
        long isrc = orc_cell(&m, src_xz[2 * is], src_xz[2 * is + 1]);
        for (long i = 0; i < m.N; i++) { s.p[i] = s.px[i] = s.pz[i] = s.vx[i] = s.vz[i] = 0.0; }
        for (int kt = 0; kt < m.fd_nt; kt++) {
            orc_step(&m, &s, kt, isrc, 1.0);
            if ((kt + 1) % m.dt_factor == 0) {
                const int k = (kt + 1) / m.dt_factor - 1;
                if (k < m.real_nt)
                    for (int g = 0; g < ng; g++) {
                        long gc = orc_cell(&m, rec_xz[2 * (rec_off + g)], rec_xz[2 * (rec_off + g) + 1]);
                        J += (double)adj[a_off + (long)g * m.real_nt + k] * s.p[gc];
                    }
            }
        }
        rec_off += ng;
        a_off += (long)ng * m.real_nt;
    }
    orc_state_free(&s);
    orc_model_free(&m);
    return J;
}

/* dJ/dvp and dJ/drho on the model grid (x-major nx*nz, FP64), plus the
 * intermediate gradients on the padded grid when the pointers are given:
 * g_kappa (dJ/dkappa), g_bu1, g_bu2 (dJ/d of the staggered buoyancies).
 *
 * ckpt_k > 0 checkpoints the forward state (px, pz, vx, vz) every ckpt_k
 * steps and recomputes each segment of ckpt_k steps before its adjoint
 * steps, so memory is 4 N doubles per checkpoint plus 3 N ckpt_k for the
 * segment; ckpt_k <= 0 (or >= fd_nt) keeps every state of the record
 * (3 N fd_nt doubles).  The recomputed states are the same bits as the
 * first sweep's (same steps, same order), so the result does not depend on
 * ckpt_k. */
int orc_gradient_k(const orc_params *P, const float *vp, const float *rho, int ns,
                   const float *src_xz, const int *nrec, const float *rec_xz, const float *adj,
                   int ckpt_k, double *g_vp, double *g_rho, double *gk_out, double *gb1_out,
                   double *gb2_out)
{
    orc_model m;
    orc_model_init(&m, P->nx, P->nz, P->dx, P->dz, vp, rho, P->rec_len, P->fd_dt,
                   P->real_dt, P->fm, P->npml);
    const int nnz = m.nnz, nnx = m.nnx, nt = m.fd_nt;
    const long N = m.N;
    const int K = (ckpt_k <= 0 || ckpt_k > nt) ? nt : ckpt_k;
    const int nseg = (nt + K - 1) / K;
    double *gk = (double *)calloc(N, 8), *gb1 = (double *)calloc(N, 8), *gb2 = (double *)calloc(N, 8);
    /* segment ring: vx, vz before step n and p after the p-phase of n */
    double *Vx = (double *)malloc(sizeof(double) * N * K);
    double *Vz = (double *)malloc(sizeof(double) * N * K);
    double *Pn = (double *)malloc(sizeof(double) * N * K);
    /* checkpoint c: px, pz, vx, vz entering step c*K (c = 0 is the zero state) */
    double *ck = (double *)calloc((size_t)4 * N * nseg, 8);
    double *lpx = (double *)calloc(N, 8), *lpz = (double *)calloc(N, 8);
    double *lvx = (double *)calloc(N, 8), *lvz = (double *)calloc(N, 8);
    double *lp = (double *)calloc(N, 8), *t1 = (double *)calloc(N, 8), *t2 = (double *)calloc(N, 8);
    double *nvx = (double *)calloc(N, 8), *nvz = (double *)calloc(N, 8);
    if (!Vx || !Vz || !Pn || !ck) return -30;
    orc_state s;
    orc_state_alloc(&s, m.N);
    long rec_off = 0, a_off = 0;
    for (int is = 0; is < ns; is++) {
        const int ng = nrec[is];
        const long isrc = orc_cell(&m, src_xz[2 * is], src_xz[2 * is + 1]);
        long *gc = (long *)malloc(sizeof(long) * (ng > 0 ? ng : 1));
        for (int g = 0; g < ng; g++)
            gc[g] = orc_cell(&m, rec_xz[2 * (rec_off + g)], rec_xz[2 * (rec_off + g) + 1]);
        /* forward sweep: checkpoints 1 .. nseg-1 */
        for (long i = 0; i < N; i++) { s.p[i] = s.px[i] = s.pz[i] = s.vx[i] = s.vz[i] = 0.0; }
        memset(ck, 0, sizeof(double) * 4 * N);
        for (int n = 0; n < (nseg - 1) * K; n++) {
            orc_step(&m, &s, n, isrc, 1.0);
            if ((n + 1) % K == 0) {
                double *c = ck + (size_t)4 * N * ((n + 1) / K);
                memcpy(c, s.px, 8 * N); memcpy(c + N, s.pz, 8 * N);
                memcpy(c + 2 * N, s.vx, 8 * N); memcpy(c + 3 * N, s.vz, 8 * N);
            }
        }
        for (long i = 0; i < N; i++) { lpx[i] = lpz[i] = lvx[i] = lvz[i] = 0.0; }
        for (int seg = nseg - 1; seg >= 0; seg--) {
            const int n0 = seg * K, n1 = (n0 + K < nt) ? n0 + K : nt;
            /* recompute the segment's states from its checkpoint */
            const double *c = ck + (size_t)4 * N * seg;
            memcpy(s.px, c, 8 * N); memcpy(s.pz, c + N, 8 * N);
            memcpy(s.vx, c + 2 * N, 8 * N); memcpy(s.vz, c + 3 * N, 8 * N);
            for (int n = n0; n < n1; n++) {
                memcpy(Vx + (long)(n - n0) * N, s.vx, sizeof(double) * N);
                memcpy(Vz + (long)(n - n0) * N, s.vz, sizeof(double) * N);
                orc_step(&m, &s, n, isrc, 1.0);
                memcpy(Pn + (long)(n - n0) * N, s.p, sizeof(double) * N);
            }
            for (int n = n1 - 1; n >= n0; n--) {
                const double *vxo = Vx + (long)(n - n0) * N, *vzo = Vz + (long)(n - n0) * N;
                const double *pn = Pn + (long)(n - n0) * N;
                /* record (cpu_record after step n) */
                #pragma omp parallel for schedule(static)
                for (long i = 0; i < N; i++) lp[i] = 0.0;
                if ((n + 1) % m.dt_factor == 0) {
                    const int k = (n + 1) / m.dt_factor - 1;
                    if (k < m.real_nt)
                        for (int g = 0; g < ng; g++) lp[gc[g]] += adj[a_off + (long)g * m.real_nt + k];
                }
                /* v-update: vx' = damp2b vx + buw2 Dx+ p ;  vz' = damp1b vz + buw1 Dz+ p */
                #pragma omp parallel for schedule(static)
                for (int ix = 0; ix < nnx; ix++)
                    for (int iz = 0; iz < nnz; iz++) {
                        const long id = iz + (long)ix * nnz;
                        t1[id] = m.buw2[id] * lvx[id];
                        t2[id] = m.buw1[id] * lvz[id];
                        gb2[id] += lvx[id] * dplus_at(pn + id, ix, nnx, nnz) * m.dtdx * m.damp2b1[ix];
                        gb1[id] += lvz[id] * dplus_at(pn + id, iz, nnz, 1) * m.dtdz * m.damp1b1[iz];
                        nvx[id] = m.damp2b[ix] * lvx[id];
                        nvz[id] = m.damp1b[iz] * lvz[id];
                    }
                #pragma omp parallel for schedule(static)
                for (int iz = 0; iz < nnz; iz++) dplus_T_scatter(t1, lp, nnx, nnz, iz);
                #pragma omp parallel for schedule(static)
                for (int ix = 0; ix < nnx; ix++) dplus_T_scatter(t2, lp, nnz, 1, (long)ix * nnz);
                /* p = px + pz, then the source on px */
                #pragma omp parallel for schedule(static)
                for (long i = 0; i < N; i++) { lpx[i] += lp[i]; lpz[i] += lp[i]; }
                if (n < m.nts) gk[isrc] += lpx[isrc] * (m.fd_dt * m.wlt[n] / (m.dx * m.dz));
                /* p-update: px' = damp2a px + kappaw2 Dx- vx ;  pz' = damp1a pz + kappaw1 Dz- vz */
                #pragma omp parallel for schedule(static)
                for (int ix = 0; ix < nnx; ix++)
                    for (int iz = 0; iz < nnz; iz++) {
                        const long id = iz + (long)ix * nnz;
                        gk[id] += lpx[id] * dminus_at(vxo + id, ix, nnx, nnz) * m.dtdx * m.damp2a1[ix]
                                + lpz[id] * dminus_at(vzo + id, iz, nnz, 1) * m.dtdz * m.damp1a1[iz];
                        t1[id] = m.kappaw2[id] * lpx[id];
                        t2[id] = m.kappaw1[id] * lpz[id];
                        lpx[id] *= m.damp2a[ix];
                        lpz[id] *= m.damp1a[iz];
                    }
                #pragma omp parallel for schedule(static)
                for (int iz = 0; iz < nnz; iz++) dminus_T_scatter(t1, nvx, nnx, nnz, iz);
                #pragma omp parallel for schedule(static)
                for (int ix = 0; ix < nnx; ix++) dminus_T_scatter(t2, nvz, nnz, 1, (long)ix * nnz);
                memcpy(lvx, nvx, sizeof(double) * N);
                memcpy(lvz, nvz, sizeof(double) * N);
            }
        }
        free(gc);
        rec_off += ng;
        a_off += (long)ng * m.real_nt;
    }
    if (gk_out) memcpy(gk_out, gk, sizeof(double) * N);
    if (gb1_out) memcpy(gb1_out, gb1, sizeof(double) * N);
    if (gb2_out) memcpy(gb2_out, gb2, sizeof(double) * N);

    /* Chain to the padded vp, rho (cpu_stagger, cpu_modelling_kernels.cpp:64-121):
     * bu2's last column copies bu1's column nnx-2; bu1's last row is constant 0;
     * bu1 = (b[iz] + b[iz+1])/2, bu2 = (b[ix] + b[ix+1])/2, b = 1/rho;
     * kappa = vp^2 rho. */
    float *vpe = (float *)calloc(N, 4), *rhoe = (float *)calloc(N, 4);
    orc_expand(vp, P->nx, P->nz, P->npml, vpe);
    orc_expand(rho, P->nx, P->nz, P->npml, rhoe);
    for (int iz = 0; iz < nnz; iz++)
        gb1[(long)(nnx - 2) * nnz + iz] += gb2[(long)(nnx - 1) * nnz + iz];
    double *gb = (double *)calloc(N, 8);
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz + 1 < nnz; iz++) {
            const long id = iz + (long)ix * nnz;
            gb[id] += 0.5 * gb1[id];
            gb[id + 1] += 0.5 * gb1[id];
        }
    for (int ix = 0; ix + 1 < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            const long id = iz + (long)ix * nnz;
            gb[id] += 0.5 * gb2[id];
            gb[id + nnz] += 0.5 * gb2[id];
        }
    for (long i = 0; i < (long)P->nx * P->nz; i++) { g_vp[i] = 0.0; g_rho[i] = 0.0; }
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            const long id = iz + (long)ix * nnz;
            const double v = vpe[id], r = rhoe[id];
            const double gv = 2.0 * v * r * gk[id];
            double gr = v * v * gk[id];
            if (rhoe[id] > 0.0f) gr += -gb[id] / (r * r);
            int mx = ix - P->npml, mz = iz - P->npml;
            mx = mx < 0 ? 0 : (mx >= P->nx ? P->nx - 1 : mx);
            mz = mz < 0 ? 0 : (mz >= P->nz ? P->nz - 1 : mz);
            g_vp[(long)mx * P->nz + mz] += gv;
            g_rho[(long)mx * P->nz + mz] += gr;
        }
    free(vpe); free(rhoe); free(gb);
    free(gk); free(gb1); free(gb2); free(Vx); free(Vz); free(Pn); free(ck);
    free(lpx); free(lpz); free(lvx); free(lvz); free(lp); free(t1); free(t2); free(nvx); free(nvz);
    orc_state_free(&s);
    orc_model_free(&m);
    return 0;
}

/* The same gradient with every forward state of the record kept (no
 * recomputation). */
int orc_gradient(const orc_params *P, const float *vp, const float *rho, int ns,
                 const float *src_xz, const int *nrec, const float *rec_xz, const float *adj,
                 double *g_vp, double *g_rho, double *gk_out, double *gb1_out, double *gb2_out)
{
    return orc_gradient_k(P, vp, rho, ns, src_xz, nrec, rec_xz, adj, 0, g_vp, g_rho, gk_out,
                          gb1_out, gb2_out);
}
