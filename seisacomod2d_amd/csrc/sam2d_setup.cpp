// sam2d_setup.cpp -- see sam2d_setup.h.
#include "sam2d_setup.h"
#include "sam2d_round.h"

#include <cmath>
#include <cstdint>
#include <cstring>

namespace sam2d {

static const double kPi = 3.141592654;   // cpu_modelling_kernels.cpp:7
static const float kSpongeFac = 0.05f;   // xxfac, cpumodelling.cpp:254

TimeAxis time_axis(const sam2d_grid &g)
{
    // Every quotient below is a float operation in the reference
    // (cpumodelling.cpp:172-188); ceil of a float and std::sqrt(6.0f) take
    // the float overloads there.
    TimeAxis t;
    const float inv_dx = 1.0f / g.dx, inv_dz = 1.0f / g.dz;
    t.fd_nt = (int)std::ceil(g.rec_len / g.fd_dt);
    t.real_nt = (int)std::ceil(g.rec_len / g.real_dt);
    t.dtdx = (double)(g.fd_dt * inv_dx);
    t.dtdz = (double)(g.fd_dt * inv_dz);
    t.dt_factor = (int)(g.real_dt / g.fd_dt);
    const double td = (double)std::sqrt(6.0f) / (kPi * (double)g.dom_freq);
    int nts = (int)(5.0f * td / g.fd_dt) + 1;
    t.nts = nts > t.fd_nt ? t.fd_nt : nts;
    return t;
}

void sponge_at(int i, int npml, int n, bool half, float &decay, float &scale)
{
    // Distance into the sponge, in cells, as the reference forms it in float.
    float x;
    if (!half) {
        // pressure grid: left i < npml at npml-1-i; right i >= npml+n at i-npml-n
        if (i < npml) x = (float)(npml - (i + 1));
        else if (i >= npml + n) x = (float)(i - npml - n);
        else { decay = 1.0f; scale = 1.0f; return; }
    } else {
        // velocity grid: left i < npml-1 at npml-1.5-i; ones on [npml-1, npml+n);
        // right i >= npml+n at i-npml-n+0.5
        if (i < npml - 1) x = (float)(npml - 1 - (i + 1)) + 0.5f;
        else if (i >= npml + n) x = (float)(i - npml - n) + 0.5f;
        else { decay = 1.0f; scale = 1.0f; return; }
    }
    const float t = kSpongeFac * x;
    const float t2 = t * t;
    decay = std::exp(-t2);                       // float overload
    scale = (float)std::exp(-0.5 * (double)t2);  // double
}

std::vector<double> ricker(float fm, float dt, int nts)
{
    std::vector<double> w(nts > 0 ? nts : 0);
    const double t0 = 1.5 * std::sqrt(6.0) / (kPi * (double)fm);
    for (int i = 0; i < nts; i++) {
        const double t = (double)((float)i * dt);
        const double a = kPi * (double)fm * (t - t0);
        const double a2 = a * a;
        w[i] = (1.0 - 2.0 * a2) * std::exp(-a2);
    }
    return w;
}

void coord_to_cell(const sam2d_grid &g, float x, float z, int &ix, int &iz)
{
    const float inv_dx = 1.0f / g.dx, inv_dz = 1.0f / g.dz;
    ix = (int)(x * inv_dx);
    iz = (int)(z * inv_dz);
}

// The coefficient fields are rounded to FP32 without bias (sam2d_round.h):
// with round-to-nearest, every cell of a homogeneous region would carry the
// same relative error, a systematic wave-speed error whose phase error
// reaches 2.4e-5 relative L2 after 8000 steps of the 3201 x 1201 salt-model
// window at 7.62 m against the reference's FP64 seismograms; with unbiased
// rounding the cells' errors average out along a wavelength and the same
// window comes to 1.3e-6.

void build_domain(const sam2d_grid &g, const TimeAxis &ta, const float *vp, const float *rho,
                  int wx0, int w, DomainHost &out)
{
    const int P = g.npml, nz = g.nz;
    const int nnz = nz + 2 * P, nnx = w + 2 * P;
    const size_t N = (size_t)nnz * nnx;
    out.wx0 = wx0; out.w = w; out.nnx = nnx; out.nnz = nnz;

    // Padded vp / rho: each padded cell takes the model value of the nearest
    // interior cell (expand's edge replication; corners come from the
    // replicated columns, which is the same nearest-cell rule).
    std::vector<float> vpe(N), rhoe(N);
    for (int ix = 0; ix < nnx; ix++) {
        int mx = ix - P;
        mx = mx < 0 ? 0 : (mx >= w ? w - 1 : mx);
        const size_t src_col = (size_t)(wx0 + mx) * nz;
        for (int iz = 0; iz < nnz; iz++) {
            int mz = iz - P;
            mz = mz < 0 ? 0 : (mz >= nz ? nz - 1 : mz);
            vpe[(size_t)ix * nnz + iz] = vp[src_col + mz];
            rhoe[(size_t)ix * nnz + iz] = rho[src_col + mz];
        }
    }

    // Bulk modulus (FP64) and buoyancy (FP32), cpu_stagger.
    out.kappa.assign(N, 0.0);
    std::vector<float> b(N);
    for (size_t i = 0; i < N; i++) {
        out.kappa[i] = ((double)vpe[i] * (double)vpe[i]) * (double)rhoe[i];
        b[i] = rhoe[i] > 0.0f ? 1.0f / rhoe[i] : 1.0e10f;
    }
    // bu1: z-staggered average; the last row of every column is 0 (the
    // reference copies it from bu2 before bu2 is filled).  bu2: x-staggered
    // average; the last column copies bu1's column nnx-2.
    std::vector<float> bu1(N, 0.0f), bu2(N, 0.0f);
    for (int ix = 0; ix < nnx; ix++) {
        const size_t c = (size_t)ix * nnz;
        for (int iz = 0; iz + 1 < nnz; iz++) bu1[c + iz] = 0.5f * (b[c + iz + 1] + b[c + iz]);
        bu1[c + nnz - 1] = 0.0f;
    }
    for (int ix = 0; ix + 1 < nnx; ix++) {
        const size_t c = (size_t)ix * nnz;
        for (int iz = 0; iz < nnz; iz++) bu2[c + iz] = 0.5f * (b[c + nnz + iz] + b[c + iz]);
    }
    std::memcpy(&bu2[(size_t)(nnx - 1) * nnz], &bu1[(size_t)(nnx - 2) * nnz], sizeof(float) * nnz);

    out.B = b;
    out.dtdx = ta.dtdx;
    out.dtdz = ta.dtdz;
    out.kz_alias = (ta.dtdx == ta.dtdz);
    out.KX.resize(N); out.BX.resize(N); out.BZ.resize(N);
    if (!out.kz_alias) out.KZ.resize(N); else out.KZ.clear();
    // The reference's FP64 coefficients (cpumodelling.cpp:270-281) without
    // the 1-D sponge factors, which the kernels apply: dtdx*kappa, dtdz*kappa,
    // dtdx*bu2, dtdz*bu1, each rounded to FP32 without bias.  The rounding
    // depends on the cell's position in its domain only, so a window of the
    // model gives the same fields as the cropped model.
    for (int ix = 0; ix < nnx; ix++)
        for (int iz = 0; iz < nnz; iz++) {
            const size_t i = (size_t)ix * nnz + iz;
            out.KX[i] = round_unbiased(ta.dtdx * out.kappa[i], cell_draw(RND_KX, ix, iz));
            if (!out.kz_alias)
                out.KZ[i] = round_unbiased(ta.dtdz * out.kappa[i], cell_draw(RND_KZ, ix, iz));
            // dtdx, dtdz are float products widened to double: dtdx * bu2 is
            // the exact product of two floats
            out.BX[i] = round_unbiased_prod_pos((float)ta.dtdx, bu2[i], cell_draw(RND_BX, ix, iz));
            out.BZ[i] = round_unbiased_prod_pos((float)ta.dtdz, bu1[i], cell_draw(RND_BZ, ix, iz));
        }

    out.prof_x.assign(4 * (size_t)nnx, 0.0f);
    out.prof_z.assign(4 * (size_t)nnz, 0.0f);
    for (int i = 0; i < nnx; i++) {
        sponge_at(i, P, w, false, out.prof_x[i], out.prof_x[nnx + i]);
        sponge_at(i, P, w, true, out.prof_x[2 * nnx + i], out.prof_x[3 * nnx + i]);
    }
    for (int i = 0; i < nnz; i++) {
        sponge_at(i, P, nz, false, out.prof_z[i], out.prof_z[nnz + i]);
        sponge_at(i, P, nz, true, out.prof_z[2 * nnz + i], out.prof_z[3 * nnz + i]);
    }
}

}  // namespace sam2d
