// modelling_module.cpp -- the reference's per-rank entry point
// (modelling.h:106, cpumodelling.cpp:158-383) over the C ABI: same globals in,
// same seismogram files out, the time loop on the GPU.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "sam2d.h"
#include "sam2d_modelling.h"

static std::string restart_file()
{
    // <tmppath><jobname>checkpoint_rank<r>.txt (cpumodelling.cpp:295-297); one
    // process, rank 0.
    return std::string(job_sp->tmppath) + job_sp->jobname + "checkpoint_rank0.txt";
}

extern "C" int modelling_module(int my_nsrc, geo2d_t *geo)
{
    const int nx = mod_sp->nx, nz = mod_sp->nz;
    sam2d_grid g = {nx, nz, mod_sp->dx, mod_sp->dz, wave_sp->rec_len, wave_sp->fd_dt_sec,
                    wave_sp->real_dt_sec, wave_sp->dom_freq, sam2d_opts.npml, job_sp->fdop};
    // New / Restart (cpumodelling.cpp:299-321)
    int istart = 0;
    if (std::strcmp(job_sp->jbtype, "New") == 0) {
        istart = 0;
    } else if (std::strcmp(job_sp->jbtype, "Restart") == 0) {
        FILE *fp = std::fopen(restart_file().c_str(), "r");
        if (!fp) {
            std::cerr << "\n Error !!! in openining log file: " << restart_file();
            return -10;
        }
        if (std::fscanf(fp, "%d", &istart) != 1) istart = 0;
        std::fclose(fp);
    } else {
        std::cerr << "\n Error !!! Unknow job status";
        return -11;
    }
    std::vector<float> vp((size_t)nx * nz), rho((size_t)nx * nz);
    for (int ix = 0; ix < nx; ix++) {
        std::memcpy(&vp[(size_t)ix * nz], mod_sp->vel2d[ix], sizeof(float) * nz);
        std::memcpy(&rho[(size_t)ix * nz], mod_sp->rho2d[ix], sizeof(float) * nz);
    }
    sam2d_ctx *ctx = nullptr;
    int rc = sam2d_create(&g, vp.data(), rho.data(), &ctx);
    if (rc) {
        std::cerr << "\n " << sam2d_last_error();
        return rc;
    }
    sam2d_dims dims;
    sam2d_get_dims(ctx, &dims);
    std::cout << "\n fd_nt: " << dims.fd_nt << "   real_nt: " << dims.real_nt
              << "   DT factor: " << dims.dt_factor << "\n nnx: " << dims.nnx
              << "   nnz: " << dims.nnz << "    nts: " << dims.nts << "\n Started Modelling.......";
    std::vector<float> seis;
    for (int is = istart; is < my_nsrc && rc == 0; is++) {
        sam2d_shot sh = {geo->src2d_sp[is].x, geo->src2d_sp[is].z, geo->nrec[is],
                         (const float *)geo->rec2d_sp[is]};
        seis.assign((size_t)sh.nrec * dims.real_nt, 0.0f);
        rc = sam2d_forward(ctx, &sh, 1, seis.data());
        if (rc) {
            std::cerr << "\n " << sam2d_last_error();
            break;
        }
        // cpumodelling.cpp:370-374: <tmppath><jobname>seismogram.bin, trace-major
        std::string out = std::string(job_sp->tmppath) + job_sp->jobname;
        if (sam2d_opts.per_shot_files) out += "shot" + std::to_string(is) + "_";
        out += "seismogram.bin";
        FILE *fp = std::fopen(out.c_str(), "wb");
        if (!fp) { rc = -2; break; }
        std::fwrite(seis.data(), sizeof(float), seis.size(), fp);
        std::fclose(fp);
        if (sam2d_opts.write_restart) {
            FILE *fr = std::fopen(restart_file().c_str(), "w");
            if (fr) { std::fprintf(fr, "%d\n", is + 1); std::fclose(fr); }
        }
        sam2d_stats st;
        sam2d_get_stats(ctx, &st);
        std::cout << "\n Shot remaining: " << my_nsrc - (is + 1) << "   ("
                  << st.cell_updates / (st.kernel_ms * 1e6) << " Gcell-updates/s)";
    }
    sam2d_destroy(ctx);
    return rc;
}
