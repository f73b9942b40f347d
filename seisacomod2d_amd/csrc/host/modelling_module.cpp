// modelling_module.cpp -- the reference's per-rank entry point
// (modelling.h:106, cpumodelling.cpp:158-383) over the C ABI: same globals in,
// same seismogram files out, the time loop on the GPU.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
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

static std::string output_base(int is)
{
    std::string out = std::string(job_sp->tmppath) + job_sp->jobname;
    if (sam2d_opts.per_shot_files) out += "shot" + std::to_string(is) + "_";
    return out + "seismogram";
}

// ------------------------------------------------------------------ SU output
// One SEG-Y trace identification header (240 bytes, native endian as SU
// keeps it), the field offsets of the trace header standard that
// src/src_cpu/segy.h:27-310 declares.
namespace {
struct TraceHeader {
    unsigned char b[240];
    TraceHeader() { std::memset(b, 0, sizeof b); }
    void i32(int off, int32_t v) { std::memcpy(b + off, &v, 4); }
    void i16(int off, int16_t v) { std::memcpy(b + off, &v, 2); }
    void u16(int off, uint16_t v) { std::memcpy(b + off, &v, 2); }
};
enum {
    TRACL = 0, TRACR = 4, FLDR = 8, TRACF = 12, TRID = 28, OFFSET = 36, GELEV = 40, SELEV = 44,
    SDEPTH = 48, SCALEL = 68, SCALCO = 70, SX = 72, GX = 80, COUNIT = 88, NS = 114, DT = 116
};
// Coordinates in centimetres with the SEG-Y scalar -100 (divide by 100).
int32_t cm(float metres) { return (int32_t)std::lround((double)metres * 100.0); }
}  // namespace

static int write_su(const std::string &path, bool create, int is, const geo2d_t *geo,
                    const float *seis, int real_nt, float real_dt, long long *ntraces)
{
    if (real_nt > 65535) {
        std::cerr << "\n Error!!! SU output holds at most 65535 samples per trace (" << real_nt << ")";
        return -6;
    }
    FILE *fp = std::fopen(path.c_str(), create ? "wb" : "ab");
    if (!fp) {
        std::cerr << "\n Error!!! Unable to open " << path;
        return -2;
    }
    if (!create && *ntraces == 0) {
        // appending to a file of an earlier run (a restarted job): continue
        // its trace numbering
        std::fseek(fp, 0, SEEK_END);
        *ntraces = std::ftell(fp) / (240 + 4 * (long)real_nt);
    }
    const crd2d_t src = geo->src2d_sp[is];
    const int nrec = geo->nrec[is];
    for (int g = 0; g < nrec; g++) {
        const crd2d_t r = geo->rec2d_sp[is][g];
        TraceHeader h;
        const long long t = ++*ntraces;
        h.i32(TRACL, (int32_t)t);
        h.i32(TRACR, (int32_t)t);
        h.i32(FLDR, is + 1);
        h.i32(TRACF, g + 1);
        h.i16(TRID, 1);
        h.i32(OFFSET, cm(r.x) - cm(src.x));
        h.i32(GELEV, -cm(r.z));
        h.i32(SELEV, -cm(src.z));
        h.i32(SDEPTH, cm(src.z));
        h.i16(SCALEL, -100);
        h.i16(SCALCO, -100);
        h.i32(SX, cm(src.x));
        h.i32(GX, cm(r.x));
        h.i16(COUNIT, 1);
        h.u16(NS, (uint16_t)real_nt);
        h.u16(DT, (uint16_t)std::lround((double)real_dt * 1e6));
        if (std::fwrite(h.b, 1, 240, fp) != 240 ||
            std::fwrite(seis + (size_t)g * real_nt, sizeof(float), real_nt, fp) != (size_t)real_nt) {
            std::fclose(fp);
            return -2;
        }
    }
    std::fclose(fp);
    return 0;
}

extern "C" int write_shot_output(int is, const geo2d_t *geo, const float *seis, int real_nt,
                                 float real_dt, int first_shot, long long *ntraces)
{
    // cpumodelling.cpp:370-374: <tmppath><jobname>seismogram.bin, trace-major
    const std::string base = output_base(is);
    FILE *fp = std::fopen((base + ".bin").c_str(), "wb");
    if (!fp) {
        std::cerr << "\n Error!!! Unable to open " << base << ".bin";
        return -2;
    }
    const size_t n = (size_t)geo->nrec[is] * real_nt;
    const bool ok = std::fwrite(seis, sizeof(float), n, fp) == n;
    std::fclose(fp);
    if (!ok) return -2;
    if (sam2d_opts.su_output) {
        long long local = 0;
        long long *count = ntraces ? ntraces : &local;
        const bool create = first_shot || sam2d_opts.per_shot_files;
        if (sam2d_opts.per_shot_files) *count = 0;
        return write_su(base + ".su", create, is, geo, seis, real_nt, real_dt, count);
    }
    return 0;
}

extern "C" int sam2d_modelling_module(int my_nsrc, geo2d_t *geo)
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
    if (sam2d_opts.reset_quirk) sam2d_set_option(ctx, "reset_quirk", 1);
    sam2d_dims dims;
    sam2d_get_dims(ctx, &dims);
    std::cout << "\n fd_nt: " << dims.fd_nt << "   real_nt: " << dims.real_nt
              << "   DT factor: " << dims.dt_factor << "\n nnx: " << dims.nnx
              << "   nnz: " << dims.nnz << "    nts: " << dims.nts << "\n Started Modelling.......";
    std::vector<float> seis;
    long long ntraces = 0;
    for (int is = istart; is < my_nsrc && rc == 0; is++) {
        sam2d_shot sh = {geo->src2d_sp[is].x, geo->src2d_sp[is].z, geo->nrec[is],
                         (const float *)geo->rec2d_sp[is]};
        seis.assign((size_t)sh.nrec * dims.real_nt, 0.0f);
        rc = sam2d_forward(ctx, &sh, 1, seis.data());
        if (rc) {
            std::cerr << "\n " << sam2d_last_error();
            break;
        }
        // A restarted job appends to the SU file the first run began.
        const int first = is == istart && std::strcmp(job_sp->jbtype, "New") == 0;
        if ((rc = write_shot_output(is, geo, seis.data(), dims.real_nt, wave_sp->real_dt_sec, first,
                                    &ntraces)))
            break;
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

// Options from the environment for a caller that cannot set sam2d_opts (a
// reference front end linked against this library).
static void options_from_environment()
{
    auto get = [](const char *name, int &field) {
        if (const char *v = std::getenv(name)) field = std::atoi(v);
    };
    get("SAM2D_NPML", sam2d_opts.npml);
    get("SAM2D_RESET_QUIRK", sam2d_opts.reset_quirk);
    get("SAM2D_PER_SHOT_FILES", sam2d_opts.per_shot_files);
    get("SAM2D_WRITE_RESTART", sam2d_opts.write_restart);
    get("SAM2D_SU_OUTPUT", sam2d_opts.su_output);
}

void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp)
{
    options_from_environment();
    const int rc = sam2d_modelling_module(my_nsrc, mygeo2d_sp);
    if (rc != 0) {
        std::cerr << "\n modelling_module: aborting with code " << rc << "\n";
        std::exit(rc & 0xff);
    }
}
