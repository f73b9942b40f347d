// main.cpp -- SeisAcoMod2D_MI355X.exe <jobcard> [--npml N] [--per-shot-files]
//             [--no-restart-file] [--reset-quirk] [--su] [--check]
// The reference's master-rank flow (main.cpp:52-90, modelling_master.cpp:10-91)
// in one process on one GPU: job card, geometry, models, stability check,
// modelling_module.  Ends with the exit status the reference's
// MPI::COMM_WORLD.Abort(code) gives under MPICH, code & 0xff (251 for a CFL
// violation, the reference's Abort(-5)); 0 on success.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "sam2d_modelling.h"

static int run(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "\n Error!!! Insufficient number of argument\n Please provide input parameter file\n";
        return -1;
    }
    bool check_only = false;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--npml") && a + 1 < argc) sam2d_opts.npml = std::atoi(argv[++a]);
        else if (!std::strcmp(argv[a], "--per-shot-files")) sam2d_opts.per_shot_files = 1;
        else if (!std::strcmp(argv[a], "--no-restart-file")) sam2d_opts.write_restart = 0;
        else if (!std::strcmp(argv[a], "--reset-quirk")) sam2d_opts.reset_quirk = 1;
        else if (!std::strcmp(argv[a], "--su")) sam2d_opts.su_output = 1;
        else if (!std::strcmp(argv[a], "--check")) check_only = true;
        else { std::cerr << "unknown option " << argv[a] << "\n"; return -6; }
    }
    const auto t0 = std::chrono::steady_clock::now();
    const char *job = argv[1];
    std::cout << "\n Job file : " << job;
    set_job_name(job);
    int rc = read_job_card(job);
    if (rc) return rc;
    rc = job_sp->geomflag == 0 ? read_recv2d(job_sp->geomfile) : read_source2d(job_sp->shotfile);
    if (rc) return rc;
    write_geometry_dump("test_geometry.txt");
    std::cout << "\n Total number of shot gathers are : " << geo2d_sp->nsrc;
    mod_sp->vel2d = new float *[mod_sp->nx];
    mod_sp->rho2d = new float *[mod_sp->nx];
    for (int ix = 0; ix < mod_sp->nx; ix++) {
        mod_sp->vel2d[ix] = new float[mod_sp->nz];
        mod_sp->rho2d[ix] = new float[mod_sp->nz];
    }
    if ((rc = read2d(mod_sp->velfile, mod_sp->vel2d, mod_sp->nx, mod_sp->nz))) return rc;
    if ((rc = read2d(mod_sp->densfile, mod_sp->rho2d, mod_sp->nx, mod_sp->nz))) return rc;
    minmax2d(mod_sp->vel2d, mod_sp->nx, mod_sp->nz, &mod_sp->vmin, &mod_sp->vmax);
    std::cout << "\n Min. Velocity : " << mod_sp->vmin << "\t Max. Velocity : " << mod_sp->vmax;
    if ((rc = check_stability())) return rc;
    if (check_only) {
        std::cout << "\n --check: inputs accepted, modelling skipped\n";
        return 0;
    }
    rc = sam2d_modelling_module(geo2d_sp->nsrc, geo2d_sp);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "\n Total time of execution : " << s / 3600 << "  .Hrs (" << s << " s)\n";
    return rc;
}

int main(int argc, char **argv)
{
    const int rc = run(argc, argv);
    if (rc) std::cerr << "\n aborting with code " << rc << "\n";
    return rc & 0xff;
}
