// model_io.cpp -- model files and the stability check (src/src_cpu/model.cpp,
// src/src_cpu/check_stability.cpp); errors return the reference's codes.
#include <cstdio>
#include <iostream>

#include "sam2d.h"
#include "sam2d_modelling.h"

// read2d (model.cpp:12-45): raw float32, nx blocks of nz values; the size
// test divides the byte count by nx*nz in float and wants exactly 4.
extern "C" int read2d(const char *file, float **arr, int nx, int nz)
{
    FILE *fp = std::fopen(file, "r");
    if (!fp) {
        std::cerr << "\n Error!!! Unable to open input mode : " << file;
        return -2;
    }
    std::fseek(fp, 0, SEEK_END);
    const float file_size = (float)std::ftell(fp);
    std::rewind(fp);
    if (file_size / (float)(nx * nz) != 4.0f) {
        std::cerr << "\n Error!!! Either input file: " << file << "   is empty or didn't contain NX*NZ element.....\n";
        std::fclose(fp);
        return -2;
    }
    for (int i = 0; i < nx; i++)
        if (std::fread(arr[i], sizeof(float), nz, fp) != (size_t)nz) {
            std::fclose(fp);
            return -2;
        }
    std::fclose(fp);
    return 0;
}

// minmax2d (model.cpp:47-64)
extern "C" void minmax2d(float **arr, int nx, int nz, float *vmin, float *vmax)
{
    *vmax = arr[0][0];
    *vmin = arr[0][0];
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            if (arr[ix][iz] < *vmin) *vmin = arr[ix][iz];
            if (arr[ix][iz] > *vmax) *vmax = arr[ix][iz];
        }
}

// check_stability (check_stability.cpp:8-72) through the library's check.
extern "C" int check_stability(void)
{
    sam2d_grid g = {mod_sp->nx, mod_sp->nz, mod_sp->dx, mod_sp->dz, wave_sp->rec_len,
                    wave_sp->fd_dt_sec, wave_sp->real_dt_sec, wave_sp->dom_freq, sam2d_opts.npml,
                    job_sp->fdop};
    const int rc = sam2d_check_stability(&g, mod_sp->vmin, mod_sp->vmax);
    if (rc != 0) std::cout << "\n " << sam2d_last_error() << "\n";
    else std::cout << "\n     STABILITY AND DISPERSION CRITERIA FULFILLED\n";
    return rc;
}
