// sam2d_setup.h -- host-side construction of one padded modelling domain:
// time axis, sponge profiles, Ricker wavelet, staggered coefficients.
// Computed in FP64 where the reference computes in FP64, in FP32 where it
// computes in FP32, and rounded once, without bias, to the FP32 fields the
// kernels read.
#pragma once
#include <vector>
#include "../../include/sam2d.h"

namespace sam2d {

struct TimeAxis {
    int fd_nt, real_nt, dt_factor, nts;
    double dtdx, dtdz;
};

// modelling_module's scalar setup (src/src_cpu/cpumodelling.cpp:168-195).
TimeAxis time_axis(const sam2d_grid &g);

// Sponge profile value at index i of an axis with n interior points and a
// sponge of npml cells (cpu_pml_coefficient_a/_b,
// src/src_cpu/cpu_modelling_kernels.cpp:123-199).  half = false: pressure
// grid; half = true: velocity grid.  decay = exp(-t^2) (FP32 expf), scale =
// exp(-t^2/2) (FP64 exp, rounded to float).
void sponge_at(int i, int npml, int n, bool half, float &decay, float &scale);

// Ricker wavelet, cpu_ricker (cpu_modelling_kernels.cpp:33-62).
std::vector<double> ricker(float fm, float dt, int nts);

// One padded domain built from model columns [wx0, wx0 + w).
struct DomainHost {
    int wx0, w;
    int nnx, nnz;
    bool kz_alias;                  // dtdz == dtdx: KZ is KX
    double dtdx, dtdz;              // the time axis's fd_dt/dx, fd_dt/dz
    std::vector<double> kappa;      // FP64 bulk modulus, padded
    std::vector<float> KX, KZ, BX, BZ;
    std::vector<float> B;           // buoyancy 1/rho (1e10 where rho <= 0), padded
    std::vector<float> prof_x;      // [4][nnx]
    std::vector<float> prof_z;      // [4][nnz]
};

// expand (cpumodelling.cpp:39-60), cpu_stagger (cpu_modelling_kernels.cpp:64-121),
// sponge profiles (cpumodelling.cpp:254-259) and the scaled coefficient
// fields (cpumodelling.cpp:270-281) for the window.  vp, rho: x-major nx*nz.
void build_domain(const sam2d_grid &g, const TimeAxis &ta, const float *vp, const float *rho,
                  int wx0, int w, DomainHost &out);

// Grid cell of a coordinate in metres (cpumodelling.cpp:288-289, 333-334 and
// cpu_set_sg, cpu_modelling_kernels.cpp:15-22): model column and row by float
// multiplication with the float reciprocal, truncated.
void coord_to_cell(const sam2d_grid &g, float x, float z, int &ix, int &iz);

}  // namespace sam2d
