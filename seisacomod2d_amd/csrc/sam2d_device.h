// sam2d_device.h -- descriptors shared by the kernels and the host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace sam2d {

// Rows owned by one wavefront, and its halo rows above / below (64 lanes).
constexpr int OWN_Z = 58;
constexpr int HALO_T = 3;
constexpr int HALO_B = 64 - OWN_Z - HALO_T;

// One resident shot: its padded domain, coefficient fields, ping-pong state
// and receiver list, all device pointers.  Kernels read it from device memory
// so a captured time loop can be replayed for another shot by rewriting the
// descriptors only.
struct ShotDev {
    int nnz, nnx;          // padded domain of this shot
    int ngroups, nchunks;  // tiles: z-strip groups x x-chunks
    int chunk;             // columns per x-chunk
    int tile_begin;        // first flat tile id of this shot
    int src_x, src_z;      // source cell (padded indices)
    int nts;               // wavelet length in steps
    int ng, real_nt;       // receivers, samples per trace
    int pad0;
    const float *amp;      // [nts] source increments (fd_dt*w*kappa_s/(dx*dz))
    const float *KX, *KZ;  // dt/dx*kappa, dt/dz*kappa   (KZ may alias KX)
    const float *BX, *BZ;  // dt/dx*bu2,  dt/dz*bu1
    const float *prof_x;   // [4][nnx]: damp2a, damp2a1, damp2b, damp2b1
    const float *prof_z;   // [4][nnz]: damp1a, damp1a1, damp1b, damp1b1
    float *st[2][4];       // [buffer set][px, pz, vx, vz]
    const int *gcell;      // [ng] receiver cells (ix*nnz + iz)
    float *seis;           // [ng][real_nt]
};

int num_variants();
void variant_shape(int v, int *waves, int *chunk);

hipError_t launch_step(int waves, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                       int step, int par, int rec_slot, hipStream_t stream);
hipError_t launch_record(const ShotDev *d_shots, int nshots, int max_ng, int par, int slot,
                         hipStream_t stream);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t stream);

}  // namespace sam2d
