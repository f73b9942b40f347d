// sam2d_kernelsn.hip -- launches of the multi-step forward kernel (k_stepn,
// sam2d_kernelsn.h: T = 2..5 time steps per pass, and its recomputation
// form) and the deferred receiver gather.
#include "sam2d_kernelsn.h"

namespace sam2d {

__global__ void k_gather(const ShotDev *__restrict__ shots, int nshots, int band_r, int4 band_s,
                         int state_slot, int state_sample)
{
    gather_deferred(shots, nshots, band_r, band_s, state_slot, state_sample,
                    blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

hipError_t launch_stepn(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                        bool kz_alias, int step, int in_slot, int out_slot, int band_mask,
                        int band_w, const int *band_samples, int state_sample, int max_ng,
                        TileOrder order, hipStream_t stream)
{
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int T = variant_steps(variant);
    const int per = per_xcd_for(ntiles, order);
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const dim3 g(grid);
    int4 bs = make_int4(-1, -1, -1, -1);
    if (max_ng > 0) {
        bs = make_int4(band_samples[0], band_samples[1], band_samples[2], band_samples[3]);
    } else {
        band_mask = 0;
        state_sample = -1;
    }
    const StepNArgs a{d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot,
                      band_mask, band_w, bs, state_sample, order.ngroups, order.xorder, 0};
    hipError_t e;
    switch (T) {
    case 2: e = launch_stepn_t2(waves, pf, kz_alias, g, stream, a); break;
    case 3: e = launch_stepn_t3(waves, pf, kz_alias, g, stream, a); break;
    case 4: e = launch_stepn_t4(waves, pf, kz_alias, g, stream, a); break;
    case 5: e = launch_stepn_t5(waves, pf, kz_alias, g, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

bool ring_supported(int variant) { return variant_steps(variant) == 3; }

hipError_t launch_stepn_ring(int variant, const ShotDev *d_shots, int nshots, int ntiles,
                             int grid_cap, bool kz_alias, int step, int in_slot, int ring_slot,
                             TileOrder order, hipStream_t stream)
{
    if (!ring_supported(variant)) return hipErrorInvalidValue;
    int waves, chunk, pf;
    variant_shape_pf(variant, &waves, &chunk, &pf);
    const int per = per_xcd_for(ntiles, order);
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const StepNArgs a{d_shots, nshots, vtotal, per, ntiles, step, in_slot, in_slot ^ 1,
                      0, 0, make_int4(-1, -1, -1, -1), -1, order.ngroups, order.xorder,
                      ring_slot};
    const hipError_t e = launch_stepn_t3_ring(waves, pf, kz_alias, dim3(grid), stream, a);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_gather(const ShotDev *d_shots, int nshots, int max_ng, int band_r,
                         const int *band_samples, int state_slot, int state_sample,
                         hipStream_t stream)
{
    const int4 bs = make_int4(band_samples[0], band_samples[1], band_samples[2], band_samples[3]);
    if (max_ng <= 0 || (bs.x < 0 && bs.y < 0 && bs.z < 0 && bs.w < 0 && state_sample < 0))
        return hipSuccess;
    const int blocks = (max_ng + 255) / 256;
    hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, stream, d_shots, nshots, band_r, bs,
                       state_slot, state_sample);
    return hipGetLastError();
}

}  // namespace sam2d
