// sam2d_kcommon.h -- device helpers shared by the forward and adjoint kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "sam2d_device.h"

namespace sam2d {

#define SAM2D_A0 1.125f
#define SAM2D_A1 (-1.0f / 24.0f)

// Coefficients of the backward-staggered first difference D- at index i of
// an n-point axis (fdoperator420/410, cpu_modelling_kernels.cpp:201-258):
// 4th order on [2, n-2], 2nd order at 1 and n-1, zero at 0 and outside.
__device__ __forceinline__ void dminus_coef(int i, int n, float &c0, float &c1)
{
    if (i >= 2 && i <= n - 2) { c0 = SAM2D_A0; c1 = SAM2D_A1; }
    else if (i == 1 || i == n - 1) { c0 = 1.0f; c1 = 0.0f; }
    else { c0 = 0.0f; c1 = 0.0f; }
}
// Forward-staggered D+ (fdoperator421/411, :260-316): 4th order on [1, n-3],
// 2nd order at 0 and n-2, zero at n-1 and outside.
__device__ __forceinline__ void dplus_coef(int i, int n, float &c0, float &c1)
{
    if (i >= 1 && i <= n - 3) { c0 = SAM2D_A0; c1 = SAM2D_A1; }
    else if (i == 0 || i == n - 2) { c0 = 1.0f; c1 = 0.0f; }
    else { c0 = 0.0f; c1 = 0.0f; }
}

// Descriptor pointers are generic; casting them into the global address space
// makes the compiler emit global_* instead of flat_* memory instructions.
// Flat accesses also count against lgkmcnt, so every lane-shuffle wait would
// otherwise drain all loads in flight and serialise the column prefetch.
#define SAM2D_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ SAM2D_GLOBAL T *gp(T *p) { return (SAM2D_GLOBAL T *)p; }

__device__ __forceinline__ float shfl_up(float v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ float shfl_dn(float v, int d) { return __shfl_down(v, d, 64); }

// Virtual block id -> tile id.  Blocks are dispatched round-robin over the 8
// XCDs (block b lands on XCD b % 8); giving each XCD a contiguous range of
// tiles keeps z-neighbouring tiles, which read each other's halo rows, in
// the same L2.
__device__ __forceinline__ int xcd_tile(int v, int per_xcd)
{
    return (v & 7) * per_xcd + (v >> 3);
}

// Shot owning a flat tile id (shots are few; descriptors are uniform).
template <class Desc>
__device__ __forceinline__ int shot_of_tile(const Desc *sh, int nshots, int tile)
{
    int lo = 0, hi = nshots - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sh[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

}  // namespace sam2d
