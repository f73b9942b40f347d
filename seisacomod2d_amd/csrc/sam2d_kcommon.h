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

// Four floats of a cell (the adjoint's coefficient cells with KZ): a native
// vector type, which global-address-space pointers accept.
typedef float fquad __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ SAM2D_GLOBAL T *gp(T *p) { return (SAM2D_GLOBAL T *)p; }
// Three floats of a cell (the state's (P, vx, vz), the adjoint's state,
// accumulators and coefficients), 4-byte aligned: one dwordx3 access per
// lane.
typedef float ftri __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ ftri ld3(const SAM2D_GLOBAL char *__restrict__ base, unsigned off)
{
    return *(const SAM2D_GLOBAL ftri *__restrict__)(base + off);
}
__device__ __forceinline__ void st3(SAM2D_GLOBAL char *__restrict__ base, unsigned off, float a,
                                    float b, float c)
{
    ftri v;
    v.x = a; v.y = b; v.z = c;
    *(SAM2D_GLOBAL ftri *__restrict__)(base + off) = v;
}
// Two floats of a cell (the forward ring's (p, E) pairs), 8-byte aligned: one
// dwordx2 access per lane.
typedef float fpair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ fpair ld2(const SAM2D_GLOBAL char *__restrict__ base, unsigned off)
{
    return *(const SAM2D_GLOBAL fpair *__restrict__)(base + off);
}
__device__ __forceinline__ void st2(SAM2D_GLOBAL char *__restrict__ base, unsigned off, float a,
                                    float b)
{
    fpair v;
    v.x = a; v.y = b;
    *(SAM2D_GLOBAL fpair *__restrict__)(base + off) = v;
}
// The forward ring's dJ/dkappa factors of one step at one cell (ring_pair,
// ring_side): Ex = dvx (dtdx skx), Ez = dvz (dtdz skz); returns E, Ex + Ez
// outside the sponge frame, Ex in it.  Every product and sum rounded as
// written, so the single- and multi-step recomputations store the same bits.
__device__ __forceinline__ float ring_factor(float dvx, float dtdx_skx, float dvz, float dtdz_skz,
                                             bool inner, float &ez)
{
#pragma clang fp contract(off)
    const float ex = dvx * dtdx_skx;
    ez = dvz * dtdz_skz;
    return inner ? ex + ez : ex;
}
// Read-only tables indexed by a wave-uniform column (the x-profiles of the
// sponge): in the constant address space the compiler reads them with scalar
// loads, off the vector memory path that streams the state.
#define SAM2D_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const SAM2D_CONST T *cp4(const T *p) { return (const SAM2D_CONST T *)p; }

__device__ __forceinline__ float shfl_up(float v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ float shfl_dn(float v, int d) { return __shfl_down(v, d, 64); }

// Whole-wavefront lane shifts by DPP (wave_shr:1 = 0x138, wave_shl:1 = 0x130):
// lane i receives lane i-1 (up) or lane i+1 (down); the lane without a source
// receives 0 (bound_ctrl).  One VALU operation, no LDS; all 64 lanes must be
// active: a disabled source lane also reads as 0, so a shift is never written
// inside a divergent branch, and that includes one arm of ?: (a convergent
// call keeps clang's branch for the arms).
__device__ __forceinline__ float lane_up1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_dn1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}

// A register copy the compiler cannot merge away.  Taking a prefetched operand
// into a sliding window through it leaves the prefetch register free for the
// next column's load, so the load can be issued into it directly instead of
// into a fresh register that the loop back edge would have to copy (a copy
// that waits for the load to land).
__device__ __forceinline__ float keep(float x)
{
    float y;
    asm("v_mov_b32 %0, %1" : "=&v"(y) : "v"(x));
    return y;
}

// Virtual block id -> tile id.  Blocks are dispatched round-robin over the 8
// XCDs (block b lands on XCD b % 8); giving each XCD a contiguous range of
// tiles keeps z-neighbouring tiles, which read each other's halo rows, in
// the same L2.
__device__ __forceinline__ int xcd_tile(int v, int per_xcd)
{
    return (v & 7) * per_xcd + (v >> 3);
}

// The same with the chunk columns of an XCD's range visited in a strided
// order (xorder > 1; per_xcd a multiple of ngroups, tiles ordered z-group
// fastest): the range's ncx columns go as 0, s, 2s, ..., 1, 1+s, ... with
// s = xorder, each column's ngroups tiles in order.  Concurrent tiles are
// spread along x, which measures up to 6 % faster for the multi-step kernel
// on MI355X; the best stride depends on the tile shape and is autotuned.
__device__ __forceinline__ int xcd_tile_ordered(int v, int per_xcd, int ngroups, int xorder)
{
    const int x = v & 7, p = v >> 3;
    if (xorder <= 1) return x * per_xcd + p;
    const int ncx = per_xcd / ngroups;
    int q = p / ngroups;
    const int g = p - q * ngroups;
    int col = q;
    for (int r = 0; r < xorder; r++) {
        const int cnt = (ncx - r + xorder - 1) / xorder;   // columns r, r+s, ... < ncx
        if (q < cnt) { col = r + q * xorder; break; }
        q -= cnt;
    }
    return x * per_xcd + col * ngroups + g;
}

// Band buffer of the receiver rows: [2 parities][BAND_SLICES][band_h][nnx].
// Slice j of parity w holds the pressure after the (j+1)-th step of a
// multi-step launch; the pointer is offset so that it is indexed by the
// padded row: band_slice(d, w, j)[row * nnx + col].
__device__ __forceinline__ SAM2D_GLOBAL float *band_slice(const ShotDev &d, int w, int j)
{
    return gp(d.band) + ((ptrdiff_t)(w * BAND_SLICES + j) * d.band_h - d.rrow_lo) *
                            (ptrdiff_t)d.nnx;
}

// Pressure of cell c = (col, row) in a state slot: P inside, P + pz in the
// sponge frame (the sum the kernels formed in registers, bit for bit).
__device__ __forceinline__ float state_pressure(const ShotDev &d, int slot, int c, int col, int row)
{
    const float P = gp(slot_cells(d, slot))[3 * (size_t)c];
    return inner_cell(d, col, row) ? P : P + gp(slot_pz(d, slot))[c];
}

// Deferred receiver gather of the previous multi-step launch (cpu_record,
// src/src_cpu/cpu_modelling_kernels.cpp:24-31, into the trace-major gather as
// after cpu_transpose, :370-381): samples bs.x .. bs.w from band slices 0..3
// of parity band_r, sample ss as the pressure of the state slot (each skipped
// when < 0).
__device__ __forceinline__ void gather_deferred(const ShotDev *sh, int nshots, int band_r, int4 bs,
                                                int state_slot, int ss, int gtid, int gstride)
{
    const int smp[BAND_SLICES] = {bs.x, bs.y, bs.z, bs.w};
    for (int s = 0; s < nshots; ++s) {
        const ShotDev &d = sh[s];
        const SAM2D_GLOBAL int *gc = gp(d.gcell);
        SAM2D_GLOBAL float *seis = gp(d.seis);
        for (int g = gtid; g < d.ng; g += gstride) {
            const int c = gc[g];
            SAM2D_GLOBAL float *tr = seis + (size_t)g * d.real_nt;
            const int col = c / d.nnz, row = c - col * d.nnz;
            const size_t at = (size_t)row * d.nnx + col;
            #pragma unroll
            for (int j = 0; j < BAND_SLICES; j++)
                if (smp[j] >= 0) tr[smp[j]] = band_slice(d, band_r, j)[at];
            if (ss >= 0) tr[ss] = state_pressure(d, state_slot, c, col, row);
        }
    }
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
