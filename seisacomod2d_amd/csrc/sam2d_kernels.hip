// sam2d_kernels.hip -- gfx950 kernels of the 2-D acoustic staggered-grid
// time step (4th order space / 2nd order time, split-field sponge).
//
// One launch advances every tile of every resident shot by one time step.
// It fuses what the reference does in nine full-grid sweeps per step
// (FDOperator_4, src/src_cpu/cpumodelling.cpp:115-156, kernels in
// src/src_cpu/cpu_modelling_kernels.cpp:201-368):
//   p-phase  Dx- vx, Dz- vz, px/pz sponge update, source, p = px + pz
//   v-phase  Dx+ p,  Dz+ p,  vx/vz sponge update
// plus the receiver gather of the previous step (cpu_record, :24-31).
//
// Layout: every field is column-major with z fastest, cell (ix, iz) at
// ix*nnz + iz (the reference's id = iz + ix*nnz).  A wavefront owns a strip
// of OWN_Z rows and carries a HALO_T-row halo above and HALO_B rows below in
// the same 64 lanes, so z-neighbours are lane shuffles.  It marches along x
// over a chunk of columns keeping a sliding register window of vx, vz and the
// new p; p is recomputed redundantly on the 3 halo rows and the 3 halo
// columns it needs, so one pass reads px, pz, vx, vz, KX, BX, BZ once and
// writes px, pz, vx, vz once: 44 bytes per cell update.  State is
// double-buffered (in -> out) because neighbouring tiles read the old state
// of the cells a tile overwrites.

#include <hip/hip_runtime.h>
#include <stdint.h>
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
__device__ __forceinline__ int shot_of_tile(const ShotDev *sh, int nshots, int tile)
{
    int lo = 0, hi = nshots - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sh[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Receiver gather of p = px + pz from the given state buffers
// (cpu_record, cpu_modelling_kernels.cpp:24-31, trace-major output
// as after cpu_transpose, :370-381).
__device__ __forceinline__ void gather_receivers(const ShotDev *sh, int nshots, int par,
                                                 int slot, int gtid, int gstride)
{
    for (int s = 0; s < nshots; ++s) {
        const ShotDev &d = sh[s];
        const float *px = d.st[par][0];
        const float *pz = d.st[par][1];
        for (int g = gtid; g < d.ng; g += gstride) {
            int c = d.gcell[g];
            d.seis[(size_t)g * d.real_nt + slot] = px[c] + pz[c];
        }
    }
}

// One time step.  par selects the input buffer set (output is 1 - par).
// rec_slot >= 0 asks for the receiver gather of the state held in the input
// buffers (the previous step's result) into sample rec_slot.
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64)
void k_step(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
            int ntiles, int step, int par, int rec_slot)
{
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;

    if (rec_slot >= 0)
        gather_receivers(shots, nshots, par, rec_slot,
                         blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);

    for (int v = blockIdx.x; v < vtotal; v += gridDim.x) {
        const int tile = xcd_tile(v, per_xcd);
        if (tile >= ntiles) continue;
        const int s = shot_of_tile(shots, nshots, tile);
        const ShotDev &d = shots[s];
        const int local = tile - d.tile_begin;
        const int group = local % d.ngroups;
        const int chunk = local / d.ngroups;
        const int nnz = d.nnz, nnx = d.nnx;
        const int zb = (group * WAVES + wv) * OWN_Z;
        if (zb >= nnz) continue;                      // wave-uniform
        const int xa = chunk * d.chunk;
        const int xb = min(xa + d.chunk, nnx);

        const int r = zb - HALO_T + lane;             // this lane's row
        const bool rvalid = (r >= 0) && (r < nnz);
        const bool rown = (lane >= HALO_T) && (lane < HALO_T + OWN_Z) && (r < nnz);
        const int rc = min(max(r, 0), nnz - 1);

        const float *__restrict__ pxi = d.st[par][0];
        const float *__restrict__ pzi = d.st[par][1];
        const float *__restrict__ vxi = d.st[par][2];
        const float *__restrict__ vzi = d.st[par][3];
        float *__restrict__ pxo = d.st[par ^ 1][0];
        float *__restrict__ pzo = d.st[par ^ 1][1];
        float *__restrict__ vxo = d.st[par ^ 1][2];
        float *__restrict__ vzo = d.st[par ^ 1][3];
        const float *__restrict__ KX = d.KX;
        const float *__restrict__ KZ = d.KZ;
        const float *__restrict__ BX = d.BX;
        const float *__restrict__ BZ = d.BZ;
        const float *__restrict__ prx = d.prof_x;     // [4][nnx]
        const float *__restrict__ prz = d.prof_z;     // [4][nnz]

        // Per-row constants: sponge profiles and z-operator coefficients.
        const float daz = prz[rc], skz = prz[nnz + rc];
        const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
        float zm0, zm1, zp0, zp1;
        dminus_coef(r, nnz, zm0, zm1);
        dplus_coef(r, nnz, zp0, zp1);
        const bool src_row = (r == d.src_z);
        const float amp = (step < d.nts) ? d.amp[step] : 0.0f;
        const int src_x = (step <= d.nts) ? d.src_x : -1;

        auto ld = [&](const float *__restrict__ f, int col) -> float {
            return (rvalid && col >= 0 && col < nnx) ? f[(size_t)col * nnz + r] : 0.0f;
        };

        // Sliding windows: vx at c-2, c-1, c; vz at c-1, c-2; p at c-3..c-1.
        float vx_m2 = ld(vxi, xa - 3), vx_m1 = ld(vxi, xa - 2), vx_0 = ld(vxi, xa - 1);
        float vz_m1 = 0.0f, vz_m2 = 0.0f;
        float p_m3 = 0.0f, p_m2 = 0.0f, p_m1 = 0.0f;

        // Prefetched operands of the first column.
        int c = xa - 1;
        float n_vx = ld(vxi, c + 1), n_vz = ld(vzi, c), n_px = ld(pxi, c), n_pz = ld(pzi, c);
        float n_kx = ld(KX, c), n_kz = ld(KZ, c), n_bx = ld(BX, c - 2), n_bz = ld(BZ, c - 2);

        for (; c <= xb + 1; ++c) {
            const float vx_p1 = n_vx, vz_0 = n_vz, px0 = n_px, pz0 = n_pz;
            const float kx = n_kx, kz = n_kz, bx = n_bx, bz = n_bz;
            if (c + 1 <= xb + 1) {
                n_vx = ld(vxi, c + 2); n_vz = ld(vzi, c + 1);
                n_px = ld(pxi, c + 1); n_pz = ld(pzi, c + 1);
                n_kx = ld(KX, c + 1); n_kz = ld(KZ, c + 1);
                n_bx = ld(BX, c - 1); n_bz = ld(BZ, c - 1);
            }

            // ---- p-phase at column c (cpu_pml_p, :335-351; source, :135-140)
            const bool cvalid = (c >= 0) && (c < nnx);
            float dax = 0.0f, skx = 0.0f;
            if (cvalid) { dax = prx[c]; skx = prx[nnx + c]; }
            float xm0, xm1;
            dminus_coef(c, nnx, xm0, xm1);
            const float dvx = fmaf(xm0, vx_0 - vx_m1, xm1 * (vx_p1 - vx_m2));
            const float vz_u2 = shfl_up(vz_0, 2), vz_u1 = shfl_up(vz_0, 1);
            const float vz_d1 = shfl_dn(vz_0, 1);
            const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
            float pxn = fmaf(kx * skx, dvx, dax * px0);
            const float pzn = fmaf(kz * skz, dvz, daz * pz0);
            if (src_row && c == src_x) pxn += amp;
            const float pn = pxn + pzn;
            if (rown && c >= xa && c < xb) {
                const size_t id = (size_t)c * nnz + r;
                pxo[id] = pxn;
                pzo[id] = pzn;
            }

            // ---- v-phase at column j = c - 2 (cpu_pml_v, :353-368)
            const int j = c - 2;
            if (j >= xa && j < xb) {
                const float dbx = prx[2 * nnx + j], sbx = prx[3 * nnx + j];
                float xp0, xp1;
                dplus_coef(j, nnx, xp0, xp1);
                const float dpx = fmaf(xp0, p_m1 - p_m2, xp1 * (pn - p_m3));
                const float q_u1 = shfl_up(p_m2, 1);
                const float q_d1 = shfl_dn(p_m2, 1), q_d2 = shfl_dn(p_m2, 2);
                const float dpz = fmaf(zp0, q_d1 - p_m2, zp1 * (q_d2 - q_u1));
                const float vxn = fmaf(bx * sbx, dpx, dbx * vx_m2);
                const float vzn = fmaf(bz * sbz, dpz, dbz * vz_m2);
                if (rown) {
                    const size_t id = (size_t)j * nnz + r;
                    vxo[id] = vxn;
                    vzo[id] = vzn;
                }
            }

            vx_m2 = vx_m1; vx_m1 = vx_0; vx_0 = vx_p1;
            vz_m2 = vz_m1; vz_m1 = vz_0;
            p_m3 = p_m2; p_m2 = p_m1; p_m1 = pn;
        }
    }
}

// Receiver gather alone (the last step's samples).
__global__ void k_record(const ShotDev *__restrict__ shots, int nshots, int par, int slot)
{
    gather_receivers(shots, nshots, par, slot, blockIdx.x * blockDim.x + threadIdx.x,
                     gridDim.x * blockDim.x);
}

// Fill with a constant.  Used instead of hipMemsetAsync so that clearing is
// an ordinary kernel node inside captured graphs.
__global__ void k_fill(float *__restrict__ p, size_t n, float v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- launchers

static const int kWaves[] = {1, 2, 4, 8};
static const int kChunks[] = {32, 64, 128, 256, 512};

int num_variants() { return (int)(sizeof(kWaves) / sizeof(int) * sizeof(kChunks) / sizeof(int)); }

void variant_shape(int v, int *waves, int *chunk)
{
    const int nc = (int)(sizeof(kChunks) / sizeof(int));
    *waves = kWaves[v / nc];
    *chunk = kChunks[v % nc];
}

hipError_t launch_step(int waves, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                       int step, int par, int rec_slot, hipStream_t stream)
{
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    dim3 block(64 * waves);
    switch (waves) {
    case 1: hipLaunchKernelGGL(k_step<1>, dim3(grid), block, 0, stream, d_shots, nshots, vtotal, per, ntiles, step, par, rec_slot); break;
    case 2: hipLaunchKernelGGL(k_step<2>, dim3(grid), block, 0, stream, d_shots, nshots, vtotal, per, ntiles, step, par, rec_slot); break;
    case 4: hipLaunchKernelGGL(k_step<4>, dim3(grid), block, 0, stream, d_shots, nshots, vtotal, per, ntiles, step, par, rec_slot); break;
    case 8: hipLaunchKernelGGL(k_step<8>, dim3(grid), block, 0, stream, d_shots, nshots, vtotal, per, ntiles, step, par, rec_slot); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_record(const ShotDev *d_shots, int nshots, int max_ng, int par, int slot,
                         hipStream_t stream)
{
    int blocks = (max_ng + 255) / 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_record, dim3(blocks), dim3(256), 0, stream, d_shots, nshots, par, slot);
    return hipGetLastError();
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

}  // namespace sam2d

namespace sam2d {
// Variant used until the autotuner has run: 4 strips per workgroup, 128-column
// chunks.
int default_variant_index()
{
    const int nc = (int)(sizeof(kChunks) / sizeof(int));
    return 2 * nc + 2;
}
}  // namespace sam2d
