// stream_probe.hip -- memory-system ceiling for the stencil kernels' access
// pattern on the 5100 x 5100 padded grid of the 5000 x 5000 workload.
//
// Each kernel moves 44 B per cell (28 B read: four state fields and three
// coefficients; 16 B written: four state fields) with no stencil arithmetic,
// in four data layouts:
//   arrays  one array per field (the library's layout)
//   pairs   (px, pz), (vx, vz) and (BX, BZ) side by side, KX alone
//   cell    the four state fields of a cell together, the three coefficients
//           together
//   linear  the cell layout, read and written as one flat stream (no tiles)
// The first three use the kernels' tiling: a wave owns 64 rows and marches
// `chunk` columns along x; `wpb` waves per block are stacked along z.
// Prints one JSON line per layout.  Usage: stream_probe [chunk] [wpb].
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

namespace {

#define PROBE_CHECK(x)                                                             \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "stream_probe: %s (line %d)\n", hipGetErrorString(e_), \
                         __LINE__);                                                \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)

struct Tiles {
    int nnz, nnx, chunk, nstrips;
};

// Row of this lane and column range of this wave; false for rows past the grid.
__device__ inline bool tile_of(const Tiles t, int &r, int &xa, int &xb)
{
    const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int strip = w % t.nstrips, ch = w / t.nstrips;
    r = strip * 64 + (threadIdx.x & 63);
    xa = ch * t.chunk;
    xb = min(xa + t.chunk, t.nnx);
    return r < t.nnz;
}

__global__ void k_arrays(const float *__restrict__ px, const float *__restrict__ pz,
                         const float *__restrict__ vx, const float *__restrict__ vz,
                         const float *__restrict__ kx, const float *__restrict__ bx,
                         const float *__restrict__ bz, float *__restrict__ opx,
                         float *__restrict__ opz, float *__restrict__ ovx,
                         float *__restrict__ ovz, Tiles t)
{
    int r, xa, xb;
    if (!tile_of(t, r, xa, xb)) return;
    for (int c = xa; c < xb; c++) {
        const size_t i = (size_t)c * t.nnz + r;
        const float k = kx[i];
        opx[i] = px[i] + k;
        opz[i] = pz[i] + k;
        ovx[i] = vx[i] + bx[i];
        ovz[i] = vz[i] + bz[i];
    }
}

__global__ void k_pairs(const float2 *__restrict__ p, const float2 *__restrict__ v,
                        const float *__restrict__ kx, const float2 *__restrict__ b,
                        float2 *__restrict__ op, float2 *__restrict__ ov, Tiles t)
{
    int r, xa, xb;
    if (!tile_of(t, r, xa, xb)) return;
    for (int c = xa; c < xb; c++) {
        const size_t i = (size_t)c * t.nnz + r;
        const float2 a = p[i], d = v[i], e = b[i];
        const float k = kx[i];
        op[i] = make_float2(a.x + k, a.y + k);
        ov[i] = make_float2(d.x + e.x, d.y + e.y);
    }
}

__device__ inline void cell_update(const float4 *__restrict__ st, const float *__restrict__ co,
                                   float4 *__restrict__ ost, size_t i)
{
    const float4 s = st[i];
    const float k = co[3 * i], f = co[3 * i + 1], h = co[3 * i + 2];
    ost[i] = make_float4(s.x + k, s.y + k, s.z + f, s.w + h);
}

__global__ void k_cell(const float4 *__restrict__ st, const float *__restrict__ co,
                       float4 *__restrict__ ost, Tiles t)
{
    int r, xa, xb;
    if (!tile_of(t, r, xa, xb)) return;
    for (int c = xa; c < xb; c++) cell_update(st, co, ost, (size_t)c * t.nnz + r);
}

__global__ void k_linear(const float4 *__restrict__ st, const float *__restrict__ co,
                         float4 *__restrict__ ost, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        cell_update(st, co, ost, i);
}

float *zeros(size_t bytes)
{
    float *p = nullptr;
    PROBE_CHECK(hipMalloc(&p, bytes));
    PROBE_CHECK(hipMemset(p, 0, bytes));
    return p;
}

}  // namespace

int main(int argc, char **argv)
{
    const int nnx = 5100, nnz = 5100;
    const size_t n = (size_t)nnx * nnz;
    const int chunk = argc > 1 ? std::atoi(argv[1]) : 64;
    const int wpb = argc > 2 ? std::atoi(argv[2]) : 4;
    if (chunk < 1 || wpb < 1 || wpb > 16) {
        std::fprintf(stderr, "usage: stream_probe [chunk >= 1] [waves per block 1..16]\n");
        return 2;
    }
    const Tiles t{nnz, nnx, chunk, (nnz + 63) / 64};
    // Every wave w < nstrips * nchunks gets a strip and a chunk; waves past
    // that (the last block's tail) find xa >= nnx and do nothing.
    const int waves = t.nstrips * ((nnx + chunk - 1) / chunk);
    const int blocks = (waves + wpb - 1) / wpb;
    float *f[11];
    for (auto &q : f) q = zeros(n * 4);
    float2 *p2 = (float2 *)zeros(n * 8), *v2 = (float2 *)zeros(n * 8), *b2 = (float2 *)zeros(n * 8);
    float2 *op2 = (float2 *)zeros(n * 8), *ov2 = (float2 *)zeros(n * 8);
    float4 *st = (float4 *)zeros(n * 16), *ost = (float4 *)zeros(n * 16);
    float *co = zeros(n * 12);
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    const char *names[] = {"arrays", "pairs", "cell", "linear"};
    for (int mode = 0; mode < 4; mode++) {
        auto run = [&]() {
            if (mode == 0)
                hipLaunchKernelGGL(k_arrays, dim3(blocks), dim3(64 * wpb), 0, 0, f[0], f[1], f[2],
                                   f[3], f[4], f[5], f[6], f[7], f[8], f[9], f[10], t);
            else if (mode == 1)
                hipLaunchKernelGGL(k_pairs, dim3(blocks), dim3(64 * wpb), 0, 0, p2, v2, f[4], b2,
                                   op2, ov2, t);
            else if (mode == 2)
                hipLaunchKernelGGL(k_cell, dim3(blocks), dim3(64 * wpb), 0, 0, st, co, ost, t);
            else
                hipLaunchKernelGGL(k_linear, dim3(8192), dim3(256), 0, 0, st, co, ost, n);
        };
        for (int it = 0; it < 3; it++) run();
        PROBE_CHECK(hipGetLastError());
        const int iters = 20;
        PROBE_CHECK(hipEventRecord(e0));
        for (int it = 0; it < iters; it++) run();
        PROBE_CHECK(hipEventRecord(e1));
        PROBE_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / iters;
        std::printf("{\"layout\": \"%s\", \"chunk\": %d, \"waves_per_block\": %d, "
                    "\"us_per_pass\": %.1f, \"tb_per_s\": %.3f}\n",
                    names[mode], chunk, wpb, us, 44.0 * n / (us * 1e-6) / 1e12);
        std::fflush(stdout);
    }
    return 0;
}
