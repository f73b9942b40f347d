// sam2d_noalloc.hip -- sam2d_forward_noalloc: the per-shot forward of
// sam2d_forward_dev with every working buffer in one static device array, for
// callers that cannot allocate device memory.  It runs the same kernels on
// the same descriptors as the allocating path, so the seismograms are the
// same bits.
#include "sam2d_internal.h"

#include <cstring>
#include <mutex>

// The working memory: descriptor, source increments, receiver cells, band
// buffer and the two state slots of one shot.  Each GPU the library runs on
// holds its own copy from the first use of the code object until the process
// ends.  In host code `scratch` names a host-side stand-in, not the device
// address: the device address comes from hipGetSymbolAddress, once per
// device (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[SAM2D_NOALLOC_SCRATCH_BYTES];

namespace sam2d {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
}  // namespace

// Device address of this device's scratch array (the current device).  Called
// by sam2d_create for its device; later calls return the cached address.
int noalloc_scratch(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(SAM2D_ERR_DEVICE, "device %d beyond the %d the no-allocation path tracks",
                    device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(SAM2D_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        }
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return SAM2D_OK;
}

}  // namespace sam2d

using namespace sam2d;

extern "C" int sam2d_forward_noalloc(sam2d_ctx *ctx, const sam2d_shot *shot, float *seis_dev)
{
    clear_error();
    if (!ctx || !shot) return fail(SAM2D_ERR_PARAM, "NULL argument");
    if (shot->nrec > 0 && !seis_dev) return fail(SAM2D_ERR_PARAM, "seis_dev is NULL");
    const double t0 = now_ms();
    ctx->stats = sam2d_stats();
    ctx->quirk_valid = false;
    HIP_TRY(hipSetDevice(ctx->device));
    void *base = nullptr;
    SAM2D_TRY(noalloc_scratch(ctx->device, &base));

    // The full model's domain, built by sam2d_create: nothing to allocate.
    const long long key = (long long)(unsigned)ctx->grid.nx;      // window (0, nx)
    auto it = ctx->domains.find(key);
    if (it == ctx->domains.end())
        return fail(SAM2D_ERR_PARAM, "the context holds no full-model domain");
    DeviceDomain *dom = it->second.get();
    ShotPlan plan;
    SAM2D_TRY(plan_shot(ctx, *shot, dom, plan));
    const int variant = pick_variant(ctx);
    const int chunk = tile_chunk(ctx, variant);
    const TimeAxis &ta = ctx->ta;
    const size_t N = (size_t)dom->host.nnx * dom->host.nnz;
    SAM2D_TRY(check_multistep_size(ctx, variant, N));

    ShotDev d;
    describe_shot(plan, variant, chunk, 0, ta.nts, ta.real_nt, d);
    // Scratch layout: descriptor | amp[nts] | gcell[ng] | band | state (2 slots)
    const size_t o_desc = 0;
    const size_t o_amp = align256(o_desc + sizeof(ShotDev));
    const size_t o_gcell = align256(o_amp + sizeof(float) * (size_t)ta.nts);
    const size_t o_band = align256(o_gcell + sizeof(int) * (size_t)plan.ng);
    const size_t band_n = (size_t)2 * BAND_SLICES * d.band_h * d.nnx;
    const size_t o_state = align256(o_band + sizeof(float) * (band_n ? band_n : 1));
    const size_t need = o_state + sizeof(float) * 4 * 2 * N;
    if (need > (size_t)SAM2D_NOALLOC_SCRATCH_BYTES)
        return fail(SAM2D_ERR_PARAM,
                    "sam2d_forward_noalloc: the shot needs %zu bytes of working memory "
                    "(%zu padded cells), above the %zu-byte scratch array",
                    need, N, (size_t)SAM2D_NOALLOC_SCRATCH_BYTES);
    char *s = (char *)base;
    d.amp = (const float *)(s + o_amp);
    d.gcell = (const int *)(s + o_gcell);
    d.band = (float *)(s + o_band);
    d.state = (float *)(s + o_state);
    d.seis = seis_dev;
    ShotDev *desc_dev = (ShotDev *)(s + o_desc);

    hipStream_t st = ctx->stream;
    // The copies read d and plan's vectors, which live until the stream has
    // drained: every return below this point waits for it first.
    auto run = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(desc_dev, &d, sizeof d, hipMemcpyHostToDevice, st));
        if (ta.nts) HIP_TRY(hipMemcpyAsync((void *)d.amp, plan.amp.data(), sizeof(float) * ta.nts,
                                           hipMemcpyHostToDevice, st));
        if (plan.ng) HIP_TRY(hipMemcpyAsync((void *)d.gcell, plan.gcell.data(),
                                            sizeof(int) * plan.ng, hipMemcpyHostToDevice, st));
        // The array keeps the last call's contents: clear both state slots
        // (the kernels read pz as 0 outside the sponge frame) and the output
        // (fill kernels, not memsets).
        HIP_TRY(launch_fill(d.state, 2 * 4 * N, 0.0f, st));
        if (plan.ng) HIP_TRY(launch_fill(seis_dev, (size_t)plan.ng * ta.real_nt, 0.0f, st));
        HIP_TRY(hipEventRecord(ctx->ev0, st));
        SAM2D_TRY(enqueue_steps_on(ctx, desc_dev, d.ngroups, 1, d.ngroups * d.nchunks, variant, 0,
                                   ta.fd_nt, plan.ng, 0, nullptr));
        HIP_TRY(hipEventRecord(ctx->ev1, st));
        return SAM2D_OK;
    };
    const int rc = run();
    const hipError_t e = hipStreamSynchronize(st);
    if (rc != SAM2D_OK) return rc;
    HIP_TRY(e);
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    const int spl = variant_steps(variant);
    ctx->stats.kernel_ms = ms;
    ctx->stats.cell_updates = (double)N * ta.fd_nt;
    ctx->stats.launches = (ta.fd_nt + spl - 1) / spl;
    ctx->stats.steps_per_launch = spl;
    ctx->stats.used_graph = 0;
    ctx->stats.variant = variant;
    ctx->stats.wall_ms = now_ms() - t0;
    return SAM2D_OK;
}
