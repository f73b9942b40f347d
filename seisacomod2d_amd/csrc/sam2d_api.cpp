// sam2d_api.cpp -- the C ABI of include/sam2d.h: contexts, resident device
// domains and shot state, the captured time loop, batched tile lists and the
// autotuner.  The gradient entry point lives in sam2d_adjoint.cpp.
#include "sam2d_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

using namespace sam2d;

namespace sam2d {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

const char *last_error() { return g_last_error.c_str(); }

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------- device memory

int dev_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return SAM2D_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(SAM2D_ERR_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return SAM2D_OK;
}

void dev_free(void *p)
{
    if (p) (void)hipFree(p);
}

int upload(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return SAM2D_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return SAM2D_OK;
}

DeviceDomain::~DeviceDomain()
{
    dev_free(KX); if (KZ != KX) dev_free(KZ);
    dev_free(BX); dev_free(BZ); dev_free(B); dev_free(prof_x); dev_free(prof_z); dev_free(ADJ);
}

ShotBuffers::~ShotBuffers() { release(); }

void ShotBuffers::release()
{
    dev_free(state); dev_free(amp); dev_free(gcell); dev_free(seis); dev_free(band);
    state = nullptr; amp = nullptr; gcell = nullptr; seis = nullptr; band = nullptr;
    cells = 0; cap_state = 0; cap_nts = 0; cap_ng = 0; cap_seis = 0; cap_band = 0;
}

int ShotBuffers::reserve_band(size_t n)
{
    if (n == 0) n = 1;
    if (n > cap_band) {
        dev_free(band); band = nullptr; cap_band = 0;
        SAM2D_TRY(dev_alloc((void **)&band, sizeof(float) * n));
        cap_band = n;
    }
    return SAM2D_OK;
}

int ShotBuffers::reserve(size_t ncells, int nslots, int nts, int ng, size_t nseis)
{
    const size_t need = ncells * 4 * (size_t)nslots;
    if (need > cap_state) {
        dev_free(state); state = nullptr; cap_state = 0;
        SAM2D_TRY(dev_alloc((void **)&state, need * sizeof(float)));
        cap_state = need;
    }
    cells = ncells;
    if (nts > cap_nts) {
        dev_free(amp); amp = nullptr;
        SAM2D_TRY(dev_alloc((void **)&amp, (size_t)nts * sizeof(float)));
        cap_nts = nts;
    }
    if (ng > cap_ng) {
        dev_free(gcell); gcell = nullptr;
        SAM2D_TRY(dev_alloc((void **)&gcell, (size_t)ng * sizeof(int)));
        cap_ng = ng;
    }
    if (nseis > cap_seis) {
        dev_free(seis); seis = nullptr;
        SAM2D_TRY(dev_alloc((void **)&seis, nseis * sizeof(float)));
        cap_seis = nseis;
    }
    return SAM2D_OK;
}

// ------------------------------------------------------------------ domains

int get_domain(sam2d_ctx *ctx, int wx0, int w, DeviceDomain **out)
{
    const long long key = ((long long)wx0 << 32) | (unsigned)w;
    auto it = ctx->domains.find(key);
    if (it != ctx->domains.end()) { *out = it->second.get(); return SAM2D_OK; }
    std::unique_ptr<DeviceDomain> d(new DeviceDomain());
    build_domain(ctx->grid, ctx->ta, ctx->vp.data(), ctx->rho.data(), wx0, w, d->host);
    const DomainHost &h = d->host;
    const size_t N = (size_t)h.nnx * h.nnz;
    SAM2D_TRY(dev_alloc((void **)&d->KX, N * 4));
    SAM2D_TRY(dev_alloc((void **)&d->BX, N * 4));
    SAM2D_TRY(dev_alloc((void **)&d->BZ, N * 4));
    SAM2D_TRY(dev_alloc((void **)&d->B, N * 4));
    if (h.kz_alias) d->KZ = d->KX;
    else SAM2D_TRY(dev_alloc((void **)&d->KZ, N * 4));
    SAM2D_TRY(dev_alloc((void **)&d->prof_x, h.prof_x.size() * 4));
    SAM2D_TRY(dev_alloc((void **)&d->prof_z, h.prof_z.size() * 4));
    SAM2D_TRY(upload(d->KX, h.KX.data(), N * 4, ctx->stream));
    if (!h.kz_alias) SAM2D_TRY(upload(d->KZ, h.KZ.data(), N * 4, ctx->stream));
    SAM2D_TRY(upload(d->BX, h.BX.data(), N * 4, ctx->stream));
    SAM2D_TRY(upload(d->BZ, h.BZ.data(), N * 4, ctx->stream));
    SAM2D_TRY(upload(d->B, h.B.data(), N * 4, ctx->stream));
    SAM2D_TRY(upload(d->prof_x, h.prof_x.data(), h.prof_x.size() * 4, ctx->stream));
    SAM2D_TRY(upload(d->prof_z, h.prof_z.data(), h.prof_z.size() * 4, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // The FP32 host copies are no longer needed; kappa stays for the source.
    d->host.KX = std::vector<float>(); d->host.KZ = std::vector<float>();
    d->host.BX = std::vector<float>(); d->host.BZ = std::vector<float>();
    d->host.B = std::vector<float>();
    *out = d.get();
    ctx->domains[key] = std::move(d);
    return SAM2D_OK;
}

// Validate a shot against its domain and fill the host part of its plan.
int plan_shot(sam2d_ctx *ctx, const sam2d_shot &sh, DeviceDomain *dom, ShotPlan &p)
{
    const sam2d_grid &g = ctx->grid;
    const DomainHost &h = dom->host;
    p.dom = dom;
    p.ng = sh.nrec;
    if (sh.nrec < 0) return fail(SAM2D_ERR_PARAM, "nrec < 0");
    if (sh.nrec > 0 && !sh.rec_xz) return fail(SAM2D_ERR_PARAM, "rec_xz is NULL");
    int ix, iz;
    coord_to_cell(g, sh.sx, sh.sz, ix, iz);
    if (ix < h.wx0 || ix >= h.wx0 + h.w || iz < 0 || iz >= g.nz)
        return fail(SAM2D_ERR_RECEIVER, "source (%g, %g) outside the modelling domain", sh.sx, sh.sz);
    p.src_x = ix - h.wx0 + g.npml;
    p.src_z = iz + g.npml;
    p.gcell.resize(sh.nrec);
    for (int k = 0; k < sh.nrec; k++) {
        coord_to_cell(g, sh.rec_xz[2 * k], sh.rec_xz[2 * k + 1], ix, iz);
        if (ix < h.wx0 || ix >= h.wx0 + h.w || iz < 0 || iz >= g.nz)
            return fail(SAM2D_ERR_RECEIVER, "receiver %d (%g, %g) outside the modelling domain", k,
                        sh.rec_xz[2 * k], sh.rec_xz[2 * k + 1]);
        p.gcell[k] = (ix - h.wx0 + g.npml) * h.nnz + (iz + g.npml);
    }
    // Source increments, cpumodelling.cpp:136-140: fd_dt*w[it]*kappa[src]/(dx*dz)
    // evaluated in FP64 and rounded once.
    const double ks = h.kappa[(size_t)p.src_x * h.nnz + p.src_z];
    const int nts = ctx->ta.nts;
    p.amp.resize(nts);
    const double area = (double)(g.dx * g.dz);
    for (int n = 0; n < nts; n++) p.amp[n] = (float)(g.fd_dt * ctx->wavelet[n] * ks / area);
    return SAM2D_OK;
}

// Tile geometry of one shot for a kernel variant.
void tile_shape(int nnz, int nnx, int waves, int chunk, int own_rows, int &ngroups, int &nchunks)
{
    const int nstrips = (nnz + own_rows - 1) / own_rows;
    ngroups = (nstrips + waves - 1) / waves;
    nchunks = (nnx + chunk - 1) / chunk;
}

// Tiles of a shot for a variant and march length chunk (d.nnz, d.nnx and the
// x-profile range xin_lo..xin_hi set): z-strip groups and chunk columns.  For
// a multi-step variant the x-frame goes into chunk columns of its own, over
// [0, x0) and [x1, nnx), with x0 and x1 chosen so that the march of every
// other tile lies in the kernel's fast range (fast_tile in
// sam2d_kernelsn.h): those tiles carry no sponge-frame columns.  The middle
// part is cut into chunk-wide columns, the frame parts into half as wide ones:
// the frame tiles take the general path, about 1.4 times the instructions per
// column, and should finish no later than the others when all tiles of a
// launch are resident at once (one small shot).
void shot_tiles(ShotDev &d, int variant, int chunk)
{
    int waves, vchunk, pf;
    variant_shape_pf(variant, &waves, &vchunk, &pf);
    const int T = variant_steps(variant);
    tile_shape(d.nnz, d.nnx, waves, chunk, variant_own_rows(variant), d.ngroups, d.nchunks);
    d.chunk = chunk;
    d.x0 = d.x1 = 0;
    if (T < 2 || d.xin_lo > d.xin_hi || chunk < 1) return;
    const int fast_lo = std::max(d.xin_lo + 3 * T - 1, 3 * T);
    const int fast_hi = std::min(d.xin_hi - 4 - pf, d.nnx - 5);
    const int x0 = fast_lo + 3 * T - 2;          // first march column = fast_lo
    const int xmax = fast_hi - 3 * T + 5;        // last group's first column = fast_hi
    if (xmax - x0 < chunk) return;
    const int m = (xmax - x0) / chunk;
    int w = (xmax - x0) - m * chunk;
    w -= (w + 6 * T - 3) & 3;                    // a march of whole four-column groups
    if (w < 0) w = 0;
    d.x0 = x0;
    d.x1 = x0 + m * chunk + w;
    d.fchunk = std::max((chunk + 1) / 2, 8);
    d.nlead = (x0 + d.fchunk - 1) / d.fchunk;
    d.nmid = m + (w > 0 ? 1 : 0);
    d.nchunks = d.nlead + d.nmid + (d.nnx - d.x1 + d.fchunk - 1) / d.fchunk;
}

int pick_variant(sam2d_ctx *ctx)
{
    if (ctx->variant >= 0) return ctx->variant;
    if (ctx->tuned_variant >= 0) return ctx->tuned_variant;
    return ctx->default_variant;
}

// Indices [lo, hi] around the middle where all four profiles of a [4][nnx]
// table are exactly 1; lo > hi when the middle index is not.
void ones_range(const std::vector<float> &prof_x, int nnx, int &lo, int &hi)
{
    auto ones = [&](int c) {
        for (int k = 0; k < 4; k++)
            if (prof_x[(size_t)k * nnx + c] != 1.0f) return false;
        return true;
    };
    const int mid = nnx / 2;
    if (!ones(mid)) { lo = 1; hi = 0; return; }
    lo = hi = mid;
    while (lo > 0 && ones(lo - 1)) lo--;
    while (hi < nnx - 1 && ones(hi + 1)) hi++;
}

int tile_chunk(const sam2d_ctx *ctx, int variant)
{
    int waves, chunk;
    variant_shape(variant, &waves, &chunk);
    const int steps = variant_steps(variant);
    if (steps == 1) return chunk;
    return ctx->chunk > 0 ? ctx->chunk : march_aligned(chunk, 6 * steps - 3);
}

// Descriptor of one shot for `variant` with march length `chunk`, its tiles
// numbered from tile_begin; the device buffers (state, amp, gcell, seis,
// band) are left for the caller.  Rows spanned by the receivers: the
// multi-step kernels store the intermediate pressure of these rows into a
// band buffer (two parities x BAND_SLICES slices of band_h x nnx) for the
// deferred gather.
void describe_shot(const ShotPlan &p, int variant, int chunk, int tile_begin, int nts, int real_nt,
                   ShotDev &d)
{
    const DeviceDomain *dom = p.dom;
    const int nnz = dom->host.nnz, nnx = dom->host.nnx;
    int rlo = 1, rhi = 0;
    for (int g = 0; g < p.ng; g++) {
        const int row = p.gcell[g] % nnz;
        if (g == 0 || row < rlo) rlo = row;
        if (g == 0 || row > rhi) rhi = row;
    }
    std::memset(&d, 0, sizeof d);
    d.nnz = nnz; d.nnx = nnx;
    ones_range(dom->host.prof_x, nnx, d.xin_lo, d.xin_hi);
    ones_range(dom->host.prof_z, nnz, d.zin_lo, d.zin_hi);
    shot_tiles(d, variant, chunk);
    d.tile_begin = tile_begin;
    d.src_x = p.src_x; d.src_z = p.src_z;
    d.nts = nts; d.ng = p.ng; d.real_nt = real_nt;
    d.rrow_lo = rlo; d.rrow_hi = rhi; d.band_h = p.ng > 0 ? rhi - rlo + 1 : 0;
    d.KX = dom->KX; d.KZ = dom->KZ; d.BX = dom->BX; d.BZ = dom->BZ; d.B = dom->B;
    d.dtdx = (float)dom->host.dtdx; d.dtdz = (float)dom->host.dtdz;
    d.prof_x = dom->prof_x; d.prof_z = dom->prof_z;
    d.ncell = (size_t)nnz * nnx;
}

// 32-bit byte offsets of the multi-step kernels: a state slot spans 16 N
// bytes (the 12-byte cells, then pz), a coefficient field 4 N.
int check_multistep_size(const sam2d_ctx *ctx, int variant, size_t ncell)
{
    if (variant_steps(variant) > 1 && ncell >= ((size_t)1 << 28))
        return fail(SAM2D_ERR_PARAM, "grid of %zu cells too large for a multi-step variant "
                    "(32-bit byte offsets); choose a single-step variant", ncell);
    return SAM2D_OK;
}

// Make `plans` resident: buffers, descriptors, zeroed state and seismograms.
int stage_batch(sam2d_ctx *ctx, std::vector<ShotPlan> &plans, int variant, int *ntiles_out,
                int nslots)
{
    const int chunk = tile_chunk(ctx, variant);
    const int ns = (int)plans.size();
    if ((int)ctx->bufs.size() < ns) ctx->bufs.resize(ns);
    std::vector<ShotDev> desc(ns);
    int tiles = 0;
    const int real_nt = ctx->ta.real_nt, nts = ctx->ta.nts;
    for (int s = 0; s < ns; s++) {
        ShotPlan &p = plans[s];
        DeviceDomain *dom = p.dom;
        const int nnx = dom->host.nnx;
        const size_t N = (size_t)dom->host.nnz * nnx;
        SAM2D_TRY(check_multistep_size(ctx, variant, N));
        if (!ctx->bufs[s]) ctx->bufs[s].reset(new ShotBuffers());
        ShotBuffers &b = *ctx->bufs[s];
        SAM2D_TRY(b.reserve(N, nslots, nts, p.ng, (size_t)p.ng * real_nt));
        SAM2D_TRY(upload(b.amp, p.amp.data(), sizeof(float) * nts, ctx->stream));
        SAM2D_TRY(upload(b.gcell, p.gcell.data(), sizeof(int) * p.ng, ctx->stream));
        ShotDev &d = desc[s];
        describe_shot(p, variant, chunk, tiles, nts, real_nt, d);
        tiles += d.ngroups * d.nchunks;
        SAM2D_TRY(b.reserve_band((size_t)2 * BAND_SLICES * d.band_h * nnx));
        // Every slot starts at zero: the kernels write pz only in the sponge
        // frame and read it as 0 elsewhere.
        HIP_TRY(launch_fill(b.slot(0), (size_t)nslots * 4 * N, 0.0f, ctx->stream));
        if (p.ng) HIP_TRY(launch_fill(b.seis, (size_t)p.ng * real_nt, 0.0f, ctx->stream));
        d.amp = b.amp;
        d.state = b.state;
        d.gcell = b.gcell; d.seis = b.seis;
        d.band = b.band;
        p.seis_dev = b.seis;
    }
    if (ns > ctx->cap_desc) {
        // Captured graphs hold the old descriptor address: drop them.
        for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
        ctx->graphs.clear();
        dev_free(ctx->d_desc); ctx->d_desc = nullptr;
        SAM2D_TRY(dev_alloc((void **)&ctx->d_desc, sizeof(ShotDev) * ns));
        ctx->cap_desc = ns;
    }
    SAM2D_TRY(upload(ctx->d_desc, desc.data(), sizeof(ShotDev) * ns, ctx->stream));
    ctx->staged_desc = desc;
    *ntiles_out = tiles;
    return SAM2D_OK;
}

// Receiver sample written by step kernel `n` (the state it reads is the
// result of step n-1): cpumodelling.cpp:355-360 records after step kt when
// (kt+1) % dt_factor == 0 into sample (kt+1)/dt_factor - 1.
static int rec_slot_for_kernel(const TimeAxis &ta, int n)
{
    if (n < 1 || n % ta.dt_factor != 0) return -1;
    const int slot = n / ta.dt_factor - 1;
    return slot < ta.real_nt ? slot : -1;
}

// Receiver sample recorded after step n, or -1.
static int rec_after(const TimeAxis &ta, int n)
{
    if ((n + 1) % ta.dt_factor != 0) return -1;
    const int k = (n + 1) / ta.dt_factor - 1;
    return k < ta.real_nt ? k : -1;
}

// Enqueue steps [n0, n1) of the time loop starting from state slot
// start_slot; *end_slot receives the slot holding the final state.
// Single-step variants record step n-1 inside kernel n (deferred gather), and
// the receivers of the last step when n1 is the last step.  Multi-step
// variants (T = 2 to 5 steps per launch) gather the receivers of each launch
// at the start of the next one (intermediate steps from the band buffer, the
// last from the stored state); fewer than T remaining steps run the
// single-step kernel; a final gather launch records what is still pending.
int enqueue_steps(sam2d_ctx *ctx, int nshots, int ntiles, int variant, int n0, int n1, int max_ng,
                  int start_slot, int *end_slot)
{
    const int ngroups = ctx->staged_desc.empty() ? 0 : ctx->staged_desc[0].ngroups;
    return enqueue_steps_on(ctx, ctx->d_desc, ngroups, nshots, ntiles, variant, n0, n1, max_ng,
                            start_slot, end_slot);
}

int enqueue_steps_on(sam2d_ctx *ctx, const ShotDev *desc, int ngroups, int nshots, int ntiles,
                     int variant, int n0, int n1, int max_ng, int start_slot, int *end_slot)
{
    const TileOrder order{ngroups, ctx->tile_order};
    const TimeAxis &ta = ctx->ta;
    int cur = start_slot;
    int n = n0;
    const int T = variant_steps(variant);
    if (T > 1) {
        const int single = variant_single_step(variant);
        const bool kza = ta.dtdx == ta.dtdz;
        // Samples still to gather: band slices of parity band_w ^ 1, and the
        // current state.
        int pend[BAND_SLICES] = {-1, -1, -1, -1};
        int pend_state = -1, band_w = 0;
        for (; n + T - 1 < n1; n += T) {
            int mask = 0, rec[BAND_SLICES] = {-1, -1, -1, -1};
            for (int j = 0; j < T - 1; j++) {
                rec[j] = rec_after(ta, n + j);
                if (rec[j] >= 0) mask |= 1 << j;
            }
            HIP_TRY(launch_stepn(variant, desc, nshots, ntiles, ctx->grid_cap, kza, n, cur, cur ^ 1,
                                 mask, band_w, pend, pend_state, max_ng, order, ctx->stream));
            for (int j = 0; j < BAND_SLICES; j++) pend[j] = rec[j];
            pend_state = rec_after(ta, n + T - 1);
            cur ^= 1;
            band_w ^= 1;
        }
        if (n < n1) {
            // Fewer than T steps left: flush, then single steps (each kernel
            // gathers the step before it).
            HIP_TRY(launch_gather(desc, nshots, max_ng, band_w ^ 1, pend, cur, pend_state,
                                  ctx->stream));
            for (int j = 0; j < BAND_SLICES; j++) pend[j] = -1;
            for (bool first = true; n < n1; n++, first = false) {
                HIP_TRY(launch_step(single, desc, nshots, ntiles, ctx->grid_cap, n, cur, cur ^ 1,
                                    first ? -1 : rec_slot_for_kernel(ta, n), ctx->stream));
                cur ^= 1;
            }
            pend_state = rec_after(ta, n1 - 1);
        }
        HIP_TRY(launch_gather(desc, nshots, max_ng, band_w ^ 1, pend, cur, pend_state,
                              ctx->stream));
    } else {
        for (; n < n1; n++) {
            HIP_TRY(launch_step(variant, desc, nshots, ntiles, ctx->grid_cap, n, cur, cur ^ 1,
                                rec_slot_for_kernel(ta, n), ctx->stream));
            cur ^= 1;
        }
        if (n1 == ta.fd_nt) {
            const int slot = rec_slot_for_kernel(ta, n1);
            if (slot >= 0 && max_ng > 0)
                HIP_TRY(launch_record(desc, nshots, max_ng, cur, slot, ctx->stream));
        }
    }
    if (end_slot) *end_slot = cur;
    return SAM2D_OK;
}

// Run the whole time loop of the staged batch, as a replayed hipGraph when
// enabled.  Adds the step-kernel time to ctx->stats.
int run_time_loop(sam2d_ctx *ctx, int nshots, int ntiles, int variant, int max_ng)
{
    int waves, chunk;
    variant_shape(variant, &waves, &chunk);
    const int fd_nt = ctx->ta.fd_nt;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    bool used_graph = false;
    if (ctx->use_graph) {
        GraphKey key{nshots, ntiles, variant, max_ng, ctx->grid_cap, ctx->tile_order,
                     tile_chunk(ctx, variant)};
        hipGraphExec_t exec = nullptr;
        auto it = ctx->graphs.find(key);
        if (it != ctx->graphs.end()) exec = it->second;
        if (!exec) {
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_steps(ctx, nshots, ntiles, variant, 0, fd_nt, max_ng, 0, nullptr);
            hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            if (rc != SAM2D_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            HIP_TRY(e);
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_TRY(e);
            // The first capture of a shape pays instantiation; the timing below
            // brackets only the replay.
            HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
            if (ctx->graphs.size() > 16) {
                for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
                ctx->graphs.clear();
            }
            ctx->graphs[key] = exec;
        }
        HIP_TRY(hipGraphLaunch(exec, ctx->stream));
        used_graph = true;
    } else {
        SAM2D_TRY(enqueue_steps(ctx, nshots, ntiles, variant, 0, fd_nt, max_ng, 0, nullptr));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms += ms;
    const int spl = variant_steps(variant);
    ctx->stats.launches += (fd_nt + spl - 1) / spl;
    ctx->stats.steps_per_launch = spl;
    ctx->stats.used_graph = used_graph ? 1 : 0;
    ctx->stats.variant = variant;
    return SAM2D_OK;
}

// Forward over a list of shot plans, `batch` shots resident at a time.
// seis: host (dev_out == false) or device destination, concatenated.
int forward_plans(sam2d_ctx *ctx, std::vector<ShotPlan> &plans, int batch, float *seis, bool dev_out)
{
    const int real_nt = ctx->ta.real_nt;
    size_t off = 0;
    const int ns = (int)plans.size();
    for (int s0 = 0; s0 < ns; s0 += batch) {
        const int s1 = std::min(ns, s0 + batch);
        std::vector<ShotPlan> sub(plans.begin() + s0, plans.begin() + s1);
        const int variant = pick_variant(ctx);
        int ntiles = 0;
        SAM2D_TRY(stage_batch(ctx, sub, variant, &ntiles));
        int max_ng = 0;
        for (auto &p : sub) {
            max_ng = std::max(max_ng, p.ng);
            ctx->stats.cell_updates += (double)p.dom->host.nnx * p.dom->host.nnz * ctx->ta.fd_nt;
        }
        SAM2D_TRY(run_time_loop(ctx, (int)sub.size(), ntiles, variant, max_ng));
        for (auto &p : sub) {
            const size_t n = (size_t)p.ng * real_nt;
            if (n) HIP_TRY(hipMemcpyAsync(seis + off, p.seis_dev, n * sizeof(float),
                                          dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                          ctx->stream));
            off += n;
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return SAM2D_OK;
}

// The reference's carry-over between shots (cpumodelling.cpp:343-348): the
// next shot starts from the previous final state with only the first nnz
// cells of p, px, pz, vx, vz cleared.  The previous shot is the last one this
// context modelled in this mode, in this call or an earlier one (the
// reference's modelling_module runs its shots one after another on the same
// arrays); any other modelling call ends the chain.  Per-shot loop only, no
// graph.
static int forward_reset_quirk(sam2d_ctx *ctx, std::vector<ShotPlan> &plans, float *seis)
{
    const int real_nt = ctx->ta.real_nt;
    size_t off = 0;
    const int variant = pick_variant(ctx);
    for (size_t s = 0; s < plans.size(); s++) {
        std::vector<ShotPlan> one(1, plans[s]);
        // Keep the previous state: stage into slot 0 without its clearing.
        ShotBuffers *b0 = ctx->bufs.empty() || !ctx->bufs[0] ? nullptr : ctx->bufs[0].get();
        const size_t N = (size_t)one[0].dom->host.nnx * one[0].dom->host.nnz;
        std::vector<float> keep;
        if (ctx->quirk_valid && ctx->quirk_cells == N && b0) {
            // Bring the final state into set 0, then clear the first nnz cells.
            keep.resize(4 * N);
            HIP_TRY(hipMemcpyAsync(keep.data(), b0->slot(ctx->quirk_slot), 4 * N * 4,
                                   hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            const int nnz = one[0].dom->host.nnz;
            // cells 0 .. nnz-1 (column 0, in the sponge frame): their
            // (px, vx, vz) and their pz
            std::fill(keep.begin(), keep.begin() + 3 * (size_t)nnz, 0.0f);
            std::fill(keep.begin() + 3 * N, keep.begin() + 3 * N + nnz, 0.0f);
        }
        int ntiles = 0;
        ctx->quirk_valid = false;
        SAM2D_TRY(stage_batch(ctx, one, variant, &ntiles));
        if (!keep.empty()) {
            ShotBuffers &b = *ctx->bufs[0];
            SAM2D_TRY(upload(b.slot(0), keep.data(), 4 * N * 4, ctx->stream));
        }
        ctx->stats.cell_updates += (double)N * ctx->ta.fd_nt;
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        int fin = 0;
        SAM2D_TRY(enqueue_steps(ctx, 1, ntiles, variant, 0, ctx->ta.fd_nt, one[0].ng, 0, &fin));
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->stats.kernel_ms += ms;
        ctx->stats.launches += (ctx->ta.fd_nt + variant_steps(variant) - 1) / variant_steps(variant);
        ctx->stats.steps_per_launch = variant_steps(variant);
        const size_t n = (size_t)one[0].ng * real_nt;
        if (n) HIP_TRY(hipMemcpy(seis + off, one[0].seis_dev, n * 4, hipMemcpyDeviceToHost));
        off += n;
        ctx->quirk_valid = true;
        ctx->quirk_slot = fin;
        ctx->quirk_cells = N;
    }
    return SAM2D_OK;
}

int make_plans(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, const int *windows,
               std::vector<ShotPlan> &plans)
{
    if (nshots < 0) return fail(SAM2D_ERR_PARAM, "nshots < 0");
    if (nshots > 0 && !shots) return fail(SAM2D_ERR_PARAM, "shots is NULL");
    plans.assign(nshots, ShotPlan());
    for (int s = 0; s < nshots; s++) {
        int wx0 = 0, w = ctx->grid.nx;
        if (windows) {
            wx0 = windows[2 * s];
            w = windows[2 * s + 1];
            if (wx0 < 0 || w < 1 || wx0 + w > ctx->grid.nx)
                return fail(SAM2D_ERR_PARAM, "window %d (%d, %d) outside the model (nx=%d)", s, wx0,
                            w, ctx->grid.nx);
            if (w + 2 * ctx->grid.npml < 4)
                return fail(SAM2D_ERR_PARAM, "window %d too narrow", s);
        }
        DeviceDomain *dom = nullptr;
        SAM2D_TRY(get_domain(ctx, wx0, w, &dom));
        SAM2D_TRY(plan_shot(ctx, shots[s], dom, plans[s]));
    }
    return SAM2D_OK;
}

static size_t seis_total(const sam2d_shot *shots, int nshots, int real_nt)
{
    size_t n = 0;
    for (int s = 0; s < nshots; s++) n += (size_t)shots[s].nrec * real_nt;
    return n;
}

}  // namespace sam2d

// ======================================================================= ABI

extern "C" {

const char *sam2d_version(void) { return "sam2d-mi355x 0.1 (gfx950)"; }

const char *sam2d_last_error(void) { return sam2d::last_error(); }

int sam2d_num_variants(void) { return num_variants(); }

const char *sam2d_variant_name(int v)
{
    static thread_local char buf[64];
    if (v < 0 || v >= num_variants()) return "invalid";
    variant_name(v, buf, sizeof buf);
    return buf;
}

int sam2d_check_stability(const sam2d_grid *g, float vmin, float vmax)
{
    clear_error();
    if (!g) return fail(SAM2D_ERR_PARAM, "grid is NULL");
    // check_stability.cpp:19-64, float arithmetic as there.
    const float fm = 2.0f * g->dom_freq;
    const float hmin = g->dx < g->dz ? g->dx : g->dz;
    const float hmax = g->dx > g->dz ? g->dx : g->dz;
    float criteria = (0.606f * hmin) / vmax;
    if (g->fd_dt > criteria)
        return fail(SAM2D_ERR_STABILITY, "stability condition violated: dt %g > %g", g->fd_dt, criteria);
    criteria = vmin / (hmax * 5.0f);
    if (fm > criteria)
        return fail(SAM2D_ERR_STABILITY, "dispersion condition violated: peak frequency %g > %g",
                    g->dom_freq, criteria / 2.0f);
    return SAM2D_OK;
}

int sam2d_create(const sam2d_grid *g, const float *vp, const float *rho, sam2d_ctx **out)
{
    clear_error();
    if (!out) return fail(SAM2D_ERR_PARAM, "out is NULL");
    *out = nullptr;
    if (!g || !vp || !rho) return fail(SAM2D_ERR_PARAM, "NULL grid or model");
    if (g->nx < 1 || g->nz < 1) return fail(SAM2D_ERR_PARAM, "bad grid size %d x %d", g->nx, g->nz);
    if (!(g->dx > 0.0f) || !(g->dz > 0.0f)) return fail(SAM2D_ERR_PARAM, "grid spacing must be > 0");
    if (!(g->fd_dt > 0.0f) || !(g->real_dt > 0.0f) || !(g->rec_len > 0.0f))
        return fail(SAM2D_ERR_PARAM, "time parameters must be > 0");
    if (g->npml < 1) return fail(SAM2D_ERR_PARAM, "npml must be >= 1");
    if (g->fdop != 4)
        return fail(SAM2D_ERR_PARAM, "only the 4th-order operator is implemented (fdop=%d)", g->fdop);
    // json_parser.cpp:189-190
    if (std::fmod(g->real_dt, g->fd_dt) != 0.0f)
        return fail(SAM2D_ERR_PARAM, "DATA DT must be a multiple of FD DT");
    if (g->nz + 2 * g->npml < 4 || g->nx + 2 * g->npml < 4)
        return fail(SAM2D_ERR_PARAM, "padded grid too small");
    const size_t n = (size_t)g->nx * g->nz;
    if ((size_t)(g->nz + 2 * g->npml) * (size_t)(g->nx + 2 * g->npml) > (size_t)0x7fffffff)
        return fail(SAM2D_ERR_PARAM, "padded grid exceeds 2^31 cells");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(SAM2D_ERR_DEVICE, "no HIP device available");
    }
    std::unique_ptr<sam2d_ctx> ctx(new sam2d_ctx());
    ctx->grid = *g;
    ctx->ta = time_axis(*g);
    if (ctx->ta.dt_factor < 1) return fail(SAM2D_ERR_PARAM, "DATA DT < FD DT");
    ctx->vp.assign(vp, vp + n);
    ctx->rho.assign(rho, rho + n);
    ctx->wavelet = ricker(g->dom_freq, g->fd_dt, ctx->ta.nts);
    HIP_TRY(hipGetDevice(&ctx->device));
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&ctx->ev0));
    HIP_TRY(hipEventCreate(&ctx->ev1));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
    ctx->num_cu = prop.multiProcessorCount;
    ctx->default_variant = default_variant_index();
    void *scratch_ptr = nullptr;                   // sam2d_forward_noalloc's array on this device
    SAM2D_TRY(noalloc_scratch(ctx->device, &scratch_ptr));
    DeviceDomain *dom = nullptr;
    SAM2D_TRY(get_domain(ctx.get(), 0, g->nx, &dom));
    *out = ctx.release();
    return SAM2D_OK;
}

int sam2d_get_dims(const sam2d_ctx *ctx, sam2d_dims *out)
{
    if (!ctx || !out) return fail(SAM2D_ERR_PARAM, "NULL argument");
    out->fd_nt = ctx->ta.fd_nt;
    out->real_nt = ctx->ta.real_nt;
    out->dt_factor = ctx->ta.dt_factor;
    out->nts = ctx->ta.nts;
    out->nnx = ctx->grid.nx + 2 * ctx->grid.npml;
    out->nnz = ctx->grid.nz + 2 * ctx->grid.npml;
    return SAM2D_OK;
}

static int forward_common(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, const int *windows,
                          float *seis, bool dev_out, bool batched)
{
    clear_error();
    if (!ctx) return fail(SAM2D_ERR_PARAM, "ctx is NULL");
    if (nshots > 0 && !seis && seis_total(shots, nshots, ctx->ta.real_nt) > 0)
        return fail(SAM2D_ERR_PARAM, "seis is NULL");
    const double t0 = now_ms();
    ctx->stats = sam2d_stats();
    std::vector<ShotPlan> plans;
    SAM2D_TRY(make_plans(ctx, shots, nshots, windows, plans));
    if (ctx->reset_quirk && !batched && !dev_out && !windows) {
        SAM2D_TRY(forward_reset_quirk(ctx, plans, seis));
    } else {
        ctx->quirk_valid = false;
        int batch = 1;
        if (batched) {
            batch = ctx->max_batch > 0 ? ctx->max_batch : nshots;
            if (batch < 1) batch = 1;
            if (ctx->variant < 0 && ctx->tuned_variant < 0 && nshots > 0)
                SAM2D_TRY(sam2d_autotune(ctx, windows, nshots, 8, nullptr));
        }
        SAM2D_TRY(forward_plans(ctx, plans, batch, seis, dev_out));
    }
    ctx->stats.wall_ms = now_ms() - t0;
    return SAM2D_OK;
}

int sam2d_forward(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, float *seis)
{
    return forward_common(ctx, shots, nshots, nullptr, seis, false, false);
}

int sam2d_forward_dev(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, float *seis_dev)
{
    return forward_common(ctx, shots, nshots, nullptr, seis_dev, true, false);
}

int sam2d_forward_windows(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, const int *windows,
                          float *seis)
{
    return forward_common(ctx, shots, nshots, windows, seis, false, false);
}

int sam2d_forward_batched(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, const int *windows,
                          float *seis)
{
    return forward_common(ctx, shots, nshots, windows, seis, false, true);
}

int sam2d_set_option(sam2d_ctx *ctx, const char *key, int value)
{
    clear_error();
    if (!ctx || !key) return fail(SAM2D_ERR_PARAM, "NULL argument");
    std::string k(key);
    if (k == "graph") ctx->use_graph = value != 0;
    else if (k == "variant") {
        if (value >= num_variants()) return fail(SAM2D_ERR_PARAM, "variant %d out of range", value);
        ctx->variant = value < 0 ? -1 : value;
    } else if (k == "reset_quirk") {
        ctx->reset_quirk = value != 0;
        ctx->quirk_valid = false;            // the next shot starts fresh
    }
    else if (k == "max_batch") ctx->max_batch = value < 0 ? 0 : value;
    else if (k == "grid_cap") ctx->grid_cap = value < 0 ? 0 : value;
    else if (k == "hbm_budget_gb") {
        if (value < 0) return fail(SAM2D_ERR_PARAM, "hbm_budget_gb must be >= 0");
        ctx->hbm_budget_gb = value;
    } else if (k == "adjoint_steps") {
        if (value != 1 && value != 2) return fail(SAM2D_ERR_PARAM, "adjoint_steps must be 1 or 2");
        ctx->adjoint_steps = value;
    } else if (k == "tile_order") {
        if (value < 0 || value > 64) return fail(SAM2D_ERR_PARAM, "tile_order must be 0..64");
        ctx->tile_order = value;
    } else if (k == "chunk") {
        if (value != 0 && (value < 8 || value > 4096))
            return fail(SAM2D_ERR_PARAM, "chunk must be 0 or 8..4096");
        ctx->chunk = value;
    } else if (k == "adjoint_tile_order") {
        if (value < 0 || value > 64) return fail(SAM2D_ERR_PARAM, "adjoint_tile_order must be 0..64");
        ctx->adjoint_tile_order = value;
    } else if (k == "adjoint_variant") {
        if (value >= num_variants()) return fail(SAM2D_ERR_PARAM, "variant %d out of range", value);
        ctx->adjoint_variant = value < 0 ? -1 : variant_single_step(value);
    } else if (k == "ring_variant") {
        if (value >= num_variants() || (value >= 0 && !ring_supported(value)))
            return fail(SAM2D_ERR_PARAM, "ring_variant %d is not a three-step variant", value);
        ctx->ring_variant = value < 0 ? -1 : value;
    } else if (k == "ring_chunk") {
        if (value != 0 && (value < 8 || value > 4096))
            return fail(SAM2D_ERR_PARAM, "ring_chunk must be 0 or 8..4096");
        ctx->ring_chunk = value;
    }
    else return fail(SAM2D_ERR_PARAM, "unknown option '%s'", key);
    return SAM2D_OK;
}

int sam2d_get_option(sam2d_ctx *ctx, const char *key, int *value)
{
    clear_error();
    if (!ctx || !key || !value) return fail(SAM2D_ERR_PARAM, "NULL argument");
    std::string k(key);
    if (k == "graph") *value = ctx->use_graph ? 1 : 0;
    else if (k == "variant") *value = ctx->variant;
    else if (k == "reset_quirk") *value = ctx->reset_quirk ? 1 : 0;
    else if (k == "max_batch") *value = ctx->max_batch;
    else if (k == "grid_cap") *value = ctx->grid_cap;
    else if (k == "hbm_budget_gb") *value = ctx->hbm_budget_gb;
    else if (k == "adjoint_steps") *value = ctx->adjoint_steps;
    else if (k == "adjoint_variant") *value = ctx->adjoint_variant;
    else if (k == "ring_variant") *value = ctx->ring_variant;
    else if (k == "ring_chunk") *value = ctx->ring_chunk;
    else if (k == "tile_order") *value = ctx->tile_order;
    else if (k == "adjoint_tile_order") *value = ctx->adjoint_tile_order;
    else if (k == "chunk") *value = ctx->chunk;
    else return fail(SAM2D_ERR_PARAM, "unknown option '%s'", key);
    return SAM2D_OK;
}

int sam2d_autotune(sam2d_ctx *ctx, const int *windows, int nshots, int steps, int *best_variant)
{
    clear_error();
    if (!ctx) return fail(SAM2D_ERR_PARAM, "ctx is NULL");
    if (steps < 1) steps = 4;
    if (nshots < 1) nshots = 1;
    // Synthetic shots at the centre of each window (or of the model), no
    // receivers: the tuner times the step kernel alone.
    std::vector<sam2d_shot> shots(nshots);
    for (int s = 0; s < nshots; s++) {
        int wx0 = windows ? windows[2 * s] : 0;
        int w = windows ? windows[2 * s + 1] : ctx->grid.nx;
        shots[s].sx = (float)(wx0 + w / 2) * ctx->grid.dx;
        shots[s].sz = 0.0f;
        shots[s].nrec = 0;
        shots[s].rec_xz = nullptr;
    }
    std::vector<ShotPlan> plans;
    SAM2D_TRY(make_plans(ctx, shots.data(), nshots, windows, plans));
    ctx->quirk_valid = false;
    int batch = ctx->max_batch > 0 ? std::min(ctx->max_batch, nshots) : nshots;
    plans.resize(batch);
    double best_ms = 1e300;
    int best = default_variant_index();
    ctx->tune_ms.assign(num_variants(), 0.0);
    ctx->tile_order = 0;                       // variants are compared in the plain order
    ctx->chunk = 0;                            // and with their own march lengths
    for (int v = 0; v < num_variants(); v++) {
        int waves, chunk;
        variant_shape(v, &waves, &chunk);
        int ntiles = 0;
        const int rc = stage_batch(ctx, plans, v, &ntiles);
        if (rc == SAM2D_ERR_PARAM && variant_steps(v) > 1) {
            clear_error();             // grid beyond the multi-step limit: not a candidate
            continue;
        }
        SAM2D_TRY(rc);
        SAM2D_TRY(enqueue_steps(ctx, batch, ntiles, v, 0, 2, 0, 0, nullptr));   // warm
        HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
        SAM2D_TRY(enqueue_steps(ctx, batch, ntiles, v, 2, 2 + steps, 0, 0, nullptr));
        HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->tune_ms[v] = ms / steps;
        if (ms < best_ms) { best_ms = ms; best = v; }
    }
    // Second pass, for the six fastest multi-step variants: the march
    // length (the variant's chunk scaled by 1/2 .. 3/2, rounded to whole
    // four-column groups of the march)
    // jointly with the chunk-column order (stride 1 = in order, 2..9), each
    // combination timed twice over four times the steps.  The four fastest
    // combinations are then timed three times over twelve times the steps
    // (they differ by a few per cent, as much as a short timing varies) and
    // the fastest is kept.
    struct Combo {
        double ms;
        int variant, chunk, order;
    };
    auto time_combo = [&](const Combo &cb, int nsteps, int reps, double *out) -> int {
        ctx->chunk = cb.chunk;
        ctx->tile_order = cb.order;
        int ntiles = 0;
        SAM2D_TRY(stage_batch(ctx, plans, cb.variant, &ntiles));
        double t = 1e300;
        for (int rep = 0; rep < reps; rep++) {
            SAM2D_TRY(enqueue_steps(ctx, batch, ntiles, cb.variant, 0, 2, 0, 0, nullptr));   // warm
            HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
            SAM2D_TRY(enqueue_steps(ctx, batch, ntiles, cb.variant, 2, 2 + nsteps, 0, 0, nullptr));
            HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
            HIP_TRY(hipEventSynchronize(ctx->ev1));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            t = std::min(t, (double)ms);
        }
        *out = t / nsteps;
        return SAM2D_OK;
    };
    std::vector<int> multi;
    for (int v = 0; v < num_variants(); v++)
        if (variant_steps(v) > 1 && ctx->tune_ms[v] > 0.0) multi.push_back(v);
    std::sort(multi.begin(), multi.end(),
              [&](int a, int b) { return ctx->tune_ms[a] < ctx->tune_ms[b]; });
    if (variant_steps(best) > 1 && !multi.empty()) {
        // At its own march length a variant can rank several places below
        // where its best length puts it (C1: fifth of the first pass, first of
        // the second, 251 against 272 us per launch).
        if (multi.size() > 6) multi.resize(6);
        std::vector<Combo> combos;
        for (int v : multi) {
            int w0, c0;
            variant_shape(v, &w0, &c0);
            std::vector<int> chunks;
            for (int eighths : {4, 5, 6, 7, 8, 9, 10, 12}) {
                const int c = march_aligned(std::max(16, (c0 * eighths / 8 + 2) / 4 * 4),
                                            6 * variant_steps(v) - 3);
                if (std::find(chunks.begin(), chunks.end(), c) == chunks.end())
                    chunks.push_back(c);
            }
            for (int c : chunks)
                for (int xo = 1; xo <= 9; xo++) {
                    Combo cb{0.0, v, c, xo};
                    SAM2D_TRY(time_combo(cb, 4 * steps, 2, &cb.ms));
                    combos.push_back(cb);
                }
        }
        std::sort(combos.begin(), combos.end(),
                  [](const Combo &a, const Combo &b) { return a.ms < b.ms; });
        if (combos.size() > 4) combos.resize(4);
        for (Combo &cb : combos) SAM2D_TRY(time_combo(cb, 12 * steps, 3, &cb.ms));
        const Combo win = *std::min_element(
            combos.begin(), combos.end(),
            [](const Combo &a, const Combo &b) { return a.ms < b.ms; });
        best = win.variant;
        ctx->chunk = win.chunk;
        ctx->tile_order = win.order;
        ctx->tune_ms[best] = win.ms;
    }
    ctx->tuned_variant = best;
    if (best_variant) *best_variant = best;
    return SAM2D_OK;
}

int sam2d_coefficients(const sam2d_grid *g, const float *vp, const float *rho, int wx0, int w,
                       float *KX, float *KZ, float *BX, float *BZ, float *prof_x, float *prof_z)
{
    clear_error();
    if (!g || !vp || !rho) return fail(SAM2D_ERR_PARAM, "NULL argument");
    if (wx0 < 0 || w < 1 || wx0 + w > g->nx) return fail(SAM2D_ERR_PARAM, "bad window");
    const TimeAxis ta = time_axis(*g);
    DomainHost h;
    build_domain(*g, ta, vp, rho, wx0, w, h);
    const size_t N = (size_t)h.nnx * h.nnz;
    if (KX) std::memcpy(KX, h.KX.data(), N * 4);
    if (KZ) std::memcpy(KZ, h.kz_alias ? h.KX.data() : h.KZ.data(), N * 4);
    if (BX) std::memcpy(BX, h.BX.data(), N * 4);
    if (BZ) std::memcpy(BZ, h.BZ.data(), N * 4);
    if (prof_x) std::memcpy(prof_x, h.prof_x.data(), h.prof_x.size() * 4);
    if (prof_z) std::memcpy(prof_z, h.prof_z.data(), h.prof_z.size() * 4);
    return SAM2D_OK;
}

int sam2d_source_increments(const sam2d_grid *g, const float *vp, const float *rho, float sx,
                            float sz, float *amp)
{
    clear_error();
    if (!g || !vp || !rho || !amp) return fail(SAM2D_ERR_PARAM, "NULL argument");
    sam2d_ctx tmp;
    tmp.grid = *g;
    tmp.ta = time_axis(*g);
    tmp.wavelet = ricker(g->dom_freq, g->fd_dt, tmp.ta.nts);
    tmp.vp.assign(vp, vp + (size_t)g->nx * g->nz);
    tmp.rho.assign(rho, rho + (size_t)g->nx * g->nz);
    DeviceDomain dom;
    build_domain(*g, tmp.ta, vp, rho, 0, g->nx, dom.host);
    sam2d_shot sh{sx, sz, 0, nullptr};
    ShotPlan p;
    SAM2D_TRY(plan_shot(&tmp, sh, &dom, p));
    std::memcpy(amp, p.amp.data(), p.amp.size() * 4);
    return SAM2D_OK;
}

int sam2d_get_tuning(const sam2d_ctx *ctx, double *ms_per_step, int n)
{
    if (!ctx || !ms_per_step) return fail(SAM2D_ERR_PARAM, "NULL argument");
    const int m = std::min(n, num_variants());
    for (int i = 0; i < m; i++)
        ms_per_step[i] = i < (int)ctx->tune_ms.size() ? ctx->tune_ms[i] : 0.0;
    return SAM2D_OK;
}

int sam2d_get_stats(const sam2d_ctx *ctx, sam2d_stats *out)
{
    if (!ctx || !out) return fail(SAM2D_ERR_PARAM, "NULL argument");
    *out = ctx->stats;
    return SAM2D_OK;
}

void sam2d_destroy(sam2d_ctx *ctx)
{
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    ctx->graphs.clear();
    ctx->bufs.clear();
    ctx->domains.clear();
    dev_free(ctx->d_desc);
    sam2d_adjoint_release(ctx);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"
