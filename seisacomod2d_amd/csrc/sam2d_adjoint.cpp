// sam2d_adjoint.cpp -- sam2d_gradient: checkpointed FWI gradient.
//
// Per batch of shots:
//   1. forward over all fd_nt steps in two ring slots, saving the state that
//      enters step m*k into checkpoint m (and the seismograms if asked for);
//   2. backward, segment by segment from the last: restore checkpoint m into
//      state slot 0, recompute steps [m*k, m*k + k) (bit-identical to the
//      first pass: the same arithmetic per cell, three steps per launch where
//      a three-step variant fits), storing for each step n the pressure
//      after it and the dJ/dkappa factors of the state before it (ring_pair,
//      ring_side in sam2d_device.h) into ring slot n - m*k,
//      then run the adjoint steps of the segment in reverse, two per pass
//      where every shot has a residual band, step n reading ring slot
//      n - m*k;
//   3. chain dJ/dkappa, dJ/dbu2, dJ/dbu1 on the device, in FP64, to dJ/dvp
//      and dJ/drho through the buoyancy staggering, kappa = vp^2 rho and the
//      sponge padding (the transpose of cpu_stagger,
//      src/src_cpu/cpu_modelling_kernels.cpp:64-121, and expand,
//      src/src_cpu/cpumodelling.cpp:39-60), shot after shot into the model's
//      gradients, which reach the host once per call.
#include "sam2d_internal.h"

#include <algorithm>
#include <cstring>

namespace sam2d {

struct AdjShotBuffers {
    float *lam = nullptr;     // 2 adjoint slots of (L_px, L_vx, L_vz) cells, 2 of L_pz,
                              // then the accumulator cells (dJ/dkappa, dJ/dbu2, dJ/dbu1)
    float *ckpt = nullptr;    // nckpt state slots (cell-interleaved)
    float *ring = nullptr;    // k ring slots of 12 bytes per cell (ShotDev::ring)
    float *srcg = nullptr;
    int *rcell = nullptr, *rptr = nullptr, *rid = nullptr;
    float *res = nullptr;
    float *rband = nullptr;
    int *bmap = nullptr;
    size_t cap[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // bytes of each buffer above
    ~AdjShotBuffers()
    {
        dev_free(lam); dev_free(ckpt); dev_free(ring); dev_free(srcg);
        dev_free(rcell); dev_free(rptr); dev_free(rid); dev_free(res); dev_free(rband);
        dev_free(bmap);
    }
};

struct AdjointState {
    std::vector<std::unique_ptr<AdjShotBuffers>> bufs;
    AdjDev *d_desc = nullptr;
    int cap_desc = 0;
    AdjDev *d_desc2 = nullptr;     // tile shapes of the two-step adjoint
    int cap_desc2 = 0;
    ShotDev *d_fwd = nullptr;      // forward-sweep descriptors (multi-step tile shapes)
    int cap_fwd = 0;
    ShotDev *d_ring = nullptr;     // recomputation descriptors (three-step tile shapes)
    int cap_ring = 0;
    // Checkpoints of the last sam2d_forward_checkpointed call, held in bufs
    // for sam2d_gradient_checkpointed (a sam2d_gradient call reuses bufs and
    // drops them).
    bool held = false;
    std::vector<ShotPlan> held_plans;
    int held_k = 0;
    // The model (vp then rho) and its FP64 gradients (dJ/dvp then dJ/drho)
    // on the device, nx*nz each.
    float *d_model = nullptr;
    double *d_grad = nullptr;
    size_t cap_model = 0, cap_grad = 0;
    ~AdjointState()
    {
        dev_free(d_desc); dev_free(d_desc2); dev_free(d_fwd); dev_free(d_ring);
        dev_free(d_model); dev_free(d_grad);
    }
};

static int grow(void **p, size_t &cap, size_t need)
{
    if (need <= cap) return SAM2D_OK;
    dev_free(*p);
    *p = nullptr;
    cap = 0;
    SAM2D_TRY(dev_alloc(p, need));
    cap = need;
    return SAM2D_OK;
}

// Receiver-sample index recorded after step n, or -1.
static int rec_sample(const TimeAxis &ta, int n)
{
    if ((n + 1) % ta.dt_factor != 0) return -1;
    const int k = (n + 1) / ta.dt_factor - 1;
    return k < ta.real_nt ? k : -1;
}

// Host chain rule of one shot's padded-grid gradients into the model grid.
// Device bytes one shot of `plan` needs for a gradient with checkpoint
// interval k: two state slots, k ring slots, the checkpoints, the adjoint
// state with its accumulators and the adjoint's coefficient cells.
static size_t shot_bytes(const ShotPlan &p, int fd_nt, int k)
{
    const size_t N = (size_t)p.dom->host.nnx * p.dom->host.nnz;
    const size_t nck = (size_t)(fd_nt + k - 1) / k;
    return N * sizeof(float) * (2 * 4 + 3 * (size_t)k + 4 * nck + 11 + 4);
}

// The three-step variant with the tile shape of v and two columns of operands
// in flight: the recomputation form needs no spills at that depth.
static int ring_form(int v)
{
    int w, c, pf;
    variant_shape_pf(v, &w, &c, &pf);
    for (int u = 0; u < num_variants(); u++) {
        int w2, c2, pf2;
        variant_shape_pf(u, &w2, &c2, &pf2);
        if (variant_steps(u) == 3 && w2 == w && c2 == c && pf2 == 2) return u;
    }
    return v;
}

// Phases of a gradient batch: the forward sweep with checkpoints, and the
// backward pass (recomputation, adjoint, chain rule) from them.
enum { PH_SWEEP = 1, PH_BACKWARD = 2 };

// One batch: the phases asked for.  adj (backward): the adjoint sources per
// shot; seis_out (sweep, optional): the seismograms per shot.  A backward
// pass without the sweep reads the checkpoints an earlier sweep of the same
// plans and k left in the batch's buffers.
static int gradient_batch(sam2d_ctx *ctx, std::vector<ShotPlan> &plans, const float *const *adj,
                          int k, float *const *seis_out, int phases = PH_SWEEP | PH_BACKWARD)
{
    const bool backward = phases & PH_BACKWARD;
    AdjointState &A = *ctx->adj;
    const TimeAxis &ta = ctx->ta;
    const int ns = (int)plans.size();
    // Tile shape of the adjoint and of the single-step recomputation: the
    // "adjoint_variant" option, else the fastest shape measured for the C4
    // gradient (32 shots, k = 100) on MI355X: two steps, 1 wave x 128
    // columns, prefetch 2 (9.78-9.83 s per gradient with prefetch 2 or 4,
    // 9.84-9.88 for 2 waves x 64-128 columns, 10.34 for 2 waves x 256, the
    // default while the forward ring held three floats per cell); one step:
    // 1 wave x 64 columns, prefetch 2.
    const int variant = ctx->adjoint_variant >= 0 ? ctx->adjoint_variant
                        : ctx->adjoint_steps == 2 ? variant_index(1, 128, 2)
                                                  : variant_index(1, 64, 2);
    int ntiles = 0;
    SAM2D_TRY(stage_batch(ctx, plans, variant, &ntiles, 2));
    const bool kza = ta.dtdx == ta.dtdz;
    if ((int)A.bufs.size() < ns) A.bufs.resize(ns);
    std::vector<AdjDev> desc(ns);
    std::vector<int> ncells(ns, 0);
    bool all_band = true;                // every shot's residuals come from a band
    const int nseg = (ta.fd_nt + k - 1) / k;
    for (int s = 0; s < ns; s++) {
        ShotPlan &p = plans[s];
        const DomainHost &h = p.dom->host;
        const size_t N = (size_t)h.nnx * h.nnz;
        // the adjoint and recomputation kernels address cells with 32-bit
        // byte offsets (up to 16 bytes per cell)
        if (N >= ((size_t)1 << 28))
            return fail(SAM2D_ERR_PARAM, "grid of %zu cells too large for the gradient", N);
        if (!A.bufs[s]) A.bufs[s].reset(new AdjShotBuffers());
        AdjShotBuffers &b = *A.bufs[s];
        const size_t ng1 = (size_t)std::max(1, p.ng);
        SAM2D_TRY(grow((void **)&b.lam, b.cap[0], 11 * N * sizeof(float)));
        SAM2D_TRY(grow((void **)&b.ckpt, b.cap[1], (size_t)nseg * 4 * N * sizeof(float)));
        SAM2D_TRY(grow((void **)&b.ring, b.cap[9], (size_t)k * 3 * N * sizeof(float)));
        ctx->staged_desc[s].ring = b.ring;
        // The adjoint's coefficient cells of this domain, built once.
        if (!p.dom->ADJ) {
            SAM2D_TRY(dev_alloc((void **)&p.dom->ADJ, (kza ? 3 : 4) * N * sizeof(float)));
            HIP_TRY(launch_pack_coef(p.dom->KX, kza ? nullptr : p.dom->KZ, p.dom->BX, p.dom->BZ, N,
                                     p.dom->ADJ, ctx->stream));
        }
        SAM2D_TRY(grow((void **)&b.srcg, b.cap[2], (size_t)std::max(1, ta.nts) * sizeof(float)));
        SAM2D_TRY(grow((void **)&b.rcell, b.cap[3], ng1 * sizeof(int)));
        SAM2D_TRY(grow((void **)&b.rptr, b.cap[4], (ng1 + 1) * sizeof(int)));
        SAM2D_TRY(grow((void **)&b.rid, b.cap[5], ng1 * sizeof(int)));
        SAM2D_TRY(grow((void **)&b.res, b.cap[6], ng1 * ta.real_nt * sizeof(float)));
        // Two adjoint slots of (L_px, L_vx, L_vz) cells, two of L_pz, then the
        // accumulator cells.
        float *lam = b.lam, *acc = b.lam + 8 * N;
        if (backward) HIP_TRY(launch_fill(lam, 11 * N, 0.0f, ctx->stream));

        // Source-weight per step and receivers by column (CSR).
        std::vector<float> srcg(std::max(1, ta.nts), 0.0f);
        const double area = (double)(ctx->grid.dx * ctx->grid.dz);
        for (int n = 0; n < ta.nts; n++)
            srcg[n] = (float)((double)ctx->grid.fd_dt * ctx->wavelet[n] / area);
        // Receivers grouped by cell (ascending cell, receiver order kept
        // within a cell): duplicates of one cell are summed in a fixed order.
        std::vector<int> order(p.ng);
        for (int g = 0; g < p.ng; g++) order[g] = g;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return p.gcell[a] < p.gcell[b]; });
        std::vector<int> rcell, rptr(1, 0), rid(std::max(1, p.ng));
        for (int e = 0; e < p.ng; e++) {
            const int g = order[e];
            if (rcell.empty() || rcell.back() != p.gcell[g]) {
                if (!rcell.empty()) rptr.push_back(e);
                rcell.push_back(p.gcell[g]);
            }
            rid[e] = g;
        }
        rptr.push_back(p.ng);
        if (rcell.empty()) rptr.assign(1, 0);
        ncells[s] = (int)rcell.size();
        SAM2D_TRY(upload(b.srcg, srcg.data(), srcg.size() * 4, ctx->stream));
        SAM2D_TRY(upload(b.rcell, rcell.data(), rcell.size() * 4, ctx->stream));
        SAM2D_TRY(upload(b.rptr, rptr.data(), rptr.size() * 4, ctx->stream));
        SAM2D_TRY(upload(b.rid, rid.data(), (size_t)p.ng * 4, ctx->stream));
        if (backward) SAM2D_TRY(upload(b.res, adj[s], (size_t)p.ng * ta.real_nt * 4, ctx->stream));
        // Dense residual band over the receiver rows, [real_nt][band_h][nnx]:
        // each receiver cell holds the sum of its receivers' residuals (in the
        // order above), every other cell -0, so adding it is exact elsewhere.
        int rlo = 1, rhi = 0;
        for (int g = 0; g < p.ng; g++) {
            const int row = p.gcell[g] % h.nnz;
            if (g == 0 || row < rlo) rlo = row;
            if (g == 0 || row > rhi) rhi = row;
        }
        const int band_h = p.ng > 0 ? rhi - rlo + 1 : 0;
        const size_t band_n = (size_t)ta.real_nt * band_h * h.nnx;
        bool use_band = p.ng > 0 && band_n * sizeof(float) <= ((size_t)1 << 30);
        if (!backward) use_band = false;      // no residuals yet
        if (use_band) {
            // Built on the device from the residuals: a map from band cell to
            // receiver cell (or -1), then one pass over the band.
            const int band_cells = band_h * h.nnx;
            std::vector<int> bmap(band_cells, -1);
            for (size_t u = 0; u < rcell.size(); u++)
                bmap[(rcell[u] % h.nnz - rlo) * h.nnx + rcell[u] / h.nnz] = (int)u;
            SAM2D_TRY(grow((void **)&b.bmap, b.cap[8], (size_t)band_cells * sizeof(int)));
            SAM2D_TRY(grow((void **)&b.rband, b.cap[7], band_n * sizeof(float)));
            SAM2D_TRY(upload(b.bmap, bmap.data(), (size_t)band_cells * sizeof(int), ctx->stream));
            HIP_TRY(launch_band(b.bmap, b.rptr, b.rid, b.res, ta.real_nt, band_cells, b.rband,
                                ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));   // the host map goes out of scope
        }
        all_band = all_band && (use_band || p.ng == 0);

        const ShotDev &fd = ctx->staged_desc[s];
        AdjDev &d = desc[s];
        std::memset(&d, 0, sizeof d);
        d.nnz = fd.nnz; d.nnx = fd.nnx; d.ngroups = fd.ngroups; d.nchunks = fd.nchunks;
        d.chunk = fd.chunk; d.tile_begin = fd.tile_begin;
        d.src_x = p.src_x; d.src_z = p.src_z; d.nts = ta.nts; d.real_nt = ta.real_nt;
        d.dtdx = (float)ta.dtdx; d.dtdz = (float)ta.dtdz;
        d.srcg = b.srcg;
        d.coef = p.dom->ADJ;
        d.prof_x = fd.prof_x; d.prof_z = fd.prof_z;
        d.ring = b.ring; d.lam = lam; d.lpz = lam + 6 * N; d.ncell = N;
        d.acc = acc;
        ones_range(h.prof_z, h.nnz, d.zin_lo, d.zin_hi);
        d.xin_lo = fd.xin_lo; d.xin_hi = fd.xin_hi;
        d.ncells = ncells[s];
        d.rcell = b.rcell; d.rptr = b.rptr; d.rid = b.rid; d.res = b.res;
        d.band_h = band_h; d.rrow_lo = rlo; d.rrow_hi = rhi;
        d.rband = use_band ? b.rband : nullptr;
    }
    if (ns > A.cap_desc) {
        dev_free(A.d_desc); A.d_desc = nullptr; A.cap_desc = 0;
        SAM2D_TRY(dev_alloc((void **)&A.d_desc, sizeof(AdjDev) * ns));
        A.cap_desc = ns;
    }
    SAM2D_TRY(upload(A.d_desc, desc.data(), sizeof(AdjDev) * ns, ctx->stream));
    // The staged descriptors now carry the rings (single-step recomputation).
    SAM2D_TRY(upload(ctx->d_desc, ctx->staged_desc.data(), sizeof(ShotDev) * ns, ctx->stream));
    // Two backward steps per pass need every residual in a band (nothing to
    // inject between the steps) and strips of OWN_Z2 rows.
    const bool two_step = ctx->adjoint_steps == 2 && all_band;
    int ntiles2 = 0, groups2 = 0;
    if (two_step) {
        int w2, c2;
        variant_shape(variant, &w2, &c2);
        c2 = march_aligned(c2, 9);
        std::vector<AdjDev> desc2(desc);
        for (auto &dd : desc2) {
            tile_shape(dd.nnz, dd.nnx, w2, c2, OWN_Z2, dd.ngroups, dd.nchunks);
            dd.chunk = c2;
            dd.tile_begin = ntiles2;
            ntiles2 += dd.ngroups * dd.nchunks;
        }
        groups2 = desc2[0].ngroups;
        if (ns > A.cap_desc2) {
            dev_free(A.d_desc2); A.d_desc2 = nullptr; A.cap_desc2 = 0;
            SAM2D_TRY(dev_alloc((void **)&A.d_desc2, sizeof(AdjDev) * ns));
            A.cap_desc2 = ns;
        }
        SAM2D_TRY(upload(A.d_desc2, desc2.data(), sizeof(AdjDev) * ns, ctx->stream));
    }
    int max_ng = 0, max_cells = 0;
    for (auto &p : plans) max_ng = std::max(max_ng, p.ng);
    for (int c : ncells) max_cells = std::max(max_cells, c);

    // Checkpoint m is a copy of the state entering step m*k; a restore copies
    // it back into state slot 0.
    auto save_or_restore = [&](int m, int slot, bool save) -> int {
        for (int s = 0; s < ns; s++) {
            ShotBuffers &fb = *ctx->bufs[s];
            float *ck = A.bufs[s]->ckpt + (size_t)m * 4 * fb.cells;
            HIP_TRY(hipMemcpyAsync(save ? ck : fb.slot(slot), save ? fb.slot(slot) : ck,
                                   4 * fb.cells * sizeof(float), hipMemcpyDeviceToDevice,
                                   ctx->stream));
        }
        return SAM2D_OK;
    };

    // Descriptors of the staged shots with the tile shapes of a multi-step
    // variant and march length, in device buffer *buf (grown as needed).
    auto multistep_desc = [&](int v, int chunk, ShotDev **buf, int *cap, int *tiles,
                              int *groups) -> int {
        std::vector<ShotDev> d2(ctx->staged_desc.begin(), ctx->staged_desc.begin() + ns);
        int t2 = 0;
        for (auto &dd : d2) {
            shot_tiles(dd, v, chunk);
            dd.tile_begin = t2;
            t2 += dd.ngroups * dd.nchunks;
        }
        if (ns > *cap) {
            dev_free(*buf); *buf = nullptr; *cap = 0;
            SAM2D_TRY(dev_alloc((void **)buf, sizeof(ShotDev) * ns));
            *cap = ns;
        }
        SAM2D_TRY(upload(*buf, d2.data(), sizeof(ShotDev) * ns, ctx->stream));
        *tiles = t2;
        *groups = d2[0].ngroups;
        return SAM2D_OK;
    };
    // Forward-sweep kernel: the context's variant when it advances T > 1
    // steps per launch, on descriptors with that variant's tile shapes (each
    // checkpoint segment runs its steps as T-step launches plus single steps
    // for the rest); otherwise the single-step variant of the recomputation.
    int fwd_variant = pick_variant(ctx);
    const ShotDev *fwd_desc = ctx->d_desc;
    int fwd_tiles = ntiles;
    int fwd_groups = ctx->staged_desc[0].ngroups;
    if (variant_steps(fwd_variant) > 1) {
        SAM2D_TRY(multistep_desc(fwd_variant, tile_chunk(ctx, fwd_variant), &A.d_fwd, &A.cap_fwd,
                                 &fwd_tiles, &fwd_groups));
        fwd_desc = A.d_fwd;
    } else {
        fwd_variant = variant;
    }
    // Recomputation kernel: the ring form of a three-step variant.  The
    // "ring_variant" option (march length "ring_chunk"); else the sweep's
    // variant when it is a three-step one (with its tuned march length and
    // tile order); else the three-step variant the autotuner timed fastest
    // (at its own march length, plain order), else 1 wave x 128 columns.
    // Without the option, at prefetch depth 2.  Every variant gives the
    // same bits.
    int ring_variant = fwd_variant;
    const ShotDev *ring_desc = fwd_desc;
    int ring_tiles = fwd_tiles, ring_groups = fwd_groups, ring_order = ctx->tile_order;
    int ring_chunk = tile_chunk(ctx, fwd_variant);
    bool own_desc = ctx->ring_variant >= 0 || !ring_supported(fwd_variant);
    if (own_desc) {
        ring_variant = ctx->ring_variant;             // the option, else
        for (int v = 0; ctx->ring_variant < 0 && v < (int)ctx->tune_ms.size(); v++)
            if (ring_supported(v) && ctx->tune_ms[v] > 0.0 &&
                (ring_variant < 0 || ctx->tune_ms[v] < ctx->tune_ms[ring_variant]))
                ring_variant = v;
        if (ring_variant < 0) ring_variant = ring_form(variant_index(1, 128, 4));
        int w, c;
        variant_shape(ring_variant, &w, &c);
        ring_chunk = ctx->ring_chunk > 0 ? ctx->ring_chunk : march_aligned(c, 15);
        ring_order = 0;
    }
    if (ctx->ring_variant < 0 && ring_form(ring_variant) != ring_variant) {
        ring_variant = ring_form(ring_variant);
        own_desc = true;
    }
    if (own_desc) {
        SAM2D_TRY(multistep_desc(ring_variant, ring_chunk, &A.d_ring, &A.cap_ring, &ring_tiles,
                                 &ring_groups));
        ring_desc = A.d_ring;
    }

    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    // 1. forward with checkpoints: checkpoint m holds the state entering
    //    step m*k
    //    (cell-interleaved state, as the forward-modelling calls)
    const bool want_seis = seis_out != nullptr;
    int cur = 0;
    for (int m = 0; m < nseg && (phases & PH_SWEEP); m++) {
        SAM2D_TRY(save_or_restore(m, cur, true));
        const int n0 = m * k, n1 = std::min(ta.fd_nt, n0 + k);
        SAM2D_TRY(enqueue_steps_on(ctx, fwd_desc, fwd_groups, ns, fwd_tiles, fwd_variant, n0, n1,
                                   want_seis ? max_ng : 0, cur, &cur));
    }
    // 2. backward by segments
    int lam = 0;
    for (int m = nseg - 1; m >= 0 && backward; m--) {
        const int n0 = m * k, n1 = std::min(ta.fd_nt, n0 + k);
        SAM2D_TRY(save_or_restore(m, 0, false));
        // Recompute the segment, step n's ring entries into ring slot n - n0:
        // three steps per launch, single steps for the rest.
        int n = n0, st = 0;
        if (ring_supported(ring_variant))
            for (; n + 2 < n1; n += 3, st ^= 1)
                HIP_TRY(launch_stepn_ring(ring_variant, ring_desc, ns, ring_tiles, ctx->grid_cap,
                                          kza, n, st, n - n0, TileOrder{ring_groups, ring_order},
                                          ctx->stream));
        for (; n < n1; n++, st ^= 1)
            HIP_TRY(launch_step_ring(variant, ctx->d_desc, ns, ntiles, ctx->grid_cap, n, st, st ^ 1,
                                     n - n0, ctx->stream));
        int n_top = n1 - 1;
        if (two_step)
            for (; n_top - 1 >= n0; n_top -= 2) {
                HIP_TRY(launch_adjoint_step2(variant, A.d_desc2, ns, ntiles2, kza, n_top, lam,
                                             lam ^ 1, n_top - n0, rec_sample(ta, n_top),
                                             rec_sample(ta, n_top - 1),
                                             TileOrder{groups2, ctx->adjoint_tile_order},
                                             ctx->stream));
                lam ^= 1;
            }
        for (int n = n_top; n >= n0; n--) {
            if (!all_band)
                HIP_TRY(launch_inject(A.d_desc, ns, max_cells, lam, rec_sample(ta, n), ctx->stream));
            HIP_TRY(launch_adjoint_step(variant, A.d_desc, ns, ntiles, kza, n, lam, lam ^ 1, n - n0,
                                        rec_sample(ta, n),
                                        TileOrder{desc[0].ngroups, ctx->adjoint_tile_order},
                                        ctx->stream));
            lam ^= 1;
        }
    }
    // 3. the chain rule, shot after shot
    const size_t nm = (size_t)ctx->grid.nx * ctx->grid.nz;
    for (int s = 0; s < ns && backward; s++) {
        const DomainHost &h = plans[s].dom->host;
        const size_t N = (size_t)h.nnx * h.nnz;
        HIP_TRY(launch_chain(A.bufs[s]->lam + 8 * N, h.nnz, h.nnx, ctx->grid.npml, h.wx0, h.w,
                             ctx->grid.nz, A.d_model, A.d_model + nm, A.d_grad, A.d_grad + nm,
                             ctx->stream));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.kernel_ms += ms;
    if (backward) ctx->stats.launches += (long long)ta.fd_nt * 2;
    if (phases & PH_SWEEP)
        ctx->stats.launches += (ta.fd_nt + variant_steps(fwd_variant) - 1) / variant_steps(fwd_variant);
    ctx->stats.variant = variant;
    ctx->stats.steps_per_launch = 1;

    for (int s = 0; s < ns; s++) {
        const DomainHost &h = plans[s].dom->host;
        ctx->stats.cell_updates += (double)h.nnx * h.nnz * ta.fd_nt;
        if (want_seis && plans[s].ng)
            HIP_TRY(hipMemcpy(seis_out[s], plans[s].seis_dev,
                              (size_t)plans[s].ng * ta.real_nt * sizeof(float), hipMemcpyDeviceToHost));
    }
    return SAM2D_OK;
}

// The model on the device and its gradients cleared, before a call's batches.
static int begin_gradients(sam2d_ctx *ctx)
{
    AdjointState &A = *ctx->adj;
    const size_t nm = (size_t)ctx->grid.nx * ctx->grid.nz;
    SAM2D_TRY(grow((void **)&A.d_model, A.cap_model, 2 * nm * sizeof(float)));
    SAM2D_TRY(grow((void **)&A.d_grad, A.cap_grad, 2 * nm * sizeof(double)));
    SAM2D_TRY(upload(A.d_model, ctx->vp.data(), nm * sizeof(float), ctx->stream));
    SAM2D_TRY(upload(A.d_model + nm, ctx->rho.data(), nm * sizeof(float), ctx->stream));
    HIP_TRY(launch_fill((float *)A.d_grad, 4 * nm, 0.0f, ctx->stream));   // +0.0 in FP64
    return SAM2D_OK;
}

// The gradients to the caller's FP32 arrays.
static int end_gradients(sam2d_ctx *ctx, float *g_vp, float *g_rho)
{
    const size_t nm = (size_t)ctx->grid.nx * ctx->grid.nz;
    std::vector<double> g(2 * nm);
    HIP_TRY(hipMemcpyAsync(g.data(), ctx->adj->d_grad, 2 * nm * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < nm; i++) { g_vp[i] = (float)g[i]; g_rho[i] = (float)g[nm + i]; }
    return SAM2D_OK;
}

}  // namespace sam2d

using namespace sam2d;

void sam2d_adjoint_release(sam2d_ctx *ctx)
{
    delete ctx->adj;
    ctx->adj = nullptr;
}

// Device bytes a gradient batch may use: the "hbm_budget_gb" option, else
// the device memory that is free or held by this context's reusable buffers,
// less a reserve of 5 % of the device (at least 4 GiB), and at most 90 % of
// the device.
static int gradient_budget(sam2d_ctx *ctx, size_t *out)
{
    if (ctx->hbm_budget_gb > 0) {
        *out = (size_t)ctx->hbm_budget_gb << 30;
        return SAM2D_OK;
    }
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    size_t held = 0;
    for (auto &b : ctx->bufs)
        if (b) held += b->cap_state * sizeof(float);
    for (auto &b : ctx->adj->bufs)
        if (b)
            for (size_t c : b->cap) held += c;
    const size_t reserve = std::max<size_t>((size_t)4 << 30, total_b / 20);
    const size_t avail = free_b + held > reserve ? free_b + held - reserve : 0;
    *out = std::min(avail, total_b / 10 * 9);
    return SAM2D_OK;
}

// Per-shot pointers into a concatenated [sum nrec][real_nt] array.
template <class T>
static std::vector<T *> per_shot(T *base, const sam2d_shot *shots, int nshots, int real_nt)
{
    std::vector<T *> out(nshots, nullptr);
    size_t off = 0;
    for (int s = 0; s < nshots; s++) {
        out[s] = base ? base + off : nullptr;
        off += (size_t)shots[s].nrec * real_nt;
    }
    return out;
}

extern "C" int sam2d_gradient(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                              const int *windows, const float *adj, int ckpt_k, float *seis,
                              float *g_vp, float *g_rho)
{
    clear_error();
    if (!ctx) return fail(SAM2D_ERR_PARAM, "ctx is NULL");
    if (!g_vp || !g_rho) return fail(SAM2D_ERR_PARAM, "gradient outputs are NULL");
    const TimeAxis &ta = ctx->ta;
    size_t total = 0;
    for (int s = 0; s < nshots; s++) total += (size_t)shots[s].nrec * ta.real_nt;
    if (total > 0 && !adj) return fail(SAM2D_ERR_PARAM, "adj is NULL");
    const double t0 = now_ms();
    ctx->stats = sam2d_stats();
    if (!ctx->adj) ctx->adj = new AdjointState();
    int k = ckpt_k < 1 ? 1 : (ckpt_k > ta.fd_nt ? ta.fd_nt : ckpt_k);

    std::vector<ShotPlan> plans;
    SAM2D_TRY(make_plans(ctx, shots, nshots, windows, plans));
    ctx->quirk_valid = false;
    ctx->adj->held = false;                  // bufs are reused: held checkpoints go
    const std::vector<const float *> adj_ptr = per_shot(adj, shots, nshots, ta.real_nt);
    const std::vector<float *> seis_ptr = per_shot(seis, shots, nshots, ta.real_nt);
    // Shots per batch: the option, else the fewest batches of equal shot
    // counts whose ring, checkpoints and adjoint state fit the budget of
    // gradient_budget.  The 288 GB of an MI355X hold a 32-shot C4 batch at
    // k = 100 (189 GB) in one piece.
    int batch = ctx->max_batch > 0 ? ctx->max_batch : nshots;
    if (ctx->max_batch <= 0 && nshots > 0) {
        size_t budget = 0;
        SAM2D_TRY(gradient_budget(ctx, &budget));
        auto fits = [&](int b) {
            for (int s0 = 0; s0 < nshots; s0 += b) {
                size_t used = 0;
                for (int s = s0; s < std::min(nshots, s0 + b); s++)
                    used += shot_bytes(plans[s], ta.fd_nt, k);
                if (used > budget && b > 1) return false;
            }
            return true;
        };
        int nb = 1;
        batch = nshots;
        while (batch > 1 && !fits(batch)) {
            nb++;
            batch = (nshots + nb - 1) / nb;
        }
    }
    SAM2D_TRY(begin_gradients(ctx));
    for (int s0 = 0; s0 < nshots; s0 += batch) {
        const int s1 = std::min(nshots, s0 + batch);
        std::vector<ShotPlan> sub(plans.begin() + s0, plans.begin() + s1);
        SAM2D_TRY(gradient_batch(ctx, sub, adj_ptr.data() + s0, k,
                                 seis ? seis_ptr.data() + s0 : nullptr));
    }
    SAM2D_TRY(end_gradients(ctx, g_vp, g_rho));
    ctx->stats.wall_ms = now_ms() - t0;
    return SAM2D_OK;
}

extern "C" int sam2d_forward_checkpointed(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                                          const int *windows, int ckpt_k, float *seis)
{
    clear_error();
    if (!ctx) return fail(SAM2D_ERR_PARAM, "ctx is NULL");
    const TimeAxis &ta = ctx->ta;
    size_t total = 0;
    for (int s = 0; s < nshots; s++) total += (size_t)shots[s].nrec * ta.real_nt;
    if (total > 0 && !seis) return fail(SAM2D_ERR_PARAM, "seis is NULL");
    const double t0 = now_ms();
    ctx->stats = sam2d_stats();
    if (!ctx->adj) ctx->adj = new AdjointState();
    AdjointState &A = *ctx->adj;
    A.held = false;
    const int k = ckpt_k < 1 ? 1 : (ckpt_k > ta.fd_nt ? ta.fd_nt : ckpt_k);
    std::vector<ShotPlan> plans;
    SAM2D_TRY(make_plans(ctx, shots, nshots, windows, plans));
    ctx->quirk_valid = false;
    // Every shot stays resident until the backward pass: one batch.
    size_t budget = 0, used = 0;
    SAM2D_TRY(gradient_budget(ctx, &budget));
    for (auto &p : plans) used += shot_bytes(p, ta.fd_nt, k);
    if (used > budget)
        return fail(SAM2D_ERR_MEMORY, "checkpoints, ring and adjoint state of %d shots need %zu "
                    "MB, over the %zu MB budget (sam2d_gradient runs them in batches)", nshots,
                    used >> 20, budget >> 20);
    const std::vector<float *> seis_ptr = per_shot(seis, shots, nshots, ta.real_nt);
    SAM2D_TRY(gradient_batch(ctx, plans, nullptr, k, seis_ptr.data(), PH_SWEEP));
    A.held_plans = plans;
    A.held_k = k;
    A.held = true;
    ctx->stats.wall_ms = now_ms() - t0;
    return SAM2D_OK;
}

extern "C" int sam2d_gradient_checkpointed(sam2d_ctx *ctx, const float *adj, float *g_vp,
                                           float *g_rho)
{
    clear_error();
    if (!ctx) return fail(SAM2D_ERR_PARAM, "ctx is NULL");
    if (!g_vp || !g_rho) return fail(SAM2D_ERR_PARAM, "gradient outputs are NULL");
    if (!ctx->adj || !ctx->adj->held)
        return fail(SAM2D_ERR_PARAM, "no checkpoints held: call sam2d_forward_checkpointed first "
                    "(a sam2d_gradient call in between drops them)");
    AdjointState &A = *ctx->adj;
    const TimeAxis &ta = ctx->ta;
    size_t total = 0;
    for (auto &p : A.held_plans) total += (size_t)p.ng * ta.real_nt;
    if (total > 0 && !adj) return fail(SAM2D_ERR_PARAM, "adj is NULL");
    const double t0 = now_ms();
    ctx->stats = sam2d_stats();
    ctx->quirk_valid = false;
    std::vector<const float *> adj_ptr(A.held_plans.size());
    size_t off = 0;
    for (size_t s = 0; s < A.held_plans.size(); s++) {
        adj_ptr[s] = adj + off;
        off += (size_t)A.held_plans[s].ng * ta.real_nt;
    }
    std::vector<ShotPlan> plans = A.held_plans;
    SAM2D_TRY(begin_gradients(ctx));
    SAM2D_TRY(gradient_batch(ctx, plans, adj_ptr.data(), A.held_k, nullptr, PH_BACKWARD));
    SAM2D_TRY(end_gradients(ctx, g_vp, g_rho));
    ctx->stats.wall_ms = now_ms() - t0;
    return SAM2D_OK;
}
