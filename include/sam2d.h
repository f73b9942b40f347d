/*
 * sam2d.h -- C ABI of the MI355X (gfx950) forward-modelling library.
 *
 * 2-D acoustic finite-difference modelling on a staggered grid: 4th order in
 * space, 2nd order in time, split-field exponential sponge, Ricker source,
 * pressure receivers.  FP32 state on the GPU, FP64-derived coefficients.
 *
 * The reference has no library boundary: its backends are swapped at link time
 * behind `void modelling_module(int my_nsrc, geo2d_t *mygeo2d_sp)`
 * (src/src_cpu/modelling.h:106, defined at src/src_cpu/cpumodelling.cpp:158
 * and src/src_gpu/gpumodelling.cu:238).  These entry points are what that
 * function needs below it; sam2d_modelling.h restores the reference-shaped
 * drop-in above them.
 *
 * Conventions
 *   - Plain C types and pointers only.  Host pointers unless the name says
 *     _dev.  The caller owns every buffer it passes; the library owns device
 *     memory, which lives in the context until sam2d_destroy.
 *   - Models are float32, x-major: value (ix, iz) at [ix*nz + iz], the layout
 *     of the reference's model files (src/src_cpu/model.cpp:37-40).
 *   - Seismograms are float32, trace-major: for each shot, ng traces of
 *     real_nt samples, shots concatenated in order -- the reference's output
 *     file layout (src/src_cpu/cpumodelling.cpp:367-374).
 *   - Every call returns 0 (SAM2D_OK) or a negative status that mirrors the
 *     reference's abort codes; sam2d_last_error() has the message.  The library
 *     never aborts the process.
 *   - One context per host thread at a time; calls on one context serialise.
 */
#ifndef SAM2D_H
#define SAM2D_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (reference abort codes: src/src_cpu/model.cpp:21-34,
 * check_stability.cpp:42,63, json_parser.cpp:88, create_geometry.cpp:381-432,
 * gpumodelling.cu:140). */
#define SAM2D_OK              0
#define SAM2D_ERR_FILE       (-2)   /* file missing or model size mismatch   */
#define SAM2D_ERR_RECEIVER   (-3)   /* receiver / source outside the model   */
#define SAM2D_ERR_STABILITY  (-5)   /* CFL or dispersion limit violated      */
#define SAM2D_ERR_PARAM      (-6)   /* missing or invalid parameter          */
#define SAM2D_ERR_DEVICE     (-20)  /* HIP runtime error                     */
#define SAM2D_ERR_MEMORY     (-30)  /* device or host allocation failed      */

/* Grid, time axis and source parameters: the job-card keys NX, NZ, DX, DZ,
 * Time, FD DT, DATA DT, Frequency and Order of FD Operator
 * (src/src_cpu/json_parser.cpp:141-265), plus the sponge width that the
 * reference fixes at compile time (npml = 32, cpumodelling.cpp:15). */
typedef struct sam2d_grid {
    int   nx, nz;        /* model grid points                               */
    float dx, dz;        /* grid spacing, m                                 */
    float rec_len;       /* record length, s                                */
    float fd_dt;         /* modelling time step, s                          */
    float real_dt;       /* seismogram sample interval, s (multiple of fd_dt) */
    float dom_freq;      /* Ricker dominant frequency, Hz                   */
    int   npml;          /* sponge width in cells (reference: 32)           */
    int   fdop;          /* FD order; only 4 is implemented (as reference)  */
} sam2d_grid;

/* One shot: source and receiver coordinates in metres, the crd2d_t pairs of
 * geo2d_t (src/src_cpu/modelling.h:35-45). */
typedef struct sam2d_shot {
    float        sx, sz;
    int          nrec;
    const float *rec_xz;  /* nrec (x, z) pairs */
} sam2d_shot;

/* Derived sizes, as modelling_module computes them
 * (src/src_cpu/cpumodelling.cpp:168-195). */
typedef struct sam2d_dims {
    int fd_nt;      /* modelling steps                  */
    int real_nt;    /* samples per trace                */
    int dt_factor;  /* steps per sample                 */
    int nts;        /* source wavelet length in steps   */
    int nnx, nnz;   /* padded grid (sponge included)    */
} sam2d_dims;

/* Timing of the last forward / batched / gradient call, measured with HIP
 * events on the library's stream. */
typedef struct sam2d_stats {
    double kernel_ms;        /* summed duration of the step kernels          */
    double wall_ms;          /* whole call, host wall clock                  */
    double cell_updates;     /* padded cells x steps, summed over shots      */
    long long launches;      /* step-kernel launches                         */
    int    used_graph;       /* 1 when the time loop ran as a hipGraph       */
    int    variant;          /* kernel variant index chosen (autotuner)      */
    int    steps_per_launch; /* time steps one step-kernel launch advances   */
} sam2d_stats;

typedef struct sam2d_ctx sam2d_ctx;

/* Replaces modelling_module's setup (cpumodelling.cpp:168-281): validates the
 * grid (sizes and spacings, fdop == 4, fd_dt divides real_dt, a padded grid
 * below 2^31 cells), uploads the model and builds the FP32 coefficient fields
 * on the device.  Model values are taken as they are, as the reference takes
 * them (a density <= 0 gives the buoyancy 1e10, cpu_modelling_kernels.cpp:64-121).
 * Does not run the CFL / dispersion check; see sam2d_check_stability.
 * Size limit of the multi-step kernels: their tiles address a shot's state
 * with 32-bit byte offsets, so a padded grid (or window) of 2^28 cells or more
 * runs on the single-step kernels only; sam2d_autotune skips the multi-step
 * variants there, and choosing one with the "variant" option makes the
 * modelling calls return SAM2D_ERR_PARAM. */
int sam2d_create(const sam2d_grid *grid, const float *vp, const float *rho,
                 sam2d_ctx **out);

/* check_stability() (src/src_cpu/check_stability.cpp:8-72):
 * SAM2D_ERR_STABILITY if dt > 0.606*hmin/vmax or 2*f > vmin/(5*hmax). */
int sam2d_check_stability(const sam2d_grid *grid, float vmin, float vmax);

int sam2d_get_dims(const sam2d_ctx *ctx, sam2d_dims *out);

/* The per-shot loop of modelling_module (cpumodelling.cpp:326-380) over the
 * full padded domain: for each shot, fresh state, fd_nt steps, pressure
 * sampled every dt_factor steps.  seis: sum_s(nrec_s) * real_nt floats. */
int sam2d_forward(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                  float *seis);

/* Same as sam2d_forward but the seismograms stay on the device: seis_dev is a
 * device pointer of sum_s(nrec_s) * real_nt floats.  For callers that keep
 * data in HBM (benchmarks, tensor frameworks). */
int sam2d_forward_dev(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                      float *seis_dev);

/* Batched multi-shot forward: all shots advance together, one fused launch
 * per time step over every shot's tiles, the time loop replayed as a hipGraph.
 * No reference counterpart (the reference loops over shots,
 * cpumodelling.cpp:326); each shot's results equal what the per-shot path
 * gives on the same domain.  windows: NULL for the full model, or nshots
 * pairs (first model column, width) that crop each shot's model to an
 * aperture window (the sponge is added around the window, as around the
 * full model).  Receivers and source must lie inside their window. */
int sam2d_forward_batched(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                          const int *windows, float *seis);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sam2d_forward_dev for one shot on the full model, with
 * every working buffer (the shot's descriptor, source increments, receiver
 * cells, receiver band and both state slots) in one static device array of
 * SAM2D_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's first
 * use until the process ends.  seis_dev (caller-owned device memory) receives
 * nrec * real_nt floats, the same bits as sam2d_forward_dev gives.  The shot
 * needs 32 bytes per padded cell plus its band and receiver lists, so the
 * padded grid is limited to about 6 M cells (the 3201 x 1201 model with a
 * 60-cell sponge fits; 5000 x 5000 does not): larger shots return
 * SAM2D_ERR_PARAM.  Calls no hipMalloc or hipFree and captures no graph.
 * Every call on a GPU shares that GPU's array, which keeps the last call's
 * contents: calls that overlap on one GPU (from several threads or
 * contexts) race. */
#define SAM2D_NOALLOC_SCRATCH_BYTES (192u << 20)
int sam2d_forward_noalloc(sam2d_ctx *ctx, const sam2d_shot *shot, float *seis_dev);

/* The per-shot loop over cropped windows (the plain baseline that
 * sam2d_forward_batched is compared with): same arguments and results. */
int sam2d_forward_windows(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                          const int *windows, float *seis);

/* FWI gradient (no reference counterpart: job_t.adj is declared but unused,
 * src/src_cpu/modelling.h:51).  Given adj = dL/dseis for every sample of the
 * shots' seismograms (layout of sam2d_forward's output), returns dL/dvp and
 * dL/drho on the model grid (x-major nx*nz, summed over shots).  The forward
 * wavefield is checkpointed every ckpt_k steps and recomputed segment by
 * segment; the adjoint is the exact transpose of the discrete FP32 scheme.
 * windows: NULL or as in sam2d_forward_batched.  seis (optional, may be NULL)
 * receives the forward seismograms.  For an L2 misfit 0.5*|d - d_obs|^2 pass
 * adj = d - d_obs. */
int sam2d_gradient(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                   const int *windows, const float *adj, int ckpt_k,
                   float *seis, float *g_vp, float *g_rho);

/* sam2d_gradient in two halves, for FWI loops that model the data before
 * they know the residual (no reference counterpart).  The first is forward
 * modelling: the seismograms of the shots (as sam2d_forward_windows gives
 * them, the same bits; windows may be NULL) from the gradient's forward
 * sweep, which also keeps the state entering every ckpt_k-th step on the
 * device.  Every shot stays resident, so the call returns SAM2D_ERR_MEMORY
 * when the shots' checkpoints, ring and adjoint state exceed the gradient
 * budget ("hbm_budget_gb"); sam2d_gradient runs such sets in batches. */
int sam2d_forward_checkpointed(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots,
                               const int *windows, int ckpt_k, float *seis);
/* The second half: the backward pass of sam2d_gradient from the checkpoints
 * of this context's last sam2d_forward_checkpointed call, for adjoint sources
 * adj of those shots (layout of its seis); g_vp, g_rho as sam2d_gradient, the
 * same bits.  Other modelling calls in between leave the checkpoints in
 * place, and the pass may be repeated with other sources; a sam2d_gradient
 * call drops them (SAM2D_ERR_PARAM then). */
int sam2d_gradient_checkpointed(sam2d_ctx *ctx, const float *adj, float *g_vp,
                                float *g_rho);

/* Tuning and options.  Keys: "graph" (0/1, default 1), "variant" (-1 = auto,
 * else a kernel variant index), "reset_quirk" (1 = each shot of the per-shot
 * host-output loop, sam2d_forward, starts from the final state of the shot
 * modelled before it in that mode on this context, in the same call or an
 * earlier one, with only the first nnz cells cleared, as the reference's
 * modelling_module does between its shots, cpumodelling.cpp:343-348; the first
 * shot after the option is set, and the first after any other modelling call,
 * starts fresh; default 0 = fresh state for every shot),
 * "max_batch" (shots resident at once in batched and gradient calls, 0 = all
 * or, for gradients, the fewest batches of equal shot counts that fit
 * "hbm_budget_gb" GiB of device memory for ring, checkpoints and adjoint
 * state (default 0 = the device memory that is free or held by the context's
 * reusable buffers, less a reserve of 5 % of the device, at most 90 %)), "adjoint_steps" (backward
 * steps per adjoint kernel pass in sam2d_gradient, 1 or 2, default 2; both
 * give the same bits), "adjoint_variant" (tile shape of the adjoint kernels,
 * the single-step form of the given variant index; -1 = automatic, the
 * default), "tile_order" (0 or 1 = chunk columns of each XCD's tile range in
 * order, s = 2..64 = with stride s, for the multi-step kernels; set by
 * sam2d_autotune, default 0; results are the same bits),
 * "adjoint_tile_order" (the same for the adjoint kernels, default 7),
 * "chunk" (columns each multi-step tile marches, 0 = the variant's own
 * rounded up so that the march, chunk + 6T - 3 columns for T steps per
 * launch, is whole groups of four columns; set by sam2d_autotune; results are
 * the same bits). */
int sam2d_set_option(sam2d_ctx *ctx, const char *key, int value);
/* Current value of an option above (after sam2d_autotune: "tile_order"). */
int sam2d_get_option(sam2d_ctx *ctx, const char *key, int *value);

/* Project-written autotuner: times every kernel variant (z-strips per
 * workgroup x march length along x) for a few steps on this context's grid
 * (or on the given windows) and selects the fastest; for a multi-step winner
 * it then times march lengths around the variant's jointly with the
 * chunk-column strides 1..9 and sets "chunk" and "tile_order".
 * best_variant may be NULL. */
int sam2d_autotune(sam2d_ctx *ctx, const int *windows, int nshots, int steps,
                   int *best_variant);

/* Milliseconds per step the last sam2d_autotune measured for each variant
 * (0 for variants not timed); fills min(n, sam2d_num_variants()) values. */
int sam2d_get_tuning(const sam2d_ctx *ctx, double *ms_per_step, int n);

/* Number of kernel variants and a printable description of one. */
int sam2d_num_variants(void);
const char *sam2d_variant_name(int variant);

int sam2d_get_stats(const sam2d_ctx *ctx, sam2d_stats *out);

/* Host-only (no device needed): the FP32 fields the kernels read for model
 * columns [wx0, wx0 + w), as sam2d_create builds them from modelling_module's
 * setup (cpumodelling.cpp:207-281).  Padded sizes nnx = w + 2*npml and
 * nnz = nz + 2*npml; fields are nnx*nnz (column-major, z fastest):
 *   KX = dt/dx*kappa, KZ = dt/dz*kappa, BX = dt/dx*bu2, BZ = dt/dz*bu1
 * and profiles prof_x[4*nnx] = damp2a, damp2a1, damp2b, damp2b1,
 * prof_z[4*nnz] = damp1a, damp1a1, damp1b, damp1b1.  Any output may be NULL. */
int sam2d_coefficients(const sam2d_grid *grid, const float *vp, const float *rho, int wx0, int w,
                       float *KX, float *KZ, float *BX, float *BZ, float *prof_x, float *prof_z);

/* Host-only: source increment per step for a source at (sx, sz) metres,
 * fd_dt*w[it]*kappa[src]/(dx*dz) (cpumodelling.cpp:136-140), nts floats. */
int sam2d_source_increments(const sam2d_grid *grid, const float *vp, const float *rho, float sx,
                            float sz, float *amp);

void sam2d_destroy(sam2d_ctx *ctx);

/* Message of the last failing call on this thread ("" if none). */
const char *sam2d_last_error(void);

/* Library version string. */
const char *sam2d_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SAM2D_H */
