// sam2d_internal.h -- context and helpers shared by the library's host files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/sam2d.h"
#include "sam2d_device.h"
#include "sam2d_setup.h"

namespace sam2d {

int fail(int code, const char *fmt, ...);
void clear_error();
const char *last_error();
double now_ms();

#define SAM2D_TRY(expr)                  \
    do {                                 \
        int _rc = (expr);                \
        if (_rc != SAM2D_OK) return _rc; \
    } while (0)

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                      \
            return ::sam2d::fail(SAM2D_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr,        \
                                 hipGetErrorString(_e), __FILE__, __LINE__);              \
        }                                                                                 \
    } while (0)

int dev_alloc(void **p, size_t bytes);
void dev_free(void *p);
int upload(void *dst, const void *src, size_t bytes, hipStream_t s);

// Coefficients of one padded domain on the device.
struct DeviceDomain {
    DomainHost host;
    float *KX = nullptr, *KZ = nullptr, *BX = nullptr, *BZ = nullptr, *B = nullptr;
    float *prof_x = nullptr, *prof_z = nullptr;
    float *ADJ = nullptr;          // the adjoint's coefficient cells (launch_pack_coef), built
                                   // by the first gradient on this domain
    ~DeviceDomain();
};

// Device state of one resident shot slot (grown on demand, reused).
struct ShotBuffers {
    float *state = nullptr;        // nslots x 4 x cells floats (slot_cells, slot_pz)
    float *amp = nullptr;
    int *gcell = nullptr;
    float *seis = nullptr;
    float *band = nullptr;         // [2][band_h][cells / nnz] receiver-row pressure
    size_t cells = 0, cap_state = 0, cap_seis = 0, cap_band = 0;
    int cap_nts = 0, cap_ng = 0;
    int reserve(size_t cells, int nslots, int nts, int ng, size_t nseis);
    int reserve_band(size_t n);
    float *slot(int s) const { return state + (size_t)s * 4 * cells; }
    void release();
    ~ShotBuffers();
};

// Host-side plan of one shot.
struct ShotPlan {
    DeviceDomain *dom = nullptr;
    int src_x = 0, src_z = 0;      // padded indices in its domain
    int ng = 0;
    std::vector<int> gcell;
    std::vector<float> amp;
    float *seis_dev = nullptr;
};

struct GraphKey {
    int nshots, ntiles, variant, max_ng, grid_cap, tile_order, chunk;
    bool operator<(const GraphKey &o) const
    {
        return std::tie(nshots, ntiles, variant, max_ng, grid_cap, tile_order, chunk) <
               std::tie(o.nshots, o.ntiles, o.variant, o.max_ng, o.grid_cap, o.tile_order,
                        o.chunk);
    }
};

struct AdjointState;

int default_variant_index();

}  // namespace sam2d

struct sam2d_ctx {
    sam2d_grid grid;
    sam2d::TimeAxis ta;
    std::vector<float> vp, rho;            // host model, x-major
    std::vector<double> wavelet;           // FP64 Ricker, nts samples
    int device = 0, num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<long long, std::unique_ptr<sam2d::DeviceDomain>> domains;
    std::vector<std::unique_ptr<sam2d::ShotBuffers>> bufs;
    sam2d::ShotDev *d_desc = nullptr;
    int cap_desc = 0;
    std::vector<sam2d::ShotDev> staged_desc;
    std::map<sam2d::GraphKey, hipGraphExec_t> graphs;
    bool use_graph = true;
    bool reset_quirk = false;
    // reset_quirk carry-over: the final state of the last shot modelled in
    // that mode (slot quirk_slot of bufs[0], a domain of quirk_cells cells)
    bool quirk_valid = false;
    int quirk_slot = 0;
    size_t quirk_cells = 0;
    int variant = -1, tuned_variant = -1, default_variant = 0;
    int max_batch = 0;
    int grid_cap = 0;
    int hbm_budget_gb = 0;                 // device memory a gradient batch may use (0: automatic)
    int adjoint_steps = 2;                 // backward steps per adjoint launch (1 or 2)
    int adjoint_variant = -1;              // single-step variant for adjoint tiles, -1 = auto
    int ring_variant = -1;                 // three-step variant of the recomputation, -1 = auto
    int ring_chunk = 0;                    // its march length, 0 = the variant's
    int tile_order = 0;                    // chunk-column stride of multi-step kernels (tuned)
    int adjoint_tile_order = 7;            // chunk-column stride of the adjoint kernels
    int chunk = 0;                         // march length of multi-step kernels, 0 = variant's
    std::vector<double> tune_ms;           // per variant, from the autotuner
    sam2d_stats stats = sam2d_stats();
    sam2d::AdjointState *adj = nullptr;
};

namespace sam2d {
int get_domain(sam2d_ctx *ctx, int wx0, int w, DeviceDomain **out);
int plan_shot(sam2d_ctx *ctx, const sam2d_shot &sh, DeviceDomain *dom, ShotPlan &p);
int make_plans(sam2d_ctx *ctx, const sam2d_shot *shots, int nshots, const int *windows,
               std::vector<ShotPlan> &plans);
int pick_variant(sam2d_ctx *ctx);
// March length (columns per tile) of `variant`: the "chunk" option for a
// multi-step variant when set, else the variant's own.
int tile_chunk(const sam2d_ctx *ctx, int variant);
int stage_batch(sam2d_ctx *ctx, std::vector<ShotPlan> &plans, int variant, int *ntiles_out,
                int nslots = 2);
int enqueue_steps(sam2d_ctx *ctx, int nshots, int ntiles, int variant, int n0, int n1, int max_ng,
                  int start_slot, int *end_slot);
// The same on an explicit device descriptor array (tile shapes of `variant`).
// desc: device descriptors of the staged shots with variant's tile shapes,
// ngroups z-groups per chunk column (the same for every shot).
int enqueue_steps_on(sam2d_ctx *ctx, const ShotDev *desc, int ngroups, int nshots, int ntiles,
                     int variant, int n0, int n1, int max_ng, int start_slot, int *end_slot);
void tile_shape(int nnz, int nnx, int waves, int chunk, int own_rows, int &ngroups, int &nchunks);
// Indices [lo, hi] around the middle of an n-point axis where all four
// sponge profiles of prof[4][n] are exactly 1 (lo > hi: none).
void ones_range(const std::vector<float> &prof, int n, int &lo, int &hi);
void shot_tiles(ShotDev &d, int variant, int chunk);
// Descriptor of one shot for `variant` with march length `chunk`, tiles from
// tile_begin; the device buffers are left for the caller to set.
void describe_shot(const ShotPlan &p, int variant, int chunk, int tile_begin, int nts, int real_nt,
                   ShotDev &d);
// SAM2D_ERR_PARAM for a multi-step variant on a grid beyond its 32-bit offsets.
int check_multistep_size(const sam2d_ctx *ctx, int variant, size_t ncell);
// Device address of the no-allocation path's scratch array on `device`
// (sam2d_noalloc.hip), looked up once per device.
int noalloc_scratch(int device, void **out);
}  // namespace sam2d

void sam2d_adjoint_release(sam2d_ctx *ctx);
