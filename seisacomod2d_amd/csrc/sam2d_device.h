// sam2d_device.h -- descriptors shared by the kernels and the host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace sam2d {

// Rows owned by one wavefront, and its halo rows above / below (64 lanes).
constexpr int OWN_Z = 58;
constexpr int HALO_T = 3;
constexpr int HALO_B = 64 - OWN_Z - HALO_T;
// Kernels that advance T > 1 time steps per pass carry 3T halo rows on each
// side and own 64 - 6T rows (T = 2: the two-step adjoint).
constexpr int OWN_Z2 = 52;
constexpr int HALO_T2 = 6;
// Band slices of the receiver-row pressure buffer: one per intermediate step
// of a launch (up to five steps per launch).
constexpr int BAND_SLICES = 4;

// One resident shot: its padded domain, coefficient fields, ping-pong state
// and receiver list, all device pointers.  Kernels read it from device memory
// so a captured time loop can be replayed for another shot by rewriting the
// descriptors only.
struct ShotDev {
    int nnz, nnx;          // padded domain of this shot
    int ngroups, nchunks;  // tiles: z-strip groups x x-chunks
    int chunk;             // columns per x-chunk
    int x0, x1;            // x0 < x1: the chunk columns split [0, x0), [x0, x1) and
    int nlead, nmid;       //   [x1, nnx) into nlead columns fchunk wide, nmid chunk
    int fchunk;            //   wide and the rest fchunk wide (the last of each part
                           //   shorter); otherwise chunk-wide columns from 0
                           //   (tile_columns)
    int tile_begin;        // first flat tile id of this shot
    int src_x, src_z;      // source cell (padded indices)
    int nts;               // wavelet length in steps
    int ng, real_nt;       // receivers, samples per trace
    int xin_lo, xin_hi;    // columns where all four x-profiles are exactly 1
    int zin_lo, zin_hi;    // rows where all four z-profiles are exactly 1
    int rrow_lo, rrow_hi;  // padded rows of the receivers (lo > hi: none)
    int band_h;            // rrow_hi - rrow_lo + 1 (0: no receivers)
    const float *amp;      // [nts] source increments (fd_dt*w*kappa_s/(dx*dz))
    const float *KX, *KZ;  // dt/dx*kappa, dt/dz*kappa   (KZ may alias KX)
    const float *BX, *BZ;  // dt/dx*bu2,  dt/dz*bu1
    const float *B;        // buoyancy 1/rho (the multi-step kernels form BX, BZ from it)
    float dtdx, dtdz;      // fd_dt/dx, fd_dt/dz (float products, as the reference forms them)
    const float *prof_x;   // [4][nnx]: damp2a, damp2a1, damp2b, damp2b1
    const float *prof_z;   // [4][nnz]: damp1a, damp1a1, damp1b, damp1b1
    float *state;          // state slots of 4*ncell floats (slot_cells, slot_pz)
    size_t ncell;
    float *ring;           // recomputation only: the forward ring, slot m (segment
                           //   step m) at ring + 3*m*ncell (ring_pair, ring_side)
    const int *gcell;      // [ng] receiver cells (ix*nnz + iz)
    float *seis;           // [ng][real_nt]
    float *band;           // [2][BAND_SLICES][band_h][nnx]: pressure of the receiver
                           // rows after the intermediate steps of a launch
};

// State slot s holds 4*ncell floats: the cells, then the frame's pz.
//  * cells: (P, vx, vz) of cell i at [3i .. 3i+2].  Inside the sponge frame
//    P is the split field px; outside it (columns xin_lo..xin_hi and rows
//    zin_lo..zin_hi, where all eight profile factors are exactly 1) the split
//    is never needed, since px and pz are only ever used as their sum there,
//    and P is the pressure p = px + pz itself.
//  * pz of cell i at [3*ncell + i], read and written only in the frame; zero
//    elsewhere.
// Per column a wave reads and writes one contiguous 768-byte run of cells
// (the frame adds a masked 256-byte run); outside the frame a cell moves
// 12 bytes each way instead of 16.
__host__ __device__ inline float *slot_cells(const ShotDev &d, int slot)
{
    return d.state + (size_t)slot * 4 * d.ncell;
}
__host__ __device__ inline float *slot_pz(const ShotDev &d, int slot)
{
    return d.state + (size_t)slot * 4 * d.ncell + 3 * d.ncell;
}
// Columns [xa, xb) of chunk column k of a shot's tiles.
__host__ __device__ inline void tile_columns(const ShotDev &d, int k, int &xa, int &xb)
{
    if (d.x0 < d.x1) {
        int lo = 0, hi = d.x0, w = d.fchunk;
        if (k >= d.nlead + d.nmid) { k -= d.nlead + d.nmid; lo = d.x1; hi = d.nnx; }
        else if (k >= d.nlead) { k -= d.nlead; lo = d.x0; hi = d.x1; w = d.chunk; }
        xa = lo + k * w;
        xb = xa + w < hi ? xa + w : hi;
    } else {
        xa = k * d.chunk;
        xb = xa + d.chunk < d.nnx ? xa + d.chunk : d.nnx;
    }
}

// The forward ring: slot m holds, for segment step m, what the backward step
// of the same index reads of the forward: for cell i the pair (p, E) at
// [2i, 2i+1] (ring_pair) and, in the sponge frame only, Ez at [2N + i]
// (ring_side).  p is the pressure after the step; with Ex = Dx- vx (dt/dx skx)
// and Ez = Dz- vz (dt/dz skz) of the state before it (the p-sponge scale
// factors skx, skz), E is Ex + Ez outside the frame and Ex in it.  These are
// the dJ/dkappa factors of the step: outside the frame L_px = L_pz there and
// dJ/dkappa takes L_px E; in the frame, L_px Ex + L_pz Ez.
__host__ __device__ inline float *ring_pair(float *ring, size_t ncell, int slot)
{
    return ring + (size_t)slot * 3 * ncell;
}
__host__ __device__ inline float *ring_side(float *ring, size_t ncell, int slot)
{
    return ring + (size_t)slot * 3 * ncell + 2 * ncell;
}

// Whether (col, row) lies outside the sponge frame, where P holds p.
__host__ __device__ inline bool inner_cell(const ShotDev &d, int col, int row)
{
    return col >= d.xin_lo && col <= d.xin_hi && row >= d.zin_lo && row <= d.zin_hi;
}

// One shot of the adjoint (backward) time loop.  Tiles as in ShotDev.
struct AdjDev {
    int nnz, nnx;
    int ngroups, nchunks, chunk, tile_begin;
    int src_x, src_z, nts, real_nt;
    int xin_lo, xin_hi;    // columns where all four x-profiles are exactly 1
    int ncells;            // distinct receiver cells
    int band_h, rrow_lo, rrow_hi;   // receiver rows (rband layout)
    float dtdx, dtdz;      // fd_dt/dx, fd_dt/dz as the coefficients use them
    const float *srcg;     // [nts] d(source increment)/d(kappa_s) = fd_dt*w[n]/(dx*dz)
    const float *coef;     // KX, BX, BZ (and KZ unless it equals KX) of cell i at
                           //   [3i .. 3i+2] ([4i .. 4i+3] with KZ last)
    const float *prof_x, *prof_z;
    const float *ring;     // forward ring (ShotDev::ring: ring_pair, ring_side)
    float *lam;            // adjoint state: 2 slots of 3*ncell floats, L_px, L_vx, L_vz
                           //   of cell i at [3i .. 3i+2]
    float *lpz;            // L_pz: 2 slots of ncell floats, read and written only in the
                           //   sponge frame; elsewhere L_pz equals L_px (both sponge
                           //   factors are 1 there and both take the same increments)
    int zin_lo, zin_hi;    // rows where all four z-profiles are exactly 1
    size_t ncell;
    float *acc;            // dJ/dkappa, dJ/dbu2 (x-staggered buoyancy), dJ/dbu1
                           //   (z-staggered) of cell i at [3i .. 3i+2]
    const int *rcell;      // [ncells] distinct receiver cells (ix*nnz + iz)
    const int *rptr;       // [ncells+1] receivers of each cell (CSR)
    const int *rid;        // [ng] receiver index of each entry, ascending per cell
    const float *res;      // [ng][real_nt] adjoint source dJ/dseis
    const float *rband;    // [real_nt][band_h][nnx] residual per receiver cell
                           // (-0 elsewhere), or NULL: injected by k_inject
};

// Order of the chunk columns within each XCD's range of tiles (multi-step
// kernels): xorder > 1 visits them with stride xorder, otherwise in order.
// ngroups: z-groups per chunk column, the same for every staged shot.
struct TileOrder {
    int ngroups;
    int xorder;
};
// March length rounded up so that a march over chunk + extra columns is
// whole groups of four: the kernels march four columns per group, and a
// partial last group runs on the per-column (edge) path.  extra: 6T - 3 for
// the T-step forward kernels, 9 for the two-step adjoint.
inline int march_aligned(int chunk, int extra)
{
    return chunk + (4 - (chunk + extra) % 4) % 4;
}
// Blocks per XCD range for a launch of ntiles tiles: a multiple of ngroups
// when the columns are reordered (the tiles must then be whole columns of
// ngroups); otherwise the order falls back to the plain one.
inline int per_xcd_for(int ntiles, TileOrder &order)
{
    if (order.xorder > 1 && order.ngroups > 0 && ntiles % order.ngroups == 0)
        return (ntiles / order.ngroups + 7) / 8 * order.ngroups;
    order.xorder = 0;
    return (ntiles + 7) / 8;
}
// One backward step n of every staged shot: adjoint state lam_in -> lam_out,
// the forward ring slot of step n is fslot; rec_k >= 0 injects residual
// sample rec_k of the shots that carry an rband.  kz_alias: every staged
// shot's KZ is its KX (three-float coefficient cells).
hipError_t launch_adjoint_step(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                               bool kz_alias, int step, int lam_in, int lam_out, int fslot,
                               int rec_k, TileOrder order, hipStream_t stream);
// Backward steps n and n-1 in one pass (all shots must carry an rband or no
// receivers): the forward ring slot of step n is fslot, of step n-1
// fslot - 1; residual samples rec1 (step n) and rec2 (step n-1).  Tiles of
// OWN_Z2-row strips.
hipError_t launch_adjoint_step2(int variant, const AdjDev *d_adj, int nshots, int ntiles,
                                bool kz_alias, int step, int lam_in, int lam_out, int fslot,
                                int rec1, int rec2, TileOrder order, hipStream_t stream);
// The adjoint's coefficient cells of one domain of n cells: KX, BX, BZ at
// [3i .. 3i+2] when KZ is KX (kz == NULL), else KX, BX, BZ, KZ at [4i .. 4i+3].
hipError_t launch_pack_coef(const float *kx, const float *kz, const float *bx, const float *bz,
                            size_t n, float *out, hipStream_t stream);
// Dense residual band [real_nt][band_cells]: element (k, i) is the sum, in
// rid order, of res[rid[e] * real_nt + k] over e in [rptr[u], rptr[u+1]) with
// u = bmap[i], starting from -0; -0 where bmap[i] < 0.
hipError_t launch_band(const int *bmap, const int *rptr, const int *rid, const float *res,
                       int real_nt, int band_cells, float *band, hipStream_t stream);
// Residual sample rec_k into adjoint slot lam before a backward step (shots
// without an rband).
hipError_t launch_inject(const AdjDev *d_adj, int nshots, int max_cells, int lam, int rec_k,
                         hipStream_t stream);

// dJ/dvp, dJ/drho of one shot added into the model's FP64 gradients
// (nx x nz, x-major): acc holds the shot's (dJ/dkappa, dJ/dbu2, dJ/dbu1)
// cells on its padded domain of nnx x nnz cells, built from model columns
// [wx0, wx0 + w) with npml sponge cells on each side; vp, rho the model.
hipError_t launch_chain(const float *acc, int nnz, int nnx, int npml, int wx0, int w, int nz,
                        const float *vp, const float *rho, double *g_vp, double *g_rho,
                        hipStream_t stream);

int num_variants();
void variant_shape(int v, int *waves, int *chunk);
// The same plus the operand columns in flight (prefetch depth).
void variant_shape_pf(int v, int *waves, int *chunk, int *pf);
void variant_name(int v, char *buf, int n);
// Time steps one launch of variant v advances (1 to 5), and the rows a
// wavefront owns for it.
int variant_steps(int v);
int variant_own_rows(int v);
// The single-step variant with the same workgroup and chunk shape.
int variant_single_step(int v);
// Single-step variant with this shape (ordinary stores), or -1.
int variant_index(int waves, int chunk, int pf);

// T = variant_steps(variant) time steps (step .. step+T-1) of every staged
// shot, state slot in_slot -> out_slot.  band_mask bit j keeps the pressure
// after step step+j (j < T-1) of the receiver rows in band slice j of parity
// band_w.  Before that, the launch gathers the previous launch's samples:
// band_samples[j] from slice j of parity band_w ^ 1 and state_sample from the
// in_slot state (each skipped when < 0).  kz_alias: every staged shot's KZ is
// its KX (dt/dz == dt/dx).
hipError_t launch_stepn(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                        bool kz_alias, int step, int in_slot, int out_slot, int band_mask,
                        int band_w, const int *band_samples, int state_sample, int max_ng,
                        TileOrder order, hipStream_t stream);
// The recomputation form of a three-step variant: state slot in_slot ->
// in_slot ^ 1, and for each of the three steps step+t the ring entries of
// every owned cell (p after it and the dJ/dkappa factors of the state before
// it, ring_pair / ring_side) into ring slot ring_slot + t (no receivers).
bool ring_supported(int variant);
hipError_t launch_stepn_ring(int variant, const ShotDev *d_shots, int nshots, int ntiles,
                             int grid_cap, bool kz_alias, int step, int in_slot, int ring_slot,
                             TileOrder order, hipStream_t stream);
// The deferred gathers as a launch of their own (after the last multi-step
// launch): band_samples[j] from slice j of parity band_r, the state slot.
hipError_t launch_gather(const ShotDev *d_shots, int nshots, int max_ng, int band_r,
                         const int *band_samples, int state_slot, int state_sample,
                         hipStream_t stream);

// One time step of every staged shot: reads state slot in_slot, writes
// out_slot; rec_sample >= 0 gathers the receivers of the in_slot state.
hipError_t launch_step(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                       int step, int in_slot, int out_slot, int rec_sample, hipStream_t stream);
// The recomputation form of one step: in_slot -> out_slot and the step's
// ring entries into ring slot ring_slot; the single-step variant's tile shape
// (stores without the non-temporal hint), no receivers.
hipError_t launch_step_ring(int variant, const ShotDev *d_shots, int nshots, int ntiles,
                            int grid_cap, int step, int in_slot, int out_slot, int ring_slot,
                            hipStream_t stream);
hipError_t launch_record(const ShotDev *d_shots, int nshots, int max_ng, int slot, int sample,
                         hipStream_t stream);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t stream);

}  // namespace sam2d
