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
// Layout: cells are column-major with z fastest, cell (ix, iz) at
// ix*nnz + iz (the reference's id = iz + ix*nnz), the state cell-interleaved
// ((P, vx, vz) of a cell together, P = px in the sponge frame and p = px + pz
// outside it, pz a field of the frame; slot_cells).  A wavefront owns a strip
// of OWN_Z rows and carries a HALO_T-row halo above and HALO_B rows below in
// the same 64 lanes, so z-neighbours are lane shuffles.  It marches along x
// over a chunk of columns keeping a sliding register window of vx, vz and the
// new p; p is recomputed redundantly on the 3 halo rows and the 3 halo
// columns it needs, so one pass reads the state, KX, BX, BZ once and writes
// the state once: 36 bytes per cell update outside the frame.  State is
// double-buffered (in -> out) because neighbouring tiles read the old state
// of the cells a tile overwrites.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <algorithm>
#include "sam2d_device.h"
#include "sam2d_kcommon.h"

namespace sam2d {

// Receiver gather of p = px + pz from a state slot (cpu_record,
// cpu_modelling_kernels.cpp:24-31, trace-major output as after
// cpu_transpose, :370-381).
__device__ __forceinline__ void gather_receivers(const ShotDev *sh, int nshots, int in_slot,
                                                 int slot, int gtid, int gstride)
{
    for (int s = 0; s < nshots; ++s) {
        const ShotDev &d = sh[s];
        const SAM2D_GLOBAL int *gc = gp(d.gcell);
        SAM2D_GLOBAL float *seis = gp(d.seis);
        for (int g = gtid; g < d.ng; g += gstride) {
            const int c = gc[g];
            const int col = c / d.nnz;
            seis[(size_t)g * d.real_nt + slot] = state_pressure(d, in_slot, c, col, c - col * d.nnz);
        }
    }
}

// Operands of one column that the march loads (one row per lane).
struct ColOps {
    ftri s;                     // cells (P, vx, vz) at column c+1, kept whole until consumed
    float pz;                   // pz at column c+1 (frame lanes; 0 elsewhere)
    float kx, kz;               // at column c
    float dax, skx;             // x-profiles at column c (uniform)
    float bx, bz;               // BX, BZ at column j = c-2
    float dbx, sbx;             // x-profiles at column j (uniform)
};

template <bool NT>
__device__ __forceinline__ void st3n(SAM2D_GLOBAL char *base, unsigned off, float a, float b,
                                     float c)
{
    ftri v;
    v.x = a; v.y = b; v.z = c;
    SAM2D_GLOBAL ftri *p = (SAM2D_GLOBAL ftri *)(base + off);
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One time step: state slot in_slot -> out_slot.  rec_slot >= 0 asks for the
// receiver gather of the input state (the previous step's result) into
// sample rec_slot.
//
// WAVES: z-strips per workgroup.  PF: columns of operands in flight (2 or 4).
// NTS: non-temporal hint on the state stores.  RING (the gradient's
// recomputation): the ring entries of every owned cell (p after the step and
// the dJ/dkappa factors of the state before it, ring_pair / ring_side) also go
// to ring slot ring_slot.
//
// The march along x is unrolled by four columns so that the sliding windows
// (vx at c-2..c+1, new p at c-3..c) rotate by compile-time index and the
// operand slots are refilled in place: no register a load is still writing
// is ever copied, which keeps PF columns of loads in flight.
//
// Loads use clamped indices and are never masked: a lane outside the grid
// (halo rows above row 0 or below nnz-1, columns outside [0, nnx)) reads a
// finite in-grid value, and every such value meets a zero stencil coefficient
// (the reference's edge rules) before it could reach an owned cell.
//
// One 12-byte load per column brings the cells of column c+1 (a frame lane
// also loads pz), whose P, pz, vz wait in windows; the v-phase at j stores
// the cell of column j (and a frame lane's pz), P and pz kept from its
// p-phase two columns earlier.
template <int WAVES, int PF, bool NTS, bool RING>
__global__ __launch_bounds__(WAVES * 64)
void k_step(const ShotDev *__restrict__ shots, int nshots, int vtotal, int per_xcd,
            int ntiles, int step, int in_slot, int out_slot, int rec_slot, int ring_slot)
{
    static_assert(PF == 1 || PF == 2 || PF == 4, "PF must divide the unroll of 4");
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;

    if (rec_slot >= 0)
        gather_receivers(shots, nshots, in_slot, rec_slot,
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
        int xa, xb;
        tile_columns(d, chunk, xa, xb);

        const int r = zb - HALO_T + lane;             // this lane's row
        const bool rown = (lane >= HALO_T) && (lane < HALO_T + OWN_Z) && (r < nnz);
        const unsigned rc = (unsigned)min(max(r, 0), nnz - 1);

        const SAM2D_GLOBAL char *__restrict__ si = (const SAM2D_GLOBAL char *)gp(slot_cells(d, in_slot));
        SAM2D_GLOBAL char *__restrict__ so = (SAM2D_GLOBAL char *)gp(slot_cells(d, out_slot));
        const SAM2D_GLOBAL float *__restrict__ pzi = gp(slot_pz(d, in_slot));
        SAM2D_GLOBAL float *__restrict__ pzo = gp(slot_pz(d, out_slot));
        SAM2D_GLOBAL char *__restrict__ ro =
            RING ? (SAM2D_GLOBAL char *)gp(ring_pair(d.ring, d.ncell, ring_slot)) : nullptr;
        SAM2D_GLOBAL float *__restrict__ rs =
            RING ? gp(ring_side(d.ring, d.ncell, ring_slot)) : nullptr;
        const SAM2D_GLOBAL float *__restrict__ KX = gp(d.KX);
        const SAM2D_GLOBAL float *__restrict__ KZ = gp(d.KZ);
        const SAM2D_GLOBAL float *__restrict__ BX = gp(d.BX);
        const SAM2D_GLOBAL float *__restrict__ BZ = gp(d.BZ);
        const SAM2D_CONST float *__restrict__ prx = cp4(d.prof_x);   // [4][nnx]
        const SAM2D_GLOBAL float *__restrict__ prz = gp(d.prof_z);   // [4][nnz]

        // Per-row constants: sponge profiles and z-operator coefficients.
        const float daz = prz[rc], skz = prz[nnz + rc];
        const float dbz = prz[2 * nnz + rc], sbz = prz[3 * nnz + rc];
        float zm0, zm1, zp0, zp1;
        dminus_coef(r, nnz, zm0, zm1);
        dplus_coef(r, nnz, zp0, zp1);
        const bool src_row = (r == d.src_z);
        const float amp = (step < d.nts) ? gp(d.amp)[step] : 0.0f;
        const int src_x = (step <= d.nts) ? d.src_x : -1;

        // Column base (uniform, in elements) of a clamped column index.
        auto colbase = [&](int col) -> unsigned {
            return (unsigned)min(max(col, 0), nnx - 1) * (unsigned)nnz;
        };
        // Sponge frame: this lane's (clamped) row in the z-frame; a (clamped)
        // column outside the x-frame.  A lane outside the frame reads pz at
        // element 0 (one cache line for all of them) and takes 0.
        const bool zfr = (int)rc < d.zin_lo || (int)rc > d.zin_hi;
        auto xin = [&](int col) -> bool {
            const int cc = min(max(col, 0), nnx - 1);
            return cc >= d.xin_lo && cc <= d.xin_hi;
        };
        auto load_pz = [&](int col, unsigned base) -> float {
            const bool fr = zfr || !xin(col);
            const float v = pzi[fr ? base + rc : 0u];
            return fr ? v : 0.0f;
        };
        auto load_col = [&](int c) -> ColOps {
            ColOps o;
            const unsigned b1 = colbase(c + 1), b0 = colbase(c), bj = colbase(c - 2);
            o.s = ld3(si + (size_t)12 * (b1 + rc), 0u);
            o.pz = load_pz(c + 1, b1);
            o.kx = KX[b0 + rc]; o.kz = KZ[b0 + rc];
            o.bx = BX[bj + rc]; o.bz = BZ[bj + rc];
            const int cc = min(max(c, 0), nnx - 1), jc = min(max(c - 2, 0), nnx - 1);
            o.dax = prx[cc]; o.skx = prx[nnx + cc];
            o.dbx = prx[2 * nnx + jc]; o.sbx = prx[3 * nnx + jc];
            return o;
        };

        // Windows, indexed modulo 4 relative to the unrolled block's base
        // column c: vxw, pxw, pzw, vzw[k & 3] = old vx, px, pz, vz at c-2+k;
        // pw, pxnw, pznw[k & 3] = new p, px, pz at c-3+k.
        float vxw[4], pw[4], pxw[4], pzw[4], vzw[4], pxnw[4], pznw[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) vxw[k] = pw[k] = pxw[k] = pzw[k] = vzw[k] = pxnw[k] = pznw[k] = 0.0f;
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const unsigned b = colbase(xa - 3 + k);
            const ftri sv = ld3(si + (size_t)12 * (b + rc), 0u);
            pxw[k] = sv.x; vxw[k] = sv.y; vzw[k] = sv.z;
            pzw[k] = load_pz(xa - 3 + k, b);
        }

        ColOps q[PF];
        #pragma unroll
        for (int u = 0; u < PF; u++) q[u] = load_col(xa - 1 + u);

        // The march ends at column xb+1 (stores need c < xb, j < xb); the
        // prefetches past it re-read column xb+1, a cache hit.
        for (int c0 = xa - 1; c0 <= xb + 1; c0 += 4) {
            #pragma unroll
            for (int u = 0; u < 4; u++) {
                const int c = c0 + u;
                if (c > xb + 1) break;
                ColOps &cur = q[u % PF];
                // ---- p-phase at column c (cpu_pml_p, :335-351; source, :135-140)
                pxw[(u + 3) & 3] = cur.s.x;
                pzw[(u + 3) & 3] = cur.pz;
                vxw[(u + 3) & 3] = cur.s.y;
                vzw[(u + 3) & 3] = cur.s.z;
                const float vx_m2 = vxw[u & 3], vx_m1 = vxw[(u + 1) & 3];
                const float vx_0 = vxw[(u + 2) & 3], vx_p1 = vxw[(u + 3) & 3];
                float xm0, xm1;
                dminus_coef(c, nnx, xm0, xm1);
                const float dvx = fmaf(xm0, vx_0 - vx_m1, xm1 * (vx_p1 - vx_m2));
                const float vz_0 = vzw[(u + 2) & 3];
                const float vz_u2 = shfl_up(vz_0, 2), vz_u1 = shfl_up(vz_0, 1);
                const float vz_d1 = shfl_dn(vz_0, 1);
                const float dvz = fmaf(zm0, vz_0 - vz_u1, zm1 * (vz_d1 - vz_u2));
                const float px_old = pxw[(u + 2) & 3];
                const float pz_old = pzw[(u + 2) & 3];
                float pxn = fmaf(cur.kx * cur.skx, dvx, cur.dax * px_old);
                const float pzn = fmaf(cur.kz * skz, dvz, daz * pz_old);
                if (src_row && c == src_x) pxn += amp;
                const float pn = pxn + pzn;
                // outside the frame P carries p and pz stays 0
                const bool inner = !zfr && xin(c);
                pw[(u + 3) & 3] = pn;
                pxnw[(u + 3) & 3] = inner ? pn : pxn;  // stored with vx, vz at j
                pznw[(u + 3) & 3] = inner ? 0.0f : pzn;
                if (RING && rown && c >= xa && c < xb) {
                    // the step's pressure and the dJ/dkappa factors of the
                    // state before it (ring_pair, ring_side)
                    float ez;
                    const float e = ring_factor(dvx, d.dtdx * cur.skx, dvz, d.dtdz * skz, inner, ez);
                    const size_t cell = (size_t)c * (size_t)nnz + (size_t)r;
                    st2(ro + cell * 8, 0u, pn, e);
                    if (!inner) rs[cell] = ez;
                }

                // ---- v-phase at column j = c - 2 (cpu_pml_v, :353-368)
                const int j = c - 2;
                if (j >= xa && j < xb) {
                    const float p_m3 = pw[u & 3], p_m2 = pw[(u + 1) & 3], p_m1 = pw[(u + 2) & 3];
                    float xp0, xp1;
                    dplus_coef(j, nnx, xp0, xp1);
                    const float dpx = fmaf(xp0, p_m1 - p_m2, xp1 * (pn - p_m3));
                    const float q_u1 = shfl_up(p_m2, 1);
                    const float q_d1 = shfl_dn(p_m2, 1), q_d2 = shfl_dn(p_m2, 2);
                    const float dpz = fmaf(zp0, q_d1 - p_m2, zp1 * (q_d2 - q_u1));
                    const float vxn = fmaf(cur.bx * cur.sbx, dpx, cur.dbx * vx_m2);
                    const float vz_j = vzw[u & 3];
                    const float vzn = fmaf(cur.bz * sbz, dpz, dbz * vz_j);
                    if (rown) {
                        const unsigned id = (unsigned)j * (unsigned)nnz + (unsigned)r;
                        st3n<NTS>(so + (size_t)12 * id, 0u, pxnw[(u + 1) & 3], vxn, vzn);
                        if (zfr || !xin(j)) pzo[id] = pznw[(u + 1) & 3];
                    }
                }
                cur = load_col(min(c + PF, xb + 1));
            }
        }
    }
}

// Receiver gather alone (the last step's samples).
__global__ void k_record(const ShotDev *__restrict__ shots, int nshots, int in_slot, int slot)
{
    gather_receivers(shots, nshots, in_slot, slot, blockIdx.x * blockDim.x + threadIdx.x,
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
static const int kChunks[] = {16, 32, 64, 128, 256};
static const int kPF[] = {2, 4};
static const int kNTS[] = {0, 1};
static const int nW = 4, nC = 5, nP = 2, nN = 2;
static const int kChunks2[] = {16, 32, 64, 128, 256};
static const int nSingle = nW * nC * nP * nN;     // single-step variants first
static const int nMultiEach = nW * 5 * nP;        // then 2-, 3-, 4- and 5-step variants
static const int kMaxSteps = 5;

int num_variants() { return nSingle + nMultiEach * (kMaxSteps - 1); }

// Variant index -> (steps per launch, waves, chunk, prefetch, nt-stores).
static void decode(int v, int &w, int &c, int &pf, int &nts, int &steps)
{
    if (v >= nSingle) {
        v -= nSingle;
        steps = 2 + v / nMultiEach;
        v %= nMultiEach;
        nts = 0;
        c = kChunks2[v % 5];
        w = kWaves[(v / 5) % nW];
        pf = kPF[v / (5 * nW)];
        return;
    }
    steps = 1;
    c = kChunks[v % nC]; v /= nC;
    w = kWaves[v % nW]; v /= nW;
    pf = kPF[v % nP]; v /= nP;
    nts = kNTS[v % nN];
}
static void decode(int v, int &w, int &c, int &pf, int &nts)
{
    int steps;
    decode(v, w, c, pf, nts, steps);
}

void variant_shape(int v, int *waves, int *chunk)
{
    int pf, nts;
    decode(v, *waves, *chunk, pf, nts);
}

void variant_shape_pf(int v, int *waves, int *chunk, int *pf)
{
    int nts;
    decode(v, *waves, *chunk, *pf, nts);
}

int variant_steps(int v)
{
    int w, c, pf, nts, steps;
    decode(v, w, c, pf, nts, steps);
    return steps;
}

int variant_own_rows(int v)
{
    const int t = variant_steps(v);
    return t == 1 ? OWN_Z : 64 - 6 * t;
}

int variant_single_step(int v)
{
    if (v < nSingle) return v;
    int w, c, pf, nts;
    decode(v, w, c, pf, nts);
    int wi = 0, ci = 1;
    for (int i = 0; i < nW; i++) if (kWaves[i] == w) wi = i;
    for (int i = 0; i < nC; i++) if (kChunks[i] == c) ci = i;
    return wi * nC + ci;   // prefetch 2, ordinary stores
}

int variant_index(int waves, int chunk, int pf)
{
    int wi = -1, ci = -1, pi = -1;
    for (int i = 0; i < nW; i++) if (kWaves[i] == waves) wi = i;
    for (int i = 0; i < nC; i++) if (kChunks[i] == chunk) ci = i;
    for (int i = 0; i < nP; i++) if (kPF[i] == pf) pi = i;
    if (wi < 0 || ci < 0 || pi < 0) return -1;
    return (pi * nW + wi) * nC + ci;   // ordinary stores
}

void variant_name(int v, char *buf, int n)
{
    int w, c, pf, nts, steps;
    decode(v, w, c, pf, nts, steps);
    static const char *kStepNames[] = {"", "", "twostep", "threestep", "fourstep", "fivestep"};
    if (steps > 1)
        snprintf(buf, n, "%s_waves%d_chunk%d%s", kStepNames[steps], w, c, pf == 4 ? "_pf4" : "");
    else snprintf(buf, n, "waves%d_chunk%d_pf%d%s", w, c, pf, nts ? "_nts" : "");
}

struct StepArgs {
    const ShotDev *d;
    int ns, vt, per, nt, step, in, out, rec, ring;
};

template <int W, int PF, bool NTS, bool RING>
static void launch_t(dim3 g, hipStream_t s, const StepArgs &a)
{
    hipLaunchKernelGGL((k_step<W, PF, NTS, RING>), g, dim3(64 * W), 0, s, a.d, a.ns, a.vt, a.per,
                       a.nt, a.step, a.in, a.out, a.rec, a.ring);
}

template <int W>
static void launch_w(int pf, int nts, bool ring, dim3 g, hipStream_t s, const StepArgs &a)
{
    if (ring) {
        if (pf == 2) launch_t<W, 2, false, true>(g, s, a);
        else launch_t<W, 4, false, true>(g, s, a);
    } else if (pf == 2 && !nts) launch_t<W, 2, false, false>(g, s, a);
    else if (pf == 2) launch_t<W, 2, true, false>(g, s, a);
    else if (!nts) launch_t<W, 4, false, false>(g, s, a);
    else launch_t<W, 4, true, false>(g, s, a);
}

static hipError_t launch_step_l(int variant, const ShotDev *d_shots, int nshots, int ntiles,
                                int grid_cap, int step, int in_slot, int out_slot, int rec_slot,
                                int ring_slot, bool ring, hipStream_t stream)
{
    int waves, chunk, pf, nts;
    decode(variant, waves, chunk, pf, nts);
    const int per = (ntiles + 7) / 8;
    const int vtotal = per * 8;
    int grid = vtotal;
    if (grid_cap > 0 && grid > grid_cap) grid = (grid_cap + 7) / 8 * 8;
    const dim3 g(grid);
    const StepArgs a{d_shots, nshots, vtotal, per, ntiles, step, in_slot, out_slot, rec_slot,
                     ring_slot};
    switch (waves) {
    case 1: launch_w<1>(pf, nts, ring, g, stream, a); break;
    case 2: launch_w<2>(pf, nts, ring, g, stream, a); break;
    case 4: launch_w<4>(pf, nts, ring, g, stream, a); break;
    case 8: launch_w<8>(pf, nts, ring, g, stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_step(int variant, const ShotDev *d_shots, int nshots, int ntiles, int grid_cap,
                       int step, int in_slot, int out_slot, int rec_slot, hipStream_t stream)
{
    return launch_step_l(variant, d_shots, nshots, ntiles, grid_cap, step, in_slot, out_slot,
                         rec_slot, 0, false, stream);
}

hipError_t launch_step_ring(int variant, const ShotDev *d_shots, int nshots, int ntiles,
                            int grid_cap, int step, int in_slot, int out_slot, int ring_slot,
                            hipStream_t stream)
{
    return launch_step_l(variant, d_shots, nshots, ntiles, grid_cap, step, in_slot, out_slot, -1,
                         ring_slot, true, stream);
}

hipError_t launch_record(const ShotDev *d_shots, int nshots, int max_ng, int in_slot, int slot,
                         hipStream_t stream)
{
    int blocks = (max_ng + 255) / 256;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_record, dim3(blocks), dim3(256), 0, stream, d_shots, nshots, in_slot, slot);
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
// Variant used until the autotuner has run: 4 strips per workgroup, 32-column
// chunks, two columns of operands in flight, ordinary stores.
int default_variant_index()
{
    return ((0 * nP + 0) * nW + 2) * nC + 1;
}
}  // namespace sam2d
