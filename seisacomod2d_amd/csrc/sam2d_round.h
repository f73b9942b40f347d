// sam2d_round.h -- unbiased FP32 rounding of the coefficient products,
// shared by the host setup (sam2d_setup.cpp) and the kernels that form the
// buoyancy coefficients themselves (sam2d_kernelsn.hip), so both give the
// same bits.
//
// A coefficient t (dt/dx * kappa, dt/dx * bu2, dt/dz * bu1, the reference's
// FP64 products) is stored as the nearest float n or as its neighbour on
// t's side of n, the neighbour with probability |t - n| / (its distance from
// n), so that the stored value is t on average; the draw is a fixed function
// of (field, cell), the same on every run.  Round-to-nearest would give every
// cell of a homogeneous region the same relative error (up to 2^-24 in each
// of the two coefficients whose product sets the wave speed), a systematic
// speed error whose phase error grows along the record.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace sam2d {

// Fields of the draw.
enum { RND_KX = 0, RND_KZ = 1, RND_BX = 2, RND_BZ = 3 };

// The draw of (field, cell) as 32 fixed-point bits of [0, 1): the
// two-dimensional Weyl sequence frac(ix * a + iz * b), with a = 1/g and
// b = 1/g^2 for the plastic number g (g^3 = g + 1; the constants of Roberts'
// R2 low-discrepancy sequence).  KX and KZ take (a, b), BX and BZ (b, a); KZ
// and BZ add one half.  The cells of any region draw values spread evenly
// over [0, 1), so the fraction rounded up follows the remainder closely, and
// the draw costs the kernels one multiply-add per cell (the iz term is fixed
// per lane).
__host__ __device__ inline uint32_t cell_draw(uint32_t field, uint32_t ix, uint32_t iz)
{
    const uint32_t a = 0xC13FA9A9u, b = 0x91E10DA6u;
    const bool swap = field >= (uint32_t)RND_BX;
    return ix * (swap ? b : a) + iz * (swap ? a : b) + ((field & 1u) ? 0x80000000u : 0u);
}

__host__ __device__ inline float f32_bits(uint32_t u)
{
    union { uint32_t u; float f; } v;
    v.u = u;
    return v.f;
}
__host__ __device__ inline uint32_t u32_bits(float f)
{
    union { uint32_t u; float f; } v;
    v.f = f;
    return v.u;
}

// The draw as a float in [0, 1) on a 2^-24 grid (exact).
__host__ __device__ inline float draw24(uint32_t h) { return (float)(h >> 8) * 0x1.0p-24f; }

// Nearest float n of a positive t and the remainder rem = t - n: returns n,
// or its neighbour towards t when draw * spacing < |rem|.  Every product
// and difference below is exact, so host and device agree bit for bit.
__host__ __device__ inline float round_towards(float n, float rem, uint32_t h)
{
    if (rem == 0.0f || !(n > 0.0f) || !(n < 3.0e38f)) return n;
    const float other = f32_bits(u32_bits(n) + (rem > 0.0f ? 1u : 0xFFFFFFFFu));
    const float spacing = fabsf(other - n);             // a power of two
    return draw24(h) * spacing < fabsf(rem) ? other : n;
}

// An FP64 value t (dt/dx * kappa, kappa = vp^2 rho in FP64).
__host__ __device__ inline float round_unbiased(double t, uint32_t h)
{
    const float n = (float)t;
    return round_towards(n, (float)(t - (double)n), h);
}

// The exact product of two floats a * b, known to be finite and not negative
// (dt/dx and dt/dz are float products, cpumodelling.cpp:180-181, and the
// staggered buoyancies are positive floats): n = a * b rounded to nearest,
// rem = fma(a, b, -n) exactly, then as round_towards.  The guards of
// round_towards can only fire with a zero remainder here, where both forms
// return n (for n = 0 the neighbour is a NaN, which the compare rejects), so
// the kernels that form the buoyancy coefficients themselves skip them.
__host__ __device__ inline float round_unbiased_prod_pos(float a, float b, uint32_t h)
{
#pragma clang fp contract(off)
    const float n = a * b;
    const float rem = fmaf(a, b, -n);
    const float other = f32_bits(u32_bits(n) + (rem > 0.0f ? 1u : 0xFFFFFFFFu));
    const float spacing = fabsf(other - n);             // a power of two (NaN when n = 0)
    return draw24(h) * spacing < fabsf(rem) ? other : n;
}

}  // namespace sam2d
