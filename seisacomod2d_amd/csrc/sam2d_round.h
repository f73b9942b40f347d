// sam2d_round.h -- unbiased FP32 rounding of the coefficient products,
// shared by the host setup (sam2d_setup.cpp) and the kernels that form the
// buoyancy coefficients themselves (sam2d_kernelsn.hip), so both give the
// same bits.
//
// A coefficient t (dt/dx * kappa, dt/dx * bu2, dt/dz * bu1, the reference's
// FP64 products) is stored as the nearest float n or as its neighbour on
// t's side of n, the neighbour with probability |t - n| / (its distance from
// n), so that the stored value is t on average; the draw is a hash of
// (field, cell), the same on every run.  Round-to-nearest would give every
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

__host__ __device__ inline uint32_t cell_hash(uint32_t field, uint32_t ix, uint32_t iz)
{
    uint32_t h = (ix * 0x9E3779B1u) ^ ((iz + 0x7F4A7C15u) * 0x85EBCA77u) ^ (field * 0xC2B2AE3Du);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
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

// The exact product of two floats a * b (dt/dx and dt/dz are float
// products, cpumodelling.cpp:180-181, and the staggered buoyancies are
// floats): n = a * b rounded to nearest, rem = fma(a, b, -n) exactly.
__host__ __device__ inline float round_unbiased_prod(float a, float b, uint32_t h)
{
#pragma clang fp contract(off)
    const float n = a * b;
    const float rem = fmaf(a, b, -n);
    return round_towards(n, rem, h);
}

}  // namespace sam2d
