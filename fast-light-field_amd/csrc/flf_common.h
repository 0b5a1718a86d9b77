// Shared device/host definitions for the light-field projector on gfx950.
//
// Conventions used throughout (see the plan builder in flf_plan.cpp):
//   N      lenslet pitch in pixels (Nnum), odd; c = N/2; h = (N+1)/2
//   f      Fourier grid (fY, fX), multiples of N;  r = f / N (reduced grid)
//   rfX    = fX/2 + 1  (half-plane columns of a real 2-D transform)
//   u      index of a stored quadrant spectrum (36 for square grids, 64 else)
//   slot   position of an output bin inside an item (4 = quad, 2 = pair)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flf_types.h"

__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// acc += s * g
__device__ __forceinline__ void cmac(float2& acc, float2 s, float2 g) {
    acc.x = fmaf(s.x, g.x, acc.x);
    acc.x = fmaf(-s.y, g.y, acc.x);
    acc.y = fmaf(s.x, g.y, acc.y);
    acc.y = fmaf(s.y, g.x, acc.y);
}
// acc += s * conj(g)
__device__ __forceinline__ void cmacc(float2& acc, float2 s, float2 g) {
    acc.x = fmaf(s.x, g.x, acc.x);
    acc.x = fmaf(s.y, g.y, acc.x);
    acc.y = fmaf(s.y, g.x, acc.y);
    acc.y = fmaf(-s.x, g.y, acc.y);
}

// e^{-2 pi i * num / den} for integers, with the argument reduced exactly.
// 32-bit form for |num| < 2^31 (the MAC epilogues: |outer| < 4N (s2 + N)).
__device__ __forceinline__ float2 phase_frac_i(int num, int den) {
    int m = num % den;
    if (m < 0) m += den;
    if (2 * m > den) m -= den;
    float s, c;
    sincospif(-2.0f * (float)m / (float)den, &s, &c);
    return make_float2(c, s);
}
__device__ __forceinline__ float2 phase_frac(long long num, long long den) {
    long long m = num % den;
    if (m < 0) m += den;
    if (2 * m > den) m -= den;            // |angle| <= pi keeps the rounding small
    float s, c;
    sincospif(-2.0f * (float)m / (float)den, &s, &c);
    return make_float2(c, s);
}
