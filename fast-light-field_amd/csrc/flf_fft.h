// Batched mixed-radix (2/3/4/5) complex FFTs staged in LDS.
//
// Stockham auto-sort formulation: pass p with radix R and Ns = product of
// the radices of the earlier passes maps, for j in [0, L/R), k = j mod Ns,
//     a_t = x[j + t*L/R] * exp(sign*2*pi*i*t*k/(Ns*R))      t in [0, R)
//     y   = DFT_R(a)
//     x'[(j/Ns)*Ns*R + k + q*Ns] = y_q                      q in [0, R)
// and the output is in natural order after the last pass.  Each pass reads
// all of its inputs into registers, synchronises, then writes in place, so one
// LDS buffer suffices.  Every thread of the block must call these routines.
#pragma once
#include "flf_common.h"

namespace flf {

// Largest radix of a pass (wave-level transforms; the block-level passes
// below use 2..5).
constexpr int FLF_RADIX_MAX = 8;
// Largest odd prime radix (a prime factor of Nnum above 5, Nnum <= 31).  Such
// a radix is always the last pass of a plan: then Ns * P = L, every
// butterfly writes its outputs onto its own inputs, and a thread can take its
// butterflies one at a time.
constexpr int FLF_PRIME_MAX = 31;

template <int R>
__device__ __forceinline__ void dft_small(float2* a, float sign);

// cos and sin of 2 pi m / N for m = 0..3: the internal twiddles of the
// two-level butterflies below (N = 6, 8).
template <int N>
__device__ __forceinline__ float2 unit_root(int m, float sign) {
    constexpr float c6[4] = {1.000000000e+00f, 5.000000000e-01f, -5.000000000e-01f, -1.000000000e+00f};
    constexpr float s6[4] = {0.000000000e+00f, 8.660254038e-01f, 8.660254038e-01f, 0.000000000e+00f};
    constexpr float c8[4] = {1.000000000e+00f, 7.071067812e-01f, 0.000000000e+00f, -7.071067812e-01f};
    constexpr float s8[4] = {0.000000000e+00f, 7.071067812e-01f, 1.000000000e+00f, 7.071067812e-01f};
    const float c = N == 6 ? c6[m] : c8[m];
    const float sn = N == 6 ? s6[m] : s8[m];
    return make_float2(c, sign * sn);
}

// DFT of size N = R1 * R2 as R2 DFTs of size R1 (inputs R2 apart), internal
// twiddles W_N^(n2 k1), then R1 DFTs of size R2 (Cooley-Tukey, outputs in
// natural order): X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) W_N^(n2 k1)
// sum_n1 W_R1^(n1 k1) x[R2 n1 + n2], W_K = exp(sign 2 pi i / K).
template <int R1, int R2>
__device__ __forceinline__ void dft_two_level(float2* a, float sign) {
    constexpr int N = R1 * R2;
    // Stage 1 in place: the size-R1 transform of inputs a[R2 n1 + n2] (n1 <
    // R1) returns output k1 to slot R2 k1 + n2.
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        float2 t[5];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[R2 * n1 + n2];
        dft_small<R1>(t, sign);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1)
            a[R2 * k1 + n2] = (n2 * k1 == 0) ? t[k1] : cmul(t[k1], unit_root<N>(n2 * k1, sign));
    }
    // Stage 2: size-R2 transforms of a[R2 k1 + n2] (n2 < R2), output k2 to
    // X[k1 + R1 k2]; the outputs are permuted into place afterwards.
    float2 y[N];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) {
        float2 t[5];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) t[n2] = a[R2 * k1 + n2];
        dft_small<R2>(t, sign);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) y[k1 + R1 * k2] = t[k2];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) a[i] = y[i];
}

template <int R>
__device__ __forceinline__ void dft_small(float2* a, float sign) {
    if constexpr (R == 2) {
        float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        float2 s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
        // j*d13 with j = sign*i
        float2 jd13 = make_float2(-sign * d13.y, sign * d13.x);
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, jd13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, jd13);
    } else if constexpr (R == 3) {
        const float c = -0.5f;
        const float s = sign * 0.86602540378443864676f;
        float2 t1 = cadd(a[1], a[2]);
        float2 t2 = csub(a[1], a[2]);
        float2 m = make_float2(a[0].x + c * t1.x, a[0].y + c * t1.y);
        float2 is = make_float2(-s * t2.y, s * t2.x);
        a[0] = cadd(a[0], t1);
        a[1] = cadd(m, is);
        a[2] = csub(m, is);
    } else if constexpr (R == 5) {
        const float c1 = 0.30901699437494742410f;   // cos(2pi/5)
        const float c2 = -0.80901699437494742410f;  // cos(4pi/5)
        const float s1 = sign * 0.95105651629515357212f;
        const float s2 = sign * 0.58778525229247312917f;
        float2 t1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
        float2 t2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
        float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        // i*(s1 d1 + s2 d2) and i*(s2 d1 - s1 d2)
        float2 u1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        float2 u2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        float2 iu1 = make_float2(-u1.y, u1.x);
        float2 iu2 = make_float2(-u2.y, u2.x);
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(m1, iu1);
        a[4] = csub(m1, iu1);
        a[2] = cadd(m2, iu2);
        a[3] = csub(m2, iu2);
    } else if constexpr (R == 6) {
        dft_two_level<2, 3>(a, sign);
    } else if constexpr (R == 8) {
        dft_two_level<2, 4>(a, sign);
    }
}

// One Stockham pass of radix R over `nbatch` sequences of length L stored at
// buf[b*bs + i*es].  MAXB bounds the butterflies a thread holds in registers:
// nbatch*L/R <= MAXB*blockDim.x is required (checked by the host planner).
template <int R, int MAXB>
__device__ __forceinline__ void stockham_pass(float2* buf, int L, int nbatch, int es, int bs,
                                              int Ns, float sign) {
    const int M = L / R;
    const int total = nbatch * M;
    const int nth = blockDim.x;
    const int tid = threadIdx.x;
    float2 v[MAXB][5];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        int idx = tid + q * nth;
        if (idx < total) {
            int b = idx / M;
            int j = idx - b * M;
            int k = j % Ns;
            float2* base = buf + (long)b * bs;
            float2 a[5];
#pragma unroll
            for (int t = 0; t < R; ++t) a[t] = base[(j + t * M) * es];
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    float s, c;
                    sincospif(sign * 2.0f * (float)(t * k) / (float)(Ns * R), &s, &c);
                    a[t] = cmul(a[t], make_float2(c, s));
                }
            }
            dft_small<R>(a, sign);
#pragma unroll
            for (int t = 0; t < R; ++t) v[q][t] = a[t];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        int idx = tid + q * nth;
        if (idx < total) {
            int b = idx / M;
            int j = idx - b * M;
            int k = j % Ns;
            float2* base = buf + (long)b * bs;
            int o = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) base[(o + t * Ns) * es] = v[q][t];
        }
    }
    __syncthreads();
}

// Last pass of an odd prime radix P (5 < P <= FLF_PRIME_MAX; Ns = L / P):
// for butterfly j < M = L/P, y_q = sum_t x[j + t M] W_L^(j t) W_P^(t q),
// written back to x[j + q M].  Twiddles by sincospif of exactly reduced
// arguments (j t < L, t q mod P).
__device__ __forceinline__ void stockham_pass_prime(float2* buf, int L, int P, int nbatch, int es, int bs,
                                                    float sign) {
    const int M = L / P;
    const int total = nbatch * M;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int b = idx / M;
        const int j = idx - b * M;
        float2* base = buf + (long)b * bs;
        float2 a[FLF_PRIME_MAX];
#pragma unroll
        for (int t = 0; t < FLF_PRIME_MAX; ++t) {
            if (t < P) {
                float2 v = base[(j + t * M) * es];
                if (t > 0 && j > 0) {
                    float s, c;
                    sincospif(sign * 2.0f * (float)(j * t) / (float)L, &s, &c);
                    v = cmul(v, make_float2(c, s));
                }
                a[t] = v;
            }
        }
        for (int q = 0; q < P; ++q) {
            float2 y = a[0];
            int m = 0;
#pragma unroll
            for (int t = 1; t < FLF_PRIME_MAX; ++t) {
                if (t < P) {
                    m += q;
                    if (m >= P) m -= P;
                    float s, c;
                    sincospif(sign * 2.0f * (float)m / (float)P, &s, &c);
                    y = cadd(y, cmul(a[t], make_float2(c, s)));
                }
            }
            base[(j + q * M) * es] = y;
        }
    }
    __syncthreads();
}

// Full transform: sign = -1 forward, +1 inverse (unnormalised).  PRIME: the
// plan may end with an odd prime radix above 5.
template <int MAXB, bool PRIME = false>
__device__ void lds_fft(float2* buf, const RadixPlan& rp, int nbatch, int es, int bs, float sign) {
    int Ns = 1;
    for (int p = 0; p < rp.n; ++p) {
        int R = rp.r[p];
        switch (R) {
            case 4: stockham_pass<4, MAXB>(buf, rp.L, nbatch, es, bs, Ns, sign); break;
            case 2: stockham_pass<2, MAXB>(buf, rp.L, nbatch, es, bs, Ns, sign); break;
            case 3: stockham_pass<3, MAXB>(buf, rp.L, nbatch, es, bs, Ns, sign); break;
            case 5: stockham_pass<5, MAXB>(buf, rp.L, nbatch, es, bs, Ns, sign); break;
            default:
                if constexpr (PRIME) stockham_pass_prime(buf, rp.L, R, nbatch, es, bs, sign);   // last pass
                break;
        }
        Ns *= R;
    }
}

}  // namespace flf
