// Wave-level mixed-radix (2/3/4/5) FFTs in LDS.
//
// One 64-lane wave transforms one sequence held in LDS (element stride es).
// Stockham passes as in flf_fft.h, but each pass is executed by the wave
// alone: all lanes read their butterflies' inputs into registers, the wave
// synchronises (LDS operations of one wave complete in order; the barrier
// stops the compiler from moving accesses across it), then writes the
// outputs in place.  Twiddles come from a table tw[m] = exp(-2 pi i m / L)
// (computed on the host in double precision), so no transcendental is
// evaluated per butterfly.
#pragma once
#include "flf_common.h"
#include "flf_fft.h"

namespace flf {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// q = n / d and r = n % d for 0 <= n < 2^22 and d >= 1, from the float
// reciprocal rd = 1/d of the pass (one correction step makes it exact:
// n * rd is within 0.5 of n / d).
__device__ __forceinline__ int div_small(int n, int d, float rd, int& r) {
    int q = (int)((float)n * rd);
    r = n - q * d;
    if (r < 0) { --q; r += d; } else if (r >= d) { ++q; r -= d; }
    return q;
}

// a[t] *= w^t for t = 1..R-1, w = tw[k * tstep] (conjugated for the inverse):
// one table read, the higher powers as running products (at most R - 2
// roundings, a few ulp), two values live.
template <int R>
__device__ __forceinline__ void twiddle_powers(float2* a, const float2* __restrict__ tw, int idx, bool inv) {
    float2 w1 = tw[idx];
    if (inv) w1.y = -w1.y;
    float2 w = w1;
#pragma unroll
    for (int t = 1; t < R; ++t) {
        a[t] = cmul(a[t], w);
        if (t + 1 < R) w = cmul(w, w1);
    }
}

template <int R, int LMAX>
__device__ __forceinline__ void wave_pass(float2* buf, int es, int L, int Ns, const float2* __restrict__ tw,
                                          bool inv, int lane) {
    constexpr int MAXQ = (LMAX / R + 63) / 64;
    const int M = L / R;
    const int tstep = L / (Ns * R);
    const float rNs = 1.0f / (float)Ns;
    const float sign = inv ? 1.f : -1.f;
    float2 v[MAXQ][R];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int j = lane + 64 * q;
        if (j < M) {
            int k;
            div_small(j, Ns, rNs, k);
            float2 a[FLF_RADIX_MAX];
#pragma unroll
            for (int t = 0; t < R; ++t) a[t] = buf[(j + t * M) * es];
            if (k != 0) twiddle_powers<R>(a, tw, k * tstep, inv);
            dft_small<R>(a, sign);
#pragma unroll
            for (int t = 0; t < R; ++t) v[q][t] = a[t];
        }
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int j = lane + 64 * q;
        if (j < M) {
            int k;
            const int o = div_small(j, Ns, rNs, k) * Ns * R + k;   // recomputed: fewer live registers
#pragma unroll
            for (int t = 0; t < R; ++t) buf[(o + t * Ns) * es] = v[q][t];
        }
    }
    wave_sync();
}

// Last pass of an odd prime radix P (5 < P <= FLF_PRIME_MAX; Ns = L / P):
// butterfly j < M = L/P reads x[j + t M], t < P, and writes y_q to
// x[j + q M]: the same positions, so each lane takes its butterflies one at
// a time.  y_q = sum_t x[j + t M] W^(j t) W^(M t q mod L), W = tw[1], the
// table tw[m] = exp(-2 pi i m / L) (conjugated for the inverse).  Both
// twiddle sequences are running products of one table entry, re-read from
// the table every 8 steps (at most 7 roundings), so that only a few table
// reads are live at a time.
__device__ __forceinline__ void wave_pass_prime(float2* buf, int es, int L, int P, const float2* __restrict__ tw,
                                                bool inv, int lane) {
    const int M = L / P;
    for (int j = lane; j < M; j += 64) {
        float2 a[FLF_PRIME_MAX];
        a[0] = buf[j * es];
        {
            float2 w1 = tw[j];                                 // W^j
            if (inv) w1.y = -w1.y;
            float2 w = w1;
#pragma unroll
            for (int t = 1; t < FLF_PRIME_MAX; ++t) {
                if (t < P) {
                    if ((t & 7) == 0) {                        // exact table value W^(j t), j t < L
                        w = tw[j * t];
                        if (inv) w.y = -w.y;
                    }
                    a[t] = cmul(buf[(j + t * M) * es], w);
                    w = cmul(w, w1);
                }
            }
        }
        for (int q = 0; q < P; ++q) {
            float2 wq = tw[q * M];                             // W_P^q
            if (inv) wq.y = -wq.y;
            float2 y = a[0], w = wq;
            int m = q;                                         // t q mod P
#pragma unroll
            for (int t = 1; t < FLF_PRIME_MAX; ++t) {
                if (t < P) {
                    if ((t & 7) == 0) {
                        w = tw[m * M];
                        if (inv) w.y = -w.y;
                    }
                    y = cadd(y, cmul(a[t], w));
                    w = cmul(w, wq);
                    m += q;
                    if (m >= P) m -= P;
                }
            }
            buf[(j + q * M) * es] = y;
        }
    }
    wave_sync();
}

// In-place transform of one sequence by one wave (inv: unnormalised inverse).
// PRIME: the plan may end with an odd prime radix above 5 (kernels for lenslet
// pitches such as 7, 13, 19 are compiled separately, so that the prime pass's
// registers do not lower the occupancy of the 5-smooth kernels).
template <int LMAX, bool PRIME = false>
__device__ void wave_fft(float2* buf, int es, const RadixPlan& rp, const float2* __restrict__ tw, bool inv,
                         int lane) {
    int Ns = 1;
    for (int p = 0; p < rp.n; ++p) {
        const int R = rp.r[p];
        switch (R) {
            case 8: wave_pass<8, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            case 6: wave_pass<6, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            case 4: wave_pass<4, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            case 2: wave_pass<2, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            case 3: wave_pass<3, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            case 5: wave_pass<5, LMAX>(buf, es, rp.L, Ns, tw, inv, lane); break;
            default:
                if constexpr (PRIME) wave_pass_prime(buf, es, rp.L, R, tw, inv, lane);   // last pass
                break;
        }
        Ns *= R;
    }
}

// Out-of-place forms for transforms longer than the register-held passes
// cover (grids above 2048 points): each pass reads one LDS buffer and writes
// the other, and a lane takes its butterflies one at a time, so no register
// array grows with the length.  Contiguous sequences (element stride 1).
template <int R>
__device__ __forceinline__ void wave_pass_oop(const float2* src, float2* dst, int L, int Ns,
                                              const float2* __restrict__ tw, bool inv, int lane) {
    const int M = L / R;
    const int tstep = L / (Ns * R);
    const float rNs = 1.0f / (float)Ns;
    const float sign = inv ? 1.f : -1.f;
    for (int j = lane; j < M; j += 64) {
        int k;
        const int g = div_small(j, Ns, rNs, k);
        float2 a[FLF_RADIX_MAX];
#pragma unroll
        for (int t = 0; t < R; ++t) a[t] = src[j + t * M];
        if (k != 0) twiddle_powers<R>(a, tw, k * tstep, inv);
        dft_small<R>(a, sign);
        const int o = g * Ns * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) dst[o + t * Ns] = a[t];
    }
    wave_sync();
}

// The prime last pass (Ns = L / P) out of place: y_q of butterfly j goes to
// dst[j + q M], the positions its inputs held in src.
__device__ __forceinline__ void wave_pass_prime_oop(const float2* src, float2* dst, int L, int P,
                                                    const float2* __restrict__ tw, bool inv, int lane) {
    const int M = L / P;
    for (int j = lane; j < M; j += 64) {
        float2 a[FLF_PRIME_MAX];
        a[0] = src[j];
        {
            float2 w1 = tw[j];
            if (inv) w1.y = -w1.y;
            float2 w = w1;
#pragma unroll
            for (int t = 1; t < FLF_PRIME_MAX; ++t) {
                if (t < P) {
                    if ((t & 7) == 0) {
                        w = tw[j * t];
                        if (inv) w.y = -w.y;
                    }
                    a[t] = cmul(src[j + t * M], w);
                    w = cmul(w, w1);
                }
            }
        }
        for (int q = 0; q < P; ++q) {
            float2 wq = tw[q * M];
            if (inv) wq.y = -wq.y;
            float2 y = a[0], w = wq;
            int m = q;
#pragma unroll
            for (int t = 1; t < FLF_PRIME_MAX; ++t) {
                if (t < P) {
                    if ((t & 7) == 0) {
                        w = tw[m * M];
                        if (inv) w.y = -w.y;
                    }
                    y = cadd(y, cmul(a[t], w));
                    w = cmul(w, wq);
                    m += q;
                    if (m >= P) m -= P;
                }
            }
            dst[j + q * M] = y;
        }
    }
    wave_sync();
}

// Whole transform between buffers a and b (length rp.L each); returns the
// buffer that holds the result (a after an even number of passes).
template <bool PRIME>
__device__ float2* wave_fft_pp(float2* a, float2* b, const RadixPlan& rp, const float2* __restrict__ tw, bool inv,
                               int lane) {
    int Ns = 1;
    float2* src = a;
    float2* dst = b;
    for (int p = 0; p < rp.n; ++p) {
        const int R = rp.r[p];
        switch (R) {
            case 8: wave_pass_oop<8>(src, dst, rp.L, Ns, tw, inv, lane); break;
            case 6: wave_pass_oop<6>(src, dst, rp.L, Ns, tw, inv, lane); break;
            case 4: wave_pass_oop<4>(src, dst, rp.L, Ns, tw, inv, lane); break;
            case 2: wave_pass_oop<2>(src, dst, rp.L, Ns, tw, inv, lane); break;
            case 3: wave_pass_oop<3>(src, dst, rp.L, Ns, tw, inv, lane); break;
            case 5: wave_pass_oop<5>(src, dst, rp.L, Ns, tw, inv, lane); break;
            default:
                if constexpr (PRIME) wave_pass_prime_oop(src, dst, rp.L, R, tw, inv, lane);   // last pass
                break;
        }
        Ns *= R;
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// NS sequences per call (sequence s at buf + s * ss, element stride es),
// nseq <= NS of them present: the butterflies of all sequences share the
// wave's lanes, which a single short sequence fills only partly.
template <int R, int LMAX, int NS>
__device__ __forceinline__ void wave_pass_ns(float2* buf, int es, int ss, int nseq, int L, int Ns,
                                             const float2* __restrict__ tw, bool inv, int lane) {
    constexpr int MAXQ = (NS * (LMAX / R) + 63) / 64;
    const int M = L / R;
    const int tot = nseq * M;
    const int tstep = L / (Ns * R);
    const float rNs = 1.0f / (float)Ns, rM = 1.0f / (float)M;
    const float sign = inv ? 1.f : -1.f;
    float2 v[MAXQ][R];
    int ov[MAXQ];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int j = lane + 64 * q;
        if (j < tot) {
            int jj, k;
            const int sq = div_small(j, M, rM, jj);
            const int g = div_small(jj, Ns, rNs, k);
            ov[q] = sq * ss + (g * Ns * R + k) * es;
            const float2* b = buf + sq * ss;
            float2 a[FLF_RADIX_MAX];
#pragma unroll
            for (int t = 0; t < R; ++t) a[t] = b[(jj + t * M) * es];
            if (k != 0) twiddle_powers<R>(a, tw, k * tstep, inv);
            dft_small<R>(a, sign);
#pragma unroll
            for (int t = 0; t < R; ++t) v[q][t] = a[t];
        }
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int j = lane + 64 * q;
        if (j < tot) {
            const int o = ov[q];
#pragma unroll
            for (int t = 0; t < R; ++t) buf[o + t * Ns * es] = v[q][t];
        }
    }
    wave_sync();
}

// Radices 2..5 only (plans from make_radix_plan without the wide set): the
// composite radices' code raised the small FFT's registers and slowed it.
template <int LMAX, int NS>
__device__ void wave_fft_ns(float2* buf, int es, int ss, int nseq, const RadixPlan& rp, const float2* __restrict__ tw,
                            bool inv, int lane) {
    int Ns = 1;
    for (int p = 0; p < rp.n; ++p) {
        const int R = rp.r[p];
        switch (R) {
            case 4: wave_pass_ns<4, LMAX, NS>(buf, es, ss, nseq, rp.L, Ns, tw, inv, lane); break;
            case 2: wave_pass_ns<2, LMAX, NS>(buf, es, ss, nseq, rp.L, Ns, tw, inv, lane); break;
            case 3: wave_pass_ns<3, LMAX, NS>(buf, es, ss, nseq, rp.L, Ns, tw, inv, lane); break;
            default: wave_pass_ns<5, LMAX, NS>(buf, es, ss, nseq, rp.L, Ns, tw, inv, lane); break;
        }
        Ns *= R;
    }
}

}  // namespace flf
