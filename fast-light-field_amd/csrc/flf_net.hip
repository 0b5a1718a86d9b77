// Learned initialiser: a lenslet-aware convolutional network that maps a
// rectified light-field image to a first guess of the volume for the
// Richardson-Lucy loop.  The reference has no counterpart; RL from a uniform
// start (lfdeconv.py:97-99 with Xguess=None) stays the fallback.
//
// Architecture (all 3x3 convolutions, zero padding 1, at lenslet resolution):
//   x   = pixel-unshuffle(sqrt(image / mean(image))) N*N channels (padded to 16k);
//         the square root stabilises the Poisson variance and keeps x of order 1
//   h0  = relu(conv0(x))                             width channels
//   h_i = h_{i-1} + conv_b(relu(conv_a(h_{i-1})))    i = 1..nblocks
//   o   = conv_head(h_n)                             nz channels (padded)
//   y   = softplus(o) * mean(image) / nz             per plane, per lenslet
// (a PSF plane sums to 1, so mean(image) / nz is the mean voxel value and the
// normalised output is of order 1).
// The volume guess is y upsampled (nearest) to full resolution.
//
// Arithmetic: one BF16 MFMA GEMM kernel (v_mfma_f32_32x32x16_bf16, FP32
// accumulation) C = A * B^T with both operands K-contiguous.  Forward and
// dgrad read the activation itself as an implicit im2col (each K step of 16
// channels lies in one tap: rows shifted by the tap, zero outside the frame);
// wgrad is fed by a transposed im2col and a transposition:
//   forward  out[p, co]   = sum_k im2col(x)[p, k] * W[co, k]          k = (tap, ci)
//   dgrad    dx[p, ci]    = sum_k im2col_flip(dy)[p, k] * Wt[ci, k]   k = (tap, co)
//   wgrad    dW[co, k]    = sum_p dy^T[co, p] * im2col^T(x)[k, p]     (split over p)
// Master weights and Adam state are FP32; activations and GEMM operands BF16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/flf_hip.h"
#include "flf_devptr.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ bf16 to_bf16(float x) { return (bf16)x; }
__device__ __forceinline__ float from_bf16(bf16 x) { return (float)x; }

// ---------------------------------------------------------------------------
// C[M x N] (+)= A[M x K] * B[N x K]^T, bf16 operands, fp32 accumulation.
// Wave tile 32 x 64 (two 32x32x16 MFMAs sharing the A fragment); workgroup
// 4 waves = 64 x 128.  Operand layout per lane l (32x32x16 bf16):
//   A[i = l % 32][k = 8 (l / 32) + e]     B[k = 8 (l / 32) + e][j = l % 32]
//   D[i = 8 (r / 4) + 4 (l / 32) + r % 4][j = l % 32]   r = 0..15
// K must be a multiple of 16.  grid (ceil(N/128), ceil(M/64), splitK).
// ---------------------------------------------------------------------------
struct GemmArgs {
    // conv != 0: A is not a matrix but an NHWC activation (convN frames of
    // convH x convW positions, row stride lda, convC channels, convC a
    // multiple of 16, channels past the real ones zero) read as its implicit
    // 3x3 pad-1 im2col: A[p][tap * convC + c] = x[shift(p, tap)][c], zero
    // outside the frame; conv == 2 mirrors the taps (dgrad).
    int conv, convH, convW, convC;
    const bf16* A; long long lda;
    const bf16* B; long long ldb;
    int M, N, K;
    int kchunk;                 // K per split (multiple of 16)
    // epilogue
    const float* bias;          // [N] or null
    const bf16* resid;          // [M][ldr] or null (added before the store)
    long long ldr;
    int relu;
    bf16* Cb; long long ldcb;   // bf16 output or null
    float* Cf; long long ldcf;  // fp32 output or null (split z writes at Cf + z * M * ldcf)
};

__device__ __forceinline__ bf16x8 load8(const bf16* p, bool ok) {
    if (ok) return *reinterpret_cast<const bf16x8*>(p);
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e) z[e] = (bf16)0.0f;
    return z;
}

// 16 bytes at p, or zeros: the load always happens (from a valid address) and
// invalid lanes are cleared by a select, so no branch surrounds it.
__device__ __forceinline__ bf16x8 load8_sel(const bf16* p, const bf16* valid_base, bool ok) {
    const bf16x8 v = *reinterpret_cast<const bf16x8*>(ok ? p : valid_base);
    bf16x8 z;
#pragma unroll
    for (int e = 0; e < 8; ++e) z[e] = ok ? v[e] : (bf16)0.0f;
    return z;
}

// C[M x N] (+)= A[M x K] * B[N x K]^T through LDS: workgroup tile BM x 128
// (four waves, 2 x 2, each BM/2 x 64 of 32x32 MFMA tiles), K in steps of 32
// staged in LDS (two buffers).  Rows are padded to 40 bf16 (80 bytes), so the
// 16-byte fragment reads of 16 consecutive rows fall on distinct banks.
// Global loads run PF steps ahead in a register ring: a step's loads are
// issued PF - 1 steps before its LDS store.  Loads are unconditional (masked
// by selects, clamped addresses) and the step count is padded to a multiple
// of PF with zero steps, so the unrolled loop has no branches and the
// compiler's waits cover only the step being stored.  Operand layout per
// lane l (32x32x16 bf16):
//   A[i = l % 32][k = 8 (l / 32) + e]     B[k = 8 (l / 32) + e][j = l % 32]
//   D[i = 8 (r / 4) + 4 (l / 32) + r % 4][j = l % 32]   r = 0..15
// K a multiple of 16 (a step may end in zeros past ke).  grid (ceil(N/128),
// ceil(M/BM), splitK).
template <int CONV, int BM>
__global__ __launch_bounds__(256) void k_gemm_nt(GemmArgs g) {
    constexpr int BN = 128, BK = 32, LDK = BK + 8, PF = 4;
    constexpr int WM = BM / 2, MT = WM / 32, NT = 2;
    constexpr int ACH = BM * BK / 8 / 256, BCH = BN * BK / 8 / 256;   // 16-byte chunks per thread and step
    __shared__ __align__(16) bf16 As[2][BM][LDK];
    __shared__ __align__(16) bf16 Bs[2][BN][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int mb = blockIdx.y * BM, nb = blockIdx.x * BN;
    const int wm0 = (wv >> 1) * WM, wn0 = (wv & 1) * 64;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);
    const int nsteps = ke > kb ? (ke - kb + BK - 1) / BK : 0;
    const int npad = (nsteps + PF - 1) / PF * PF;
    // this thread's load chunks: row c / 4, k offset 8 (c % 4) of the step
    int arow[ACH], akc[ACH], ay[ACH], ax[ACH];
#pragma unroll
    for (int h = 0; h < ACH; ++h) {
        const int c = tid + 256 * h;
        arow[h] = c >> 2;
        akc[h] = (c & 3) * 8;
        const int am = mb + arow[h];
        ax[h] = CONV ? am % g.convW : 0;
        ay[h] = CONV ? (am / g.convW) % g.convH : 0;
    }
    int brow[BCH], bkc[BCH];
#pragma unroll
    for (int h = 0; h < BCH; ++h) {
        const int c = tid + 256 * h;
        brow[h] = c >> 2;
        bkc[h] = (c & 3) * 8;
    }
    bf16x8 ra[PF][ACH], rb[PF][BCH];
    auto gload = [&](int slot, int st) {      // step st (past the last: zeros)
        const int k0 = kb + st * BK;
#pragma unroll
        for (int h = 0; h < ACH; ++h) {
            const int am = mb + arow[h], k = k0 + akc[h];
            bool ok = am < g.M && k < ke;
            const bf16* src;
            if constexpr (CONV == 0) {
                src = g.A + (long long)am * g.lda + k;
            } else {
                const int tap = k / g.convC, c = k - tap * g.convC;
                const int tt = CONV == 2 ? 8 - tap : tap;
                const int dy = tt / 3 - 1, dx = tt - 3 * (tt / 3) - 1;
                ok = ok && (unsigned)(ay[h] + dy) < (unsigned)g.convH && (unsigned)(ax[h] + dx) < (unsigned)g.convW;
                src = g.A + ((long long)am + dy * g.convW + dx) * g.lda + c;
            }
            ra[slot][h] = load8_sel(src, g.A, ok);
        }
#pragma unroll
        for (int h = 0; h < BCH; ++h) {
            const int bn = nb + brow[h], k = k0 + bkc[h];
            const bool ok = bn < g.N && k < ke;
            rb[slot][h] = load8_sel(g.B + (long long)bn * g.ldb + k, g.B, ok);
        }
    };
    auto lstore = [&](int slot, int buf) {
#pragma unroll
        for (int h = 0; h < ACH; ++h) *reinterpret_cast<bf16x8*>(&As[buf][arow[h]][akc[h]]) = ra[slot][h];
#pragma unroll
        for (int h = 0; h < BCH; ++h) *reinterpret_cast<bf16x8*>(&Bs[buf][brow[h]][bkc[h]]) = rb[slot][h];
    };
    f32x16 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x16{};
    const int li = lane & 31, lk = (lane >> 5) * 8;
    if (npad > 0) {
#pragma unroll
        for (int d = 0; d < PF; ++d) gload(d, d);
        // ring rule: once slot s (step t) is stored, it takes step t + PF
        lstore(0, 0);
        gload(0, PF);
        __syncthreads();
        for (int st0 = 0; st0 < npad; st0 += PF) {
#pragma unroll
            for (int d = 0; d < PF; ++d) {
                const int cur = d & 1;            // PF even: buffer of step st0 + d
#pragma unroll
                for (int kk = 0; kk < BK / 16; ++kk) {
                    bf16x8 af[MT], bfr[NT];
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        af[mt] = *reinterpret_cast<const bf16x8*>(&As[cur][wm0 + mt * 32 + li][kk * 16 + lk]);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        bfr[nt] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn0 + nt * 32 + li][kk * 16 + lk]);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int nt = 0; nt < NT; ++nt)
                            acc[mt][nt] =
                                __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[mt], bfr[nt], acc[mt][nt], 0, 0, 0);
                }
                // step st0 + d + 1 (loaded PF - 1 steps ago) into the other
                // buffer, last read one step ago; then its ring slot takes
                // step st0 + d + 1 + PF
                lstore((d + 1) % PF, cur ^ 1);
                gload((d + 1) % PF, st0 + d + 1 + PF);
                __syncthreads();
            }
        }
    }
    // epilogue
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const f32x16& c = acc[mt][nt];
            const int j = nb + wn0 + nt * 32 + li;
            if (j >= g.N) continue;
            const float bj = g.bias ? g.bias[j] : 0.f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = mb + wm0 + mt * 32 + 8 * (r / 4) + 4 * (lane >> 5) + (r & 3);
                if (i >= g.M) continue;
                float v = c[r] + bj;
                if (g.relu) v = fmaxf(v, 0.f);
                if (g.resid) v += from_bf16(g.resid[(long long)i * g.ldr + j]);
                if (g.Cf) g.Cf[(long long)blockIdx.z * g.M * g.ldcf + (long long)i * g.ldcf + j] = v;
                if (g.Cb) g.Cb[(long long)i * g.ldcb + j] = to_bf16(v);
            }
        }
}

// Transposed im2col for wgrad: out[tap*Cp + c][p] (row length ldo >= P, zero
// padded to a multiple of 16), from NHWC x (ldx a multiple of 8; channels
// c >= C read as zero).  grid (ceil(ldo / 64), ceil(Cp / 64)), block 256: a
// tile of 64 positions x 64 channels goes through LDS once per tap, read as
// 16-byte channel runs of x and written as 16-byte position runs of out.
__global__ __launch_bounds__(256) void k_im2col_t(const bf16* x, int n, int H, int W, int C, int ldx, int Cp, long long ldo,
                                                  bf16* out) {
    __shared__ __align__(16) bf16 tile[64][72];   // [channel][position], rows 144 B apart
    const long long P = (long long)n * H * W;
    const long long p0 = (long long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int t = threadIdx.x;
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        for (int e = t; e < 512; e += 256) {
            const int pl = e >> 3, ch = (e & 7) * 8;
            const long long p = p0 + pl;
            bf16x8 v;
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (bf16)0.0f;
            if (p < P && c0 + ch < C) {
                const int xx = (int)(p % W), yy = (int)((p / W) % H);
                if ((unsigned)(yy + dy) < (unsigned)H && (unsigned)(xx + dx) < (unsigned)W) {
                    const bf16* src = x + (p + dy * W + dx) * ldx + c0 + ch;
                    if (c0 + ch + 8 <= C) {
                        v = *reinterpret_cast<const bf16x8*>(src);
                    } else {
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                            if (c0 + ch + k < C) v[k] = src[k];
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) tile[ch + k][pl] = v[k];
        }
        __syncthreads();
        for (int e = t; e < 512; e += 256) {
            const int cl = e >> 3, pc = (e & 7) * 8;
            const int c = c0 + cl;
            const long long p = p0 + pc;
            if (c < Cp && p < ldo)
                *reinterpret_cast<bf16x8*>(out + ((long long)tap * Cp + c) * ldo + p) =
                    *reinterpret_cast<const bf16x8*>(&tile[cl][pc]);
        }
        __syncthreads();
    }
}

// y[c][p] = x[p][c] for c < C, p < P; zero padding up to ldy.
__global__ void k_transpose(const bf16* x, long long P, int C, int ldx, long long ldy, int Crows, bf16* y) {
    const long long total = (long long)Crows * ldy;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long p = idx % ldy;
        const int c = (int)(idx / ldy);
        y[idx] = (p < P && c < C) ? x[p * ldx + c] : (bf16)0.0f;
    }
}

// Pixel-unshuffle with per-frame normalisation:
// x[(t*Ly + ly)*Lx + lx][b*N + a] = sqrt(max(img[t][ly*N + b][lx*N + a], 0) / scale[t]).
__global__ void k_unshuffle(const float* img, int T, int Ly, int Lx, int N, const float* scale, int Cp, bf16* x) {
    const int NN = N * N;
    const long long total = (long long)T * Ly * Lx * Cp;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int ch = (int)(idx % Cp);
        const long long p = idx / Cp;
        const int lx = (int)(p % Lx), ly = (int)((p / Lx) % Ly);
        const int t = (int)(p / ((long long)Lx * Ly));
        float v = 0.f;
        if (ch < NN) {
            const int b = ch / N, a = ch % N;
            v = sqrtf(fmaxf(img[((long long)t * Ly * N + ly * N + b) * ((long long)Lx * N) + lx * N + a], 0.f) /
                      scale[t]);
        }
        x[idx] = to_bf16(v);
    }
}

// Per-frame mean of the image: FP64 partial sums over MEAN_BLOCKS blocks per
// frame (grid (MEAN_BLOCKS, T)), then one block per frame adds them in a
// fixed order (deterministic).
constexpr int MEAN_BLOCKS = 128;
__global__ __launch_bounds__(256) void k_frame_sum_partial(const float* img, long long n, double* part) {
    __shared__ double sh[256];
    const float* a = img + (long long)blockIdx.y * n;
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        s += a[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)blockIdx.y * gridDim.x + blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(MEAN_BLOCKS) void k_frame_mean_final(const double* part, long long n, float* out) {
    __shared__ double sh[MEAN_BLOCKS];
    sh[threadIdx.x] = part[(long long)blockIdx.x * MEAN_BLOCKS + threadIdx.x];
    __syncthreads();
    for (int k = MEAN_BLOCKS / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = (float)fmax(sh[0] / (double)n, 1e-12);
}

__device__ __forceinline__ float softplusf(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoidf(float x) { return 1.f / (1.f + expf(-x)); }

// Head output -> volume guess at lenslet resolution, out[t][z][ly][lx].
__global__ void k_head_out(const float* o, int P, int ldo, int nz, int Ly, int Lx, const float* scale, float oscale,
                           float* out) {
    const long long total = (long long)P * nz;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(idx % nz);
        const long long p = idx / nz;
        const int lx = (int)(p % Lx), ly = (int)((p / Lx) % Ly);
        const int t = (int)(p / ((long long)Lx * Ly));
        out[(((long long)t * nz + z) * Ly + ly) * Lx + lx] = softplusf(o[p * ldo + z]) * (scale[t] * oscale);
    }
}

// MSE loss on the normalised output and its gradient w.r.t. the head
// pre-activation: target tgt[t][z][ly][lx] (volume units).
__global__ __launch_bounds__(256) void k_loss(const float* o, int P, int ldo, int nz, int Ly, int Lx,
                                              const float* scale, float oscale, const float* tgt, bf16* dO, int ldd,
                                              double* part) {
    __shared__ double sh[256];
    double acc = 0.0;
    const long long total = (long long)P * ldd;
    const double inv = 1.0 / ((double)P * nz);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(idx % ldd);
        const long long p = idx / ldd;
        float gsc = 0.f;
        if (z < nz) {
            const int lx = (int)(p % Lx), ly = (int)((p / Lx) % Ly);
            const int t = (int)(p / ((long long)Lx * Ly));
            const float x = o[p * ldo + z];
            const float y = softplusf(x);
            const float tv = tgt[(((long long)t * nz + z) * Ly + ly) * Lx + lx] / (scale[t] * oscale);
            const float d = y - tv;
            acc += (double)d * d;
            gsc = (float)(2.0 * inv) * d * sigmoidf(x);
        }
        dO[p * ldd + z] = to_bf16(gsc);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] * inv;
}

__global__ void k_sum_parts(const double* part, int n, double* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        *out = s;
    }
}

// dx *= (act > 0)  (ReLU backward on the stored post-activation).
__global__ void k_relu_bwd(bf16* dx, const bf16* act, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!(from_bf16(act[i]) > 0.f)) dx[i] = (bf16)0.0f;
}

// Sum split-K partials: out[i] = sum_z part[z][i] (fixed order).
__global__ void k_reduce_splits(const float* part, int splits, long long n, float* out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < splits; ++z) s += part[(long long)z * n + i];
        out[i] = s;
    }
}
// db[c] = sum_p dyt[c][p] over the transposed gradient (rows of ldt >= P
// entries, zero past P, ldt a multiple of 8): contiguous 16-byte loads.
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* dyt, long long ldt, float* db) {
    __shared__ float sh[256];
    const int c = blockIdx.x;
    const bf16* row = dyt + (long long)c * ldt;
    float s = 0.f;
    for (long long p = 8LL * threadIdx.x; p < ldt; p += 8LL * blockDim.x) {
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(row + p);
#pragma unroll
        for (int e = 0; e < 8; ++e) s += from_bf16(v[e]);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = sh[0];
}

// Adam on fp32 master weights (the bf16 copies are refreshed afterwards).
__global__ void k_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}
__global__ void k_weights_bf16(const float* w, int Co, int Ci, int Cip, int Cop, bf16* Wb, bf16* Wt) {
    // w: [Co][9][Ci] fp32 (unpadded);  Wb: [Co][9][Cip];  Wt: [Ci][9][Cop], the
    // dgrad operand (the tap mirroring is done by the dgrad's mirrored implicit im2col).
    const long long total = (long long)Co * 9 * Ci;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int ci = (int)(idx % Ci);
        const int tap = (int)((idx / Ci) % 9);
        const int co = (int)(idx / (9LL * Ci));
        const bf16 b = to_bf16(w[idx]);
        Wb[((long long)co * 9 + tap) * Cip + ci] = b;
        Wt[((long long)ci * 9 + tap) * Cop + co] = b;
    }
}

// Nearest upsampling of the lenslet-resolution guess to the volume:
// vol[z][t][y][x] = g[t][z][y / N][x / N].
// One workgroup per output row (z, t, y): 32-bit index math, coalesced
// row writes, the row's lenslet values read through the cache.
__global__ void k_upsample(const float* g, int T, int nz, int Ly, int Lx, int N, float* vol) {
    const int Y = Ly * N, X = Lx * N;
    const int rows = nz * T * Y;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int y = r % Y, zt = r / Y;
        const int t = zt % T, z = zt / T;
        const float* gr = g + (((long long)t * nz + z) * Ly + y / N) * Lx;
        float* out = vol + (long long)r * X;
        for (int x = threadIdx.x; x < X; x += blockDim.x) out[x] = gr[x / N];
    }
}

__global__ void k_fill0(float* p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

int blocks_for(long long n) { return (int)std::max<long long>(1, std::min<long long>(65536, (n + 255) / 256)); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct NetLayer {
    int Ci, Co, Cip, Cop;        // channels and their padded sizes (multiples of 16)
    long long woff, boff;        // offsets into the flat parameter vector (W [Co][9][Ci], b [Co])
    bf16 *Wb = nullptr, *Wt = nullptr;
};

struct flf_net {
    int device = 0;
    std::string err;
    int N = 15, nz = 0, width = 128, nblocks = 4;
    std::vector<NetLayer> L;
    long long nparam = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int step = 0;
    // activation buffers (sized for the largest batch seen)
    long long cap_P = 0;
    bf16 *x = nullptr, *colt = nullptr, *dyt = nullptr, *dbuf0 = nullptr, *dbuf1 = nullptr;
    std::vector<bf16*> h, t;     // h[0..nblocks], t[1..nblocks] (t[0] unused)
    float *o = nullptr, *scale = nullptr, *wpart = nullptr;
    double *lpart = nullptr, *lout = nullptr, *mpart = nullptr;
    bf16* dO = nullptr;
    std::vector<void*> acts;
};

namespace {

int nfail(flf_net* n, int code, const std::string& msg) {
    if (n) n->err = msg;
    return code;
}

#define NCHK(call)                                                                             \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return nfail(net, FLF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int pad16(int c) { return (c + 15) / 16 * 16; }

template <class T>
int nalloc(flf_net* net, T** p, size_t count, std::vector<void*>& list) {
    if (hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return nfail(net, FLF_ENOMEM, "hipMalloc failed in the network");
    }
    list.push_back((void*)*p);
    return FLF_OK;
}

void free_acts(flf_net* net) {
    for (void* p : net->acts) (void)hipFree(p);
    net->acts.clear();
    net->h.clear();
    net->t.clear();
    net->cap_P = 0;
}

int ensure_acts(flf_net* net, long long P) {
    if (P <= net->cap_P) return FLF_OK;
    free_acts(net);
    const int Cin = net->L[0].Cip, Wd = net->width, Cop = net->L.back().Cop;
    const int maxC = std::max(std::max(Cin, Wd), Cop);
    const long long Pp = (P + 15) / 16 * 16;
    auto& A = net->acts;
    int rc;
    if ((rc = nalloc(net, &net->x, (size_t)P * Cin, A))) return rc;
    if ((rc = nalloc(net, &net->colt, (size_t)9 * maxC * Pp, A))) return rc;
    if ((rc = nalloc(net, &net->dyt, (size_t)std::max(Wd, Cop) * Pp, A))) return rc;
    if ((rc = nalloc(net, &net->dbuf0, (size_t)P * std::max(Wd, Cop), A))) return rc;
    if ((rc = nalloc(net, &net->dbuf1, (size_t)P * std::max(Wd, Cop), A))) return rc;
    net->h.assign(net->nblocks + 1, nullptr);
    net->t.assign(net->nblocks + 1, nullptr);
    for (int i = 0; i <= net->nblocks; ++i)
        if ((rc = nalloc(net, &net->h[i], (size_t)P * Wd, A))) return rc;
    for (int i = 1; i <= net->nblocks; ++i)
        if ((rc = nalloc(net, &net->t[i], (size_t)P * Wd, A))) return rc;
    if ((rc = nalloc(net, &net->o, (size_t)P * Cop, A))) return rc;
    if ((rc = nalloc(net, &net->dO, (size_t)P * Cop, A))) return rc;
    if ((rc = nalloc(net, &net->scale, (size_t)1024, A))) return rc;
    if ((rc = nalloc(net, &net->lpart, (size_t)1024, A))) return rc;
    if ((rc = nalloc(net, &net->mpart, (size_t)MEAN_BLOCKS * 1024, A))) return rc;
    if ((rc = nalloc(net, &net->lout, (size_t)1, A))) return rc;
    // split-K partials for wgrad: splits x Co x 9*Ci
    if ((rc = nalloc(net, &net->wpart, (size_t)33 * 9 * maxC * std::max(Wd, Cop), A))) return rc;
    net->cap_P = P;
    return FLF_OK;
}

// Workgroup tile rows: 128 when that still gives every CU a workgroup,
// else 64 (FLF_GEMM_BM overrides).
int gemm_bm(const GemmArgs& a, int splits) {
    static const int env = std::getenv("FLF_GEMM_BM") ? std::atoi(std::getenv("FLF_GEMM_BM")) : 0;
    if (env == 64 || env == 128) return env;
    const long long wg128 = (long long)((a.N + 127) / 128) * ((a.M + 127) / 128) * splits;
    return wg128 >= 256 ? 128 : 64;
}

template <int CONV>
void launch_gemm(const GemmArgs& a, int splits, hipStream_t s) {
    if (gemm_bm(a, splits) == 128)
        hipLaunchKernelGGL((k_gemm_nt<CONV, 128>), dim3((a.N + 127) / 128, (a.M + 127) / 128, splits), dim3(256), 0, s,
                           a);
    else
        hipLaunchKernelGGL((k_gemm_nt<CONV, 64>), dim3((a.N + 127) / 128, (a.M + 63) / 64, splits), dim3(256), 0, s, a);
}

hipError_t gemm(const GemmArgs& a, int splits, hipStream_t s) {
    if (a.K % 16 || a.kchunk % 16 || a.lda % 8 || a.ldb % 8) return hipErrorInvalidValue;
    if (a.conv) {
        // the implicit im2col reads A while the epilogue writes: no aliasing
        if (a.convC % 16 || a.K != 9 * a.convC || (const void*)a.A == (const void*)a.Cb) return hipErrorInvalidValue;
        if (a.conv == 2) launch_gemm<2>(a, splits, s);
        else launch_gemm<1>(a, splits, s);
    } else {
        launch_gemm<0>(a, splits, s);
    }
    return hipGetLastError();
}

// out = act(conv(x) + b) [+ resid]; x NHWC with ldx, out NHWC [P][ldo].
int conv_fwd(flf_net* net, const NetLayer& l, const bf16* x, int ldx, int n, int H, int W, int relu, const bf16* resid,
             bf16* outb, float* outf, int ldo, hipStream_t s) {
    const long long P = (long long)n * H * W;
    GemmArgs a{};
    a.conv = 1; a.convH = H; a.convW = W; a.convC = l.Cip;
    a.A = x; a.lda = ldx;
    a.B = l.Wb; a.ldb = 9LL * l.Cip;
    a.M = (int)P; a.N = l.Co; a.K = 9 * l.Cip; a.kchunk = a.K;
    a.bias = net->w + l.boff;
    a.resid = resid; a.ldr = ldo;
    a.relu = relu;
    a.Cb = outb; a.ldcb = ldo;
    a.Cf = outf; a.ldcf = ldo;
    NCHK(gemm(a, 1, s));
    return FLF_OK;
}

// dx = dgrad(dy) (+ add if accumulate); dy NHWC [P][ldy] with l.Co channels.
int conv_dgrad(flf_net* net, const NetLayer& l, const bf16* dy, int ldy, int n, int H, int W, const bf16* add,
               bf16* dx, int ldx, hipStream_t s) {
    const long long P = (long long)n * H * W;
    GemmArgs a{};
    a.conv = 2; a.convH = H; a.convW = W; a.convC = l.Cop;
    a.A = dy; a.lda = ldy;
    a.B = l.Wt; a.ldb = 9LL * l.Cop;
    a.M = (int)P; a.N = l.Ci; a.K = 9 * l.Cop; a.kchunk = a.K;
    a.resid = add; a.ldr = ldx;
    a.Cb = dx; a.ldcb = ldx;
    NCHK(gemm(a, 1, s));
    return FLF_OK;
}

// dW (fp32, [Co][9][Ci] in the gradient vector) and db from x and dy.
int conv_wgrad(flf_net* net, const NetLayer& l, const bf16* x, int ldx, const bf16* dy, int ldy, int n, int H, int W,
               hipStream_t s) {
    const long long P = (long long)n * H * W;
    const long long Pp = (P + 15) / 16 * 16;
    if (ldx % 8) return nfail(net, FLF_EINVAL, "wgrad: activation row stride must be a multiple of 8");
    hipLaunchKernelGGL(k_im2col_t, dim3((unsigned)((Pp + 63) / 64), (l.Cip + 63) / 64), dim3(256), 0, s, x, n, H, W,
                       l.Ci, ldx, l.Cip, Pp, net->colt);
    hipLaunchKernelGGL(k_transpose, dim3(blocks_for((long long)l.Co * Pp)), dim3(256), 0, s, dy, P, l.Co, ldy, Pp,
                       l.Co, net->dyt);
    NCHK(hipGetLastError());
    int splits = (int)std::min<long long>(32, std::max<long long>(1, Pp / 2048));
    int kchunk = (int)((Pp + splits - 1) / splits);
    kchunk = (kchunk + 15) / 16 * 16;
    splits = (int)((Pp + kchunk - 1) / kchunk);
    GemmArgs a{};
    a.A = net->dyt; a.lda = Pp;
    a.B = net->colt; a.ldb = Pp;
    a.M = l.Co; a.N = 9 * l.Cip; a.K = (int)Pp; a.kchunk = kchunk;
    a.Cf = net->wpart; a.ldcf = 9LL * l.Cip;
    NCHK(gemm(a, splits, s));
    // reduce splits into the padded layout, then compact into [Co][9][Ci]
    const long long n_pad = (long long)l.Co * 9 * l.Cip;
    hipLaunchKernelGGL(k_reduce_splits, dim3(blocks_for(n_pad)), dim3(256), 0, s, net->wpart, splits, n_pad,
                       net->wpart + (long long)splits * n_pad);
    NCHK(hipGetLastError());
    // compaction on the host side of the layout: copy rows of Ci out of Cip
    NCHK(hipMemcpy2DAsync(net->g + l.woff, sizeof(float) * l.Ci, net->wpart + (long long)splits * n_pad,
                          sizeof(float) * l.Cip, sizeof(float) * l.Ci, (size_t)l.Co * 9, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_bias_grad, dim3(l.Co), dim3(256), 0, s, net->dyt, Pp, net->g + l.boff);
    NCHK(hipGetLastError());
    return FLF_OK;
}

int refresh_bf16(flf_net* net, hipStream_t s) {
    for (auto& l : net->L) {
        hipLaunchKernelGGL(k_weights_bf16, dim3(blocks_for((long long)l.Co * 9 * l.Ci)), dim3(256), 0, s,
                           net->w + l.woff, l.Co, l.Ci, l.Cip, l.Cop, l.Wb, l.Wt);
        NCHK(hipGetLastError());
    }
    return FLF_OK;
}

int forward(flf_net* net, const float* img, int T, int Ly, int Lx, hipStream_t s) {
    const int N = net->N, Wd = net->width;
    const long long P = (long long)T * Ly * Lx;
    int rc = ensure_acts(net, P);
    if (rc) return rc;
    if (T > 1024) return nfail(net, FLF_EINVAL, "too many frames for the network");
    hipLaunchKernelGGL(k_frame_sum_partial, dim3(MEAN_BLOCKS, T), dim3(256), 0, s, img, (long long)Ly * N * Lx * N,
                       net->mpart);
    hipLaunchKernelGGL(k_frame_mean_final, dim3(T), dim3(MEAN_BLOCKS), 0, s, net->mpart, (long long)Ly * N * Lx * N,
                       net->scale);
    hipLaunchKernelGGL(k_unshuffle, dim3(blocks_for(P * net->L[0].Cip)), dim3(256), 0, s, img, T, Ly, Lx, N, net->scale,
                       net->L[0].Cip, net->x);
    NCHK(hipGetLastError());
    if ((rc = conv_fwd(net, net->L[0], net->x, net->L[0].Cip, T, Ly, Lx, 1, nullptr, net->h[0], nullptr, Wd, s))) return rc;
    for (int i = 1; i <= net->nblocks; ++i) {
        const NetLayer& la = net->L[2 * i - 1];
        const NetLayer& lb = net->L[2 * i];
        if ((rc = conv_fwd(net, la, net->h[i - 1], Wd, T, Ly, Lx, 1, nullptr, net->t[i], nullptr, Wd, s))) return rc;
        if ((rc = conv_fwd(net, lb, net->t[i], Wd, T, Ly, Lx, 0, net->h[i - 1], net->h[i], nullptr, Wd, s))) return rc;
    }
    const NetLayer& lh = net->L.back();
    if ((rc = conv_fwd(net, lh, net->h[net->nblocks], Wd, T, Ly, Lx, 0, nullptr, nullptr, net->o, lh.Cop, s))) return rc;
    return FLF_OK;
}

int backward(flf_net* net, const float* tgt, int T, int Ly, int Lx, hipStream_t s) {
    const int Wd = net->width, nb = net->nblocks;
    const long long P = (long long)T * Ly * Lx;
    const NetLayer& lh = net->L.back();
    const int nlb = 256;
    hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, s, net->o, (int)P, lh.Cop, net->nz, Ly, Lx, net->scale,
                       1.0f / (float)net->nz, tgt, net->dO, lh.Cop, net->lpart);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(1), 0, s, net->lpart, nlb, net->lout);
    NCHK(hipGetLastError());
    int rc;
    // head
    if ((rc = conv_wgrad(net, lh, net->h[nb], Wd, net->dO, lh.Cop, T, Ly, Lx, s))) return rc;
    bf16* dh = net->dbuf0;
    bf16* dt = net->dbuf1;
    if ((rc = conv_dgrad(net, lh, net->dO, lh.Cop, T, Ly, Lx, nullptr, dh, Wd, s))) return rc;
    for (int i = nb; i >= 1; --i) {
        const NetLayer& la = net->L[2 * i - 1];
        const NetLayer& lb = net->L[2 * i];
        // h_i = h_{i-1} + conv_b(t_i)
        if ((rc = conv_wgrad(net, lb, net->t[i], Wd, dh, Wd, T, Ly, Lx, s))) return rc;
        if ((rc = conv_dgrad(net, lb, dh, Wd, T, Ly, Lx, nullptr, dt, Wd, s))) return rc;
        hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(P * Wd)), dim3(256), 0, s, dt, net->t[i], P * Wd);
        NCHK(hipGetLastError());
        if ((rc = conv_wgrad(net, la, net->h[i - 1], Wd, dt, Wd, T, Ly, Lx, s))) return rc;
        // dh_{i-1} = dh_i + dgrad_a(dt), accumulated in place in dh (each
        // element is read and written by the same thread; dt is only read)
        if ((rc = conv_dgrad(net, la, dt, Wd, T, Ly, Lx, dh, dh, Wd, s))) return rc;
    }
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(P * Wd)), dim3(256), 0, s, dh, net->h[0], P * Wd);
    NCHK(hipGetLastError());
    if ((rc = conv_wgrad(net, net->L[0], net->x, net->L[0].Cip, dh, Wd, T, Ly, Lx, s))) return rc;
    return FLF_OK;
}

}  // namespace

extern "C" {

int flf_net_create(int device, int nnum, int nz, int width, int nblocks, flf_net** out) {
    if (!out) return FLF_EINVAL;
    *out = nullptr;
    if (nnum < 3 || nz < 1 || nz > 512 || width < 16 || width % 16 || nblocks < 0 || nblocks > 64) return FLF_EINVAL;
    flf_net* net = new flf_net();
    net->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete net; return FLF_EHIP; }
    net->N = nnum; net->nz = nz; net->width = width; net->nblocks = nblocks;
    auto add = [&](int Ci, int Co) {
        NetLayer l;
        l.Ci = Ci; l.Co = Co; l.Cip = pad16(Ci); l.Cop = pad16(Co);
        l.woff = net->nparam;
        net->nparam += (long long)Co * 9 * Ci;
        l.boff = net->nparam;
        net->nparam += Co;
        net->L.push_back(l);
    };
    add(nnum * nnum, width);
    for (int i = 0; i < nblocks; ++i) { add(width, width); add(width, width); }
    add(width, nz);
    std::vector<void*> keep;
    int rc = 0;
    rc |= nalloc(net, &net->w, (size_t)net->nparam, keep);
    rc |= nalloc(net, &net->g, (size_t)net->nparam, keep);
    rc |= nalloc(net, &net->m, (size_t)net->nparam, keep);
    rc |= nalloc(net, &net->v, (size_t)net->nparam, keep);
    for (auto& l : net->L) {
        rc |= nalloc(net, &l.Wb, (size_t)l.Co * 9 * l.Cip, keep);
        rc |= nalloc(net, &l.Wt, (size_t)l.Ci * 9 * l.Cop, keep);
        if (!rc) {
            (void)hipMemset(l.Wb, 0, (size_t)l.Co * 9 * l.Cip * sizeof(bf16));
            (void)hipMemset(l.Wt, 0, (size_t)l.Ci * 9 * l.Cop * sizeof(bf16));
        }
    }
    if (rc) {
        for (void* p : keep) (void)hipFree(p);
        delete net;
        return FLF_ENOMEM;
    }
    (void)hipMemset(net->m, 0, net->nparam * sizeof(float));
    (void)hipMemset(net->v, 0, net->nparam * sizeof(float));
    (void)hipMemset(net->g, 0, net->nparam * sizeof(float));
    (void)hipMemset(net->w, 0, net->nparam * sizeof(float));
    net->acts.clear();
    // keep the parameter allocations separate from the activation list
    net->err.clear();
    *out = net;
    return FLF_OK;
}

void flf_net_destroy(flf_net* net) {
    if (!net) return;
    (void)hipSetDevice(net->device);
    free_acts(net);
    (void)hipFree(net->w); (void)hipFree(net->g); (void)hipFree(net->m); (void)hipFree(net->v);
    for (auto& l : net->L) { (void)hipFree(l.Wb); (void)hipFree(l.Wt); }
    delete net;
}

const char* flf_net_last_error(const flf_net* net) { return net ? net->err.c_str() : "null network"; }

long long flf_net_num_params(const flf_net* net) { return net ? net->nparam : 0; }

int flf_net_set_params(flf_net* net, const float* host) {
    if (!net || !host) return FLF_EINVAL;
    NCHK(hipMemcpy(net->w, host, net->nparam * sizeof(float), hipMemcpyHostToDevice));
    (void)hipMemset(net->m, 0, net->nparam * sizeof(float));
    (void)hipMemset(net->v, 0, net->nparam * sizeof(float));
    net->step = 0;
    int rc = refresh_bf16(net, 0);
    if (rc) return rc;
    NCHK(hipDeviceSynchronize());
    return FLF_OK;
}

int flf_net_get_params(flf_net* net, float* host) {
    if (!net || !host) return FLF_EINVAL;
    NCHK(hipDeviceSynchronize());
    NCHK(hipMemcpy(host, net->w, net->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return FLF_OK;
}

int flf_net_get_grads(flf_net* net, float* host) {
    if (!net || !host) return FLF_EINVAL;
    NCHK(hipDeviceSynchronize());
    NCHK(hipMemcpy(host, net->g, net->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return FLF_OK;
}

int flf_net_forward(flf_net* net, const float* img, int T, int Ly, int Lx, float* out, void* stream) {
    if (!net || !img || !out || T < 1 || Ly < 1 || Lx < 1) return nfail(net, FLF_EINVAL, "bad network input");
    std::string why;
    const long long npx = (long long)T * Ly * net->N * Lx * net->N;
    if (!flf::device_span(net->device, img, (size_t)npx * sizeof(float), "image", why) ||
        !flf::device_span(net->device, out, (size_t)T * net->nz * Ly * Lx * sizeof(float), "output", why))
        return nfail(net, FLF_EINVAL, why);
    hipStream_t s = (hipStream_t)stream;
    int rc = forward(net, img, T, Ly, Lx, s);
    if (rc) return rc;
    const NetLayer& lh = net->L.back();
    const long long P = (long long)T * Ly * Lx;
    hipLaunchKernelGGL(k_head_out, dim3(blocks_for(P * net->nz)), dim3(256), 0, s, net->o, (int)P, lh.Cop, net->nz, Ly,
                       Lx, net->scale, 1.0f / (float)net->nz, out);
    NCHK(hipGetLastError());
    return FLF_OK;
}

int flf_net_train_step(flf_net* net, const float* img, const float* target, int T, int Ly, int Lx, float lr,
                       double* loss_host, void* stream) {
    if (!net || !img || !target || T < 1 || Ly < 1 || Lx < 1) return nfail(net, FLF_EINVAL, "bad training input");
    std::string why;
    const long long npx = (long long)T * Ly * net->N * Lx * net->N;
    if (!flf::device_span(net->device, img, (size_t)npx * sizeof(float), "image", why) ||
        !flf::device_span(net->device, target, (size_t)T * net->nz * Ly * Lx * sizeof(float), "target", why))
        return nfail(net, FLF_EINVAL, why);
    hipStream_t s = (hipStream_t)stream;
    int rc = forward(net, img, T, Ly, Lx, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill0, dim3(blocks_for(net->nparam)), dim3(256), 0, s, net->g, net->nparam);
    if ((rc = backward(net, target, T, Ly, Lx, s))) return rc;
    if (lr > 0.f) {
        net->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float bc1 = 1.f - std::pow(b1, (float)net->step), bc2 = 1.f - std::pow(b2, (float)net->step);
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(net->nparam)), dim3(256), 0, s, net->w, net->g, net->m, net->v,
                           net->nparam, lr, b1, b2, 1e-8f, bc1, bc2);
        NCHK(hipGetLastError());
        if ((rc = refresh_bf16(net, s))) return rc;
    }
    if (loss_host) {
        NCHK(hipMemcpyAsync(loss_host, net->lout, sizeof(double), hipMemcpyDeviceToHost, s));
        NCHK(hipStreamSynchronize(s));
    }
    return FLF_OK;
}

int flf_net_upsample(const float* g, int T, int nz, int Ly, int Lx, int nnum, float* vol, void* stream) {
    flf::tls_error().clear();
    if (!g || !vol || T < 1 || nz < 1 || Ly < 1 || Lx < 1 || nnum < 1) return FLF_EINVAL;
    const long long rows = (long long)nz * T * Ly * nnum;
    if (rows > INT32_MAX || (long long)Lx * nnum > INT32_MAX) return FLF_EINVAL;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FLF_EHIP;
    std::string why;
    if (!flf::device_span(dev, g, (size_t)T * nz * Ly * Lx * sizeof(float), "guess", why) ||
        !flf::device_span(dev, vol, (size_t)rows * Lx * nnum * sizeof(float), "volume", why)) {
        flf::tls_error() = why;
        return FLF_EINVAL;
    }
    hipLaunchKernelGGL(k_upsample, dim3((unsigned)std::min<long long>(rows, 65536)), dim3(256), 0, (hipStream_t)stream,
                       g, T, nz, Ly, Lx, nnum, vol);
    return hipGetLastError() == hipSuccess ? FLF_OK : FLF_EHIP;
}

// C = A * B^T for bf16 row-major A [M][K], B [N][K]; fp32 C [M][N].  Exposed
// for the MFMA layout test.
int flf_mfma_gemm_nt(const void* A, const void* B, float* C, int M, int N, int K, void* stream) {
    flf::tls_error().clear();
    if (!A || !B || !C || M < 1 || N < 1 || K < 16 || K % 16) return FLF_EINVAL;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return FLF_EHIP;
    std::string why;
    if (!flf::device_span(dev, A, (size_t)M * K * 2, "A", why) || !flf::device_span(dev, B, (size_t)N * K * 2, "B", why) ||
        !flf::device_span(dev, C, (size_t)M * N * sizeof(float), "C", why)) {
        flf::tls_error() = why;
        return FLF_EINVAL;
    }
    GemmArgs a{};
    a.A = (const bf16*)A; a.lda = K;
    a.B = (const bf16*)B; a.ldb = K;
    a.M = M; a.N = N; a.K = K; a.kchunk = K;
    a.Cf = C; a.ldcf = N;
    return gemm(a, 1, (hipStream_t)stream) == hipSuccess ? FLF_OK : FLF_EHIP;
}

}  // extern "C"
