// HIP kernels for the light-field projector and Richardson-Lucy loop (gfx950).
//
// Data flow of one projection (forward: volume -> image, backward: image ->
// volume), for each Fourier grid group g (planes sharing (fY, fX)):
//   k_small_fft    S[p][t][off][k0][k1]  = FFT_r(src[b::N, a::N])   (half plane)
//   k_mac          A[...][k0][k1]        = sum over 225 offsets of S * spectra
//   k_ifft_cols    T[...][y][k1]         = column inverse FFT, rows y < Y kept
//   k_ifft_rows_*  real rows, 'same' crop folded into A as a phase, summed over
//                  groups (forward) or written / RL-updated per plane (backward)
// and once per PSF set: k_psf_rows + k_psf_cols (pruned 2-D FFT of each
// quadrant PSF) and k_spec_pack (gather into the MAC's orbit-major layout).
#include "flf_common.h"
#include "flf_fft.h"
#include "flf_kernels.h"

using namespace flf;

namespace {

constexpr int FFT_THREADS = 512;
constexpr int FFT_MAXB = 16;

__device__ __forceinline__ int centred_res(int k, int r) {
    // Centred representative of residue k in Z_r, with r/2 (r even) -> 0.
    if (2 * k < r) return k;
    if (2 * k > r) return k - r;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Small FFTs of the lenslet-subsampled source (py_light_field.cpp:1001-1027).
// grid (N*N, T, nsrc); dynamic LDS rY*rX float2.
// Output half spectrum: S[z][t][off][k0][k1], k1 in [0, rX/2].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FFT_THREADS) void k_small_fft(SmallFFTArgs a) {
    extern __shared__ float2 sh[];
    const int off = blockIdx.x;
    const int t = blockIdx.y;
    const int z = blockIdx.z;
    const int N = a.N;
    const int b = off / N, aa = off % N;
    const int rY = a.rpY.L, rX = a.rpX.L, rXh = rX / 2 + 1;
    const float* src = a.src + a.src_off[z] + (long long)t * a.src_tstride;
    const int Ys = (a.Y - b + N - 1) / N;
    const int Xs = (a.X - aa + N - 1) / N;
    for (int i = threadIdx.x; i < rY * rX; i += blockDim.x) sh[i] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int i = threadIdx.x; i < Ys * Xs; i += blockDim.x) {
        int y2 = i / Xs, x2 = i - y2 * Xs;
        sh[y2 * rX + x2] = make_float2(src[(long long)(b + N * y2) * a.src_pitch + aa + N * x2], 0.f);
    }
    __syncthreads();
    lds_fft<FFT_MAXB>(sh, a.rpX, Ys, 1, rX, -1.f);          // rows (only non-zero ones)
    lds_fft<FFT_MAXB>(sh, a.rpY, rXh, rX, 1, -1.f);         // columns of the half plane
    float2* S = a.S + a.S_off[z] + (long long)t * a.S_tstride + (long long)off * rY * rXh;
    for (int i = threadIdx.x; i < rY * rXh; i += blockDim.x) {
        int k0 = i / rXh, k1 = i - k0 * rXh;
        S[i] = sh[k0 * rX + k1];
    }
}

// ---------------------------------------------------------------------------
// PSF spectra, stage 1: row FFTs of the s2 non-zero rows of each quadrant PSF
// (zero padded to fX), half plane kept.  (py_light_field.cpp:757-812 computes
// the same pruned 2-D transform column-first on the full plane.)
// grid (ceil(s2/RB), nu); LDS RB*fX float2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FFT_THREADS) void k_psf_rows(PsfRowArgs a) {
    extern __shared__ float2 sh[];
    const int u = blockIdx.y;
    const int y0 = blockIdx.x * a.RB;
    const int nrows = min(a.RB, a.s2 - y0);
    const int fX = a.rpX.L, rfX = fX / 2 + 1, s2 = a.s2, h = a.h;
    const int bu = a.u_bu[u], au = a.u_au[u];
    const bool tr = bu > au;   // Q[a][b] for a > b is stored as the transpose of Q[b][a]
    const float* Q = a.Q + ((long long)(tr ? au : bu) * h + (tr ? bu : au)) * s2 * s2;
    for (int i = threadIdx.x; i < nrows * fX; i += blockDim.x) {
        int r = i / fX, x = i - r * fX;
        int y = y0 + r;
        float v = 0.f;
        if (x < s2) v = tr ? Q[(long long)x * s2 + y] : Q[(long long)y * s2 + x];
        sh[i] = make_float2(v, 0.f);
    }
    __syncthreads();
    lds_fft<FFT_MAXB>(sh, a.rpX, nrows, 1, fX, -1.f);
    float2* out = a.tmp + (long long)u * s2 * rfX;
    for (int i = threadIdx.x; i < nrows * rfX; i += blockDim.x) {
        int r = i / rfX, k1 = i - r * rfX;
        out[(long long)(y0 + r) * rfX + k1] = sh[r * fX + k1];
    }
}

// ---------------------------------------------------------------------------
// PSF spectra, stage 2: column FFTs (length fY, input rows >= s2 are zero).
// grid (ceil(rfX/CB), nu); LDS CB*fY float2 (column-major).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FFT_THREADS) void k_psf_cols(PsfColArgs a) {
    extern __shared__ float2 sh[];
    const int u = blockIdx.y;
    const int fY = a.rpY.L, rfX = a.rfX, s2 = a.s2;
    const int c0 = blockIdx.x * a.CB;
    const int ncol = min(a.CB, rfX - c0);
    const float2* in = a.tmp + (long long)u * s2 * rfX;
    for (int i = threadIdx.x; i < fY * a.CB; i += blockDim.x) {
        int y = i / a.CB, c = i - y * a.CB;
        float2 v = make_float2(0.f, 0.f);
        if (y < s2 && c < ncol) v = in[(long long)y * rfX + c0 + c];
        sh[c * fY + y] = v;
    }
    __syncthreads();
    lds_fft<FFT_MAXB>(sh, a.rpY, a.CB, 1, fY, -1.f);
    float2* out = a.spec + (long long)u * fY * rfX;
    for (int i = threadIdx.x; i < fY * a.CB; i += blockDim.x) {
        int y = i / a.CB, c = i - y * a.CB;
        if (c < ncol) out[(long long)y * rfX + c0 + c] = sh[c * fY + y];
    }
}

// ---------------------------------------------------------------------------
// Gather the natural-layout spectra of one plane into the orbit-major layout
// G[chunk][plane][u][item][slot], each value pre-multiplied by
// w^(bu*m2 + au*n2), w = exp(-2 pi i / (2N)) (m2, n2: half-unit aliases).
// grid (nchunks, nu), block FLF_CHUNK_MAX.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FLF_CHUNK_MAX) void k_spec_pack(SpecPackArgs a) {
    const ChunkDesc ch = a.chunks[blockIdx.x];
    const int u = blockIdx.y;
    const int i = threadIdx.x;
    if (i >= ch.n_items) return;
    const ItemDesc it = a.items[ch.item_base + i];
    const int bu = a.u_bu[u], au = a.u_au[u];
    const float2* spec = a.spec + (long long)u * a.fY * a.rfX;
    float2* G = a.G + ch.item_base * a.slots * a.nu * a.nP +
                (((long long)a.p * a.nu + u) * ch.n_items + i) * a.slots;
    for (int j = 0; j < a.slots; ++j) {
        float2 v = spec[it.pos[j]];
        float2 w = phase_frac((long long)bu * it.m2[j] + (long long)au * it.n2[j], 2 * a.N);
        G[j] = cmul(v, w);
    }
}

// ---------------------------------------------------------------------------
// The multiply-accumulate over the 225 lenslet offsets.
//
// For an output bin k with residue kappa and (half-unit) alias (m2, n2),
//   A(k) = C0 + w^(sig*m2) Cm + w^(sig*n2) Cn + w^(sig*(m2+n2)) Cmn,
//   sig = s2 - 2 + N,
// where each class sum runs over the stored spectra u and their uses
// (identity, flips, transposes), a use contributing s_off(kappa) * G'_u(q)
// or s_off(kappa) * conj(G'_u(q)) with q one of the item's slots.  s_off is
// the small FFT of offset (b, a) times exp(-2 pi i (beta_b kt0/fY +
// alpha_a kt1/fX)), beta/alpha the offset index plus s2-1 when flipped.
// This is the accumulation of py_light_field.cpp:303-341, 1029-1087 with the
// mirror multipliers of projector.py:137-151 and the 'same' crop
// (projector.py:792-796) folded into phases.
// ---------------------------------------------------------------------------
namespace {

// Use types: 0 id, 1 both-flip, 2 row-flip, 3 col-flip, 4..7 the same for the
// transposed PSF.  Per use: (which read of the output's bin set, conj, class).
//   read 0: k, 1: (-k0, k1), 2: (k1, k0), 3: (-k1, k0)
__device__ constexpr int kUseRead[8] = {0, 0, 1, 1, 2, 2, 3, 3};
__device__ constexpr int kUseConj[8] = {0, 1, 0, 1, 0, 1, 1, 0};
__device__ constexpr int kUseClass[8] = {0, 3, 1, 2, 0, 3, 1, 2};
// Quad slots P0=(p0,p1) P1=(-p0,p1) P2=(p1,p0) P3=(-p1,p0): for output slot j,
// the slot holding read r, and whether Hermitian symmetry adds a conj.
__device__ constexpr int kQuadSrc[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
__device__ constexpr int kQuadHc[4][4] = {{0, 0, 0, 0}, {0, 0, 1, 1}, {0, 0, 0, 0}, {0, 0, 1, 1}};

template <int SLOTS, int TYPE>
__device__ __forceinline__ void mac_use(float2 (&acc)[4][4], const float2 (&g)[4],
                                        const float2* __restrict__ s_tab, const int (&rb)[4], int off) {
#pragma unroll
    for (int j = 0; j < SLOTS; ++j) {
        constexpr int rd = kUseRead[TYPE];
        const int src = (SLOTS == 4) ? kQuadSrc[j][rd] : (rd == 0 ? j : 1 - j);
        const int cj = kUseConj[TYPE] ^ ((SLOTS == 4) ? kQuadHc[j][rd] : 0);
        const float2 s = s_tab[rb[j] + off];
        if (cj) cmacc(acc[j][kUseClass[TYPE]], s, g[src]);
        else    cmac(acc[j][kUseClass[TYPE]], s, g[src]);
    }
}

typedef float v4f __attribute__((ext_vector_type(4)));

// Spectra are streamed once per projection: non-temporal loads keep them from
// evicting the reused small tables from the caches.
template <int SLOTS>
__device__ __forceinline__ void load_g(float2 (&g)[4], const float2* p) {
    const v4f* q = reinterpret_cast<const v4f*>(p);
    v4f x = __builtin_nontemporal_load(q);
    g[0] = make_float2(x.x, x.y);
    g[1] = make_float2(x.z, x.w);
    if constexpr (SLOTS == 4) {
        v4f y = __builtin_nontemporal_load(q + 1);
        g[2] = make_float2(y.x, y.y);
        g[3] = make_float2(y.z, y.w);
    } else {
        g[2] = g[3] = make_float2(0.f, 0.f);
    }
}

}  // namespace

template <int SLOTS>
__global__ __launch_bounds__(FLF_CHUNK_MAX) void k_mac(MacArgs a) {
    extern __shared__ float2 s_tab[];
    const ChunkDesc ch = a.chunks[blockIdx.x];
    const int tid = threadIdx.x;
    const bool valid = tid < ch.n_items;
    const int N = a.N, NN = N * N, c = N / 2, twoN = 2 * N;
    const int t = blockIdx.z;
    ItemDesc it;
    if (valid) it = a.items[ch.item_base + tid];
    else {
        for (int j = 0; j < 4; ++j) { it.pos[j] = 0; it.res[j] = 0; it.m2[j] = 0; it.n2[j] = 0; }
        it.writemask = 0;
    }
    int rb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) rb[j] = it.res[j] * NN;
    int k0s[4], k1s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { k0s[j] = it.pos[j] / a.rfX; k1s[j] = it.pos[j] - k0s[j] * a.rfX; }

    float2 tot[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) tot[j] = make_float2(0.f, 0.f);

    const int pb = a.plane_loop ? 0 : blockIdx.y;
    const int pe = a.plane_loop ? a.nP : blockIdx.y + 1;
    const int rXh = a.rX / 2 + 1;
    const long long S_offstride = (long long)a.rY * rXh;
    const int nu = a.nu;

    for (int p = pb; p < pe; ++p) {
        if (!a.enable[p]) continue;
        const int s2 = a.s2[p];
        __syncthreads();
        // ---- s-table: n_res residues x N*N offsets, phases for this plane ----
        const float2* Sp = a.S + a.S_off[p] + (long long)t * a.S_tstride;
        for (int idx = tid; idx < ch.n_res * NN; idx += blockDim.x) {
            const int rho = idx / NN, off = idx - rho * NN;
            const int bo = off / N, ao = off - bo * N;
            const int k0 = ch.res0[rho], k1 = ch.res1[rho];
            float2 v;
            if (2 * k1 <= a.rX) v = Sp[off * S_offstride + (long long)k0 * rXh + k1];
            else v = cconj(Sp[off * S_offstride + (long long)((a.rY - k0) % a.rY) * rXh + (a.rX - k1)]);
            const long long beta = bo <= c ? bo : bo + s2 - 1;
            const long long alpha = ao <= c ? ao : ao + s2 - 1;
            const long long kt0 = centred_res(k0, a.rY), kt1 = centred_res(k1, a.rX);
            long long n0 = (beta * kt0) % a.fY; if (n0 < 0) n0 += a.fY;
            long long n1 = (alpha * kt1) % a.fX; if (n1 < 0) n1 += a.fX;
            float fr = (float)n0 / (float)a.fY + (float)n1 / (float)a.fX;
            fr -= rintf(fr);
            float sn, cs;
            sincospif(-2.0f * fr, &sn, &cs);
            s_tab[idx] = cmul(v, make_float2(cs, sn));
        }
        __syncthreads();

        float2 acc[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[j][q] = make_float2(0.f, 0.f);

        const float2* Gp = a.G + ch.item_base * SLOTS * nu * a.nP +
                           ((long long)p * nu) * ch.n_items * SLOTS + (long long)tid * SLOTS;
        const long long ustride = (long long)ch.n_items * SLOTS;
        float2 g[4];

        if constexpr (SLOTS == 4) {
            // u = 0: (c, c) -> identity only
            if (valid) {
                load_g<4>(g, Gp);
                mac_use<4, 0>(acc, g, s_tab, rb, c * N + c);
            }
            // u in [1, 1+c): (b, b) -> id, both, row, col
            for (int u = 1; u < 1 + c; ++u) {
                const int bu = u - 1;
                if (valid) {
                    load_g<4>(g, Gp + u * ustride);
                    const int bb = N - 1 - bu;
                    mac_use<4, 0>(acc, g, s_tab, rb, bu * N + bu);
                    mac_use<4, 1>(acc, g, s_tab, rb, bb * N + bb);
                    mac_use<4, 2>(acc, g, s_tab, rb, bb * N + bu);
                    mac_use<4, 3>(acc, g, s_tab, rb, bu * N + bb);
                }
            }
            // u in [1+c, 1+2c): (b, c) -> id, row, Tid, Tcol
            for (int u = 1 + c; u < 1 + 2 * c; ++u) {
                const int bu = u - 1 - c;
                if (valid) {
                    load_g<4>(g, Gp + u * ustride);
                    const int bb = N - 1 - bu;
                    mac_use<4, 0>(acc, g, s_tab, rb, bu * N + c);
                    mac_use<4, 2>(acc, g, s_tab, rb, bb * N + c);
                    mac_use<4, 4>(acc, g, s_tab, rb, c * N + bu);
                    mac_use<4, 7>(acc, g, s_tab, rb, c * N + bb);
                }
            }
            // u in [1+2c, nu): (b, a), b < a < c -> all eight uses
            for (int u = 1 + 2 * c; u < nu; ++u) {
                const int bu = a.u_bu[u], au = a.u_au[u];
                if (valid) {
                    load_g<4>(g, Gp + u * ustride);
                    const int bb = N - 1 - bu, ab = N - 1 - au;
                    mac_use<4, 0>(acc, g, s_tab, rb, bu * N + au);
                    mac_use<4, 1>(acc, g, s_tab, rb, bb * N + ab);
                    mac_use<4, 2>(acc, g, s_tab, rb, bb * N + au);
                    mac_use<4, 3>(acc, g, s_tab, rb, bu * N + ab);
                    mac_use<4, 4>(acc, g, s_tab, rb, au * N + bu);
                    mac_use<4, 5>(acc, g, s_tab, rb, ab * N + bb);
                    mac_use<4, 6>(acc, g, s_tab, rb, ab * N + bu);
                    mac_use<4, 7>(acc, g, s_tab, rb, au * N + bb);
                }
            }
        } else {
            // Non-square grids: 64 stored spectra (transposes stored), pairs.
            if (valid) {
                load_g<2>(g, Gp);
                mac_use<2, 0>(acc, g, s_tab, rb, c * N + c);
            }
            // (b, c): id, row
            for (int u = 1; u < 1 + c; ++u) {
                const int bu = u - 1;
                if (valid) {
                    load_g<2>(g, Gp + u * ustride);
                    mac_use<2, 0>(acc, g, s_tab, rb, bu * N + c);
                    mac_use<2, 2>(acc, g, s_tab, rb, (N - 1 - bu) * N + c);
                }
            }
            // (c, a): id, col
            for (int u = 1 + c; u < 1 + 2 * c; ++u) {
                const int au = u - 1 - c;
                if (valid) {
                    load_g<2>(g, Gp + u * ustride);
                    mac_use<2, 0>(acc, g, s_tab, rb, c * N + au);
                    mac_use<2, 3>(acc, g, s_tab, rb, c * N + (N - 1 - au));
                }
            }
            // (b, a), b, a < c: id, both, row, col
            for (int u = 1 + 2 * c; u < nu; ++u) {
                const int bu = a.u_bu[u], au = a.u_au[u];
                if (valid) {
                    load_g<2>(g, Gp + u * ustride);
                    const int bb = N - 1 - bu, ab = N - 1 - au;
                    mac_use<2, 0>(acc, g, s_tab, rb, bu * N + au);
                    mac_use<2, 1>(acc, g, s_tab, rb, bb * N + ab);
                    mac_use<2, 2>(acc, g, s_tab, rb, bb * N + au);
                    mac_use<2, 3>(acc, g, s_tab, rb, bu * N + ab);
                }
            }
        }

        // ---- epilogue: class constants and the 'same'-crop shift ----
        const int sig = s2 - 2 + N;
        const int c2 = (s2 - 1) / 2;
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const int m2 = it.m2[j], n2 = it.n2[j];
            float2 v = acc[j][0];
            v = cadd(v, cmul(phase_frac((long long)sig * m2, twoN), acc[j][1]));
            v = cadd(v, cmul(phase_frac((long long)sig * n2, twoN), acc[j][2]));
            v = cadd(v, cmul(phase_frac((long long)sig * (m2 + n2), twoN), acc[j][3]));
            // exp(+2 pi i c2 (k0/fY + k1/fX))
            long long e0 = ((long long)c2 * k0s[j]) % a.fY;
            long long e1 = ((long long)c2 * k1s[j]) % a.fX;
            float fr = (float)e0 / (float)a.fY + (float)e1 / (float)a.fX;
            fr -= rintf(fr);
            float sn, cs;
            sincospif(2.0f * fr, &sn, &cs);
            tot[j] = cadd(tot[j], cmul(v, make_float2(cs, sn)));
        }
    }

    if (valid) {
        float2* A = a.A + (long long)t * a.A_tstride + (a.plane_loop ? 0 : (long long)blockIdx.y * a.A_pstride);
#pragma unroll
        for (int j = 0; j < SLOTS; ++j)
            if ((it.writemask >> j) & 1) A[it.pos[j]] = tot[j];
    }
}

// ---------------------------------------------------------------------------
// Inverse transform, stage 1: column inverse FFTs (length fY) of the half
// plane, keeping output rows y < Yout.  grid (ceil(rfX/CB), nbatch).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FFT_THREADS) void k_ifft_cols(IfftColArgs a) {
    extern __shared__ float2 sh[];
    const int fY = a.rpY.L, rfX = a.rfX;
    const int c0 = blockIdx.x * a.CB;
    const int ncol = min(a.CB, rfX - c0);
    const float2* in = a.A + (long long)blockIdx.y * a.A_bstride;
    for (int i = threadIdx.x; i < fY * a.CB; i += blockDim.x) {
        int y = i / a.CB, cc = i - y * a.CB;
        float2 v = make_float2(0.f, 0.f);
        if (cc < ncol) v = in[(long long)y * rfX + c0 + cc];
        sh[cc * fY + y] = v;
    }
    __syncthreads();
    lds_fft<FFT_MAXB>(sh, a.rpY, a.CB, 1, fY, +1.f);
    float2* out = a.out + (long long)blockIdx.y * a.out_bstride;
    for (int i = threadIdx.x; i < a.Yout * a.CB; i += blockDim.x) {
        int y = i / a.CB, cc = i - y * a.CB;
        if (cc < ncol) out[(long long)y * rfX + c0 + cc] = sh[cc * fY + y];
    }
}

// c2r of one row held in LDS: Hermitian-extend the rfX inputs to fX, inverse
// complex FFT; the real part is the c2r result (imaginary parts of the
// self-conjugate bins drop out, as in a c2r transform).
__device__ void row_c2r(float2* sh, const float2* row, int fX, int rfX, const RadixPlan& rp) {
    for (int i = threadIdx.x; i < fX; i += blockDim.x) {
        float2 v;
        if (i < rfX) v = row[i];
        else v = cconj(row[fX - i]);
        sh[i] = v;
    }
    __syncthreads();
    lds_fft<FFT_MAXB>(sh, rp, 1, 1, fX, +1.f);
}

// Forward projection, stage 2: per output row, c2r each grid group's row and
// sum the groups in a fixed order.  grid (Y, T).
__global__ __launch_bounds__(FFT_THREADS) void k_ifft_rows_sum(IfftRowSumArgs a) {
    extern __shared__ float2 sh[];
    float* accum = reinterpret_cast<float*>(sh + a.maxfX);
    const int y = blockIdx.x, t = blockIdx.y;
    for (int x = threadIdx.x; x < a.X; x += blockDim.x) accum[x] = 0.f;
    for (int g = 0; g < a.ngroups; ++g) {
        const IfftRowGroup& G = a.grp[g];
        if (!G.active) continue;
        __syncthreads();
        const float2* row = G.T + (long long)t * G.tstride + (long long)y * G.rfX;
        row_c2r(sh, row, G.rp.L, G.rfX, G.rp);
        const float sc = G.scale;
        for (int x = threadIdx.x; x < a.X; x += blockDim.x) accum[x] += sh[x].x * sc;
    }
    __syncthreads();
    float* out = a.img + (long long)t * a.img_tstride + (long long)y * a.X;
    for (int x = threadIdx.x; x < a.X; x += blockDim.x) out[x] = accum[x];
}

// Backward projection, stage 2: per (row, frame, plane) c2r; then either store
// the plane, or apply the Richardson-Lucy update in place
// (lfdeconv.py:112-126 Prevedel: X *= Htf / v;  lfdeconv.py:151-164 standard:
// X *= v [/ trust]; NaN -> 0 after the update).  grid (Y, T, nplanes).
__global__ __launch_bounds__(FFT_THREADS) void k_ifft_rows_plane(IfftRowPlaneArgs a) {
    extern __shared__ float2 sh[];
    const int y = blockIdx.x, t = blockIdx.y;
    const int z = a.plane_ids[blockIdx.z];
    const RowPlaneDesc pd = a.pd[z];
    const float2* row = a.Tbuf + pd.T_off + (long long)t * pd.T_tstride + (long long)y * pd.rfX;
    row_c2r(sh, row, pd.rp.L, pd.rfX, pd.rp);
    const float sc = pd.scale;
    const long long o = (((long long)z * a.T + t) * a.Y + y) * a.X;
    for (int x = threadIdx.x; x < a.X; x += blockDim.x) {
        float v = sh[x].x * sc;
        if (a.mode == 0) {
            a.vol[o + x] = v;
        } else if (a.mode == 1) {
            float r = a.Htf[o + x] / v;
            float xn = a.vol[o + x] * r;
            a.vol[o + x] = __builtin_isnan(xn) ? 0.f : xn;
        } else {
            float xn = a.vol[o + x] * v;
            if (a.trust) xn = xn / a.trust[((long long)z * a.Y + y) * a.X + x];
            a.vol[o + x] = __builtin_isnan(xn) ? 0.f : xn;
        }
    }
}

// ---------------------------------------------------------------------------
// Data-fit sums in FP64: per frame sum((img-im)^2) and sum(im^2).  Two stages
// with a fixed reduction order (deterministic).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fit_partial(const float* img, const float* im, long long n_per_frame,
                                                     double* part, int nblk) {
    __shared__ double sd[256], si[256];
    const int t = blockIdx.y;
    const float* a = img + (long long)t * n_per_frame;
    const float* b = im + (long long)t * n_per_frame;
    double d2 = 0.0, i2 = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_per_frame;
         i += (long long)nblk * blockDim.x) {
        double d = (double)a[i] - (double)b[i];
        d2 += d * d;
        i2 += (double)b[i] * (double)b[i];
    }
    sd[threadIdx.x] = d2;
    si[threadIdx.x] = i2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sd[threadIdx.x] += sd[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((long long)t * nblk + blockIdx.x) * 2 + 0] = sd[0];
        part[((long long)t * nblk + blockIdx.x) * 2 + 1] = si[0];
    }
}

__global__ __launch_bounds__(256) void k_fit_final(const double* part, int nblk, double* out) {
    __shared__ double sd[256], si[256];
    const int t = blockIdx.x;
    double d2 = 0.0, i2 = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        d2 += part[((long long)t * nblk + i) * 2 + 0];
        i2 += part[((long long)t * nblk + i) * 2 + 1];
    }
    sd[threadIdx.x] = d2;
    si[threadIdx.x] = i2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sd[threadIdx.x] += sd[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[t * 2 + 0] = sd[0];
        out[t * 2 + 1] = si[0];
    }
}

// Fill with a constant (used instead of memset so that work stays capturable).
__global__ void k_fill(float* p, long long n, float v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------
namespace flf {

hipError_t launch_small_fft(const SmallFFTArgs& a, int nsrc, int T, hipStream_t s) {
    size_t lds = (size_t)a.rpY.L * a.rpX.L * sizeof(float2);
    hipLaunchKernelGGL(k_small_fft, dim3(a.N * a.N, T, nsrc), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_psf_rows(const PsfRowArgs& a, int nu, hipStream_t s) {
    size_t lds = (size_t)a.RB * a.rpX.L * sizeof(float2);
    hipLaunchKernelGGL(k_psf_rows, dim3((a.s2 + a.RB - 1) / a.RB, nu), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_psf_cols(const PsfColArgs& a, int nu, hipStream_t s) {
    size_t lds = (size_t)a.CB * a.rpY.L * sizeof(float2);
    hipLaunchKernelGGL(k_psf_cols, dim3((a.rfX + a.CB - 1) / a.CB, nu), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_spec_pack(const SpecPackArgs& a, int nchunks, hipStream_t s) {
    hipLaunchKernelGGL(k_spec_pack, dim3(nchunks, a.nu), dim3(FLF_CHUNK_MAX), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_mac(const MacArgs& a, int slots, int nchunks, int nplanes_grid, int T, int max_res, hipStream_t s) {
    size_t lds = (size_t)max_res * a.N * a.N * sizeof(float2);
    dim3 grid(nchunks, nplanes_grid, T);
    if (slots == 4) hipLaunchKernelGGL(k_mac<4>, grid, dim3(FLF_CHUNK_MAX), lds, s, a);
    else            hipLaunchKernelGGL(k_mac<2>, grid, dim3(FLF_CHUNK_MAX), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_ifft_cols(const IfftColArgs& a, int nbatch, hipStream_t s) {
    size_t lds = (size_t)a.CB * a.rpY.L * sizeof(float2);
    hipLaunchKernelGGL(k_ifft_cols, dim3((a.rfX + a.CB - 1) / a.CB, nbatch), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_ifft_rows_sum(const IfftRowSumArgs& a, int Y, int T, hipStream_t s) {
    size_t lds = (size_t)a.maxfX * sizeof(float2) + (size_t)a.X * sizeof(float);
    hipLaunchKernelGGL(k_ifft_rows_sum, dim3(Y, T), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_ifft_rows_plane(const IfftRowPlaneArgs& a, int nplanes, int maxfX, hipStream_t s) {
    size_t lds = (size_t)maxfX * sizeof(float2);
    hipLaunchKernelGGL(k_ifft_rows_plane, dim3(a.Y, a.T, nplanes), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_fit(const float* img, const float* im, long long n_per_frame, int T, double* part, int nblk,
                      double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fit_partial, dim3(nblk, T), dim3(256), 0, s, img, im, n_per_frame, part, nblk);
    hipLaunchKernelGGL(k_fit_final, dim3(T), dim3(256), 0, s, part, nblk, out);
    return hipGetLastError();
}
hipError_t configure_kernels(size_t max_lds) {
    const void* ks[] = {(const void*)k_small_fft, (const void*)k_psf_rows, (const void*)k_psf_cols,
                        (const void*)k_ifft_cols, (const void*)k_ifft_rows_sum, (const void*)k_ifft_rows_plane,
                        (const void*)k_mac<4>, (const void*)k_mac<2>};
    for (const void* k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_fill(float* p, long long n, float v, hipStream_t s) {
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace flf
