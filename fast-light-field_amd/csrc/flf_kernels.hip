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
#include <algorithm>

#include "flf_common.h"
#include "flf_fft.h"
#include "flf_kernels.h"
#include "flf_wfft.h"

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
// Small FFTs of the lenslet-subsampled source (py_light_field.cpp:1001-1027):
// S_{b,a} = FFT_r(src[b::N, a::N] zero-padded to rY x rX), half spectrum.
// grid (N*N, T, nsrc), block 64 * SFFT_WAVES.  Rows are real, so two rows are
// transformed as one complex sequence, in place in those two rows' storage
// (2 (rX/2+1) >= rX entries; one spare row at the end), and separated
// afterwards; each wave transforms two such row pairs at once, then pairs of
// columns of the half spectrum.  The r x (r/2+1) spectrum takes 60 KB of LDS
// at r = 120, so two workgroups share a CU; at 80 VGPRs a SIMD holds 6
// waves, i.e. 12 waves per workgroup (1455^2 forward projection: small FFT
// part 1.5% faster than with 8; the kernel is bound by LDS latency).
// Output S[z][t][off][k0][k1], k1 in [0, rX/2].
// ---------------------------------------------------------------------------
#ifndef FLF_SFFT_WAVES
#define FLF_SFFT_WAVES 12
#endif
constexpr int SFFT_WAVES = FLF_SFFT_WAVES;

// Workgroup id -> work index such that the workgroups of one XCD take a
// contiguous range: the dispatcher deals consecutive workgroup ids of a grid
// row to the 8 XCDs in turn, so ids with one value of (id mod 8) share an
// XCD and its L2.  Neighbouring work items (adjacent residues, adjacent
// lenslet offsets) write neighbouring bytes, which then combine in one L2.
// A bijection on [0, n): the last n mod 8 ids map to themselves.
#ifndef FLF_XCD_REMAP
#define FLF_XCD_REMAP 1
#endif
__device__ __forceinline__ int xcd_block(int bid, int n) {
#if FLF_XCD_REMAP
    const int q = n >> 3;
    return bid < 8 * q ? (bid & 7) * q + (bid >> 3) : bid;
#else
    (void)n;
    return bid;
#endif
}
// Grouped launches: the grid groups of one projection in one dispatch.  A
// flattened block index b falls in group g for start[g] <= b < start[g + 1];
// the group is found with static indices only (selects on kernel-argument
// values); its arguments are then read from the kernel arguments with g as
// a scalar index.
struct GroupPick {
    int g, rem, n, nch, nr;                       // n: blocks of the group (start[g + 1] - start[g])
};
__device__ __forceinline__ GroupPick pick_group(const GroupedGrid& gg, int b) {
    GroupPick p{0, b, gg.start[1], gg.nchunks[0], gg.nranges[0]};
    int s0 = 0;
#pragma unroll
    for (int k = 1; k < FLF_MAX_GROUPS; ++k)
        if (k < gg.ngroups && b >= gg.start[k]) {
            p.g = k;
            s0 = gg.start[k];
            p.n = gg.start[k + 1] - gg.start[k];
            p.nch = gg.nchunks[k];
            p.nr = gg.nranges[k];
        }
    p.rem = b - s0;
    return p;
}

// One (lenslet offset, frame, source plane) of a grid group; src0/src2/
// tsplit: the source buffers of the call.
template <bool BIG>
__device__ __forceinline__ void small_fft_body(const SmallFFTArgs& a, const float* src0, const float* src2,
                                               int tsplit, int off, int t, int z) {
    constexpr int RMAX = BIG ? 256 : 128;          // reduced grid length (BIG: 129..256)
    extern __shared__ float2 lds[];
    const int N = a.N;
    const int b = off / N, aa = off % N;
    const int rY = a.rpY.L, rX = a.rpX.L, rXh = rX / 2 + 1;
    float2* H = lds;                              // [rY + 1][rXh]
    float2* twY = H + (rY + 1) * rXh;             // [rY]
    float2* twX = twY + rY;                       // [rX]
    float2* phY = twX + rX;                       // [rY]
    float2* phX = phY + rY;                       // [rX]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < rY; i += blockDim.x) twY[i] = a.twY[i];
    for (int i = threadIdx.x; i < rX; i += blockDim.x) twX[i] = a.twX[i];
    const bool second = src2 && t >= tsplit;      // two-source calls: frames from two buffers
    const float* src = (second ? src2 : src0) + a.src_off[z] + (long long)(second ? t - tsplit : t) * a.src_tstride;
    // frame t's valid region (frames of one batch may be smaller than the plan)
    const int Yv = a.fdims ? a.fdims[t].x : a.Y, Xv = a.fdims ? a.fdims[t].y : a.X;
    const int Ys = max(0, (Yv - b + N - 1) / N);
    const int Xs = max(0, (Xv - aa + N - 1) / N);
    for (int i = Ys * rXh + threadIdx.x; i < rY * rXh; i += blockDim.x) H[i] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int q = wv; 4 * q < Ys; q += SFFT_WAVES) {
        const int y0 = 4 * q;
        const int np = (y0 + 2 < Ys) ? 2 : 1;      // row pairs in this group of four rows
        float2* P = H + (long long)y0 * rXh;      // pair s at P + 2 s rXh
        // All loads of the group's four rows are issued before any is used.
        constexpr int XC = RMAX / 64;              // 64-lane column chunks per row
        float v[2][XC][2];
#pragma unroll
        for (int sq = 0; sq < 2; ++sq) {
            const int ya = y0 + 2 * sq, yb = ya + 1;
            const float* r0 = src + (long long)(b + N * ya) * a.src_pitch + aa;
            const float* r1 = src + (long long)(b + N * yb) * a.src_pitch + aa;
#pragma unroll
            for (int c = 0; c < XC; ++c) {
                const int x = lane + 64 * c;
                const bool in = sq < np && x < Xs;
                v[sq][c][0] = in ? r0[N * x] : 0.f;
                v[sq][c][1] = (in && yb < Ys) ? r1[N * x] : 0.f;
            }
        }
#pragma unroll
        for (int sq = 0; sq < 2; ++sq)
#pragma unroll
            for (int c = 0; c < XC; ++c) {
                const int x = lane + 64 * c;
                if (sq < np && x < rX) P[2 * sq * rXh + x] = make_float2(v[sq][c][0], v[sq][c][1]);
            }
        wave_sync();
        wave_fft_ns<RMAX, 2>(P, 1, 2 * rXh, np, a.rpX, twX, false, lane);
        // Separate: row ya = (Z + Zc)/2, row yb = (Z - Zc)/(2i), Zc = conj Z(-k);
        // all reads before the in-place writes.
        float2 ev[2][2], od[2][2];
#pragma unroll
        for (int sq = 0; sq < 2; ++sq)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int k = lane + 64 * m;
                if (sq < np && k < rXh) {
                    const float2* Z = P + 2 * sq * rXh;
                    const float2 zk = Z[k];
                    const float2 zc = cconj(Z[(rX - k) % rX]);
                    ev[sq][m] = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
                    od[sq][m] = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
                }
            }
        wave_sync();
#pragma unroll
        for (int sq = 0; sq < 2; ++sq)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int k = lane + 64 * m;
                if (sq < np && k < rXh) {
                    const int ya = y0 + 2 * sq;
                    H[ya * rXh + k] = ev[sq][m];
                    H[(ya + 1) * rXh + k] = (ya + 1 < Ys) ? od[sq][m] : make_float2(0.f, 0.f);
                }
            }
        wave_sync();
    }
    __syncthreads();
    for (int c = 2 * wv; c < rXh; c += 2 * SFFT_WAVES)
        wave_fft_ns<RMAX, 2>(H + c, rXh, 1, min(2, rXh - c), a.rpY, twY, false, lane);
    // Sub-sampling phases of this offset: exp(-2 pi i b kt0 / fY) and
    // exp(-2 pi i a kt1 / fX), kt the centred residue (0 for r/2).
    __syncthreads();
    for (int i = threadIdx.x; i < rY + rX; i += blockDim.x) {
        const bool isY = i < rY;
        const int k = isY ? i : i - rY;
        const int r = isY ? rY : rX;
        const int f = isY ? a.fY : a.fX;
        const int kt = (2 * k < r) ? k : ((2 * k > r) ? k - r : 0);
        const int m = ((isY ? b : aa) * kt) % f;
        float sn, cs;
        sincospif(-2.0f * (float)m / (float)f, &sn, &cs);
        (isY ? phY : phX)[k] = make_float2(cs, sn);
    }
    __syncthreads();
    // Orbit-ordered, phase-folded output S[chunk][off][j] for the first NH
    // residues of each chunk's residue list (the MAC reads one contiguous
    // block per chunk).
    float2* S = a.S + a.S_off[z] + (long long)t * a.S_tstride;
    // OU entries per thread at a time: their residue loads (chunk
    // descriptors, from the caches) are all in flight before the first is
    // used.
    const int NH = a.NH, NN = N * N;
    const int total = a.nchunks * NH;
    constexpr int OU = 4;
    for (int i0 = threadIdx.x; i0 < total; i0 += OU * blockDim.x) {
        int k0[OU], k1[OU];
#pragma unroll
        for (int u = 0; u < OU; ++u) {
            const int i = min(i0 + u * (int)blockDim.x, total - 1);
            const int w = i / NH, j = i - w * NH;
            k0[u] = a.chunks[w].res0[j];
            k1[u] = a.chunks[w].res1[j];
        }
#pragma unroll
        for (int u = 0; u < OU; ++u) {
            const int i = i0 + u * (int)blockDim.x;
            if (i < total) {
                const int w = i / NH, j = i - w * NH;
                float2 v = (2 * k1[u] <= rX) ? H[k0[u] * rXh + k1[u]]
                                             : cconj(H[((rY - k0[u]) % rY) * rXh + (rX - k1[u])]);
                v = cmul(v, cmul(phY[k0[u]], phX[k1[u]]));
                S[((long long)w * NN + off) * NH + j] = v;
            }
        }
    }
}

template <bool BIG>
__global__ __launch_bounds__(64 * SFFT_WAVES) void k_small_fft(SmallFFTArgs a) {
    small_fft_body<BIG>(a, a.src, a.src2, a.tsplit, xcd_block(blockIdx.x, a.N * a.N), blockIdx.y, blockIdx.z);
}
// grid (N*N, T, sum over groups of their source planes): z flattened over
// (group, plane).
__global__ __launch_bounds__(64 * SFFT_WAVES) void k_small_fft_grouped(SmallFFTGrouped ga) {
    const GroupPick p = pick_group(ga.gg, blockIdx.z);
    const SmallFFTArgs& a = ga.a[p.g];
    small_fft_body<false>(a, a.src, a.src2, a.tsplit, xcd_block(blockIdx.x, a.N * a.N), blockIdx.y, p.rem);
}

// ---------------------------------------------------------------------------
// PSF spectra, stage 1: row FFTs of the s2 non-zero rows of each quadrant PSF
// (zero padded to fX), half plane kept.  (py_light_field.cpp:757-812 computes
// the same pruned 2-D transform column-first on the full plane.)
// grid (ceil(s2/RB), nu); LDS RB*fX float2.
// ---------------------------------------------------------------------------
template <bool PRIME>
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
    lds_fft<FFT_MAXB, PRIME>(sh, a.rpX, nrows, 1, fX, -1.f);
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
template <bool PRIME>
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
    lds_fft<FFT_MAXB, PRIME>(sh, a.rpY, a.CB, 1, fY, -1.f);
    float2* out = a.spec + (long long)u * fY * rfX;
    for (int i = threadIdx.x; i < fY * a.CB; i += blockDim.x) {
        int y = i / a.CB, c = i - y * a.CB;
        if (c < ncol) out[(long long)y * rfX + c0 + c] = sh[c * fY + y];
    }
}

// ---------------------------------------------------------------------------
// Gather the natural-layout spectra of one plane into the orbit-major layout
// G[chunk][plane][u][slot pair][item][2], each value pre-multiplied by
// w^(bu*m2 + au*n2), w = exp(-2 pi i / (2N)) (m2, n2: half-unit aliases).
// grid (nchunks, nu), block FLF_CHUNK_MAX.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(FLF_CHUNK_MAX) void k_spec_pack(SpecPackArgs a) {
    const int cidx = xcd_block(blockIdx.x, gridDim.x);
    const ChunkDesc* ch = a.chunks + cidx;
    const int u = blockIdx.y;
    const int i = threadIdx.x;
    const int n_items = ch->n_items;
    if (i >= n_items) return;
    const long long item_base = ch->item_base;
    const ItemDesc it = a.items[item_base + i];
    const int bu = a.u_bu[u], au = a.u_au[u];
    const int sig = a.s2 - 2 + a.N;
    const int m2 = it.m2[0], n2 = it.n2[0];
    // rho_i^{-1/2}: the phase ratio of the two uses reading slot i, the same
    // for all output slots of the item (exponents in 1/(4N) turns).
    int rho[4];
    rho[0] = sig * (m2 + n2);
    rho[1] = sig * (n2 - m2);
    rho[2] = sig * (m2 + n2);
    rho[3] = sig * (m2 - n2);
    const float2* spec = a.spec + (long long)u * a.fY * a.rfX;
    // Slots (0,1) of all items, then slots (2,3): a wave's 16-byte loads of
    // one slot pair cover whole cache lines.
    float2* G = a.G + item_base * a.slots * a.nu * a.nP + ((long long)a.p * a.nu + u) * n_items * a.slots;
    for (int j = 0; j < a.slots; ++j) {
        float2 v = spec[it.pos[j]];
        long long e = 2LL * ((long long)bu * it.m2[j] + (long long)au * it.n2[j]) - rho[j];
        G[(long long)(j >> 1) * n_items * 2 + 2 * i + (j & 1)] = cmul(v, phase_frac(e, 4 * a.N));
    }
}

// ---------------------------------------------------------------------------
// The multiply-accumulate over the N*N lenslet offsets (k_mac below).
// ---------------------------------------------------------------------------
namespace {

// Use types (order of the per-spectrum use table): 0 id, 1 both flips,
// 2 row flip, 3 column flip, 4..7 the same for the transposed PSF.
// The two uses that read the same slot of an item always pair a plain and a
// conjugated read; their pair types (plain, conj) are
//   quad:  0 (id, both)  1 (row, col)  2 (Tid, Tboth)  3 (Tcol, Trow)
//   pair:  0 (id, both)  1 (row, col)
__device__ constexpr int kPairPlain[4] = {0, 2, 4, 7};
__device__ constexpr int kPairConj[4] = {1, 3, 5, 6};
// Output slot j, source slot i -> pair type and sign of the M term.
__device__ constexpr int kQuadPT[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
__device__ constexpr int kQuadSG[4][4] = {{1, 1, 1, 1}, {1, 1, -1, -1}, {1, 1, 1, 1}, {-1, -1, 1, 1}};
__device__ constexpr int kPairPT[2][2] = {{0, 1}, {1, 0}};

typedef float v4f __attribute__((ext_vector_type(4)));

// Spectra are streamed once per projection.  Plain loads measured 10% faster
// than non-temporal ones on MI355X (FLF_MAC_NT=1 selects the latter).
#ifndef FLF_MAC_NT
#define FLF_MAC_NT 0
#endif
__device__ __forceinline__ v4f load_stream(const v4f* q) {
#if FLF_MAC_NT
    return __builtin_nontemporal_load(q);
#else
    return *q;
#endif
}
// Item ti of one (chunk, plane, spectrum) block of n_items items: slots
// (0,1) at ti, slots (2,3) n_items entries further.
template <int SLOTS>
__device__ __forceinline__ void load_g(float2 (&g)[4], const float2* blk, int ti, int n_items) {
    const v4f* q = reinterpret_cast<const v4f*>(blk) + ti;
    v4f x = load_stream(q);
    g[0] = make_float2(x.x, x.y);
    g[1] = make_float2(x.z, x.w);
    if constexpr (SLOTS == 4) {
        v4f y = load_stream(q + n_items);
        g[2] = make_float2(y.x, y.y);
        g[3] = make_float2(y.z, y.w);
    } else {
        g[2] = g[3] = make_float2(0.f, 0.f);
    }
}

}  // namespace

// acc[a][j] += P * Re(g[a][i]) + sg * M * Im(g[a][i]) for every output slot
// j, source slot i and item a of the thread (the items share residues, so
// each 16-byte LDS entry (P, M) serves all of them).
// P = s_plain + s_conj, M = i (s_plain - s_conj).
template <int SLOTS, int IPT, int FB>
__device__ __forceinline__ void mac_spectrum(float2 (&acc)[FB][IPT][4], const float2 (&g)[IPT][4],
                                             const float4* __restrict__ pm, int tab, const int (&rb)[4], int u) {
    constexpr int NPT = SLOTS == 4 ? 4 : 2;
#pragma unroll
    for (int f = 0; f < FB; ++f) {
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) {
            const float4* row = pm + f * tab + rb[j] + u * NPT;
#pragma unroll
            for (int i = 0; i < SLOTS; ++i) {
                const int pt = SLOTS == 4 ? kQuadPT[j][i] : kPairPT[j][i];
                const float sg = SLOTS == 4 ? (float)kQuadSG[j][i] : 1.0f;
                const float4 q = row[pt];          // (P.x, P.y, M.x, M.y)
                const float mx = sg * q.z, my = sg * q.w;
#pragma unroll
                for (int a = 0; a < IPT; ++a) {
                    acc[f][a][j].x = fmaf(q.x, g[a][i].x, acc[f][a][j].x);
                    acc[f][a][j].y = fmaf(q.y, g[a][i].x, acc[f][a][j].y);
                    acc[f][a][j].x = fmaf(mx, g[a][i].y, acc[f][a][j].x);
                    acc[f][a][j].y = fmaf(my, g[a][i].y, acc[f][a][j].y);
                }
            }
        }
    }
}

#ifndef FLF_MAC_UNR
#define FLF_MAC_UNR 2
#endif
#ifndef FLF_MAC_UNR2
#define FLF_MAC_UNR2 1
#endif
#ifndef FLF_MAC_UNR2_FB1
#define FLF_MAC_UNR2_FB1 2
#endif
#ifndef FLF_MAC_WAVES
#define FLF_MAC_WAVES 4
#endif

// Wave-uniform read of a per-plane value through the scalar cache (the array
// is read-only while the kernel runs).  Scalar loads keep such reads off the
// vector-memory counter, so they never make the accumulation wait for the
// spectrum loads in flight.
// The value is moved to scalar registers at once (readfirstlane), so a load
// that had to go through a vector register is consumed on every path.
__device__ __forceinline__ int const_load(const int* base, int i) {
    return __builtin_amdgcn_readfirstlane(((const __attribute__((address_space(4))) int*)base)[i]);
}
__device__ __forceinline__ long long const_load(const long long* base, int i) {
    const long long v = ((const __attribute__((address_space(4))) long long*)base)[i];
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffLL));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}
// s_waitcnt vmcnt(0) issued through the builtin, so that the compiler's own
// wait counting knows every vector-memory access has completed.
__device__ __forceinline__ void wait_vmem_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Copy of the chunk's small-FFT blocks of plane p, frames t0 .. t0+nf-1, into
// s_stage ([f][NN][NH]) by global-to-LDS loads: every load of the block is in
// flight at once and no registers are held; completion is awaited (vmcnt) by
// mac_tables.  (Started earlier, during the previous plane's accumulation,
// the copy would make the compiler drain all vector loads before every LDS
// read of the accumulation.)  The block size is a multiple of 64 threads.
template <int SLOTS>
__device__ __forceinline__ void mac_stage_async(const MacArgs& a, int cidx, int p, int t0, int nf, float2* s_stage) {
    constexpr int NH = SLOTS == 4 ? 4 : 2;
    const int blk = a.N * a.N * NH;              // float2 per frame block
    const int nd = 2 * blk;                       // dwords per frame block
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const long long sbase = const_load(a.S_off, p) + (long long)t0 * a.S_tstride + (long long)cidx * blk;
    for (int f = 0; f < nf; ++f) {
        const float* src = reinterpret_cast<const float*>(a.S + sbase + f * a.S_tstride);
        float* dst = reinterpret_cast<float*>(s_stage + f * blk);
        for (int c = wv * 64; c < nd; c += nw * 64)
            if (c + lane < nd)
                __builtin_amdgcn_global_load_lds(src + c + lane, (__attribute__((address_space(3))) void*)(dst + c), 4,
                                                 0, 0);
    }
}

// Per-plane set-up of a MAC workgroup: flip factors, then per frame f < nf
// the P/M table of the chunk's small-FFT block (staged in LDS beforehand)
//   pm[f][rho * rs + u * NPT + pt] = (P, M),  P = s_plain + s_conj,
//   M = i (s_plain - s_conj)
// for residue rho, stored spectrum u and pair type pt (all threads; contains
// barriers; frames past nf keep the tables of earlier planes and their
// accumulators are never stored).
template <int SLOTS>
__device__ __forceinline__ void mac_tables(const MacArgs& a, const ChunkDesc* ch, int cidx, int p, int t0, int nf,
                                           float4* pm_tab, float2* s_stage, float2* psi, int rs) {
    constexpr int NUSE = SLOTS == 4 ? 8 : 4;
    constexpr int NPT = SLOTS == 4 ? 4 : 2;
    constexpr int NRES = SLOTS == 4 ? 8 : 4;
    constexpr int NH = NRES / 2;
    const int tid = threadIdx.x;
    const int N = a.N, NN = N * N, c = N / 2;
    const int nu = a.nu;
    const int tab = NRES * rs;
    const int s2 = const_load(a.s2, p);
    // flip factors per residue: exp(-2 pi i (s2-1) kt / f) (psi was last read
    // before the closing barrier of the previous plane's set-up)
    if (tid < 2 * NRES) {
        const int rho = tid >> 1;
        const bool isY = (tid & 1) == 0;
        const int k = isY ? ch->res0[rho] : ch->res1[rho];
        const int r = isY ? a.rY : a.rX;
        const int f = isY ? a.fY : a.fX;
        const int kt = (2 * k < r) ? k : ((2 * k > r) ? k - r : 0);
        int m = (int)(((long long)(s2 - 1) * kt) % f);
        if (m < 0) m += f;
        float sn, cs;
        sincospif(-2.0f * (float)m / (float)f, &sn, &cs);
        psi[tid] = make_float2(cs, sn);
    }
    // Copy the small-FFT blocks of all nf frames into s_stage, wait for the
    // copy, then build every frame's table (the barrier also orders these
    // writes of pm_tab after every wave's reads of the previous plane's
    // table).
    const int blk = NN * NH;
    mac_stage_async<SLOTS>(a, cidx, p, t0, nf, s_stage);
    wait_vmem_all();
    __syncthreads();
    {
        // Thread -> table column q = (u, pt) and a share of the rows
        // (f, rho): the column's two source offsets and flip flags are
        // decoded once, the rows then cost two LDS reads and the P/M
        // arithmetic each (no integer division per entry).
        const int Q = nu * NPT;
        const int nt = blockDim.x;
        const bool fits = Q <= nt;
        const int rstep = fits ? nt / Q : 1;
        const int r0 = fits ? tid / Q : 0;
        const int rows = nf * NRES;
        for (int q = (fits ? (r0 < rstep ? tid - r0 * Q : Q) : tid); q < Q; q += nt) {
            const int pt = q % NPT, u = q / NPT;
            int src[2], flip[2];
#pragma unroll
            for (int w = 0; w < 2; ++w) {
                const int off = a.uoff[u * NUSE + (w == 0 ? kPairPlain[pt] : kPairConj[pt])];
                const int bo = off / N, ao = off - bo * N;
                src[w] = off < NN ? off * NH : -1;
                flip[w] = (bo > c ? 1 : 0) | (ao > c ? 2 : 0);
            }
            for (int row = r0; row < rows; row += rstep) {
                const int f = row / NRES, rho = row % NRES;
                const float2* sf = s_stage + f * blk;
                const int rb0 = rho < NH ? rho : rho - NH;
                float2 sv[2];
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    float2 r = make_float2(0.f, 0.f);
                    if (src[w] >= 0) {
                        r = sf[src[w] + rb0];
                        if (rho >= NH) r = cconj(r);
                        if (flip[w] & 1) r = cmul(r, psi[2 * rho]);
                        if (flip[w] & 2) r = cmul(r, psi[2 * rho + 1]);
                    }
                    sv[w] = r;
                }
                const float2 P = cadd(sv[0], sv[1]);
                const float2 d = csub(sv[0], sv[1]);
                pm_tab[f * tab + rho * rs + u * NPT + pt] = make_float4(P.x, P.y, -d.y, d.x);   // M = i * d
            }
        }
    }
    __syncthreads();
}

// For an item with base alias (m2, n2) (half units) and output slots
// P0..P3, every output accumulates
//   A_j = w4^(outer_j) * sum_u sum_i [ P_{j,i,u} Re g_i + M_{j,i,u} Im g_i ]
// with g_i the stored spectrum value of slot i (pre-multiplied by
// w2^(bu m2_i + au n2_i) and rho_i^{-1/2}), w4 = exp(-2 pi i/(4N)),
// outer = sig*(m2+n2), sig*(n2-m2), sig*(n2+m2), sig*(m2-n2), sig = s2-2+N.
// P and M combine the small-FFT values s of the two uses reading slot i at
// output j, s = S_{b,a}(kappa) exp(-2 pi i (beta_b kt0/fY + alpha_a kt1/fX)),
// beta = b (b <= N/2) or b + s2 - 1 (row-flipped offsets), likewise alpha.
// This is the accumulation of py_light_field.cpp:303-341, 1029-1087 with the
// mirror multipliers of projector.py:137-151 and the 'same' crop
// (projector.py:792-796) folded into phases.  IPT = 1: one item per
// thread; IPT = 2 (4): a lane pair (quad) owns two (four) consecutive items
// that share their residues and splits the spectra between its lanes.  FB frames per
// workgroup: every stored spectrum value loaded serves FB frames, and the
// frame blocks of one chunk run side by side on one XCD (block ids below).
// One workgroup: chunk cidx, plane range `range` of nranges, frame block
// fblock.
template <int SLOTS, int IPT, int FB>
__device__ __forceinline__ void mac_body(const MacArgs& a, int cidx, int range, int nranges, int fblock) {
    extern __shared__ float4 pm_tab[];            // [FB][NRES][rs: nu*NPT + pad], then staging
    constexpr int NRES = SLOTS == 4 ? 8 : 4;      // residues per orbit (canonical list)
    constexpr int NH = NRES / 2;                  // residues stored by the small-FFT kernel
    // spectra whose loads are issued together
    constexpr int UNR = IPT >= 2 ? (FB == 1 ? FLF_MAC_UNR2_FB1 : FLF_MAC_UNR2) : FLF_MAC_UNR;
    const ChunkDesc* ch = a.chunks + cidx;
    const long long item_base = ch->item_base;
    const int n_items = ch->n_items;
    const int tid = threadIdx.x;
    // IPT = 2 or 4: a group of IPT lanes owns IPT items that share their
    // residues and splits the stored spectra between its lanes (lane lsub
    // takes u = lsub, lsub + IPT, ...), so every P/M table read serves IPT
    // items while the workgroup keeps one thread per item slot; the group
    // adds its partial sums before each plane's epilogue.
    static_assert(IPT == 1 || IPT == 2 || IPT == 4, "items per lane group: 1, 2 or 4");
    constexpr int USTEP = IPT;
    const int lsub = tid % USTEP;
    const int slot = tid / USTEP;
    const bool valid = slot * IPT < n_items;
    const int tsafe = valid ? slot : 0;          // idle lanes read item 0 (in bounds) and never write
    const int N = a.N, NN = N * N, fourN = 4 * N;
    const int t0 = fblock * FB;
    const int nf = min(FB, a.T - t0);            // frames of this block (block-uniform)
    const int nu = a.nu;
    // Residue rows are a.rs = nu*NPT + pad entries apart, rs = 1 (mod 16):
    // threads of one wave read the same (u, pair type) entry of up to 8
    // different residues, which then fall on distinct LDS banks.
    const int rs = a.rs;
    const int tab = NRES * rs;                    // table entries per frame
    float2* s_stage = reinterpret_cast<float2*>(pm_tab + FB * tab);   // [FB][NN][NH]
    float2* psi = s_stage + FB * NN * NH;                                   // [NRES][2]
    // The item descriptors are re-read (from the caches) where they are
    // needed instead of living in registers across the spectrum loop.
    const ItemDesc* itp = a.items + item_base + (long long)tsafe * IPT;
    int rb[4];
    {
        const ItemDesc i0 = itp[0];
#pragma unroll
        for (int j = 0; j < 4; ++j) rb[j] = i0.res[j] * rs;
    }
    // The descriptor pointer plus an opaque zero, so that the compiler does
    // not hoist the epilogue's loads out of the plane loop.  The pointer
    // itself stays transparent: the compiler knows it addresses global
    // memory and emits global loads (FLAT loads would make every later
    // vector-memory wait drain all loads in flight, the next spectrum batch
    // included).
    auto items_now = [&]() {
        int z = 0;
        asm volatile("" : "+v"(z));
        return itp + z;
    };

    constexpr int QT = 1;                         // items whose totals a lane keeps
    float2 tot[FB][QT][4];
#pragma unroll
    for (int f = 0; f < FB; ++f)
#pragma unroll
        for (int q = 0; q < QT; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) tot[f][q][j] = make_float2(0.f, 0.f);

    // Planes of this workgroup: an even share of the group's nP planes for
    // nranges plane ranges (nranges = nP gives one plane per workgroup).
    const int pb = (int)(((long long)range * a.nP) / nranges);
    const int pe = (int)(((long long)(range + 1) * a.nP) / nranges);
    const long long ustride = (long long)n_items * SLOTS;
    const int nul = (nu + USTEP - 1) / USTEP;     // spectra per lane and plane
    const int nb = (nul + UNR - 1) / UNR;         // spectrum batches per plane

    // Per-plane set-up: flip factors, then per frame stage the chunk's
    // small-FFT block and build its P/M table (all threads; contains
    // barriers; frames past nf keep the tables of earlier planes and their
    // accumulators are never stored).
    auto prologue = [&](int p) {
        mac_tables<SLOTS>(a, ch, cidx, p, t0, nf, pm_tab, s_stage, psi, rs);
    };

    auto gbase = [&](int p) { return a.G + item_base * SLOTS * nu * a.nP + ((long long)p * nu) * ustride; };
    // Loads of one batch; spectra past nu re-read the last one (in bounds, unused).
    auto load_batch = [&](float2 (&g)[UNR][IPT][4], int p, int b) {
        const float2* Gp = gbase(p);
#pragma unroll
        for (int k = 0; k < UNR; ++k) {
            const int u = min((b * UNR + k) * USTEP + lsub, nu - 1);
#pragma unroll
            for (int q = 0; q < IPT; ++q) load_g<SLOTS>(g[k][q], Gp + u * ustride, tsafe * IPT + q, n_items);
        }
    };

    // Store the totals at A + slice * A_pstride (each lane of a pair stores
    // its own item), frames of this block only.
    auto store = [&](int slice) {
        const auto ip = items_now();
#pragma unroll
        for (int f = 0; f < FB; ++f) {
            if (f >= nf) break;
            float2* A = a.A + (long long)(t0 + f) * a.A_tstride + (long long)slice * a.A_pstride;
#pragma unroll
            for (int q = 0; q < QT; ++q) {
                const ItemDesc& it = ip[IPT >= 2 ? lsub : q];
#pragma unroll
                for (int j = 0; j < SLOTS; ++j)
                    if ((it.writemask >> j) & 1) A[it.pos[j]] = tot[f][q][j];
            }
        }
    };

    float2 acc[FB][IPT][4];
    // Epilogue of plane p: outer phase and the 'same'-crop shift.
    auto epilogue = [&](int p) {
        const int s2 = const_load(a.s2, p);
        const int sig = s2 - 2 + N;
        const int c2 = (s2 - 1) / 2;
        // IPT = 2: each lane keeps its own item (q = lsub): it sends the
        // partner's item and adds what the partner sends for its own.
        // IPT = 4: the same in two steps, first between lanes lsub and
        // lsub ^ 2 (each keeps the two items whose bit 1 is its own), then
        // between lsub and lsub ^ 1.
        float2 own[FB][4];
        if constexpr (IPT == 2) {
#pragma unroll
            for (int f = 0; f < FB; ++f)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float2 mine = lsub ? acc[f][1][j] : acc[f][0][j];
                    const float2 give = lsub ? acc[f][0][j] : acc[f][1][j];
                    own[f][j] = make_float2(mine.x + __shfl_xor(give.x, 1), mine.y + __shfl_xor(give.y, 1));
                }
        } else if constexpr (IPT == 4) {
            const bool hi = (lsub & 2) != 0, odd = (lsub & 1) != 0;
#pragma unroll
            for (int f = 0; f < FB; ++f)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float2 k[2];
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const float2 keep = hi ? acc[f][2 + c][j] : acc[f][c][j];
                        const float2 give = hi ? acc[f][c][j] : acc[f][2 + c][j];
                        k[c] = make_float2(keep.x + __shfl_xor(give.x, 2), keep.y + __shfl_xor(give.y, 2));
                    }
                    const float2 mine = odd ? k[1] : k[0];
                    const float2 give = odd ? k[0] : k[1];
                    own[f][j] = make_float2(mine.x + __shfl_xor(give.x, 1), mine.y + __shfl_xor(give.y, 1));
                }
        }
        const auto ip = items_now();
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            const ItemDesc itq = ip[IPT >= 2 ? lsub : q];
            const int m2 = itq.m2[0], n2 = itq.n2[0];
            int outer[4];
            outer[0] = sig * (m2 + n2);
            outer[1] = sig * (n2 - m2);
            outer[2] = sig * (n2 + m2);
            outer[3] = sig * (m2 - n2);
#pragma unroll
            for (int j = 0; j < SLOTS; ++j) {
                const int k0 = itq.pos[j] / a.rfX, k1 = itq.pos[j] - k0 * a.rfX;
                const float2 po = phase_frac_i(outer[j], fourN);
                // exp(+2 pi i c2 (k0/fY + k1/fX))
                // c2 < f / 2 and k < f <= FLF_MAX_FOURIER: 32-bit products
                const int e0 = (c2 * k0) % a.fY;
                const int e1 = (c2 * k1) % a.fX;
                float fr = (float)e0 / (float)a.fY + (float)e1 / (float)a.fX;
                fr -= rintf(fr);
                float sn, cs;
                sincospif(2.0f * fr, &sn, &cs);
                const float2 pc = make_float2(cs, sn);
#pragma unroll
                for (int f = 0; f < FB; ++f) {
                    const float2 av = IPT >= 2 ? own[f][j] : acc[f][q][j];
                    tot[f][q][j] = cadd(tot[f][q][j], cmul(cmul(av, po), pc));
                }
            }
        }
        if (a.per_plane) {                       // back projection: one output per plane
            if (valid) store(p);
#pragma unroll
            for (int f = 0; f < FB; ++f)
#pragma unroll
                for (int q = 0; q < QT; ++q)
#pragma unroll
                    for (int j = 0; j < 4; ++j) tot[f][q][j] = make_float2(0.f, 0.f);
        }
    };

    // Software pipeline over (plane, batch): the loads of the next batch (also
    // across a plane boundary, i.e. before the next plane's set-up) are in
    // flight while the current batch is accumulated.
    int p = pb;
    while (p < pe && !a.enable.on(p)) ++p;
    if (p < pe) {
        float2 gA[UNR][IPT][4], gB[UNR][IPT][4];
        load_batch(gA, p, 0);
        int b = 0;
        auto step = [&](float2 (&cur)[UNR][IPT][4], float2 (&nxt)[UNR][IPT][4]) -> bool {
            // The current batch (and anything older) has landed: the only
            // loads in flight during the accumulation below are the next
            // batch's and the staging copy (which writes no registers).
            wait_vmem_all();
            if (b == 0) {
                prologue(p);
#pragma unroll
                for (int f = 0; f < FB; ++f)
#pragma unroll
                    for (int q = 0; q < IPT; ++q)
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[f][q][j] = make_float2(0.f, 0.f);
            }
            int pn = p, bn = b + 1;
            if (bn == nb) {
                bn = 0;
                pn = p + 1;
                while (pn < pe && !a.enable.on(pn)) ++pn;
            }
            if (pn < pe) load_batch(nxt, pn, bn);
#pragma unroll
            for (int k = 0; k < UNR; ++k) {
                const int u = (b * UNR + k) * USTEP + lsub;
                if (u < nu) mac_spectrum<SLOTS, IPT, FB>(acc, cur[k], pm_tab, tab, rb, u);
            }
            if (b == nb - 1) epilogue(p);
            p = pn;
            b = bn;
            return pn < pe;
        };
        while (step(gA, gB) && step(gB, gA)) {
        }
    }

    if (valid && !a.per_plane) store(range);
}

// Block id w in [0, R * nfb) -> frame block and the rest r in [0, R):
// frame blocks in super-blocks of S (the last one possibly shorter), and
// within a super-block the frame block fastest.  S = nfb: all frame blocks of
// an r consecutive; S = 1: frame block slowest.
__device__ __forceinline__ void split_frames(int w, int R, int nfb, int S, int& fblock, int& r) {
    S = (S <= 0 || S > nfb) ? nfb : S;
    const int sb = w / (S * R);
    const int sc = min(S, nfb - sb * S);
    const int rem = w - sb * S * R;
    fblock = sb * S + rem % sc;
    r = rem / sc;
}

template <int SLOTS, int IPT, int FB>
__global__ __launch_bounds__(FLF_CHUNK_MAX) __attribute__((amdgpu_waves_per_eu(IPT * FB >= 3 ? 2 : (IPT * FB == 2 ? 3 : FLF_MAC_WAVES))))
void k_mac(MacArgs a) {
    // grid x = chunk x frame block, frame blocks of a chunk consecutive (in
    // super-blocks of a.fgroup): after the XCD remap they share one XCD's L2
    // for the spectra they all read
    const int nfb = (a.T + FB - 1) / FB;
    const int w = xcd_block(blockIdx.x, gridDim.x);
    int fblock, cidx;
    split_frames(w, gridDim.x / nfb, nfb, a.fgroup, fblock, cidx);
    mac_body<SLOTS, IPT, FB>(a, cidx, blockIdx.y, gridDim.y, fblock);
}

// All active grid groups of one projection: grid x flattened over (group,
// plane range, chunk, frame block), frame block fastest.  Each group's range
// starts at a multiple of 8 (padded with idle workgroups), and the XCD remap
// applies within the group: every XCD then takes an eighth of every group,
// whose workgroups cost about the same, where one remap over the whole grid
// would give some XCDs only a cheap group's workgroups and leave them idle
// early.  After the remap the frame blocks of one (chunk, plane range) are
// consecutive ids of one XCD, dispatched together: they read the same
// stored spectra at about the same time, which the XCD's L2 then serves
// once from HBM for all of them.  fgroup < nfb dispatches them in
// super-blocks of fgroup instead, so that more chunks run side by side
// (their scattered accumulator stores then combine in L2 more often).
template <int SLOTS, int IPT, int FB>
__global__ __launch_bounds__(FLF_CHUNK_MAX) __attribute__((amdgpu_waves_per_eu(IPT * FB >= 3 ? 2 : (IPT * FB == 2 ? 3 : FLF_MAC_WAVES))))
void k_mac_grouped(MacGrouped ga) {
    const GroupPick p = pick_group(ga.gg, blockIdx.x);
    const int w = xcd_block(p.rem, p.n);
    const int R = p.nch * p.nr;
    if (w >= R * ga.nfb) return;                  // padding
    int fblock, r;
    split_frames(w, R, ga.nfb, ga.a[p.g].fgroup, fblock, r);
    mac_body<SLOTS, IPT, FB>(ga.a[p.g], r % p.nch, r / p.nch, p.nr, fblock);
}

// Matrix-core multiply-accumulate, TC_FB = 8 frames per workgroup
// (v_mfma_f32_16x16x4_f32: FP32 operands and accumulation, as the vector
// kernel).  The pair type of source slot i at output slot j is i XOR j
// (kQuadPT, kPairPT), so with the P/M table of one residue rho as the B
// operand,
//   out_j[item][(f, c)] = sum_{u, pt, r} A_j[item][(u, pt, r)] B_rho[(u, pt, r)][(f, c)],
//   B_rho[(u, pt, Re)][(f, c)] = P_c,   B_rho[(u, pt, Im)][(f, c)] = M_c,
//   A_j[item][(u, pt, r)] = component r of the stored value of slot pt XOR j
//                           (times the M sign of (j, i) for r = Im),
// one B per residue serves every output slot and every item of that
// residue; the per-slot permutation is a choice between two registers of
// the A operand.  Items come in tiles of <= 16 of one residue class
// (TileDesc), so each slot's residue is uniform within a tile.  B lives in
// LDS in the MFMA's operand order (lane l reads float l of a 64-float block:
// no bank conflicts), built per batch of TC_UB stored spectra from the
// cached small-FFT values; lane l supplies A[l % 16][l / 16] (its tile row's
// item, read from the cached spectra: each value once per frame block) and
// holds D[4 (l / 16) + r][l % 16] (items 4 (l / 16) + r, column (f, c) =
// l % 16).  Wave w owns tiles w, w + TC_WAVES, ...
constexpr int TC_FB = 8;
constexpr int TC_WAVES = 8;
constexpr int TC_UB = 12;
template <int SLOTS>
__global__ __launch_bounds__(64 * TC_WAVES) void k_mac_tc(MacArgs a) {
    extern __shared__ float4 tc_lds[];
    constexpr int NRES = SLOTS == 4 ? 8 : 4;
    constexpr int NH = NRES / 2;
    constexpr int NPT = SLOTS == 4 ? 4 : 2;
    constexpr int KS = SLOTS / 2;                 // 4-deep k steps per stored spectrum
    constexpr int NUSE = SLOTS == 4 ? 8 : 4;
    constexpr int MAXT = (FLF_MAX_TILES + TC_WAVES - 1) / TC_WAVES;
    float* Bt = reinterpret_cast<float*>(tc_lds);                          // [TC_UB][NRES][KS][64]
    float2* psi = reinterpret_cast<float2*>(Bt + TC_UB * NRES * KS * 64);  // [NRES][2]
    float2* s_stage = psi + 2 * NRES;                                      // [TC_FB][N*N][NH]
    // grid x = chunk x frame block, frame block fastest (as k_mac)
    const int nfb = (a.T + TC_FB - 1) / TC_FB;
    const int wblk = xcd_block(blockIdx.x, gridDim.x);
    const int cidx = wblk / nfb;
    const ChunkDesc* ch = a.chunks + cidx;
    const long long item_base = ch->item_base;
    const int n_items = ch->n_items;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int col = row, f_col = col >> 1, c_col = col & 1;
    const int N = a.N, NN = N * N, c = N / 2, fourN = 4 * N;
    const int t0 = (wblk - cidx * nfb) * TC_FB;
    const int nf = min(TC_FB, a.T - t0);
    const int nu = a.nu;
    const TileDesc* tiles = a.tiles + ch->tile_base;
    const int ntiles = ch->n_tiles;
    const long long ustride = (long long)n_items * SLOTS;

    int it_of[MAXT];                              // this lane's A-row item (chunk-relative), -1 if none
#pragma unroll
    for (int q = 0; q < MAXT; ++q) {
        const int tq = wv + TC_WAVES * q;
        it_of[q] = -1;
        if (tq < ntiles) {
            const TileDesc td = tiles[tq];
            if (row < td.count) it_of[q] = td.start + row;
        }
    }
    v4f tot[MAXT][SLOTS];
#pragma unroll
    for (int q = 0; q < MAXT; ++q)
#pragma unroll
        for (int j = 0; j < SLOTS; ++j) tot[q][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // Store: lane (f, c) writes component c of its items' slot-j bins.
    auto store = [&](int slice) {
        if (f_col >= nf) return;
        float* A = reinterpret_cast<float*>(a.A + (long long)(t0 + f_col) * a.A_tstride +
                                            (long long)slice * a.A_pstride);
#pragma unroll
        for (int q = 0; q < MAXT; ++q) {
            const int tq = wv + TC_WAVES * q;
            if (tq >= ntiles) continue;                                   // wave-uniform
            const TileDesc td = tiles[tq];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ir = 4 * kq + r;
                if (ir < td.count) {
                    const ItemDesc& it = a.items[item_base + td.start + ir];
#pragma unroll
                    for (int j = 0; j < SLOTS; ++j)
                        if ((it.writemask >> j) & 1) A[2LL * it.pos[j] + c_col] = tot[q][j][r];
                }
            }
        }
    };

    const int pb = (int)(((long long)blockIdx.y * a.nP) / gridDim.y);
    const int pe = (int)(((long long)(blockIdx.y + 1) * a.nP) / gridDim.y);
    for (int p = pb; p < pe; ++p) {
        if (!a.enable.on(p)) continue;
        const int s2 = const_load(a.s2, p);
        __syncthreads();                          // the previous plane is done with Bt, psi and s_stage
        if (tid < 2 * NRES) {                     // flip factors per residue: exp(-2 pi i (s2-1) kt / f)
            const int rho = tid >> 1;
            const bool isY = (tid & 1) == 0;
            const int k = isY ? ch->res0[rho] : ch->res1[rho];
            const int r = isY ? a.rY : a.rX;
            const int f = isY ? a.fY : a.fX;
            const int kt = (2 * k < r) ? k : ((2 * k > r) ? k - r : 0);
            int m = (int)(((long long)(s2 - 1) * kt) % f);
            if (m < 0) m += f;
            float sn, cs;
            sincospif(-2.0f * (float)m / (float)f, &sn, &cs);
            psi[tid] = make_float2(cs, sn);
        }
        v4f acc[MAXT][SLOTS];
#pragma unroll
        for (int q = 0; q < MAXT; ++q)
#pragma unroll
            for (int j = 0; j < SLOTS; ++j) acc[q][j] = v4f{0.f, 0.f, 0.f, 0.f};
        const float2* Gp = a.G + item_base * SLOTS * nu * a.nP + ((long long)p * nu) * ustride;
        // the chunk's small-FFT block of every frame, staged once per plane by
        // global-to-LDS loads all in flight at once (awaited here; the
        // barrier at the top of the batch loop makes it visible)
        mac_stage_async<SLOTS>(a, cidx, p, t0, nf, s_stage);
        wait_vmem_all();
        for (int u0 = 0; u0 < nu; u0 += TC_UB) {
            const int ubn = min(TC_UB, nu - u0);
            __syncthreads();                      // psi written; the previous batch's products are done
            // B of spectra u0 .. u0 + ubn - 1: one (P, M) pair per (spectrum,
            // residue, pair type, frame).  Thread -> column (ub, pt), decoded
            // once, and a share of the (residue, frame) rows.
            {
                constexpr int Q = TC_UB * NPT;
                constexpr int ROWS = NRES * TC_FB;
                const int rstep = (int)blockDim.x / Q;
                const int r0 = tid / Q;
                const int col = tid - r0 * Q;
                const int pt = col % NPT, ub = col / NPT;
                int src[2] = {-1, -1}, flip[2] = {0, 0};
                if (r0 < rstep && ub < ubn) {
                    const int u = u0 + ub;
#pragma unroll
                    for (int w = 0; w < 2; ++w) {
                        const int off = a.uoff[u * NUSE + (w == 0 ? kPairPlain[pt] : kPairConj[pt])];
                        const int bo = off / N, ao = off - bo * N;
                        src[w] = off < NN ? off * NH : -1;
                        flip[w] = (bo > c ? 1 : 0) | (ao > c ? 2 : 0);
                    }
                }
                for (int rw = r0; r0 < rstep && rw < ROWS; rw += rstep) {
                    const int f = rw % TC_FB, rho = rw / TC_FB;
                    float4 pm = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (ub < ubn && f < nf) {
                        const int rb0 = rho < NH ? rho : rho - NH;
                        const float2* sf = s_stage + f * NN * NH;
                        float2 sv[2];
#pragma unroll
                        for (int w = 0; w < 2; ++w) {
                            float2 r = make_float2(0.f, 0.f);
                            if (src[w] >= 0) {
                                r = sf[src[w] + rb0];
                                if (rho >= NH) r = cconj(r);
                                if (flip[w] & 1) r = cmul(r, psi[2 * rho]);
                                if (flip[w] & 2) r = cmul(r, psi[2 * rho + 1]);
                            }
                            sv[w] = r;
                        }
                        const float2 P = cadd(sv[0], sv[1]);
                        const float2 d = csub(sv[0], sv[1]);
                        pm = make_float4(P.x, P.y, -d.y, d.x);                    // M = i * d
                    }
                    float* b = Bt + ((ub * NRES + rho) * KS + (pt >> 1)) * 64 + 2 * (pt & 1) * 16 + 2 * f;
                    b[0] = pm.x;                      // k row 2 (pt & 1): Re g -> P
                    b[1] = pm.y;
                    b[16] = pm.z;                     // k row 2 (pt & 1) + 1: Im g -> M
                    b[17] = pm.w;
                }
            }
            __syncthreads();
            // Spectra outer, tiles inner: the loads of spectrum ub + 1 for all
            // of this wave's tiles are in flight during the products of ub.
            int res_q[MAXT][SLOTS];
#pragma unroll
            for (int q = 0; q < MAXT; ++q) {
                const int tq = min(wv + TC_WAVES * q, ntiles - 1);
                const TileDesc td = tiles[tq];
#pragma unroll
                for (int j = 0; j < SLOTS; ++j) res_q[q][j] = td.res[j];
            }
            auto load_u = [&](float2 (&gq)[MAXT][4], int ub) {
#pragma unroll
                for (int q = 0; q < MAXT; ++q) {
                    const int it = it_of[q];
                    load_g<SLOTS>(gq[q], Gp + (long long)(u0 + ub) * ustride, it >= 0 ? it : 0, n_items);
                }
            };
            float2 gA[MAXT][4], gB[MAXT][4];
            load_u(gA, 0);
            auto products = [&](float2 (&gq)[MAXT][4], int ub) {
                const float* bu = Bt + ub * NRES * KS * 64 + lane;
#pragma unroll
                for (int q = 0; q < MAXT; ++q) {
                    if (wv + TC_WAVES * q >= ntiles) continue;                // wave-uniform
                    const bool live = it_of[q] >= 0;
#pragma unroll
                    for (int j = 0; j < SLOTS; ++j) {
#pragma unroll
                        for (int s = 0; s < KS; ++s) {
                            const int i0 = (2 * s) ^ j, i1 = (2 * s + 1) ^ j;
                            const float sg0 = SLOTS == 4 ? (float)kQuadSG[j][i0] : 1.0f;
                            const float sg1 = SLOTS == 4 ? (float)kQuadSG[j][i1] : 1.0f;
                            const float2 gi = (kq >> 1) ? gq[q][i1] : gq[q][i0];
                            const float sg = (kq >> 1) ? sg1 : sg0;
                            const float av = live ? ((kq & 1) ? sg * gi.y : gi.x) : 0.f;
                            const float bv = bu[(res_q[q][j] * KS + s) * 64];
                            acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[q][j], 0, 0, 0);
                        }
                    }
                }
            };
            // The loads are unconditional (past the batch they re-read its
            // last spectrum, in bounds and unused): with a branch around
            // them the compiler's single wait before the products would
            // have to drain the loads just issued.
            for (int ub = 0; ub < ubn; ub += 2) {
                load_u(gB, min(ub + 1, ubn - 1));
                products(gA, ub);
                load_u(gA, min(ub + 2, ubn - 1));
                if (ub + 1 < ubn) products(gB, ub + 1);
            }
        }
        // Epilogue of plane p: outer phase and 'same'-crop shift per item and
        // slot (the same for every frame); (re, im) of a column pair meet by
        // one lane swap.
        const int sig = s2 - 2 + N;
        const int c2 = (s2 - 1) / 2;
#pragma unroll
        for (int q = 0; q < MAXT; ++q) {
            const int tq = wv + TC_WAVES * q;
            if (tq >= ntiles) continue;                                       // wave-uniform
            const TileDesc td = tiles[tq];
#pragma unroll
            for (int j = 0; j < SLOTS; ++j) {
                float mine[4], other[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    mine[r] = acc[q][j][r];
                    other[r] = __shfl_xor(mine[r], 1);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ir = 4 * kq + r;
                    if (ir < td.count) {
                        const ItemDesc& it = a.items[item_base + td.start + ir];
                        const int m2 = it.m2[0], n2 = it.n2[0];
                        const int outer = j == 0 ? sig * (m2 + n2)
                                        : j == 1 ? sig * (n2 - m2)
                                        : j == 2 ? sig * (n2 + m2) : sig * (m2 - n2);
                        const int pos = it.pos[j];
                        const int k0 = pos / a.rfX, k1 = pos - k0 * a.rfX;
                        const float2 po = phase_frac_i(outer, fourN);
                        const int e0 = (c2 * k0) % a.fY;
                        const int e1 = (c2 * k1) % a.fX;
                        float fr = (float)e0 / (float)a.fY + (float)e1 / (float)a.fX;
                        fr -= rintf(fr);
                        float sn, cs;
                        sincospif(2.0f * fr, &sn, &cs);
                        const float2 z = c_col == 0 ? make_float2(mine[r], other[r]) : make_float2(other[r], mine[r]);
                        const float2 v = cmul(cmul(z, po), make_float2(cs, sn));
                        tot[q][j][r] += c_col == 0 ? v.x : v.y;
                    }
                }
            }
        }
        if (a.per_plane) {                        // back projection: one output per plane
            store(p);
#pragma unroll
            for (int q = 0; q < MAXT; ++q)
#pragma unroll
                for (int j = 0; j < SLOTS; ++j) tot[q][j] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    }
    if (!a.per_plane) store(blockIdx.y);
}

// ---------------------------------------------------------------------------
// Inverse transform, stage 1: column inverse FFTs (length fY) of the half
// plane, keeping output rows y < Yout.  The input is the sum of nsum partial
// accumulators A_sstride apart (the forward projection's plane ranges).
// grid (ceil(rfX/CB), nbatch), one wave per column (blockDim = 64*CB); LDS
// CB*fY + fY (twiddles).
// ---------------------------------------------------------------------------
template <bool PRIME, bool BIG>
__device__ __forceinline__ void ifft_cols_body(const IfftColArgs& a, int colblk, int batch) {
    extern __shared__ float2 lds[];
    const int fY = a.rpY.L, rfX = a.rfX, CB = a.CB;
    const int cs = ifft_col_stride(fY, CB);
    float2* C = lds;                 // [CB][cs]
    float2* C2 = C + CB * cs;        // BIG: [CB][cs], the second buffer of the out-of-place passes
    float2* tw = BIG ? C2 + CB * cs : C2;   // [fY]
    const int c0 = colblk * CB;
    const int ncol = min(CB, rfX - c0);
    const float2* in = a.A + (long long)batch * a.A_bstride;
    for (int i = threadIdx.x; i < fY; i += blockDim.x) tw[i] = a.twY[i];
    constexpr int B = 8;                        // loads in flight per thread
    for (int i0 = threadIdx.x; i0 < fY * CB; i0 += blockDim.x * B) {
        float2 v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int i = i0 + u * blockDim.x;
            const int y = i / CB, cc = i - y * CB;
            v[u] = make_float2(0.f, 0.f);
            if (i < fY * CB && cc < ncol) v[u] = in[(long long)y * rfX + c0 + cc];
        }
        for (int q = 1; q < a.nsum; ++q) {     // partial accumulators, fixed order
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int i = i0 + u * blockDim.x;
                const int y = i / CB, cc = i - y * CB;
                if (i < fY * CB && cc < ncol) v[u] = cadd(v[u], in[(long long)y * rfX + c0 + cc + q * a.A_sstride]);
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int i = i0 + u * blockDim.x;
            const int y = i / CB, cc = i - y * CB;
            if (i < fY * CB) C[cc * cs + y] = v[u];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if constexpr (BIG) {
        if (wv < ncol) wave_fft_pp<PRIME>(C + wv * cs, C2 + wv * cs, a.rpY, tw, true, lane);
    } else {
        if (wv < ncol) wave_fft<2048, PRIME>(C + wv * cs, 1, a.rpY, tw, true, lane);
    }
    __syncthreads();
    const float2* Res = (BIG && (a.rpY.n & 1)) ? C2 : C;   // after an odd pass count: the second buffer
    float2* out = a.out + (long long)batch * a.out_bstride;
    for (int i = threadIdx.x; i < a.Yout * CB; i += blockDim.x) {
        const int y = i / CB, cc = i - y * CB;
        if (cc < ncol) out[(long long)y * rfX + c0 + cc] = Res[cc * cs + y];
    }
}
template <bool PRIME, bool BIG>
__global__ __launch_bounds__(512) void k_ifft_cols(IfftColArgs a) {
    ifft_cols_body<PRIME, BIG>(a, blockIdx.x, blockIdx.y);
}
// All active grid groups: grid x flattened over (group, batch, column
// block), nchunks[g] = column blocks of group g; one CB for every group.
template <bool PRIME>
__global__ __launch_bounds__(512) void k_ifft_cols_grouped(IfftColGrouped ga) {
    const GroupPick p = pick_group(ga.gg, blockIdx.x);
    ifft_cols_body<PRIME, false>(ga.a[p.g], p.rem % p.nch, p.rem / p.nch);
}

// c2r of one row by one wave.  row: rfX = L/2+1 half-spectrum values (global);
// buf: LDS work area (L/2 entries for even L, L for odd L; BIG, for L above
// 2048: buf2 is a second such area for the out-of-place passes).  For even L the
// length-L/2 complex transform of Z[k] = E[k] + i O[k],
//   E = X[k] + conj(X[L/2-k]),  O = (X[k] - conj(X[L/2-k])) exp(+2 pi i k/L),
// gives x[2m] = Re z[m], x[2m+1] = Im z[m]; odd L uses the Hermitian
// extension and a full-length transform.  The imaginary parts of X[0] and
// X[L/2] are dropped, as a c2r transform does.  Returns with the result in
// buf (interleaved real samples for even L, real parts for odd L).
template <bool PRIME, bool BIG>
__device__ __forceinline__ float2* wave_c2r(float2* buf, float2* buf2, const float2* __restrict__ row, int L,
                                            const RadixPlan& rph, const float2* twh,
                                            const RadixPlan& rpf, const float2* twf, int lane) {
    // Loads are issued C2R_BATCH per lane at a time, so that their latencies
    // overlap instead of adding up.
    constexpr int B = 8;
    if ((L & 1) == 0) {
        const int L2 = L / 2;
        for (int k0 = lane; k0 < L2; k0 += 64 * B) {
            float2 xs[B], xrs[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int k = k0 + 64 * u;
                if (k < L2) {
                    xs[u] = row[k];
                    xrs[u] = row[L2 - k];
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int k = k0 + 64 * u;
                if (k < L2) {
                    float2 x = xs[u];
                    float2 xr = cconj(xrs[u]);
                    if (k == 0) { x.y = 0.f; xr.y = 0.f; }
                    const float2 E = cadd(x, xr);
                    const float2 d = csub(x, xr);
                    const float2 w = cconj(twf[k]);          // exp(+2 pi i k / L)
                    const float2 O = cmul(d, w);
                    buf[k] = make_float2(E.x - O.y, E.y + O.x);  // E + i O
                }
            }
        }
        wave_sync();
        if constexpr (BIG) return wave_fft_pp<PRIME>(buf, buf2, rph, twh, true, lane);
        wave_fft<1024, PRIME>(buf, 1, rph, twh, true, lane);
        return buf;
    } else {
        const int rfX = L / 2 + 1;
        for (int k0 = lane; k0 < L; k0 += 64 * B) {
            float2 vs[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int k = k0 + 64 * u;
                if (k < L) vs[u] = k < rfX ? row[k] : cconj(row[L - k]);
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int k = k0 + 64 * u;
                if (k < L) {
                    float2 v = vs[u];
                    if (k == 0) v.y = 0.f;
                    buf[k] = v;
                }
            }
        }
        wave_sync();
        if constexpr (BIG) return wave_fft_pp<PRIME>(buf, buf2, rpf, twf, true, lane);
        wave_fft<2048, PRIME>(buf, 1, rpf, twf, true, lane);
        return buf;
    }
}

__device__ __forceinline__ float c2r_sample(const float2* buf, int L, int x) {
    if ((L & 1) == 0) {
        const float2 v = buf[x >> 1];
        return (x & 1) ? v.y : v.x;
    }
    return buf[x].x;
}

// Forward projection, stage 2: per output row, c2r each grid group's row and
// sum the groups in a fixed order.  One wave per (row, frame);
// grid (ceil(Y/4), T), block 256.
template <bool PRIME, bool BIG>
__global__ __launch_bounds__(256) void k_ifft_rows_sum(IfftRowSumArgs a) {
    extern __shared__ float2 lds[];
    constexpr int RPB = BIG ? 2 : 4;            // rows (waves) per workgroup
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int y = blockIdx.x * RPB + wv, t = blockIdx.y;
    float2* twf = lds;                          // [maxfX]
    float2* twh = twf + a.maxfX;                // [maxfX/2]
    float2* buf = twh + a.maxfX / 2 + 1 + wv * (BIG ? 2 : 1) * a.maxfX;
    float2* buf2 = buf + a.maxfX;               // BIG only
    constexpr int XPL = (BIG ? 4096 : 2048) / 64;   // output samples per lane (X <= 2048, BIG: 4096)
    float acc[XPL];
#pragma unroll
    for (int q = 0; q < XPL; ++q) acc[q] = 0.f;
    for (int g = 0; g < a.ngroups; ++g) {
        const IfftRowGroup& G = a.grp[g];
        if (!G.active) continue;
        const int L = G.rp.L;
        __syncthreads();
        for (int i = threadIdx.x; i < L; i += blockDim.x) twf[i] = G.twf[i];
        if ((L & 1) == 0)
            for (int i = threadIdx.x; i < L / 2; i += blockDim.x) twh[i] = G.twh[i];
        __syncthreads();
        if (y < a.Y) {
            const float2* row = G.T + (long long)t * G.tstride + (long long)y * G.rfX;
            const float2* res = wave_c2r<PRIME, BIG>(buf, buf2, row, L, G.rph, twh, G.rp, twf, lane);
            const float sc = G.scale;
#pragma unroll
            for (int q = 0; q < XPL; ++q) {
                const int x = lane + 64 * q;
                if (x < a.X) acc[q] += c2r_sample(res, L, x) * sc;
            }
            wave_sync();
        }
    }
    if (y < a.Y) {
        const int Yv = a.fdims ? a.fdims[t].x : a.Y, Xv = a.fdims ? a.fdims[t].y : a.X;
        const long long ro = (long long)t * a.img_tstride + (long long)y * a.X;
        if (a.img) {
#pragma unroll
            for (int q = 0; q < XPL; ++q) {
                const int x = lane + 64 * q;
                if (x < a.X) a.img[ro + x] = (y < Yv && x < Xv) ? acc[q] : 0.f;
            }
        }
        if (a.ratio) {
            // standard RL's error image im / HX (lfdeconv.py:191) with IEEE
            // division: 0/0 gives NaN and x/0 inf, as in numpy
#pragma unroll
            for (int q = 0; q < XPL; ++q) {
                const int x = lane + 64 * q;
                if (x < a.X) a.ratio[ro + x] = (y < Yv && x < Xv) ? a.im[ro + x] / acc[q] : 0.f;
            }
        }
    }
}

// Backward projection, stage 2: per (row, frame, plane) c2r; then either store
// the plane, or apply the Richardson-Lucy update in place
// (lfdeconv.py:112-126 Prevedel: X *= Htf / v;  lfdeconv.py:151-164 standard:
// X *= v [/ trust]; NaN -> 0 after the update).  One wave per row;
// grid (ceil(Y/4), T, nplanes), block 256.  Twiddles of every grid group are
// read from global memory (they stay in the caches).
template <bool PRIME, bool BIG>
__global__ __launch_bounds__(256) void k_ifft_rows_plane(IfftRowPlaneArgs a) {
    extern __shared__ float2 lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int y = blockIdx.x * 4 + wv, t = blockIdx.y;
    const int z = a.plane_ids[blockIdx.z];      // kernel argument (scalar load)
    const RowPlaneDesc& pd = a.pd[z];           // read in place (no private copy)
    const int L = pd.rp.L;
    float2* buf = lds + wv * (BIG ? 2 : 1) * a.wbuf;
    if (y >= a.Y) return;                       // no block-level barrier below
    const float2* row = a.Tbuf + pd.T_off + (long long)t * pd.T_tstride + (long long)y * pd.rfX;
    const float2* res = wave_c2r<PRIME, BIG>(buf, buf + a.wbuf, row, L, pd.rph, pd.twh, pd.rp, pd.twf, lane);
    buf = const_cast<float2*>(res);
    const float sc = pd.scale;
    const long long o = (((long long)z * a.T + t) * a.Y + y) * a.X;
    const int Yv = a.fdims ? a.fdims[t].x : a.Y, Xv = a.fdims ? a.fdims[t].y : a.X;
    constexpr int B = 8;                        // loads in flight per lane
    for (int x0 = lane; x0 < a.X; x0 += 64 * B) {
        float v[B], xo[B], hd[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int x = x0 + 64 * u;
            if (x < a.X) {
                v[u] = c2r_sample(buf, L, x) * sc;
                if (a.mode != 0) xo[u] = a.vol[o + x];
                if (a.mode == 1) hd[u] = a.Htf[o + x];
                else if (a.mode == 2) hd[u] = a.trust ? a.trust[((long long)z * a.Y + y) * a.X + x] : 1.f;
            }
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int x = x0 + 64 * u;
            if (x < a.X && (y >= Yv || x >= Xv)) {
                a.vol[o + x] = 0.f;                 // outside this frame's valid region
            } else if (x < a.X) {
                if (a.mode == 0) {
                    a.vol[o + x] = v[u];
                } else if (a.mode == 1) {
                    const float r = hd[u] / v[u];
                    const float xn = xo[u] * r;
                    a.vol[o + x] = __builtin_isnan(xn) ? 0.f : xn;
                } else {
                    float xn = xo[u] * v[u];
                    if (a.trust) xn = xn / hd[u];
                    a.vol[o + x] = __builtin_isnan(xn) ? 0.f : xn;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Data-fit sums in FP64: per frame sum((img-im)^2) and sum(im^2).  Two stages
// with a fixed reduction order (deterministic).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fit_partial(const float* img, const float* im, long long n_per_frame,
                                                     int im_frames, double* part, int nblk) {
    __shared__ double sd[256], si[256];
    const int t = blockIdx.y;
    const float* a = img + (long long)t * n_per_frame;
    const float* b = im + (long long)(t % im_frames) * n_per_frame;
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

// Prevedel update with a precomputed denominator (lfdeconv.py:112-126):
// X *= Htf / v, NaN -> 0, the same arithmetic as k_ifft_rows_plane mode 1.
// Used for the first plain iterate from a uniform start, X1 = Htf / Ht(H 1),
// whose denominator does not depend on the image (one slab per plane serves
// every frame: v_frames = 1).  ratio: X = Htf / v (the same update of X = 1,
// without reading X).  grid (x blocks, nz * T slabs), block RL_SCALE_THREADS;
// each thread takes RL_SCALE_E pixels RL_SCALE_THREADS apart, all loads in
// flight before the first use (indices clamped into the slab).
constexpr int RL_SCALE_THREADS = 256, RL_SCALE_E = 16;
__global__ __launch_bounds__(RL_SCALE_THREADS) void k_rl_scale(float* X, const float* Htf, const float* v, int T,
                                                               long long npix, int v_frames, int ratio) {
    const int slab = blockIdx.y;
    const int z = slab / T, t = slab - z * T;
    const long long o = (long long)slab * npix;
    const float* vs = v + ((long long)z * v_frames + (v_frames == 1 ? 0 : t)) * npix;
    constexpr long long CH = (long long)RL_SCALE_THREADS * RL_SCALE_E;
    for (long long b0 = (long long)blockIdx.x * CH + threadIdx.x; b0 < npix; b0 += (long long)gridDim.x * CH) {
        float h[RL_SCALE_E], d[RL_SCALE_E], x[RL_SCALE_E];
#pragma unroll
        for (int e = 0; e < RL_SCALE_E; ++e) {
            const long long i = min(b0 + (long long)e * RL_SCALE_THREADS, npix - 1);
            h[e] = Htf[o + i];
            d[e] = vs[i];
            x[e] = ratio ? 1.f : X[o + i];
        }
#pragma unroll
        for (int e = 0; e < RL_SCALE_E; ++e) {
            const long long i = b0 + (long long)e * RL_SCALE_THREADS;
            if (i < npix) {
                const float r = h[e] / d[e];
                const float xn = ratio ? r : x[e] * r;
                X[o + i] = __builtin_isnan(xn) ? 0.f : xn;
            }
        }
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

// Whether a plan ends with an odd prime radix above 5 (7, 11, ..., 31; the
// kernels built with that pass are launched only then).  The other radices
// are 2..6 and 8.
bool plan_has_prime(const RadixPlan& rp) {
    if (rp.n == 0) return false;
    const int r = rp.r[rp.n - 1];
    return r == 7 || r > 8;
}
static bool has_prime(const RadixPlan& rp) { return plan_has_prime(rp); }

size_t small_fft_lds(int rY, int rX) {
    return ((size_t)(rY + 1) * (rX / 2 + 1) + 2 * (size_t)(rY + rX)) * sizeof(float2);
}
// Reduced grids above 128 points take the 256-point wave transforms.
bool small_fft_big(int rY, int rX) { return rY > 128 || rX > 128; }
hipError_t launch_small_fft(const SmallFFTArgs& a, int nsrc, int T, hipStream_t s) {
    size_t lds = small_fft_lds(a.rpY.L, a.rpX.L);
    if (a.rpY.L > 256 || a.rpX.L > 256) return hipErrorInvalidValue;
    if (small_fft_big(a.rpY.L, a.rpX.L))
        hipLaunchKernelGGL(k_small_fft<true>, dim3(a.N * a.N, T, nsrc), dim3(64 * SFFT_WAVES), lds, s, a);
    else
        hipLaunchKernelGGL(k_small_fft<false>, dim3(a.N * a.N, T, nsrc), dim3(64 * SFFT_WAVES), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_small_fft_grouped(const SmallFFTGrouped& ga, int N, int T, size_t lds, hipStream_t s) {
    if (ga.gg.ngroups < 1 || ga.gg.ngroups > FLF_MAX_GROUPS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_small_fft_grouped, dim3(N * N, T, ga.gg.start[ga.gg.ngroups]), dim3(64 * SFFT_WAVES), lds, s,
                       ga);
    return hipGetLastError();
}
hipError_t launch_psf_rows(const PsfRowArgs& a, int nu, hipStream_t s) {
    size_t lds = (size_t)a.RB * a.rpX.L * sizeof(float2);
    if (has_prime(a.rpX)) hipLaunchKernelGGL(k_psf_rows<true>, dim3((a.s2 + a.RB - 1) / a.RB, nu), dim3(FFT_THREADS), lds, s, a);
    else hipLaunchKernelGGL(k_psf_rows<false>, dim3((a.s2 + a.RB - 1) / a.RB, nu), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_psf_cols(const PsfColArgs& a, int nu, hipStream_t s) {
    size_t lds = (size_t)a.CB * a.rpY.L * sizeof(float2);
    if (has_prime(a.rpY)) hipLaunchKernelGGL(k_psf_cols<true>, dim3((a.rfX + a.CB - 1) / a.CB, nu), dim3(FFT_THREADS), lds, s, a);
    else hipLaunchKernelGGL(k_psf_cols<false>, dim3((a.rfX + a.CB - 1) / a.CB, nu), dim3(FFT_THREADS), lds, s, a);
    return hipGetLastError();
}
hipError_t launch_spec_pack(const SpecPackArgs& a, int nchunks, hipStream_t s) {
    hipLaunchKernelGGL(k_spec_pack, dim3(nchunks, a.nu), dim3(FLF_CHUNK_MAX), 0, s, a);
    return hipGetLastError();
}
size_t mac_tc_lds(int slots, int N) {
    const int nres = slots == 4 ? 8 : 4;
    return (size_t)TC_UB * nres * (slots / 2) * 64 * sizeof(float) + 2 * (size_t)nres * sizeof(float2) +
           (size_t)TC_FB * N * N * (nres / 2) * sizeof(float2);
}
int mac_row_stride(int slots, int nu) {
    const int n = nu * (slots == 4 ? 4 : 2);
    return n + ((1 - n % 16) + 16) % 16;
}
size_t mac_lds(int slots, int nu, int N, int fb) {
    const int nres = slots == 4 ? 8 : 4;
    return (size_t)fb * nres * mac_row_stride(slots, nu) * sizeof(float4) +
           ((size_t)fb * N * N * (nres / 2) + 2 * nres) * sizeof(float2);
}
template <int SLOTS, int IPT>
static void launch_mac_fb(const MacArgs& a, int fb, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
    if (fb == 1) hipLaunchKernelGGL((k_mac<SLOTS, IPT, 1>), grid, block, lds, s, a);
    else if (fb == 2) hipLaunchKernelGGL((k_mac<SLOTS, IPT, 2>), grid, block, lds, s, a);
    else if (fb == 3) hipLaunchKernelGGL((k_mac<SLOTS, IPT, 3>), grid, block, lds, s, a);
    else hipLaunchKernelGGL((k_mac<SLOTS, IPT, 4>), grid, block, lds, s, a);
}
hipError_t launch_mac(const MacArgs& a, int slots, int ipt, int fb, int nchunks, int nplanes_grid, bool mfma,
                      hipStream_t s) {
    if (fb < 1 || fb > 4 || (ipt != 1 && ipt != 2 && ipt != 4) || a.T < 1 || a.rs != mac_row_stride(slots, a.nu))
        return hipErrorInvalidValue;
    const size_t lds = mac_lds(slots, a.nu, a.N, fb);
    dim3 grid(nchunks * ((a.T + fb - 1) / fb), nplanes_grid);
    if (mfma && a.tiles) {
        // matrix-core kernel: TC_FB frames per workgroup
        dim3 tgrid(nchunks * ((a.T + TC_FB - 1) / TC_FB), nplanes_grid);
        const size_t tlds = mac_tc_lds(slots, a.N);
        if (slots == 4) hipLaunchKernelGGL((k_mac_tc<4>), tgrid, dim3(64 * TC_WAVES), tlds, s, a);
        else hipLaunchKernelGGL((k_mac_tc<2>), tgrid, dim3(64 * TC_WAVES), tlds, s, a);
        return hipGetLastError();
    }
    dim3 block(FLF_CHUNK_MAX);                   // ipt 2 / 4: lane groups share ipt items
    if (slots == 4) {
        if (ipt == 1) launch_mac_fb<4, 1>(a, fb, grid, block, lds, s);
        else if (ipt == 2) launch_mac_fb<4, 2>(a, fb, grid, block, lds, s);
        else launch_mac_fb<4, 4>(a, fb, grid, block, lds, s);
    } else {
        if (ipt == 1) launch_mac_fb<2, 1>(a, fb, grid, block, lds, s);
        else if (ipt == 2) launch_mac_fb<2, 2>(a, fb, grid, block, lds, s);
        else launch_mac_fb<2, 4>(a, fb, grid, block, lds, s);
    }
    return hipGetLastError();
}
template <int SLOTS, int IPT>
static void launch_mac_grouped_fb(const MacGrouped& ga, int fb, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
    if (fb == 1) hipLaunchKernelGGL((k_mac_grouped<SLOTS, IPT, 1>), grid, block, lds, s, ga);
    else if (fb == 2) hipLaunchKernelGGL((k_mac_grouped<SLOTS, IPT, 2>), grid, block, lds, s, ga);
    else if (fb == 3) hipLaunchKernelGGL((k_mac_grouped<SLOTS, IPT, 3>), grid, block, lds, s, ga);
    else hipLaunchKernelGGL((k_mac_grouped<SLOTS, IPT, 4>), grid, block, lds, s, ga);
}
hipError_t launch_mac_grouped(const MacGrouped& ga, int slots, int ipt, int fb, int nfb, size_t lds, hipStream_t s) {
    const GroupedGrid& gg = ga.gg;
    if (fb < 1 || fb > 4 || (ipt != 1 && ipt != 2 && ipt != 4) || (slots != 2 && slots != 4) || nfb < 1 || gg.ngroups < 1 ||
        gg.ngroups > FLF_MAX_GROUPS)
        return hipErrorInvalidValue;
    if (ga.nfb != nfb) return hipErrorInvalidValue;
    for (int k = 0; k < gg.ngroups; ++k)
        if (gg.start[k + 1] - gg.start[k] < gg.nchunks[k] * gg.nranges[k] * nfb) return hipErrorInvalidValue;
    dim3 grid(gg.start[gg.ngroups]);
    dim3 block(FLF_CHUNK_MAX);
    if (slots == 4) {
        if (ipt == 1) launch_mac_grouped_fb<4, 1>(ga, fb, grid, block, lds, s);
        else if (ipt == 2) launch_mac_grouped_fb<4, 2>(ga, fb, grid, block, lds, s);
        else launch_mac_grouped_fb<4, 4>(ga, fb, grid, block, lds, s);
    } else {
        if (ipt == 1) launch_mac_grouped_fb<2, 1>(ga, fb, grid, block, lds, s);
        else if (ipt == 2) launch_mac_grouped_fb<2, 2>(ga, fb, grid, block, lds, s);
        else launch_mac_grouped_fb<2, 4>(ga, fb, grid, block, lds, s);
    }
    return hipGetLastError();
}
// Transforms longer than 2048 points run out of place (two LDS buffers).
bool wave_big(int L) { return L > 2048; }
size_t ifft_cols_lds(int CB, int fY) {
    return ((size_t)CB * ifft_col_stride(fY, CB) * (wave_big(fY) ? 2 : 1) + fY) * sizeof(float2);
}
hipError_t launch_ifft_cols_grouped(const IfftColGrouped& ga, int CB, bool prime, size_t lds, hipStream_t s) {
    const GroupedGrid& gg = ga.gg;
    if (CB < 1 || CB > 8 || gg.ngroups < 1 || gg.ngroups > FLF_MAX_GROUPS) return hipErrorInvalidValue;
    if (prime) hipLaunchKernelGGL(k_ifft_cols_grouped<true>, dim3(gg.start[gg.ngroups]), dim3(64 * CB), lds, s, ga);
    else hipLaunchKernelGGL(k_ifft_cols_grouped<false>, dim3(gg.start[gg.ngroups]), dim3(64 * CB), lds, s, ga);
    return hipGetLastError();
}
hipError_t launch_ifft_cols(const IfftColArgs& a, int nbatch, hipStream_t s) {
    size_t lds = ifft_cols_lds(a.CB, a.rpY.L);
    const dim3 grid((a.rfX + a.CB - 1) / a.CB, nbatch), block(64 * a.CB);
    const bool big = wave_big(a.rpY.L);
    if (has_prime(a.rpY)) {
        if (big) hipLaunchKernelGGL((k_ifft_cols<true, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_cols<true, false>), grid, block, lds, s, a);
    } else {
        if (big) hipLaunchKernelGGL((k_ifft_cols<false, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_cols<false, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}
// Twiddles (full and half length), then per wave one buffer (two for rows
// longer than 2048, which take two waves per workgroup instead of four).
size_t ifft_rows_sum_lds(int maxfX) {
    const bool big = wave_big(maxfX);
    return ((size_t)maxfX + maxfX / 2 + 1 + (big ? 2 * 2 : 4) * (size_t)maxfX) * sizeof(float2);
}
hipError_t launch_ifft_rows_sum(const IfftRowSumArgs& a, int Y, int T, hipStream_t s) {
    size_t lds = ifft_rows_sum_lds(a.maxfX);
    bool prime = false;
    for (int g = 0; g < a.ngroups; ++g) prime = prime || (a.grp[g].active && (has_prime(a.grp[g].rp) || has_prime(a.grp[g].rph)));
    if (wave_big(a.maxfX)) {
        const dim3 grid((Y + 1) / 2, T), block(128);
        if (prime) hipLaunchKernelGGL((k_ifft_rows_sum<true, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_rows_sum<false, true>), grid, block, lds, s, a);
    } else {
        const dim3 grid((Y + 3) / 4, T), block(256);
        if (prime) hipLaunchKernelGGL((k_ifft_rows_sum<true, false>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_rows_sum<false, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}
// c2r of an even length runs as a half-length complex FFT in place.
int c2r_wave_buffer(int fX) { return (fX & 1) ? fX : fX / 2; }
size_t ifft_rows_plane_lds(int wbuf, bool big) { return 4 * (size_t)wbuf * (big ? 2 : 1) * sizeof(float2); }
hipError_t launch_ifft_rows_plane(const IfftRowPlaneArgs& a, int nplanes, hipStream_t s) {
    size_t lds = ifft_rows_plane_lds(a.wbuf, a.big != 0);
    const dim3 grid((a.Y + 3) / 4, a.T, nplanes), block(256);
    if (a.prime) {
        if (a.big) hipLaunchKernelGGL((k_ifft_rows_plane<true, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_rows_plane<true, false>), grid, block, lds, s, a);
    } else {
        if (a.big) hipLaunchKernelGGL((k_ifft_rows_plane<false, true>), grid, block, lds, s, a);
        else hipLaunchKernelGGL((k_ifft_rows_plane<false, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}
hipError_t launch_fit(const float* img, const float* im, long long n_per_frame, int T, int im_frames, double* part,
                      int nblk, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_fit_partial, dim3(nblk, T), dim3(256), 0, s, img, im, n_per_frame, im_frames, part, nblk);
    hipLaunchKernelGGL(k_fit_final, dim3(T), dim3(256), 0, s, part, nblk, out);
    return hipGetLastError();
}
hipError_t configure_kernels(size_t max_lds) {
    const void* ks[] = {(const void*)k_small_fft<false>, (const void*)k_small_fft<true>,
                        (const void*)k_psf_rows<false>, (const void*)k_psf_cols<false>,
                        (const void*)k_psf_rows<true>, (const void*)k_psf_cols<true>,
                        (const void*)k_ifft_cols<false, false>, (const void*)k_ifft_cols<true, false>,
                        (const void*)k_ifft_cols<false, true>, (const void*)k_ifft_cols<true, true>,
                        (const void*)k_ifft_rows_sum<false, false>, (const void*)k_ifft_rows_sum<true, false>,
                        (const void*)k_ifft_rows_sum<false, true>, (const void*)k_ifft_rows_sum<true, true>,
                        (const void*)k_ifft_rows_plane<false, false>, (const void*)k_ifft_rows_plane<true, false>,
                        (const void*)k_ifft_rows_plane<false, true>, (const void*)k_ifft_rows_plane<true, true>,
                        (const void*)k_mac<4, 1, 1>, (const void*)k_mac<4, 1, 2>, (const void*)k_mac<4, 1, 3>,
                        (const void*)k_mac<4, 1, 4>, (const void*)k_mac<4, 2, 1>, (const void*)k_mac<4, 2, 2>,
                        (const void*)k_mac<4, 2, 3>, (const void*)k_mac<4, 2, 4>, (const void*)k_mac<2, 1, 1>,
                        (const void*)k_mac<2, 1, 2>, (const void*)k_mac<2, 1, 3>, (const void*)k_mac<2, 1, 4>,
                        (const void*)k_mac<2, 2, 1>, (const void*)k_mac<2, 2, 2>, (const void*)k_mac<2, 2, 3>,
                        (const void*)k_mac<2, 2, 4>,
                        (const void*)k_mac_tc<4>, (const void*)k_mac_tc<2>,
                        (const void*)k_small_fft_grouped, (const void*)k_ifft_cols_grouped<false>,
                        (const void*)k_ifft_cols_grouped<true>,
                        (const void*)k_mac_grouped<4, 1, 1>, (const void*)k_mac_grouped<4, 1, 2>,
                        (const void*)k_mac_grouped<4, 1, 3>, (const void*)k_mac_grouped<4, 1, 4>,
                        (const void*)k_mac_grouped<4, 2, 1>, (const void*)k_mac_grouped<4, 2, 2>,
                        (const void*)k_mac_grouped<4, 2, 3>, (const void*)k_mac_grouped<4, 2, 4>,
                        (const void*)k_mac_grouped<2, 1, 1>, (const void*)k_mac_grouped<2, 1, 2>,
                        (const void*)k_mac_grouped<2, 1, 3>, (const void*)k_mac_grouped<2, 1, 4>,
                        (const void*)k_mac_grouped<2, 2, 1>, (const void*)k_mac_grouped<2, 2, 2>,
                        (const void*)k_mac_grouped<2, 2, 3>, (const void*)k_mac_grouped<2, 2, 4>};
    for (const void* k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_rl_scale(float* X, const float* Htf, const float* v, int nz, int T, long long npix, int v_frames,
                           hipStream_t s, bool ratio) {
    const long long ch = (long long)RL_SCALE_THREADS * RL_SCALE_E;
    long long blocks = (npix + ch - 1) / ch;
    if (blocks > 65535) blocks = 65535;
    if (blocks < 1) blocks = 1;
    // slabs on y (nz * T <= 65535 is checked by the caller), pixels on x
    hipLaunchKernelGGL(k_rl_scale, dim3((unsigned)blocks, (unsigned)(nz * T)), dim3(RL_SCALE_THREADS), 0, s, X, Htf,
                       v, T, npix, v_frames, ratio ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_fill(float* p, long long n, float v, hipStream_t s) {
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace flf
