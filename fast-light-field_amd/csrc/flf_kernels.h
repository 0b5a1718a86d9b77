// Kernel argument blocks and host launchers (internal to libflf_hip).
#pragma once
#include "flf_common.h"

#define FLF_MAX_GROUPS 8
#define FLF_MAX_PLANES 256       // planes per plan: the plane selection travels in kernel arguments

// Planes of a grid group selected for one projection, as a bitmask carried
// by value in the kernel arguments (a captured graph replays its own
// selection, whatever later calls select).
struct PlaneMask {
    unsigned int w[FLF_MAX_PLANES / 32];
    __host__ __device__ bool on(int p) const { return (w[p >> 5] >> (p & 31)) & 1u; }
};

// Block ranges of a grouped launch (one dispatch for every active grid group
// of a projection).  start[g] is group g's first flattened block,
// start[ngroups] the total; nchunks/nranges split a group's blocks (MAC:
// chunks x plane ranges, each group's range padded to a multiple of 8;
// column transform: column blocks x batches).
struct GroupedGrid {
    int ngroups;
    int start[FLF_MAX_GROUPS + 1];
    int nchunks[FLF_MAX_GROUPS];
    int nranges[FLF_MAX_GROUPS];
};

struct SmallFFTArgs {
    const float* src;            // source base
    const long long* src_off;    // per source plane z: element offset (device array)
    long long src_tstride;       // per frame
    int src_pitch;               // row pitch (elements)
    int Y, X, N;
    RadixPlan rpY, rpX;          // lengths rY, rX
    const float2* twY;           // exp(-2 pi i m / rY), m < rY
    const float2* twX;           // exp(-2 pi i m / rX)
    const ChunkDesc* chunks;     // MAC chunks of the grid (output order)
    int nchunks, NH;             // chunks; residues stored per chunk (4 quad / 2 pair)
    int fY, fX;
    float2* S;                   // [chunk][N*N][NH] per source plane and frame
    const long long* S_off;      // per source plane z: element offset into S (device array)
    long long S_tstride;         // per frame
    const int2* fdims;           // per frame valid (Y, X) <= (Y, X), or null (all full size)
    const float* src2;           // frames t >= tsplit read from src2 (frame t - tsplit), or null
    int tsplit;
};

struct PsfRowArgs {
    const float* Q;              // quadrant (h, h, s2, s2)
    int h, s2, RB;
    const int* u_bu;
    const int* u_au;
    RadixPlan rpX;               // length fX
    float2* tmp;                 // (nu, s2, rfX)
};

struct PsfColArgs {
    const float2* tmp;           // (nu, s2, rfX)
    int s2, rfX, CB;
    RadixPlan rpY;               // length fY
    float2* spec;                // (nu, fY, rfX)
};

struct SpecPackArgs {
    const float2* spec;          // (nu, fY, rfX) natural layout, one plane
    const ChunkDesc* chunks;
    const ItemDesc* items;
    const int* u_bu;
    const int* u_au;
    int fY, rfX, N, slots, nu, nP, p;
    int s2;                      // PSF side of the plane (sets the folded phase ratios)
    float2* G;                   // orbit-major spectra of the grid group
};

struct MacArgs {
    const float2* G;
    const ChunkDesc* chunks;
    const ItemDesc* items;
    const int* u_bu;
    const int* u_au;
    const int* uoff;             // [nu][8 or 4]: s-table column of each use, N*N if absent
    const float2* S;
    const long long* S_off;      // per plane in group
    long long S_tstride;
    const int* s2;               // per plane in group
    PlaneMask enable;            // planes of the group in this projection
    int nP;
    float2* A;
    long long A_tstride, A_pstride;
    int N, fY, fX, rY, rX, rfX, nu;
    int T;                       // frames (grid z covers them in blocks of FB)
    const TileDesc* tiles;       // residue-class tiles (matrix-core kernel)
    int rs;                      // P/M table residue-row stride (mac_row_stride)
    int per_plane;               // 1: store each plane's accumulator at A + p * A_pstride (back
                                 // projection); 0: sum the planes of a range, store at
                                 // A + range * A_pstride (forward projection)
    int fgroup;                  // frame blocks dispatched side by side per chunk (0: all)
};

struct IfftColArgs {
    const float2* A;
    long long A_bstride;
    int nsum;                    // partial accumulators summed on load (>= 1)
    long long A_sstride;
    int rfX, CB, Yout;
    RadixPlan rpY;
    const float2* twY;           // exp(-2 pi i m / fY)
    float2* out;
    long long out_bstride;
};

// Every active grid group's arguments of one stage, by value in the kernel
// arguments (pointers read from kernel arguments are known to be global, so
// the kernels load through them with global rather than flat loads).
struct SmallFFTGrouped {
    SmallFFTArgs a[FLF_MAX_GROUPS];
    GroupedGrid gg;
};
struct MacGrouped {
    MacArgs a[FLF_MAX_GROUPS];
    GroupedGrid gg;                               // blocks per group: nchunks x nranges x nfb, padded to 8
    int nfb;                                      // frame blocks
};
struct IfftColGrouped {
    IfftColArgs a[FLF_MAX_GROUPS];
    GroupedGrid gg;
};

struct IfftRowGroup {
    const float2* T;             // (T, Y, rfX)
    long long tstride;
    int rfX;
    int active;
    float scale;
    RadixPlan rp;                // length fX
    RadixPlan rph;               // length fX/2 (even fX)
    const float2* twf;           // exp(-2 pi i m / fX)
    const float2* twh;           // exp(-2 pi i m / (fX/2))
};

struct IfftRowSumArgs {
    IfftRowGroup grp[FLF_MAX_GROUPS];
    int ngroups;
    int maxfX;
    int X, Y;
    float* img;                  // forward image, or null (not stored)
    long long img_tstride;
    const int2* fdims;           // per frame valid size; outside it the image is 0
    const float* im;             // with ratio: the measured image (T, Y, X)
    float* ratio;                // if set: ratio = im / image (standard RL, lfdeconv.py:191), 0 outside
};

struct RowPlaneDesc {
    long long T_off;             // element offset of this plane's (T, Y, rfX) block
    long long T_tstride;
    int rfX;
    float scale;
    RadixPlan rp;                // length fX
    RadixPlan rph;               // length fX/2 (even fX)
    const float2* twf;
    const float2* twh;
};

struct IfftRowPlaneArgs {
    const RowPlaneDesc* pd;      // per plane z
    unsigned short plane_ids[FLF_MAX_PLANES];   // planes to process (grid z), by value
    const float2* Tbuf;
    float* vol;                  // (nz, T, Y, X)
    const float* Htf;            // mode 1
    const float* trust;          // mode 2 (optional, (nz, 1, Y, X))
    int mode;                    // 0 store, 1 Prevedel update, 2 standard update
    int T, Y, X;
    int wbuf;                    // per-wave LDS buffer (float2): max over planes of fX/2 (even) or fX
    const int2* fdims;           // per frame valid size; outside it the volume is set to 0
    int prime;                   // some plane's row plan ends with an odd prime radix above 5
    int big;                     // some plane's rows are longer than 2048 (out-of-place transforms)
};

namespace flf {
size_t small_fft_lds(int rY, int rX);
size_t ifft_cols_lds(int CB, int fY);
// LDS stride of one column of the column transform: fY padded so that the CB
// columns of one row, written and read by consecutive threads, fall on
// distinct banks (FLF_COLS_BANKS banks of 4 bytes; 0: unpadded).
#ifndef FLF_COLS_BANKS
#define FLF_COLS_BANKS 64
#endif
__host__ __device__ inline int ifft_col_stride(int fY, int CB) {
    if (FLF_COLS_BANKS == 0 || CB < 2) return fY;
    const int m = FLF_COLS_BANKS / 2;      // float2 entries per sweep of the banks
    const int r = (m / CB) % m;
    return fY + ((r - fY) % m + m) % m;
}
size_t ifft_rows_sum_lds(int maxfX);
size_t ifft_rows_plane_lds(int wbuf, bool big);
bool wave_big(int L);
bool small_fft_big(int rY, int rX);
int c2r_wave_buffer(int fX);
hipError_t launch_small_fft(const SmallFFTArgs& a, int nsrc, int T, hipStream_t s);
// Grouped launches (the groups' arguments by value).
hipError_t launch_small_fft_grouped(const SmallFFTGrouped& ga, int N, int T, size_t lds, hipStream_t s);
hipError_t launch_mac_grouped(const MacGrouped& ga, int slots, int ipt, int fb, int nfb, size_t lds, hipStream_t s);
hipError_t launch_ifft_cols_grouped(const IfftColGrouped& ga, int CB, bool prime, size_t lds, hipStream_t s);
hipError_t launch_psf_rows(const PsfRowArgs& a, int nu, hipStream_t s);
hipError_t launch_psf_cols(const PsfColArgs& a, int nu, hipStream_t s);
hipError_t launch_spec_pack(const SpecPackArgs& a, int nchunks, hipStream_t s);
int mac_row_stride(int slots, int nu);
size_t mac_tc_lds(int slots, int N);
size_t mac_lds(int slots, int nu, int N, int fb);
hipError_t launch_mac(const MacArgs& a, int slots, int ipt, int fb, int nchunks, int nplanes_grid, bool mfma,
                      hipStream_t s);
hipError_t launch_ifft_cols(const IfftColArgs& a, int nbatch, hipStream_t s);
hipError_t launch_ifft_rows_sum(const IfftRowSumArgs& a, int Y, int T, hipStream_t s);
hipError_t launch_ifft_rows_plane(const IfftRowPlaneArgs& a, int nplanes, hipStream_t s);
// Data-fit sums: frame t of img against frame t % im_frames of im; part holds
// 2 * nblk * T doubles of partial sums, out 2 * T.
hipError_t launch_fit(const float* img, const float* im, long long n_per_frame, int T, int im_frames, double* part,
                      int nblk, double* out, hipStream_t s);
hipError_t launch_fill(float* p, long long n, float v, hipStream_t s);
// X = [X *] Htf / v over nslab (plane, frame) slabs of npix values; v holds
// one slab per plane (v_frames = 1, shared by the T frames) or per slab
// (v_frames = T).
hipError_t launch_rl_scale(float* X, const float* Htf, const float* v, int nz, int T, long long npix, int v_frames,
                           hipStream_t s, bool ratio = false);
// Whether a radix plan ends with an odd prime radix above 5.
bool plan_has_prime(const RadixPlan& rp);
// Raise the dynamic-LDS limit of the FFT kernels to max_lds bytes.
hipError_t configure_kernels(size_t max_lds);
}  // namespace flf
