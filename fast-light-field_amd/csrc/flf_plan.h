// Host-side planning: Fourier grid shapes, FFT radix plans and the orbit/item
// tables that drive the multiply-accumulate kernel.
#pragma once
#include <string>
#include <vector>
#include "flf_types.h"

namespace flf {

// Smallest m >= n with no prime factor above 5.
int smooth5_at_least(int n);

// Fourier grid length for one axis: N * smooth5(ceil((s1 + s2 - 1) / N)).
// Same linear convolution as convolutionShape (special_fftconvolve.py:142-174);
// only the zero padding differs (5-smooth instead of 7-smooth or minimal).
int fourier_len(int s1, int s2, int N);

// Radix plan (4s first, then 2, 3, 5).  Returns false if L is not 5-smooth.
// Pass orders of a plan (the autotuner's radix_order): 0 largest radices
// first; 1 the same passes reversed; 2 (wave-level plans) without radix 8;
// 3 (wave-level plans) the odd radices first.  An odd prime factor above 5
// is always the last pass.
constexpr int FLF_RADIX_ORDERS = 4;
bool make_radix_plan(int L, RadixPlan& rp, bool wide = false, int order = 0);

// Index tables for one Fourier grid (fY, fX).
struct GridTables {
    int fY = 0, fX = 0, rY = 0, rX = 0, rfX = 0;
    bool square = false;
    int slots = 0;               // 4 (square) or 2
    int nu = 0;                  // stored spectra per plane: h(h+1)/2 or h*h
    int max_res = 0;             // largest residue list of any chunk
    int ipt = 1;                 // items per MAC thread
    long long real_items = 0;    // items excluding padding copies
    std::vector<int> u_bu, u_au; // stored spectrum u -> quadrant (b, a) (b > a: transposed)
    std::vector<int> uoff;       // [nu][uses]: offset b*N+a of each use of u, N*N if absent
    int nuse = 0;                // uses per stored spectrum: 8 (quads) or 4 (pairs)
    std::vector<ChunkDesc> chunks;
    std::vector<ItemDesc> items;
    std::vector<TileDesc> tiles;
};

// Builds the tables; returns an empty string or an error message.
// items_per_thread > 1 groups the items of each chunk so that every run of
// items_per_thread consecutive items shares its residues (one MAC thread
// handles such a run); classes whose size is not a multiple are padded with
// copies that never write.
std::string build_grid_tables(int N, int fY, int fX, GridTables& gt, int items_per_thread = 1);

}  // namespace flf
