// Plain table types shared by the host planner and the kernels (no HIP
// dependency, so the planner also builds with a host-only compiler).
#pragma once

#define FLF_MAX_RADIX_PASSES 16
#define FLF_MAX_RES 8          // residues per orbit (D4 orbit of a reduced bin)
#define FLF_CHUNK_MAX 256      // items per MAC workgroup
#define FLF_MAX_FOURIER 16384  // Fourier grid length per axis (keeps phase products in 32 bits)
#define FLF_MAX_N 31           // largest supported Nnum (s-table = N*N entries)
#define FLF_PRIME_RADIX_MAX 31 // largest odd prime radix of a transform (a prime factor of Nnum)
#define FLF_TILE 16            // items per matrix-core tile (one residue class)
#define FLF_MAX_TILES 24       // tiles per chunk: <= FLF_CHUNK_MAX / FLF_TILE + FLF_MAX_RES

struct RadixPlan {
    int L;                      // transform length
    int n;                      // number of passes
    int r[FLF_MAX_RADIX_PASSES];
};

// One MAC workgroup: a chunk of the items whose bins fall on one residue orbit.
struct ChunkDesc {
    long long item_base;        // first item (global index within the grid group)
    int n_items;
    int n_res;
    int tile_base;              // first tile (global index within the grid group)
    int n_tiles;
    int res0[FLF_MAX_RES];      // residue row    (kappa0 in [0, rY))
    int res1[FLF_MAX_RES];      // residue column (kappa1 in [0, rX))
};

// A run of at most FLF_TILE consecutive items of one chunk that share their
// slot residues (one residue class): the unit of the matrix-core MAC.
struct TileDesc {
    int start;                  // first item, relative to the chunk
    int count;                  // items in the tile (<= FLF_TILE)
    unsigned char res[4];       // slot residues of every item in the tile
};

// One item: 4 (square grid) or 2 (non-square) half-plane bins related by
// row negation / transposition; the MAC computes all of them together.
struct ItemDesc {
    int pos[4];                 // k0 * rfX + k1 of each slot
    unsigned char res[4];       // residue index of each slot's bin in the chunk's residue list
    unsigned char m2[4];        // alias row index, in half-units mod 2N
    unsigned char n2[4];        // alias column index, in half-units mod 2N
    int writemask;              // bit j set: slot j writes its bin (first occurrence only)
};
