// ozaki.h -- work descriptors of the split-precision (INT8 / CRT) path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sipgpu {

constexpr int kOzMaxModuli = 16;
constexpr int kOzBM = 128, kOzBN = 128; // INT8 GEMM workgroup tile
constexpr int kOzBigTile = 256;         // the large-task GEMM's tile (both dims)
// tile code of the 256 x 256 GEMM with 8 waves of 64 x 128 (two per SIMD)
constexpr int kOzWideTile = 257;
// tile code (and edge) of the packed FP64 GEMM's 64 x 64 tile
constexpr int kOzSmallTile = 64;
// doubles per workspace slot of a split tile's K segment (any tile edge)
constexpr int kOzWorkSlot = kOzBM * kOzBN;
// edge of the output tile a GEMM tile code stands for
constexpr int oz_tile_size(int code) { return code == kOzWideTile ? kOzBigTile : code; }
constexpr int kOzChunk = 16;            // bytes of k per panel chunk
constexpr int kOzStage = 128;           // bytes of k per GEMM pipeline stage
constexpr int kOzStageChunks = kOzStage / kOzChunk;
constexpr int kOzZero = 256;            // zero bytes at the start of every plane
constexpr int kOzTileX = 32, kOzTileK = 128; // conversion tile (rows x k)
constexpr int kOzRecChunk = 2048;       // result elements per reconstruction workgroup

// Moduli and reconstruction constants (host-computed).
struct OzConst {
    int n;                        // number of moduli
    int P;                        // bits of the scaled integers (|A_int| <= 2^(P-1))
    int p[kOzMaxModuli];          // pairwise coprime, <= 255
    int u[kOzMaxModuli];          // (M / p_i)^-1 mod p_i
    int c17[kOzMaxModuli];        // 2^17 mod p_i
    int c34[kOzMaxModuli];        // 2^34 mod p_i
    int off[kOzMaxModuli];        // p_i * 2^19: makes the limb sum positive
    uint32_t bar[kOzMaxModuli];   // floor(2^32 / p_i) (Barrett reciprocal)
    double rp[kOzMaxModuli];      // 1 / p_i (rounded)
    uint64_t mi[kOzMaxModuli][2]; // M / p_i exactly, {low, high} 64-bit halves
    uint64_t m[2];                // M = prod p_i exactly (< 2^112)
};

// One operand block as the split path sees it: K x X through its tables,
// converted once per execution into X row-major residue rows of Kp bytes
// (K rounded up to a whole chunk, zero filled).
struct OzBlock {
    const double *data;
    int K, X, Kp;
    int ktab, xtab;        // offsets of its k / x tables in the pool
    int exp_off;           // first exponent of its row component
    int xfast;             // 1: consecutive rows are adjacent in memory
    int pad;
    long long panel_off;   // byte offset inside each modulus plane
};

// 16 bytes of a task's virtual K: where the left and right operand rows
// hold them (plane offsets of row 0 -- bytes in a residue plane, 16-byte
// units in a plane of doubles; row r adds r * stride bytes).
struct OzChunkRef {
    uint32_t a, b;
    int32_t stride;
    int32_t pad;
};

// One result block of a merged task: where its elements go.
struct OzSub {
    double *D;
    int mD, nD;           // destination tables
    int beta;             // 1: accumulate into D
    int nfast;
};

// A GEMM task: one result block, or a grid of result blocks that share a
// virtual K (rmap >= 0).  Row m of a grid belongs to block row
// tab[rmap + 2m] at row tab[rmap + 2m + 1] inside it (columns likewise
// through cmap); block (g, h) is subs[sub0 + g * ncp + h].
struct OzTask {
    double *D;            // single block: destination and its tables
    int M, N;
    int mD, nD;
    int chunk0, nchunks;  // virtual K: chunks [chunk0, chunk0 + nchunks), whole stages
    int beta;
    int expA, expB;       // first row / column exponent
    int nfast;            // 1: result elements ordered with n fastest
    long long res_off;    // element offset inside each result plane
    int rmap, cmap;       // grid: row / column maps in the table pool, else -1
    int sub0, ncp;        // grid: first block record, block columns
};

// Destination offset of element (m, n) of a task.
__device__ __forceinline__ double *oz_dest(const OzTask &t, const OzSub *subs, const int32_t *tab,
                                           int m, int n, int &beta) {
    if (t.rmap < 0) {
        beta = t.beta;
        return t.D + tab[t.mD + m] + tab[t.nD + n];
    }
    const int g = tab[t.rmap + 2 * m], lm = tab[t.rmap + 2 * m + 1];
    const int h = tab[t.cmap + 2 * n], ln = tab[t.cmap + 2 * n + 1];
    const OzSub s = subs[t.sub0 + g * t.ncp + h];
    beta = s.beta;
    return s.D + tab[s.mD + lm] + tab[s.nD + ln];
}

// A GEMM work unit: tile (mt, nt) of a task over stages [s0, s1) of its
// virtual K (whole stages).  part < 0: the whole K, the product is
// stored; part >= 0: one K segment of a split tile, its accumulators go to
// workspace slot `part` and the fixup pass sums the tile's segments.  In the
// fixup list, part is the tile's first slot and s0 its number of segments.
struct OzTile {
    int task, mt, nt, part;
    int s0, s1, pad0, pad1;
};

// Conversion work item: rows [x0, x0 + kOzTileX) and k [k0, k0 + kOzTileK) of a block.
struct OzConvTile {
    int block, x0, k0, pad;
};

// Reconstruction work item: elements [e0, e0 + kOzRecChunk) of a task.
struct OzRecItem {
    int task, e0;
};

struct OzGemmArgs {
    const OzTile *tiles;
    const OzTask *tasks;
    const OzChunkRef *chunks;
    const int8_t *planes;
    long long plane_bytes;
    uint8_t *res_planes;
    long long res_plane_bytes;
    int ntiles;
    int tile; // tile code: kOzBM (= kOzBN), kOzBigTile, kOzWideTile, kOzSmallTile (FP64)
    const int32_t *tab; // destination tables (packed FP64 path)
    const OzSub *subs;  // result blocks of merged tasks (packed FP64 path)
    double *work;       // split-tile partial products, kOzWorkSlot per slot
    const OzTile *fix;  // split tiles to sum and store
    int nfix;
    const int *wg_first; // persistent launch: units [wg_first[w], wg_first[w+1]) of
    int nwg;             // workgroup w (nullptr: one tile per workgroup)
    OzConst c;
};

struct OzConvArgs {
    const OzBlock *blocks;
    const OzConvTile *items;
    int nitems;
    const int32_t *tab;
    int *expo;
    int8_t *planes;
    long long plane_bytes;
    OzConst c;
};

struct OzRecArgs {
    const OzTask *tasks;
    const OzRecItem *items;
    int nitems;
    const int32_t *tab;
    const OzSub *subs;
    const int *expo;
    const uint8_t *res_planes;
    long long res_plane_bytes;
    OzConst c;
};

hipError_t oz_fill_int(int *p, int n, int v, hipStream_t s);
// Zeroes the kOzZero bytes at the start of each of `nplanes` planes.
hipError_t oz_zero_heads(int8_t *planes, long long plane_bytes, int nplanes, hipStream_t s);
// Row maxima -> exponents, one workgroup per kOzTileX-row tile of a block
// (`row_items`, all with k0 = 0); each sweeps the block's whole K.
hipError_t oz_row_exponents(const OzConvArgs &a, const OzConvTile *row_items, int nrow_items,
                            hipStream_t s);
hipError_t oz_to_residues(const OzConvArgs &a, hipStream_t s);
hipError_t oz_gemm(const OzGemmArgs &a, hipStream_t s);
hipError_t oz_gemm_big(const OzGemmArgs &a, hipStream_t s);
// Packed FP64 path: operand panels of doubles, then the GEMM over them.
hipError_t oz_pack(const OzConvArgs &a, hipStream_t s);
hipError_t oz_gemm_f64(const OzGemmArgs &a, hipStream_t s);
// Sums the K segments of split tiles (in segment order) and stores them.
hipError_t oz_fixup_f64(const OzGemmArgs &a, hipStream_t s);
hipError_t oz_reconstruct(const OzRecArgs &a, hipStream_t s);

} // namespace sipgpu
