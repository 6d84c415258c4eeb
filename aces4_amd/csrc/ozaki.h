// ozaki.h -- work descriptors of the split-precision (INT8 / CRT) path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sipgpu {

constexpr int kOzMaxModuli = 16;
constexpr int kOzBM = 128, kOzBN = 128; // INT8 GEMM workgroup tile
constexpr int kOzKAlign = 64;           // K of one v_mfma_i32_16x16x64_i8

// Moduli and reconstruction constants (host-computed).
struct OzConst {
    int n;                       // number of moduli
    int P;                       // bits of the scaled integers (|A_int| <= 2^(P-1))
    int p[kOzMaxModuli];         // pairwise coprime, <= 255
    int u[kOzMaxModuli];         // (M / p_i)^-1 mod p_i
    int t24[kOzMaxModuli];       // 2^24 mod p_i
    double rp[kOzMaxModuli];     // 1 / p_i (rounded)
    uint64_t mi[kOzMaxModuli][2]; // M / p_i exactly, {low, high} 64-bit halves
    uint64_t m[2];               // M = prod p_i exactly (< 2^112)
};

// One operand block as the split path sees it: K x X through its tables,
// converted once per execution into X row-major residue rows of Kp bytes.
struct OzBlock {
    const double *data;
    int K, X, Kp;
    int ktab, xtab;        // offsets of its k / x tables in the pool
    int exp_off;           // first exponent of its row component
    int pad;
    long long panel_off;   // byte offset inside each modulus plane
};

struct OzPair {
    int a, b; // OzBlock indices of the left and right operand
    int Kp;
    int pad;
};

struct OzTask {
    double *D;
    int M, N;
    int mD, nD;           // destination tables
    int first, npairs;
    int beta;
    int expA, expB;       // exponent offsets of its row / column components
    int pad;
    long long res_off;    // element offset inside each result plane
};

struct OzTile {
    int task, mt, nt, pad;
};

struct OzGemmArgs {
    const OzTile *tiles;
    const OzTask *tasks;
    const OzPair *pairs;
    const OzBlock *blocks;
    const int8_t *planes;
    long long plane_bytes;
    int8_t *res_planes;
    long long res_plane_bytes;
    int ntiles;
    OzConst c;
};

struct OzRecArgs {
    const OzTask *tasks;
    int ntasks;
    const long long *elem_start; // prefix of M*N over tasks (ntasks + 1)
    long long nelems;
    const int32_t *tab;
    const int *expo;
    const int8_t *res_planes;
    long long res_plane_bytes;
    OzConst c;
};

hipError_t oz_fill_int(int *p, int n, int v, hipStream_t s);
hipError_t oz_row_exponents(const OzBlock *blocks, int nblocks, const int32_t *tab, int *expo,
                            const int *row_start, int nrows, hipStream_t s);
hipError_t oz_to_residues(const OzBlock *blocks, int nblocks, const int32_t *tab, const int *expo,
                          const int *quad_start, int nquads, int8_t *planes,
                          long long plane_bytes, const OzConst &c, hipStream_t s);
hipError_t oz_gemm(const OzGemmArgs &a, int grid, hipStream_t s);
hipError_t oz_reconstruct(const OzRecArgs &a, hipStream_t s);

} // namespace sipgpu
