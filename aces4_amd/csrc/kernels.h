// kernels.h -- device work descriptors and launch entry points (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.h"

namespace sipgpu {

// ---------------------------------------------------------------- permute --
constexpr int kPermMaxDims = 8;

struct PermArgs {
    const double *in;
    double *out;
    int nU, nV;                   // tile-row dims (leading input dims), other dims
    int extU[kPermMaxDims];       // extents of U dims (input order)
    int soutU[kPermMaxDims];      // their output strides
    int extV[kPermMaxDims];       // extents of V dims (output order)
    int sinV[kPermMaxDims];       // their input strides
    int soutV[kPermMaxDims];      // their output strides
    int sizeU, sizeV;
    int TU, TV, pitch;            // tile extents and LDS row pitch
    int tilesU;                   // tiles along U
    int row_mode;                 // U is contiguous in the output as well
};

struct PermutePlan;
hipError_t launch_permute(const PermutePlan &plan, const double *in, double *out,
                          hipStream_t stream);
// Batched permutes: fills args[0..n), the workgroup offsets first[0..n]
// (first[n] the grid), the launch's LDS needs (largest staged tile in
// doubles, largest tile v extent) and its elements per thread (8 or 16)
// from the plans; returns the total workgroup count.
int64_t permute_batch_args(const PermutePlan *plans, const double *const *in,
                           double *const *out, int n, PermArgs *args, int *first,
                           int *tile_cap, int *tv_cap, int *items);
// One launch of `grid` workgroups over the device copies of args, first and
// the per-workgroup entry table entry[0..grid) (entry[b] = the e with
// first[e] <= b < first[e + 1]).
hipError_t launch_permute_batch(const PermArgs *args, const int *first, const int *entry,
                                int grid, int tile_cap, int tv_cap, int items,
                                hipStream_t stream);

// ------------------------------------------------------------- contraction --
// One operand pair of a task.  Table offsets index the plan's int32 pool.
struct GemmPair {
    const double *L;
    const double *R;
    int K;
    int kL, mL, kR, nR;           // offsets of koffL, moffL, koffR, noffR
    int flags;                    // bit0: L k-fast, bit1: R k-fast
};

// One destination block: D(m, n) at D[moffD[m] + noffD[n]].
struct GemmTask {
    double *D;
    int M, N;
    int mD, nD;                   // offsets of moffD, noffD
    int first, npairs;            // pair range
    int beta;                     // 0: assign, 1: accumulate
    int pad;
};

struct GemmTile {
    int task, mt, nt, pad;
};

// When `tiles` is null the launch is a single contraction: task0 with its
// one pair0, tile t at (t % tilesM, t / tilesM).  This lets a lone
// contraction run without uploading a work list.
struct GemmArgs {
    const GemmTile *tiles;
    const GemmTask *tasks;
    const GemmPair *pairs;
    const int32_t *tab;
    int ntiles;
    int tilesM;
    GemmTask task0;
    GemmPair pair0;
    // single contraction split along K (ksplit > 1, 64 x 64 kernel only):
    // tile t computes segment s = t / (ntiles / ksplit) of kseg k values;
    // segment 0 goes to task0.D, segment s > 0 is assigned to ws + (s - 1) *
    // ws_stride with task0's destination tables
    int ksplit, kseg;
    double *ws;
    long long ws_stride;
};

// Tile configurations of the grouped FP64-MFMA kernel (BM x BN, waves, BK):
//   0: 64x64 2x2 16   1: 128x128 2x2 16   2: 64x128 2x2 16   3: 64x64 2x2 32
//   4: 128x128 4x2 16 5: 128x64 2x2 16    6: 32x64 1x2 16
// double-buffered pair-flattened kernel:
//   7: 128x128 2x2 8  8: 128x128 2x2 16   9: 128x64 2x2 16   10: 64x64 2x2 16
// first kernel with a register budget (waves per SIMD):
//   11: 64x64 (4)     12: 128x64 (2)      13: 128x128 (2)
// direct-to-LDS kernel (2x2 waves, BK 16):
//   14: 64x64  15: 128x64  16: 128x128  17: 64x128
enum GemmConfig { kGemm64x64 = 0, kGemm128x128 = 1, kGemm64x128 = 2, kGemmConfigs = 18 };
int gemm_tile_m(int cfg);
int gemm_tile_n(int cfg);
hipError_t launch_gemm(int cfg, const GemmArgs &args, int grid, hipStream_t stream);

// Small kernels for the non-GEMM contraction classes of the reference
// driver (tensor_dil_omp.F90:766-777).
// D[moffD[m]] = 0 + X[moffX[m]] * s', s' = 1 when |s-1| <= 1e-13.
hipError_t launch_scale_scatter(double *D, const int32_t *moffD, const double *X,
                                const int32_t *moffX, const double *s, int M,
                                hipStream_t stream);
// D[0] = 0 + sum_k L[koffL[k]] * R[koffR[k]]; partial sums in `scratch`
// (at least kDotBlocks doubles); fixed summation order.
constexpr int kDotBlocks = 64;
hipError_t launch_dot(double *D, const double *L, const int32_t *koffL, const double *R,
                      const int32_t *koffR, int K, double *scratch, hipStream_t stream);
// D[0] = L[0] * R[0]
hipError_t launch_scalar_mul(double *D, const double *L, const double *R, hipStream_t stream);

// Calibration kernels (calib.hip).
hipError_t launch_mfma_peak(double *out, long long *ticks, int blocks, int iters, hipStream_t s);
// LDS read bandwidth loop, 8 or 16 waves per workgroup (calib.hip).
hipError_t launch_lds_read(int waves, int blocks, int *out, long long *ticks, int iters,
                           hipStream_t s);
hipError_t launch_copy(const double *in, double *out, size_t n, hipStream_t s);
// Orbital-energy denominators: rank <= kEnergyMaxRank block, one energy
// vector and one sign (+1 occupied, -1 virtual) per index.
constexpr int kEnergyMaxRank = 6;
constexpr int kEnergyBlocks = 4096; // 16 waves per CU: the reduction's gathers need them
struct EnergyArgs {
    int rank;
    int dims[kEnergyMaxRank];
    const double *eps[kEnergyMaxRank];
    int sign[kEnergyMaxRank];
    long long n;
};
hipError_t launch_denominator_divide(const EnergyArgs &a, double *x, hipStream_t s);
// out[0] += factor * sum_e (x[e] / D(e)) * y[e]; `part` holds kEnergyBlocks doubles.
hipError_t launch_energy_sum(const EnergyArgs &a, const double *x, const double *y,
                             double factor, double *out, double *part, hipStream_t s);
hipError_t launch_fill_random(double *out, long long n, unsigned long long seed,
                              unsigned long long stream, int dist, double a, double b,
                              hipStream_t s);

// ------------------------------------------------------------- block ops --
// Elementwise block operations (tensor_dil_omp.F90:144-228, :394-436):
// fill a = s, scale a *= s, scaled add a = a + b * s (product and sum rounded
// separately), add a = a + b; n elements, nothing done for n <= 0.
enum BlockOp { kOpFill = 0, kOpScale = 1, kOpAddScaled = 2, kOpAdd = 3 };
hipError_t launch_block_op(int op, double *a, const double *b, long long n, double s,
                           hipStream_t st);
// t = ((t + p[0]) + p[1]) + ... over n elements, at most kAddManyMax sources.
constexpr int kAddManyMax = 8;
struct AddManyArgs {
    const double *p[kAddManyMax];
    int n;
};
hipError_t launch_block_add_many(double *t, const AddManyArgs &a, long long n, hipStream_t st);
// Result blocks of a batch split into pair groups: t = ((t + p[0]) + p[1]) +
// ... + p[nparts - 1] over n elements, part q at p + q n.
struct AddPartsEntry {
    double *t;
    const double *p;
    long long n;
    int nparts, pad;
};
// One launch over `count` device entries (count <= 65535), each of at most
// max_n elements.
hipError_t launch_add_parts(const AddPartsEntry *entries, int count, long long max_n,
                            hipStream_t st);
// launch_block_add_many, then launch_energy_sum with x = y = t, in one pass
// (the same result bits as the two calls).
hipError_t launch_energy_sum_parts(const EnergyArgs &a, double *t, const AddManyArgs &ps,
                                   double factor, double *out, double *part, hipStream_t s);
// out[0] = sum x[i]^2 over n >= 0 elements: per-workgroup partials (at most
// kNormBlocks, in `part`) then one workgroup; the order depends on n only.
constexpr int kNormBlocks = 1024;
int norm2_blocks(long long n);
hipError_t launch_block_norm2(const double *x, long long n, double *out, double *part,
                              hipStream_t st);
// A slice plan in kernel form: n slice elements, rank >= 1 dims.
struct SliceArgs {
    int rank, n;
    FastDiv ext[kSliceMaxRank];
    int tstride[kSliceMaxRank];
};
SliceArgs make_slice_args(const SlicePlan &p);
// insert = false: slc[l] = tens[off(l)]; true: tens[off(l)] = slc[l], where
// `tens` points at the slice's first element (the plan's base).
hipError_t launch_slice(const SliceArgs &a, bool insert, double *tens, double *slc,
                        hipStream_t st);

// ---------------------------------------------------------------- triples --
// The (T) energy with W kept on chip (triples.hip).  Indices are cut into
// 5-blocks within their segments; a box is one 5-block per W index.
constexpr int kTriplesBox = 5;
constexpr int kTriplesSums = 4096; // partial sums of the final reduction
struct TriplesBlk {
    int seg;    // segment
    int off;    // offset of the 5-block inside the segment
    int ext;    // extent, 1..5
    int gstart; // global orbital index of its first element
};
// Packs PA (side 0: sets (x, y1, y2) over vblk x oblk x oblk) or PB (side 1:
// sets (x2, x3, y3) over vblk x vblk x oblk) from the blocks of T(a,d,i,j),
// V1(b,c,d,k) and V2(l,c,j,k), given as device pointer tables in segment
// order (first index slowest).
struct TriplesPackArgs {
    double *out;
    long long total; // doubles of the panel array
    int side, nkp, kd, kl;
    int n1, n2;      // 5-block counts of the second and third set index
    const TriplesBlk *blk0, *blk1, *blk2;
    int nv, no;      // segment counts
    const int *vext, *oext;             // segment extents
    const int *vseg_of, *oseg_of;       // segment of each global index
    const int *vseg_start, *oseg_start; // global index of each segment's start
    const double *const *T, *const *V1, *const *V2;
};
struct TriplesArgs {
    const double *PA, *PB;
    int nkp, nvb, nob;
    const TriplesBlk *vblk, *oblk;
    const double *e_vir, *e_occ;
    double *partial;
    // global mode: linear box ids [box0, box0 + nbox) of the super-box order
    long long box0, nbox, per_xcd;
    int super[6], nblk[6];
    // tuple mode: the boxes first[d] .. first[d] + count[d] - 1
    int tuple_mode;
    int first[6], count[6];
    double *wout;   // tuple mode, optional: W of the tuple (column-major)
    int wdims[6];
};
hipError_t launch_triples_pack(const TriplesPackArgs &a, hipStream_t s);
// grid workgroups of 512; hipErrorInvalidConfiguration at 2^23 workgroups or more
hipError_t launch_triples_boxes(const TriplesArgs &a, long long grid, hipStream_t s);
// out[0] += factor * sum of partial[0 .. n) in a fixed order (sums: kTriplesSums doubles)
hipError_t launch_triples_reduce(const double *partial, long long n, double *sums, double factor,
                                 double *out, hipStream_t s);

} // namespace sipgpu
