// plan.h -- host-side planning of permutes and contractions.
//
// A contraction D = L (x)_ptrn R is executed as the matrix product
//   D(m, n) = sum_k L(k, m) * R(k, n)
// of the reference driver (tensor_dil_omp.F90:799-859, :762), but the
// operand permutations are never materialised.  Instead every operand gets
// separable index-offset tables: the element (k, m) of the permuted left
// operand lives at L[koffL[k] + moffL[m]], and the result element (m, n)
// goes to D[moffD[m] + noffD[n]].  The kernels read through these tables, so
// the permutations cost no extra HBM pass.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fastdiv.h"

namespace sipgpu {

constexpr int kMaxRank = 6;     // MAX_RANK of the SIP (aces_defs.h:13)
constexpr int kRefMaxRank = 32; // max_tensor_rank of the reference kernels

// get_contraction_ptrn: ACES labels -> +k / -k pattern.
int convert_aces_pattern(int drank, int lrank, int rrank, const int *aces,
                         int *my_ptrn);

// contr_ptrn_ok: every destination dim hit once, contracted extents equal.
bool pattern_valid(const int *ptrn, int lrank, const int *ldims, int rrank,
                   const int *rdims, int drank, const int *ddims);

// Everything the kernels need for one contraction shape.
struct ContractionShape {
    int lrank = 0, rrank = 0, drank = 0;
    int64_t M = 1, N = 1, K = 1;          // lld, lrd, lcd
    int64_t lsize = 1, rsize = 1, dsize = 1;
    // O2N maps in the reference's convention (1-based new positions).
    int lo2n[kRefMaxRank + 1] = {0};
    int ro2n[kRefMaxRank + 1] = {0};
    int do2n[kRefMaxRank + 1] = {0};
    bool ltransp = false, rtransp = false, dtransp = false;
    // Offset tables (element offsets, int32).
    std::vector<int32_t> koffL, moffL, koffR, noffR, moffD, noffD;
    // True if consecutive k are adjacent in memory for that operand at the
    // start of the table (the operand's fastest dimension is contracted).
    bool l_kfast = false, r_kfast = false;
    // True if the destination's fastest dimension belongs to M only: the
    // kernel then computes D^T = R^T L so that lanes write along it.
    bool swap = false;
};

// determine_index_permutations + table construction.  The contracted dims
// are summed in the reference's order (the right operand's) except that a
// unit-stride contracted dim of an operand is moved first, so that operand
// can be read with consecutive lanes on consecutive addresses; only the
// summation order changes, never the set of products.  Returns 0, or the
// reference's ierr (1 bad pattern, -1 bad ranks), or SIPGPU_ESIZE.
int plan_contraction(const int *ptrn, int lrank, const int *ldims, int rrank,
                     const int *rdims, int drank, const int *ddims,
                     ContractionShape &out, bool build_tables);

// Canonical form of a permutation (extent-1 dims dropped, dims that stay
// adjacent merged), as consumed by the permute kernel.
struct PermutePlan {
    int rank = 0;                 // canonical rank
    int64_t size = 1;             // total elements
    int ext[kRefMaxRank] = {0};   // canonical input extents
    int64_t sout[kRefMaxRank] = {0}; // output stride of each canonical input dim
    int nU = 0;                   // leading input dims forming the tile rows
    int vdim[kRefMaxRank] = {0};  // remaining dims, in output order
    int nV = 0;
    int64_t sizeU = 1, sizeV = 1;
    bool identity = false;        // plain copy
    bool row_mode = false;        // input-fastest dim is also output-fastest
};

// Validates dim_transp (1-based O2N with sign slot) and builds the plan.
// Returns 0, SIPGPU_EPERM or SIPGPU_ESIZE.
int plan_permute(int rank, const int *dims, const int *dim_transp,
                 PermutePlan &out);

// Slice extraction / insertion (tensor_dil_omp.F90:271-392) in canonical
// form: slice dims of extent 1 dropped, and a dim merged into the previous
// one where the slice's element order stays contiguous in the tensor block.
// Slice element l (column-major digits d_i over ext) is tensor element
// base + sum_i d_i * tstride[i].
constexpr int kSliceMaxRank = 8;
struct SlicePlan {
    int rank = 0;                         // canonical dims (0: one element)
    int ext[kSliceMaxRank] = {0};         // their slice extents
    int64_t tstride[kSliceMaxRank] = {0}; // their tensor strides
    int64_t base = 0;                     // tensor offset of slice element 0
    int64_t count = 0;                    // slice elements (0: empty slice)
    int64_t span = 0;                     // tensor elements base .. last touched, inclusive
    int64_t tsize = 0;                    // tensor elements
};

// Rank 0 is the reference's scalar case (one element each side).  Checks
// every dim: 0 <= ext_beg, 0 <= slice_ext and ext_beg + slice_ext <=
// tens_ext (the reference checks none of this and would read or write out
// of bounds).  Returns 0, SIPGPU_ERANK (rank < 0 or > kRefMaxRank, or more
// than kSliceMaxRank canonical dims), SIPGPU_EARG (a bound violated) or
// SIPGPU_ESIZE (a block of 2^31 elements or more).
int plan_slice(int rank, const int *tens_ext, const int *slice_ext, const int *ext_beg,
               SlicePlan &out);

} // namespace sipgpu
