// permute.hip -- the permute super-instruction on gfx950.
//
// Semantics: tensor_block_copy_ of the reference (tensor_dil_omp.F90:438-660):
// out-of-place N-d transpose of a dense column-major block, old dim i going
// to new position dim_transp[i].  Bit-exact (pure data movement).
//
// Design.  After canonicalisation (plan_permute: extent-1 dims dropped, dims
// that stay adjacent merged) the block is viewed as a 2-D array (u, v):
//   u = the leading input dims (contiguous in the input, never the output's
//       fastest dim), v = the other dims in OUTPUT order (so consecutive v are
//       contiguous in the output).
// A 256-thread workgroup moves one TU x TV tile: it reads rows of u with
// consecutive lanes on consecutive input addresses, stages the tile in LDS
// (odd row pitch, so the transposed read is free of bank conflicts), and
// writes columns of v with consecutive lanes on consecutive output
// addresses.  The per-tile u/v offset tables are decoded once per tile into
// LDS, so the per-element cost is two LDS reads.  When the input's fastest
// dim is also the output's fastest ("row mode") no transpose is needed and
// the tile is copied straight through registers.
#include <algorithm>
#include <cstdint>

#include "kernels.h"
#include "plan.h"

namespace sipgpu {

namespace {

constexpr int kPermThreads = 256;
constexpr int kPermMaxItems = 16;          // elements per thread (at most)
constexpr int kPermMaxTV = 1024;
constexpr int kPermMaxTU = 64;             // transpose mode only

// Dynamic LDS of a launch: the staged tile (tile_cap doubles; none in row
// mode), then the tile's v input and output offsets (tv_cap ints each) and
// its u output offsets (kPermMaxTU ints).  Sized to the largest tile of the
// launch, so small tiles leave room for more workgroups per CU.
__host__ __device__ constexpr size_t perm_lds_bytes(int tile_cap, int tv_cap) {
    return (size_t)tile_cap * sizeof(double) + (size_t)(2 * tv_cap + kPermMaxTU) * sizeof(int);
}

// Tile `blk` of the permute `a` (at most ITEMS * kPermThreads elements),
// staged through the workgroup's LDS.
template <int ITEMS>
__device__ __forceinline__ void permute_tile(const PermArgs &a, int blk, double *lds,
                                             int tile_cap, int tv_cap) {
    double *tile = lds;
    int *sh_inV = reinterpret_cast<int *>(lds + tile_cap);
    int *sh_outV = sh_inV + tv_cap, *sh_outU = sh_outV + tv_cap;
    const int bu = blk % a.tilesU;
    const int bv = blk / a.tilesU;
    const int u0 = bu * a.TU, v0 = bv * a.TV;
    const int tu = min(a.TU, a.sizeU - u0);
    const int tv = min(a.TV, a.sizeV - v0);
    const int tid = threadIdx.x;

    // Decode the tile's v offsets (input and output) and u output offsets.
    for (int i = tid; i < tv; i += kPermThreads) {
        int v = v0 + i, oin = 0, oout = 0;
        for (int d = 0; d < a.nV; ++d) {
            const int e = a.extV[d];
            const int q = v / e;
            const int r = v - q * e;
            oin += r * a.sinV[d];
            oout += r * a.soutV[d];
            v = q;
        }
        sh_inV[i] = oin;
        sh_outV[i] = oout;
    }
    if (!a.row_mode) {
        for (int i = tid; i < tu; i += kPermThreads) {
            int u = u0 + i, oout = 0;
            for (int d = 0; d < a.nU; ++d) {
                const int e = a.extU[d];
                const int q = u / e;
                oout += (u - q * e) * a.soutU[d];
                u = q;
            }
            sh_outU[i] = oout;
        }
    }
    __syncthreads();

    const int n = tu * tv;
    const double *__restrict__ in = a.in;
    double *__restrict__ out = a.out;
    double val[ITEMS];
    // Element e = tid + i * kPermThreads of the tile, as (v, u) = (e / tu,
    // e % tu) or (u, v) = (e / tv, e % tv): one division per thread, then
    // the stride kPermThreads is added in mixed radix (no division per item).
    const int qu = kPermThreads / tu, ru = kPermThreads - qu * tu;
    const int vu0 = tid / tu, uu0 = tid - vu0 * tu;
    if (a.row_mode) {
        // u is contiguous in both the input and the output.
        int vv = vu0, uu = uu0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (tid + i * kPermThreads < n) val[i] = in[u0 + uu + sh_inV[vv]];
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
        vv = vu0;
        uu = uu0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (tid + i * kPermThreads < n)
                __builtin_nontemporal_store(val[i], &out[u0 + uu + sh_outV[vv]]);
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
        return;
    }
    const int pitch = a.pitch;
    {
        int vv = vu0, uu = uu0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (tid + i * kPermThreads < n) val[i] = in[u0 + uu + sh_inV[vv]];
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
        vv = vu0;
        uu = uu0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            if (tid + i * kPermThreads < n) tile[vv * pitch + uu] = val[i];
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
    }
    __syncthreads();
    const int qv = kPermThreads / tv, rv = kPermThreads - qv * tv;
    const int uv0 = tid / tv, vv0 = tid - uv0 * tv;
    int uu = uv0, vv = vv0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        if (tid + i * kPermThreads < n) val[i] = tile[vv * pitch + uu];
        vv += rv;
        uu += qv;
        if (vv >= tv) { vv -= tv; ++uu; }
    }
    uu = uv0;
    vv = vv0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        if (tid + i * kPermThreads < n)
            __builtin_nontemporal_store(val[i], &out[sh_outU[uu] + sh_outV[vv]]);
        vv += rv;
        uu += qv;
        if (vv >= tv) { vv -= tv; ++uu; }
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kPermThreads) void permute_kernel(PermArgs a, int tile_cap,
                                                               int tv_cap) {
    extern __shared__ double perm_lds[];
    permute_tile<ITEMS>(a, blockIdx.x, perm_lds, tile_cap, tv_cap);
}

// Many permutes in one launch: workgroup b runs tile b - first[e] of entry
// e = entry[b] (a table of one int per workgroup, so a workgroup's set-up is
// two dependent loads, not a search).
template <int ITEMS>
__global__ __launch_bounds__(kPermThreads) void permute_batch_kernel(const PermArgs *args,
                                                                     const int *first,
                                                                     const int *entry,
                                                                     int tile_cap, int tv_cap) {
    extern __shared__ double perm_lds[];
    const int b = blockIdx.x;
    const int e = entry[b];
    const PermArgs a = args[e];
    permute_tile<ITEMS>(a, b - first[e], perm_lds, tile_cap, tv_cap);
}

// Tile extent for `size` elements with at most `cap` per tile: the count of
// tiles a cap-sized split needs, with the elements spread evenly over them
// (450 splits as 8 x 57, not 7 x 64 + 2).
int even_tile(int64_t size, int64_t cap) {
    const int64_t n = (size + cap - 1) / cap;
    return (int)((size + n - 1) / n);
}

// Tile shape of one permute: elements per thread (8 or 16, so a tile holds
// at most 2048 or 4096 elements) and the cap on the tile's u extent in
// transpose mode.
struct TileShape {
    int items = kPermMaxItems;
    int tu_cap = kPermMaxTU;
};

// Measured on MI355X over 2-D to 6-D transposes (64^4 reversal, 40^4,
// 20^6, 120^3, 30^5): for a single launch, tiles of 2048 elements with u at
// most 32 wide beat 4096-element tiles on every shape (more workgroups per
// CU keep more loads in flight: 64^4 6.3 against 4.7 TB/s).  In a batched
// launch a workgroup first loads its entry and the entry's arguments from
// memory (two dependent loads that a single launch has in its kernel
// arguments), and the larger tiles amortise that: 144 ring blocks of up to
// 450 x 450 run at 3.0 TB/s with 4096-element tiles, 1.9 with 2048.
TileShape tile_shape(const PermutePlan &p, bool batched) {
    (void)p;
    TileShape t;
    t.items = batched ? 16 : 8;
    t.tu_cap = batched ? 64 : 32;
    return t;
}

// Kernel arguments and workgroup count of one permute (an identity as a
// single contiguous run in row mode).
void fill_args(const PermutePlan &p, const double *in, double *out, PermArgs &a,
               int64_t &grid, const TileShape &ts) {
    const int tile_elems = ts.items * kPermThreads;
    a = PermArgs{};
    a.in = in;
    a.out = out;
    if (p.identity) {
        a.nU = 1;
        a.extU[0] = (int)p.size;
        a.soutU[0] = 1;
        a.nV = 0;
        a.sizeU = (int)p.size;
        a.sizeV = 1;
        a.row_mode = 1;
        a.TU = even_tile(p.size, tile_elems);
        a.TV = 1;
        a.pitch = a.TU;
        a.tilesU = (a.sizeU + a.TU - 1) / a.TU;
        grid = a.tilesU;
        return;
    }
    a.nU = p.nU;
    a.nV = p.nV;
    for (int i = 0; i < p.nU; ++i) {
        a.extU[i] = p.ext[i];
        a.soutU[i] = (int)p.sout[i];
    }
    // Input strides of the canonical dims are their column-major products.
    int64_t sin[kPermMaxDims];
    int64_t acc = 1;
    for (int i = 0; i < p.rank; ++i) { sin[i] = acc; acc *= p.ext[i]; }
    for (int i = 0; i < p.nV; ++i) {
        const int d = p.vdim[i];
        a.extV[i] = p.ext[d];
        a.sinV[i] = (int)sin[d];
        a.soutV[i] = (int)p.sout[d];
    }
    a.sizeU = (int)p.sizeU;
    a.sizeV = (int)p.sizeV;
    a.row_mode = p.row_mode ? 1 : 0;
    a.TU = even_tile(p.sizeU, p.row_mode ? tile_elems : ts.tu_cap);
    a.TV = even_tile(p.sizeV, std::max<int64_t>(1, std::min<int64_t>(tile_elems / a.TU,
                                                                      kPermMaxTV)));
    a.pitch = p.row_mode ? a.TU : (a.TU | 1);
    a.tilesU = (a.sizeU + a.TU - 1) / a.TU;
    const int64_t tilesV = (a.sizeV + a.TV - 1) / a.TV;
    grid = (int64_t)a.tilesU * tilesV;
}

} // namespace

hipError_t launch_permute(const PermutePlan &p, const double *in, double *out,
                          hipStream_t stream) {
    if (p.size <= 0) return hipSuccess;
    if (p.identity)
        return hipMemcpyAsync(out, in, (size_t)p.size * sizeof(double),
                              hipMemcpyDeviceToDevice, stream);
    PermArgs a;
    int64_t grid = 0;
    const TileShape ts = tile_shape(p, false);
    fill_args(p, in, out, a, grid, ts);
    if (grid <= 0 || grid >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    const int tile_cap = a.row_mode ? 0 : a.TV * a.pitch;
    const size_t lds = perm_lds_bytes(tile_cap, a.TV);
    if (ts.items <= 8)
        hipLaunchKernelGGL(permute_kernel<8>, dim3((unsigned)grid), dim3(kPermThreads), lds,
                           stream, a, tile_cap, a.TV);
    else
        hipLaunchKernelGGL(permute_kernel<16>, dim3((unsigned)grid), dim3(kPermThreads), lds,
                           stream, a, tile_cap, a.TV);
    return hipGetLastError();
}

int64_t permute_batch_args(const PermutePlan *plans, const double *const *in,
                           double *const *out, int n, PermArgs *args, int *first,
                           int *tile_cap, int *tv_cap, int *items) {
    int64_t total = 0;
    *tile_cap = 0;
    *tv_cap = 1;
    *items = 8;
    for (int e = 0; e < n; ++e) {
        int64_t g = 0;
        if (plans[e].size > 0) {
            const TileShape ts = tile_shape(plans[e], true);
            *items = std::max(*items, ts.items);
            fill_args(plans[e], in[e], out[e], args[e], g, ts);
            if (!args[e].row_mode) *tile_cap = std::max(*tile_cap, args[e].TV * args[e].pitch);
            *tv_cap = std::max(*tv_cap, args[e].TV);
        } else {
            args[e] = PermArgs{};
        }
        first[e] = (int)std::min<int64_t>(total, INT32_MAX);
        total += g;
    }
    first[n] = (int)std::min<int64_t>(total, INT32_MAX);
    return total;
}

hipError_t launch_permute_batch(const PermArgs *args, const int *first, const int *entry,
                                int grid, int tile_cap, int tv_cap, int items,
                                hipStream_t stream) {
    if (grid <= 0) return hipSuccess;
    const size_t lds = perm_lds_bytes(tile_cap, tv_cap);
    if (items <= 8)
        hipLaunchKernelGGL(permute_batch_kernel<8>, dim3((unsigned)grid), dim3(kPermThreads), lds,
                           stream, args, first, entry, tile_cap, tv_cap);
    else
        hipLaunchKernelGGL(permute_batch_kernel<16>, dim3((unsigned)grid), dim3(kPermThreads),
                           lds, stream, args, first, entry, tile_cap, tv_cap);
    return hipGetLastError();
}

} // namespace sipgpu
