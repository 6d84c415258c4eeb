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

#include "kernels.h"
#include "plan.h"

namespace sipgpu {

namespace {

constexpr int kPermThreads = 256;
constexpr int kPermTileElems = 4096;       // elements per tile
constexpr int kPermItems = kPermTileElems / kPermThreads;
constexpr int kPermMaxTV = 1024;
constexpr int kPermMaxTU = 64;             // transpose mode only

__global__ __launch_bounds__(kPermThreads) void permute_kernel(PermArgs a) {
    __shared__ double tile[kPermTileElems + kPermMaxTV];
    __shared__ int sh_inV[kPermMaxTV];
    __shared__ int sh_outV[kPermMaxTV];
    __shared__ int sh_outU[kPermMaxTU];

    const int bu = blockIdx.x % a.tilesU;
    const int bv = blockIdx.x / a.tilesU;
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
    double val[kPermItems];
    // Element e = tid + i * kPermThreads of the tile, as (v, u) = (e / tu,
    // e % tu) or (u, v) = (e / tv, e % tv): one division per thread, then
    // the stride kPermThreads is added in mixed radix (no division per item).
    const int qu = kPermThreads / tu, ru = kPermThreads - qu * tu;
    const int vu0 = tid / tu, uu0 = tid - vu0 * tu;
    if (a.row_mode) {
        // u is contiguous in both the input and the output.
        int vv = vu0, uu = uu0;
#pragma unroll
        for (int i = 0; i < kPermItems; ++i) {
            if (tid + i * kPermThreads < n) val[i] = in[u0 + uu + sh_inV[vv]];
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
        vv = vu0;
        uu = uu0;
#pragma unroll
        for (int i = 0; i < kPermItems; ++i) {
            if (tid + i * kPermThreads < n) out[u0 + uu + sh_outV[vv]] = val[i];
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
        for (int i = 0; i < kPermItems; ++i) {
            if (tid + i * kPermThreads < n) val[i] = in[u0 + uu + sh_inV[vv]];
            uu += ru;
            vv += qu;
            if (uu >= tu) { uu -= tu; ++vv; }
        }
        vv = vu0;
        uu = uu0;
#pragma unroll
        for (int i = 0; i < kPermItems; ++i) {
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
    for (int i = 0; i < kPermItems; ++i) {
        if (tid + i * kPermThreads < n) val[i] = tile[vv * pitch + uu];
        vv += rv;
        uu += qv;
        if (vv >= tv) { vv -= tv; ++uu; }
    }
    uu = uv0;
    vv = vv0;
#pragma unroll
    for (int i = 0; i < kPermItems; ++i) {
        if (tid + i * kPermThreads < n) out[sh_outU[uu] + sh_outV[vv]] = val[i];
        vv += rv;
        uu += qv;
        if (vv >= tv) { vv -= tv; ++uu; }
    }
}

} // namespace

hipError_t launch_permute(const PermutePlan &p, const double *in, double *out,
                          hipStream_t stream) {
    if (p.size <= 0) return hipSuccess;
    if (p.identity)
        return hipMemcpyAsync(out, in, (size_t)p.size * sizeof(double),
                              hipMemcpyDeviceToDevice, stream);
    PermArgs a{};
    a.in = in;
    a.out = out;
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
    if (p.row_mode) {
        a.TU = (int)std::min<int64_t>(p.sizeU, kPermTileElems);
        a.TV = (int)std::max<int64_t>(1, std::min<int64_t>({p.sizeV, kPermTileElems / a.TU,
                                                            kPermMaxTV}));
        a.pitch = a.TU;
    } else {
        a.TU = (int)std::min<int64_t>(p.sizeU, kPermMaxTU);
        a.TV = (int)std::max<int64_t>(1, std::min<int64_t>({p.sizeV, kPermTileElems / a.TU,
                                                            kPermMaxTV}));
        a.pitch = a.TU | 1;
    }
    a.tilesU = (a.sizeU + a.TU - 1) / a.TU;
    const int64_t tilesV = (a.sizeV + a.TV - 1) / a.TV;
    const int64_t grid = (int64_t)a.tilesU * tilesV;
    if (grid <= 0 || grid >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(permute_kernel, dim3((unsigned)grid), dim3(kPermThreads), 0, stream, a);
    return hipGetLastError();
}

} // namespace sipgpu
