// blockops.hip -- the memory-bound block kernels of the reference's tensor
// module (tensor_dil_omp.F90:144-436): fill, scale, scaled add, squared
// Frobenius norm, and slice extraction / insertion.
//
// All of them stream HBM once: the elementwise kernels move 8 (fill), 16
// (scale) or 24 (add) bytes per element, the norm reads 8, and a slice moves
// 16 (one gather read and one contiguous write, or the reverse for insertion).
// Elementwise kernels use 16-byte accesses when every pointer allows it; the
// slice kernels turn a slice index into a tensor offset with one high
// multiply per dimension (fastdiv.h) instead of a division.
#include <type_traits>

#include "kernels.h"

namespace sipgpu {

namespace {

constexpr int kOpThreads = 256;
constexpr int kOpUnroll = 4;        // independent accesses in flight per thread
constexpr int kOpMaxBlocks = 8192;  // 32 workgroups per CU

int op_blocks(long long units) {
    const long long per = (long long)kOpThreads * kOpUnroll;
    long long b = (units + per - 1) / per;
    return (int)(b < 1 ? 1 : b > kOpMaxBlocks ? kOpMaxBlocks : b);
}

// One element of the elementwise operations.  No multiply-add is fused: the
// reference evaluates t0 + t1 * s as a rounded product and a rounded sum.
template <int OP>
__device__ __forceinline__ double apply(double a, double b, double s) {
#pragma clang fp contract(off)
    if (OP == kOpFill) return s;
    if (OP == kOpScale) return a * s;
    if (OP == kOpAddScaled) return a + b * s;
    return a + b;
}

template <int OP, int VEC>
__global__ __launch_bounds__(kOpThreads) void elementwise_kernel(double *a, const double *b,
                                                                 long long n, double s) {
    using V = typename std::conditional<VEC == 2, double2, double>::type;
    V *av = reinterpret_cast<V *>(a);
    const V *bv = reinterpret_cast<const V *>(b);
    const long long units = n / VEC;
    const long long stride = (long long)gridDim.x * kOpThreads;
    const long long t = (long long)blockIdx.x * kOpThreads + threadIdx.x;
    for (long long i0 = t; i0 < units; i0 += kOpUnroll * stride) {
        V x[kOpUnroll], y[kOpUnroll];
#pragma unroll
        for (int u = 0; u < kOpUnroll; ++u) {
            const long long i = i0 + u * stride;
            if (i < units) {
                if (OP != kOpFill) x[u] = av[i];
                if (OP == kOpAddScaled || OP == kOpAdd) y[u] = bv[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kOpUnroll; ++u) {
            const long long i = i0 + u * stride;
            if (i < units) {
                V r;
                if constexpr (VEC == 2) {
                    r.x = apply<OP>(OP != kOpFill ? x[u].x : 0.0,
                                    OP >= kOpAddScaled ? y[u].x : 0.0, s);
                    r.y = apply<OP>(OP != kOpFill ? x[u].y : 0.0,
                                    OP >= kOpAddScaled ? y[u].y : 0.0, s);
                } else {
                    r = apply<OP>(OP != kOpFill ? x[u] : 0.0, OP >= kOpAddScaled ? y[u] : 0.0, s);
                }
                av[i] = r;
            }
        }
    }
    // the odd last element of the 16-byte form
    if (VEC == 2 && (n & 1) && t == 0)
        a[n - 1] = apply<OP>(OP != kOpFill ? a[n - 1] : 0.0,
                             OP >= kOpAddScaled ? b[n - 1] : 0.0, s);
}

template <int OP>
hipError_t launch_op(double *a, const double *b, long long n, double s, hipStream_t st) {
    const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
    if (vec)
        hipLaunchKernelGGL((elementwise_kernel<OP, 2>), dim3(op_blocks(n / 2)), dim3(kOpThreads),
                           0, st, a, b, n, s);
    else
        hipLaunchKernelGGL((elementwise_kernel<OP, 1>), dim3(op_blocks(n)), dim3(kOpThreads), 0,
                           st, a, b, n, s);
    return hipGetLastError();
}

// t = ((t + p[0]) + p[1]) + ... elementwise, the sources added in order.
template <int VEC>
__global__ __launch_bounds__(kOpThreads) void add_many_kernel(double *t, AddManyArgs a,
                                                              long long n) {
    using V = typename std::conditional<VEC == 2, double2, double>::type;
    V *tv = reinterpret_cast<V *>(t);
    const long long units = n / VEC;
    const long long stride = (long long)gridDim.x * kOpThreads;
    const long long i0 = (long long)blockIdx.x * kOpThreads + threadIdx.x;
    constexpr int U = 2; // units per thread per round: 2 x (1 + nparts) loads in flight
    for (long long i = i0; i < units; i += U * stride) {
        V acc[U], x[U][kAddManyMax];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long iu = i + u * stride;
            if (iu < units) {
                acc[u] = tv[iu];
#pragma unroll
                for (int q = 0; q < kAddManyMax; ++q)
                    if (q < a.n) x[u][q] = reinterpret_cast<const V *>(a.p[q])[iu];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long iu = i + u * stride;
            if (iu < units) {
#pragma unroll
                for (int q = 0; q < kAddManyMax; ++q)
                    if (q < a.n) acc[u] += x[u][q];
                tv[iu] = acc[u];
            }
        }
    }
    if (VEC == 2 && (n & 1) && i0 == 0) {
        double v = t[n - 1];
        for (int q = 0; q < a.n; ++q) v += a.p[q][n - 1];
        t[n - 1] = v;
    }
}

// Sum of the kOpThreads values of `v` over the workgroup (fixed tree).
__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = kOpThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// Workgroup g sums the squares of its grid-stride share; the grid depends on
// n only, so the summation order is fixed for a given block size.
template <int VEC>
__global__ __launch_bounds__(kOpThreads) void norm2_partial_kernel(const double *x, long long n,
                                                                   double *part) {
    __shared__ double red[kOpThreads];
    using V = typename std::conditional<VEC == 2, double2, double>::type;
    const V *xv = reinterpret_cast<const V *>(x);
    const long long units = n / VEC;
    const long long stride = (long long)gridDim.x * kOpThreads;
    const long long t = (long long)blockIdx.x * kOpThreads + threadIdx.x;
    double acc[kOpUnroll] = {0.0, 0.0, 0.0, 0.0};
    for (long long i0 = t; i0 < units; i0 += kOpUnroll * stride) {
#pragma unroll
        for (int u = 0; u < kOpUnroll; ++u) {
            const long long i = i0 + u * stride;
            if (i < units) {
                const V v = xv[i];
                if constexpr (VEC == 2)
                    acc[u] = fma(v.y, v.y, fma(v.x, v.x, acc[u]));
                else
                    acc[u] = fma(v, v, acc[u]);
            }
        }
    }
    double v = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    if (VEC == 2 && (n & 1) && t == 0) v = fma(x[n - 1], x[n - 1], v);
    v = block_sum(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// One workgroup: thread t adds partials t, t + kOpThreads, ... in order,
// then the tree.
__global__ __launch_bounds__(kOpThreads) void norm2_final_kernel(const double *part, int np,
                                                                 double *out) {
    __shared__ double red[kOpThreads];
    double v = 0.0;
    for (int i = threadIdx.x; i < np; i += kOpThreads) v += part[i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) out[0] = v;
}

// Tensor offset (relative to the slice's first element) of slice element l.
__device__ __forceinline__ uint32_t slice_offset(const SliceArgs &a, uint32_t l) {
    uint32_t off = 0;
#pragma unroll
    for (int i = 0; i < kSliceMaxRank; ++i) {
        if (i < a.rank - 1) {
            const uint32_t q = fastdiv(l, a.ext[i]);
            off += (l - q * a.ext[i].d) * a.tstride[i];
            l = q;
        } else if (i == a.rank - 1) {
            off += l * a.tstride[i];
        }
    }
    return off;
}

// INSERT = false: slice[l] = tens[off(l)]; true: tens[off(l)] = slice[l].
// Every offset is below the plan's span, which the host checked against the
// tensor block.
template <bool INSERT>
__global__ __launch_bounds__(kOpThreads) void slice_kernel(SliceArgs a, double *tens,
                                                           double *slc) {
    const uint32_t stride = gridDim.x * kOpThreads;
    const uint32_t n = a.n;
    for (uint32_t l0 = blockIdx.x * kOpThreads + threadIdx.x; l0 < n; l0 += kOpUnroll * stride) {
        double v[kOpUnroll];
        uint32_t off[kOpUnroll];
#pragma unroll
        for (int u = 0; u < kOpUnroll; ++u) {
            const uint32_t l = l0 + u * stride;
            if (l < n) {
                off[u] = slice_offset(a, l);
                v[u] = INSERT ? slc[l] : tens[off[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < kOpUnroll; ++u) {
            const uint32_t l = l0 + u * stride;
            if (l < n) {
                if (INSERT)
                    tens[off[u]] = v[u];
                else
                    slc[l] = v[u];
            }
        }
    }
}

// Workgroup (x, y): one element per thread, elements [x 256, (x + 1) 256)
// of entry y; the parts are loaded eight at a time (in flight together) and
// added in order.
constexpr int kAddPartsChunk = kOpThreads;

__global__ __launch_bounds__(kOpThreads) void add_parts_kernel(const AddPartsEntry *entries) {
    const AddPartsEntry e = entries[blockIdx.y];
    const long long i = (long long)blockIdx.x * kAddPartsChunk + threadIdx.x;
    if (i >= e.n) return;
    double v = e.t[i];
    for (int q0 = 0; q0 < e.nparts; q0 += 8) {
        double x[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            x[q] = q0 + q < e.nparts ? e.p[(long long)(q0 + q) * e.n + i] : 0.0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q0 + q < e.nparts) v = v + x[q];
    }
    e.t[i] = v;
}

} // namespace

hipError_t launch_block_op(int op, double *a, const double *b, long long n, double s,
                           hipStream_t st) {
    if (n <= 0) return hipSuccess;
    switch (op) {
    case kOpFill: return launch_op<kOpFill>(a, a, n, s, st);
    case kOpScale: return launch_op<kOpScale>(a, a, n, s, st);
    case kOpAddScaled: return launch_op<kOpAddScaled>(a, b, n, s, st);
    case kOpAdd: return launch_op<kOpAdd>(a, b, n, s, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_block_add_many(double *t, const AddManyArgs &a, long long n, hipStream_t st) {
    if (n <= 0 || a.n <= 0) return hipSuccess;
    uintptr_t bits = reinterpret_cast<uintptr_t>(t);
    for (int q = 0; q < a.n; ++q) bits |= reinterpret_cast<uintptr_t>(a.p[q]);
    if ((bits & 15) == 0)
        hipLaunchKernelGGL(add_many_kernel<2>, dim3(op_blocks(n / 2)), dim3(kOpThreads), 0, st, t, a,
                           n);
    else
        hipLaunchKernelGGL(add_many_kernel<1>, dim3(op_blocks(n)), dim3(kOpThreads), 0, st, t, a, n);
    return hipGetLastError();
}

hipError_t launch_add_parts(const AddPartsEntry *entries, int count, long long max_n,
                            hipStream_t st) {
    if (count <= 0 || max_n <= 0) return hipSuccess;
    if (count > 65535) return hipErrorInvalidValue;
    const long long gx = (max_n + kAddPartsChunk - 1) / kAddPartsChunk;
    if (gx * kOpThreads >= (1LL << 32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(add_parts_kernel, dim3((unsigned)gx, (unsigned)count), dim3(kOpThreads), 0,
                       st, entries);
    return hipGetLastError();
}

int norm2_blocks(long long n) {
    const long long per = (long long)kOpThreads * kOpUnroll * 2;
    long long b = (n + per - 1) / per;
    return (int)(b < 1 ? 1 : b > kNormBlocks ? kNormBlocks : b);
}

hipError_t launch_block_norm2(const double *x, long long n, double *out, double *part,
                              hipStream_t st) {
    const int g = norm2_blocks(n);
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0)
        hipLaunchKernelGGL((norm2_partial_kernel<2>), dim3(g), dim3(kOpThreads), 0, st, x, n,
                           part);
    else
        hipLaunchKernelGGL((norm2_partial_kernel<1>), dim3(g), dim3(kOpThreads), 0, st, x, n,
                           part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(norm2_final_kernel, dim3(1), dim3(kOpThreads), 0, st, part, g, out);
    return hipGetLastError();
}

SliceArgs make_slice_args(const SlicePlan &p) {
    SliceArgs a{};
    a.rank = p.rank > 0 ? p.rank : 1;
    a.n = (int)p.count;
    for (int i = 0; i < kSliceMaxRank; ++i) {
        const bool live = i < p.rank;
        a.ext[i] = make_fastdiv(live ? (uint32_t)p.ext[i] : 1u);
        a.tstride[i] = live ? (int)p.tstride[i] : 0;
    }
    return a;
}

hipError_t launch_slice(const SliceArgs &a, bool insert, double *tens, double *slc,
                        hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
    const int blocks = op_blocks(a.n);
    if (insert)
        hipLaunchKernelGGL(slice_kernel<true>, dim3(blocks), dim3(kOpThreads), 0, st, a, tens, slc);
    else
        hipLaunchKernelGGL(slice_kernel<false>, dim3(blocks), dim3(kOpThreads), 0, st, a, tens,
                           slc);
    return hipGetLastError();
}

} // namespace sipgpu
