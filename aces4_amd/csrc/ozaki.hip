// ozaki.hip -- split-precision contraction: FP64 products emulated exactly
// with INT8 MFMA (Ozaki scheme II, integer/CRT form).
//
// For D(m, n) = sum_k A(k, m) B(k, n) (A = the left operand, B = the right,
// read through the contraction's offset tables):
//   1. every row m of A (column n of B) gets a power-of-two scale so that
//      A_int = rint(A * 2^sA_m) is an integer with |A_int| <= 2^(P-1);
//   2. the integer product C = sum_k A_int B_int is computed modulo N
//      pairwise-coprime moduli p_i <= 255, one INT8 GEMM each, with exact
//      INT32 accumulation (|residue| <= 127, so K * 127^2 < 2^31);
//   3. C is recovered exactly from its residues (explicit CRT, 128-bit
//      integers) and D = C * 2^-(sA_m + sB_n), rounded once to FP64.
// The only rounding is step 1 (relative 2^-P per element); P follows from N
// and the largest K so that |C| < M/2, M = prod p_i.
//
// Kernels, in launch order: exponent fill, row/column exponent reduction,
// FP64 -> residue panels (the operand permutation is applied here, once per
// block), the INT8 GEMM per modulus (int8 residues of the tile out), and the
// reconstruction that scatters into the destination blocks.
#include "kernels.h"
#include "ozaki.h"

#include <climits>

namespace sipgpu {

namespace {

typedef int int32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) int32_t *gitab_t;

__global__ void fill_int_kernel(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// One thread per (block, row): exponent of the row's max |value| over K,
// folded into the row's component with atomicMax (order-independent).
__global__ void row_exponent_kernel(const OzBlock *blocks, int nblocks, const int32_t *tab,
                                    int *expo, const int *row_start, int nrows_total) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nrows_total;
         g += gridDim.x * blockDim.x) {
        // binary search the block owning global row g
        int lo = 0, hi = nblocks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (row_start[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const OzBlock b = blocks[lo];
        const int x = g - row_start[lo];
        const int xo = tab[b.xtab + x];
        double mx = 0.0;
        for (int k = 0; k < b.K; ++k) mx = fmax(mx, fabs(b.data[tab[b.ktab + k] + xo]));
        if (mx > 0.0) {
            int e;
            frexp(mx, &e); // mx = f * 2^e, f in [0.5, 1)
            atomicMax(&expo[b.exp_off + x], e);
        }
    }
}

// Balanced residue of an int32 in (-p/2, p/2].
__device__ __forceinline__ int bal_mod(int v, int p) {
    int r = v % p;
    if (r > p / 2) r -= p;
    if (r < -((p - 1) / 2)) r += p;
    return r;
}

// Balanced residue of |v| < 2^52 held as hi * 2^24 + lo (0 <= lo < 2^24),
// with int32 arithmetic only.
__device__ __forceinline__ int bal_mod_wide(int hi, int lo, int p, int t24) {
    return bal_mod((hi % p) * t24 + lo % p, p);
}

// One thread per (block, row x, 4 consecutive k): residues of the scaled
// integers, packed 4 per word, into every modulus plane.  Row-major panels,
// K padded with zeros to Kp.
__global__ void to_residues_kernel(const OzBlock *blocks, int nblocks, const int32_t *tab,
                                   const int *expo, const int *quad_start, int nquads_total,
                                   int8_t *planes, long long plane_bytes, OzConst c) {
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < nquads_total;
         g += gridDim.x * blockDim.x) {
        int bl = 0, bh = nblocks - 1;
        while (bl < bh) {
            const int mid = (bl + bh + 1) >> 1;
            if (quad_start[mid] <= g) bl = mid; else bh = mid - 1;
        }
        const OzBlock b = blocks[bl];
        const int q = g - quad_start[bl];
        const int qpr = b.Kp / 4;            // quads per row
        const int x = q / qpr, k0 = (q - x * qpr) * 4;
        const int e = expo[b.exp_off + x];
        const int shift = e == INT_MIN ? 0 : (c.P - 1) - e;
        const int xo = tab[b.xtab + x];
        int hi[4], lo[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + j;
            double v = 0.0;
            if (k < b.K && e != INT_MIN) v = rint(ldexp(b.data[tab[b.ktab + k] + xo], shift));
            // v = hi * 2^24 + lo exactly (|v| <= 2^51)
            const double h = floor(v * 0x1p-24);
            hi[j] = (int)h;
            lo[j] = (int)(v - h * 0x1p24);
        }
        const long long off = b.panel_off + (long long)x * b.Kp + k0;
        for (int i = 0; i < c.n; ++i) {
            uint32_t word = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                word |= (uint32_t)(uint8_t)(int8_t)bal_mod_wide(hi[j], lo[j], c.p[i], c.t24[i])
                        << (8 * j);
            *reinterpret_cast<uint32_t *>(planes + i * plane_bytes + off) = word;
        }
    }
}

// INT8 GEMM over residue panels, one modulus per blockIdx.y.  Workgroup tile
// BM x BN (4 waves, 2 x 2), stages of KS bytes of k through LDS; MFMA
// v_mfma_i32_16x16x64_i8: lane l holds row l%16, k = 16*(l/16) .. +15 of an
// operand block; accumulator register r of lane l is C(4*(l/16) + r, l%16).
template <int BM, int BN, int KS>
__global__ __launch_bounds__(256) void gemm_i8_kernel(OzGemmArgs a) {
    constexpr int NT = 256, WM = 2, WN = 2;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int PITCH = KS + 16;                  // bytes per LDS row
    constexpr int CPR = KS / 16;                    // 16-byte chunks per row
    constexpr int LA = BM * CPR / NT, LB = BN * CPR / NT;
    static_assert(LA >= 1 && LB >= 1 && (BM * CPR) % NT == 0 && (BN * CPR) % NT == 0, "bad");
    __shared__ __attribute__((aligned(16))) int8_t As[BM * PITCH];
    __shared__ __attribute__((aligned(16))) int8_t Bs[BN * PITCH];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int mod = blockIdx.y;
    const int p = a.c.p[mod];
    const int8_t *plane = a.planes + (long long)mod * a.plane_bytes;

    for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const OzTile tile = a.tiles[t];
        if (tile.task < 0) continue;
        const OzTask task = a.tasks[tile.task];
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;

        int32x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = int32x4{0, 0, 0, 0};

        for (int pi = task.first; pi < task.first + task.npairs; ++pi) {
            const OzPair pr = a.pairs[pi];
            const int8_t *pa = plane + a.blocks[pr.a].panel_off;
            const int8_t *pb = plane + a.blocks[pr.b].panel_off;
            const int Kp = pr.Kp;
            int32x4 ra[LA], rb[LB];
            auto load = [&](int k0) {
#pragma unroll
                for (int i = 0; i < LA; ++i) {
                    const int e = tid + i * NT, row = e / CPR, ch = e % CPR;
                    const int m = min(m0 + row, task.M - 1), k = k0 + ch * 16;
                    ra[i] = k < Kp ? *reinterpret_cast<const int32x4 *>(pa + (long long)m * Kp + k)
                                   : int32x4{0, 0, 0, 0};
                    if (m0 + row >= task.M) ra[i] = int32x4{0, 0, 0, 0};
                }
#pragma unroll
                for (int i = 0; i < LB; ++i) {
                    const int e = tid + i * NT, row = e / CPR, ch = e % CPR;
                    const int n = min(n0 + row, task.N - 1), k = k0 + ch * 16;
                    rb[i] = k < Kp ? *reinterpret_cast<const int32x4 *>(pb + (long long)n * Kp + k)
                                   : int32x4{0, 0, 0, 0};
                    if (n0 + row >= task.N) rb[i] = int32x4{0, 0, 0, 0};
                }
            };
            load(0);
            for (int k0 = 0; k0 < Kp; k0 += KS) {
                __syncthreads();
#pragma unroll
                for (int i = 0; i < LA; ++i) {
                    const int e = tid + i * NT;
                    *reinterpret_cast<int32x4 *>(As + (e / CPR) * PITCH + (e % CPR) * 16) = ra[i];
                }
#pragma unroll
                for (int i = 0; i < LB; ++i) {
                    const int e = tid + i * NT;
                    *reinterpret_cast<int32x4 *>(Bs + (e / CPR) * PITCH + (e % CPR) * 16) = rb[i];
                }
                __syncthreads();
                if (k0 + KS < Kp) load(k0 + KS);
                const int ksteps = min(KS, Kp - k0) / 64;
#pragma unroll
                for (int ks = 0; ks < KS / 64; ++ks) {
                    if (ks >= ksteps) break;
                    int32x4 fa[FM], fb[FN];
#pragma unroll
                    for (int i = 0; i < FM; ++i)
                        fa[i] = *reinterpret_cast<const int32x4 *>(
                            As + (wm * TM + i * 16 + li) * PITCH + ks * 64 + lk * 16);
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        fb[j] = *reinterpret_cast<const int32x4 *>(
                            Bs + (wn * TN + j * 16 + li) * PITCH + ks * 64 + lk * 16);
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j],
                                                                              acc[i][j], 0, 0, 0);
                }
            }
        }

        // residues of the tile, balanced, into this modulus' result plane
        int8_t *res = a.res_planes + (long long)mod * a.res_plane_bytes + task.res_off;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn * TN + j * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * TM + i * 16 + 4 * lk + r;
                    if (m < task.M && n < task.N)
                        res[(long long)n * task.M + m] = (int8_t)bal_mod(acc[i][j][r], p);
                }
            }
        __syncthreads();
    }
}

// 128-bit helpers on {low, high} 64-bit halves (two's complement, mod 2^128).
struct U128 {
    unsigned long long lo, hi;
};

__device__ __forceinline__ void mad_small(U128 &acc, unsigned long long t, const uint64_t m[2]) {
    const unsigned long long plo = t * m[0];
    const unsigned long long phi = __umul64hi(t, m[0]) + t * m[1];
    const unsigned long long lo = acc.lo + plo;
    acc.hi += phi + (lo < plo ? 1ull : 0ull);
    acc.lo = lo;
}

// |v| of a signed 128-bit value, correctly rounded to FP64 (sign returned).
__device__ __forceinline__ double to_double(U128 v) {
    const bool neg = (long long)v.hi < 0;
    if (neg) {
        v.lo = ~v.lo + 1;
        v.hi = ~v.hi + (v.lo == 0 ? 1ull : 0ull);
    }
    double d;
    if (v.hi == 0) {
        d = (double)v.lo;
    } else {
        // top 64 bits, with a sticky bit for everything below them: one
        // rounding of the 64-bit value is then the correct rounding of v
        const int sh = __clzll((long long)v.hi);
        unsigned long long top = sh ? (v.hi << sh) | (v.lo >> (64 - sh)) : v.hi;
        const unsigned long long rest = sh ? v.lo << sh : v.lo;
        if (rest) top |= 1ull;
        d = ldexp((double)top, 64 - sh);
    }
    return neg ? -d : d;
}

// One thread per destination element (all tasks): explicit CRT.  With
// t_i = r_i u_i mod p_i (u_i = (M/p_i)^-1 mod p_i), C = sum_i t_i (M/p_i) - q M
// for the q that puts C in [-M/2, M/2): q = round(sum_i t_i / p_i), exact in
// FP64 because |C| <= M/4 keeps that sum a quarter away from the rounding
// boundary.  C itself is formed exactly modulo 2^128 (|C| < 2^111), then
// rounded once to FP64 and scaled by the operands' powers of two, so the
// result is the correctly rounded product of the P-bit operands.
__global__ void reconstruct_kernel(OzRecArgs a) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < a.nelems;
         g += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = a.ntasks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.elem_start[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const OzTask task = a.tasks[lo];
        const int e = (int)(g - a.elem_start[lo]);
        const int m = e % task.M, n = e / task.M;
        U128 acc{0ull, 0ull};
        double s = 0.0;
        for (int i = 0; i < a.c.n; ++i) {
            const int p = a.c.p[i];
            int r = a.res_planes[(long long)i * a.res_plane_bytes + task.res_off + e];
            if (r < 0) r += p;
            const int t = (r * a.c.u[i]) % p;
            s += (double)t * a.c.rp[i];
            mad_small(acc, (unsigned long long)t, a.c.mi[i]);
        }
        // acc -= q M
        U128 qm{0ull, 0ull};
        mad_small(qm, (unsigned long long)rint(s), a.c.m);
        const unsigned long long dlo = acc.lo - qm.lo;
        const U128 C{dlo, acc.hi - qm.hi - (acc.lo < qm.lo ? 1ull : 0ull)};
        const int ea = a.expo[task.expA + m], eb = a.expo[task.expB + n];
        double v = 0.0;
        if (ea != INT_MIN && eb != INT_MIN) {
            const int sa = (a.c.P - 1) - ea, sb = (a.c.P - 1) - eb;
            v = ldexp(to_double(C), -(sa + sb));
        }
        const int off = a.tab[task.mD + m] + a.tab[task.nD + n];
        double *D = task.D;
        if (task.beta) v += D[off];
        D[off] = v;
    }
}

} // namespace

hipError_t oz_fill_int(int *p, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3(min(4096, (n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t oz_row_exponents(const OzBlock *blocks, int nblocks, const int32_t *tab, int *expo,
                            const int *row_start, int nrows, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(row_exponent_kernel, dim3(min(8192, (nrows + 255) / 256)), dim3(256), 0, s,
                       blocks, nblocks, tab, expo, row_start, nrows);
    return hipGetLastError();
}

hipError_t oz_to_residues(const OzBlock *blocks, int nblocks, const int32_t *tab, const int *expo,
                          const int *quad_start, int nquads, int8_t *planes,
                          long long plane_bytes, const OzConst &c, hipStream_t s) {
    if (nquads <= 0) return hipSuccess;
    hipLaunchKernelGGL(to_residues_kernel, dim3(min(65536, (nquads + 255) / 256)), dim3(256), 0, s,
                       blocks, nblocks, tab, expo, quad_start, nquads, planes, plane_bytes, c);
    return hipGetLastError();
}

hipError_t oz_gemm(const OzGemmArgs &a, int grid, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL((gemm_i8_kernel<kOzBM, kOzBN, 128>), dim3(grid, a.c.n), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_reconstruct(const OzRecArgs &a, hipStream_t s) {
    if (a.ntasks <= 0 || a.nelems <= 0) return hipSuccess;
    const long long blocks = (a.nelems + 255) / 256;
    hipLaunchKernelGGL(reconstruct_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)),
                       dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace sipgpu
