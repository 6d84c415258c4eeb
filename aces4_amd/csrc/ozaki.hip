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
// and the largest K so that |C| <= M/4, M = prod p_i.  Inputs must be finite.
//
// Kernels, in launch order: exponent fill; row/column maxima -> exponents;
// FP64 -> residue panels (the operand permutation is applied here, once per
// block, through 64 x 64 LDS tiles read along whichever index is adjacent in
// memory); the INT8 GEMM per modulus over each task's virtual K (16-byte
// chunks of the pairs' panels, so K is padded per pair to 16 only), whose
// epilogue reduces the tile to t_i = (C mod p_i) u_i mod p_i; and the
// reconstruction, which walks each result block in its memory order.
#include "kernels.h"
#include "ozaki.h"

#include <climits>

namespace sipgpu {

namespace {

typedef int int32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) int8_t gi8_t;

__global__ void fill_int_kernel(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

// Rows [x0, x0 + 64) x k [k0, k0 + 64) of block b into S[x][k], reading
// along the index that is adjacent in memory; entries outside the block
// read as 0.
__device__ __forceinline__ void load_subtile(const OzBlock &b, const int32_t *tab, int x0, int k0,
                                             double (*S)[kOzTileK + 1], int tid) {
    if (b.xfast) {
        const int x = x0 + (tid & 63);
        const bool xin = x < b.X;
        const int xo = xin ? tab[b.xtab + x] : 0;
#pragma unroll 4
        for (int j = 0; j < 16; ++j) {
            const int kk = (tid >> 6) + 4 * j, k = k0 + kk;
            double v = 0.0;
            if (xin && k < b.K) v = b.data[tab[b.ktab + k] + xo];
            S[tid & 63][kk] = v;
        }
    } else {
        const int k = k0 + (tid & 63);
        const bool kin = k < b.K;
        const int ko = kin ? tab[b.ktab + k] : 0;
#pragma unroll 4
        for (int j = 0; j < 16; ++j) {
            const int xx = (tid >> 6) + 4 * j, x = x0 + xx;
            double v = 0.0;
            if (kin && x < b.X) v = b.data[ko + tab[b.xtab + x]];
            S[xx][tid & 63] = v;
        }
    }
}

// One workgroup per 64-row tile of a block: max |value| of each row over
// the block's K, as a binary exponent folded into the row's component with
// atomicMax (order-independent, so the result is deterministic).
__global__ __launch_bounds__(256) void row_exponent_kernel(OzConvArgs a, const OzConvTile *items,
                                                           int nitems) {
    __shared__ double S[kOzTileX][kOzTileK + 1];
    __shared__ double part[4][kOzTileX];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const OzConvTile w = items[it];
        const OzBlock b = a.blocks[w.block];
        double mx = 0.0;
        for (int k0 = 0; k0 < b.K; k0 += kOzTileK) {
            __syncthreads();
            load_subtile(b, a.tab, w.x0, k0, S, tid);
            __syncthreads();
            const int x = tid & 63, kg = (tid >> 6) * 16;
#pragma unroll
            for (int j = 0; j < 16; ++j) mx = fmax(mx, fabs(S[x][kg + j]));
        }
        part[tid >> 6][tid & 63] = mx;
        __syncthreads();
        if (tid < kOzTileX) {
            const double m = fmax(fmax(part[0][tid], part[1][tid]), fmax(part[2][tid], part[3][tid]));
            const int x = w.x0 + tid;
            if (x < b.X && m > 0.0) {
                int e;
                frexp(m, &e); // m = f * 2^e, f in [0.5, 1)
                atomicMax(&a.expo[b.exp_off + x], e);
            }
        }
        __syncthreads();
    }
}

// Residue of an exact integer |v| < 2^52 held in a double, in [0, p):
// q = floor(v / p) from the rounded reciprocal is off by at most one, and
// v - p q is exact (fma of integers below 2^53).
__device__ __forceinline__ int mod_exact(double v, double p, double rp) {
    const double q = floor(v * rp);
    double r = fma(-p, q, v);
    if (r < 0.0) r += p;
    if (r >= p) r -= p;
    return (int)r;
}

// One workgroup per 64 x 64 conversion tile: scaled integers of the tile,
// reduced to balanced residues (|r| <= 127) for every modulus, written as
// 16-byte row pieces of the block's panel in each modulus plane.  Panel
// bytes for k in [K, Kp) are written as zeros.
__global__ __launch_bounds__(256) void to_residues_kernel(OzConvArgs a) {
    __shared__ double S[kOzTileX][kOzTileK + 1];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const OzConvTile w = a.items[it];
        const OzBlock b = a.blocks[w.block];
        __syncthreads();
        load_subtile(b, a.tab, w.x0, w.k0, S, tid);
        __syncthreads();
        const int xx = tid >> 2, kc = (tid & 3) * 16;
        const int x = w.x0 + xx, k = w.k0 + kc;
        if (x >= b.X || k >= b.Kp) continue;
        const int e = a.expo[b.exp_off + x];
        const int shift = e == INT_MIN ? 0 : (a.c.P - 1) - e;
        double v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = e == INT_MIN ? 0.0 : rint(ldexp(S[xx][kc + j], shift));
        int8_t *dst = a.planes + b.panel_off + (long long)x * b.Kp + k;
        for (int i = 0; i < a.c.n; ++i) {
            const int p = a.c.p[i];
            const double pd = p, rp = a.c.rp[i];
            uint32_t word[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                int r = mod_exact(v[j], pd, rp);
                if (r > p / 2) r -= p;
                word[j >> 2] |= (uint32_t)(uint8_t)(int8_t)r << (8 * (j & 3));
            }
            *reinterpret_cast<int32x4 *>(dst + (long long)i * a.plane_bytes) =
                int32x4{(int)word[0], (int)word[1], (int)word[2], (int)word[3]};
        }
    }
}

// INT8 GEMM over residue panels, one modulus per blockIdx.y.  Workgroup tile
// BM x BN (4 waves, 2 x 2, each (BM/2) x (BN/2)); the task's virtual K runs
// in stages of 128 bytes = 8 chunks, double-buffered in LDS: the loads of
// stage s + 1 (16 bytes per lane into registers) are in flight while stage
// s is multiplied, then written to the other buffer.  LDS rows hold 8 chunk
// slots, chunk c of row r in slot c ^ (r & 7), so both the stores (one wave
// = 8 whole rows) and the fragment reads (16 rows x 4 chunks) spread over
// all banks.  Rows past the task's M / N read row M-1 /
// N-1 and are never stored; chunks past the pairs' K read the plane's zero
// bytes.  MFMA v_mfma_i32_16x16x64_i8: lane l holds row l%16, bytes
// 16*(l/16) .. +15 of a 64-byte k step; accumulator register r of lane l is
// C(4*(l/16) + r, l%16).
template <int BM, int BN>
__global__ __launch_bounds__(256, 2) void gemm_i8_kernel(OzGemmArgs a) {
    constexpr int WM = 2, WN = 2;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int KS = kOzStage, CH = kOzStageChunks;
    constexpr int IA = BM / 32, IB = BN / 32; // 8-row load instructions per wave per stage
    static_assert(CH == 8 && KS == 128, "a stage is 8 chunks of 16 bytes");
    static_assert(BM % 32 == 0 && BN % 32 == 0 && TM % 16 == 0 && TN % 16 == 0, "tile shape");
    __shared__ __attribute__((aligned(1024))) int8_t As[2][BM * KS];
    __shared__ __attribute__((aligned(1024))) int8_t Bs[2][BN * KS];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int g = (lane & 7) ^ (lane >> 3); // chunk of the stage this lane loads
    const int mod = blockIdx.y;
    const int p = a.c.p[mod], u = a.c.u[mod];
    const double pd = p, rp = a.c.rp[mod];
    gi8_t *plane = (gi8_t *)(a.planes + (long long)mod * a.plane_bytes);

    for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const OzTile tile = a.tiles[t];
        if (tile.task < 0) continue;
        const OzTask task = a.tasks[tile.task];
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        int ra[IA], rb[IB];
#pragma unroll
        for (int i = 0; i < IA; ++i) ra[i] = min(m0 + 8 * (wave * IA + i) + (lane >> 3), task.M - 1);
#pragma unroll
        for (int i = 0; i < IB; ++i) rb[i] = min(n0 + 8 * (wave * IB + i) + (lane >> 3), task.N - 1);

        int32x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = int32x4{0, 0, 0, 0};

        const OzChunkRef *ch = a.chunks + task.chunk0;
        const int nst = task.nchunks / CH;
        int32x4 va[IA], vb[IB];
        // stage registers: this lane's 16 bytes of 8-row groups of A and B
        auto fetch = [&](const OzChunkRef &e) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                va[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.a + ra[i] * e.stride);
#pragma unroll
            for (int i = 0; i < IB; ++i)
                vb[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.b + rb[i] * e.stride);
        };
        auto put = [&](int buf) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                *reinterpret_cast<int32x4 *>(&As[buf][(wave * IA + i) * 1024 + lane * 16]) = va[i];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                *reinterpret_cast<int32x4 *>(&Bs[buf][(wave * IB + i) * 1024 + lane * 16]) = vb[i];
        };

        // loads are unconditional (the last stage is fetched twice) so the
        // prefetched registers need no merge copies, which would wait on them
        OzChunkRef e = ch[g];
        fetch(e);
        e = ch[min(1, nst - 1) * CH + g];
        put(0);
        __syncthreads();
        int buf = 0;
        for (int s = 0; s < nst; ++s) {
            fetch(e);
            e = ch[min(s + 2, nst - 1) * CH + g];
            __builtin_amdgcn_sched_barrier(0); // keep the loads ahead of the MFMAs
            const int8_t *sa = As[buf], *sb = Bs[buf];
#pragma unroll
            for (int ks = 0; ks < KS / 64; ++ks) {
                const int slot = ((ks * 4 + lk) ^ (li & 7)) * 16;
                int32x4 fa[FM], fb[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i)
                    fa[i] = *reinterpret_cast<const int32x4 *>(sa + (wm * TM + i * 16 + li) * KS + slot);
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    fb[j] = *reinterpret_cast<const int32x4 *>(sb + (wn * TN + j * 16 + li) * KS + slot);
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j],
                                                                          0, 0, 0);
            }
            // stage s + 1 into the other buffer: every wave finished reading
            // it before the barrier that ended stage s - 1 (after the last
            // stage this writes a copy of it that nothing reads)
            put(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }

        // t = (C mod p) u mod p of the tile into this modulus' result plane,
        // in the task's element order
        uint8_t *res = a.res_planes + (long long)mod * a.res_plane_bytes + task.res_off;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int n = n0 + wn * TN + j * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * TM + i * 16 + 4 * lk + r;
                    if (m < task.M && n < task.N) {
                        const int c = mod_exact((double)acc[i][j][r], pd, rp);
                        const int tt = mod_exact((double)(c * u), pd, rp);
                        const long long idx = task.nfast ? (long long)m * task.N + n
                                                         : (long long)n * task.M + m;
                        res[idx] = (uint8_t)tt;
                    }
                }
            }
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

// A signed 128-bit value correctly rounded to FP64.
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

// One workgroup per kOzRecChunk result elements of a task, walked in the
// task's element order (the destination's memory order where it has a
// unit-stride index).  Explicit CRT: with t_i = r_i u_i mod p_i,
// C = sum_i t_i (M/p_i) - q M for q = round(sum_i t_i / p_i), exact in FP64
// because |C| <= M/4 keeps that sum a quarter away from the rounding
// boundary.  C is formed exactly modulo 2^128 (|C| < 2^111), rounded once
// to FP64 and scaled by the operands' powers of two.
__global__ __launch_bounds__(256) void reconstruct_kernel(OzRecArgs a) {
    for (int it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const OzRecItem w = a.items[it];
        const OzTask task = a.tasks[w.task];
        const int total = task.M * task.N;
        const int e1 = min(w.e0 + kOzRecChunk, total);
        const uint8_t *res = a.res_planes + task.res_off;
        for (int e = w.e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
            int m, n;
            if (task.nfast) {
                m = e / task.N;
                n = e - m * task.N;
            } else {
                n = e / task.M;
                m = e - n * task.M;
            }
            U128 acc{0ull, 0ull};
            double s = 0.0;
            for (int i = 0; i < a.c.n; ++i) {
                const unsigned t = res[(long long)i * a.res_plane_bytes + e];
                s += (double)t * a.c.rp[i];
                mad_small(acc, (unsigned long long)t, a.c.mi[i]);
            }
            U128 qm{0ull, 0ull};
            mad_small(qm, (unsigned long long)rint(s), a.c.m);
            const U128 C{acc.lo - qm.lo, acc.hi - qm.hi - (acc.lo < qm.lo ? 1ull : 0ull)};
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
}

unsigned grid_for(long long items, long long cap) {
    return (unsigned)(items < cap ? items : cap);
}

} // namespace

hipError_t oz_fill_int(int *p, int n, int v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3(grid_for((n + 255) / 256, 4096)), dim3(256), 0, s, p,
                       n, v);
    return hipGetLastError();
}

hipError_t oz_row_exponents(const OzConvArgs &a, const OzConvTile *row_items, int nrow_items,
                            hipStream_t s) {
    if (nrow_items <= 0) return hipSuccess;
    hipLaunchKernelGGL(row_exponent_kernel, dim3(grid_for(nrow_items, 65535)), dim3(256), 0, s, a,
                       row_items, nrow_items);
    return hipGetLastError();
}

hipError_t oz_to_residues(const OzConvArgs &a, hipStream_t s) {
    if (a.nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(to_residues_kernel, dim3(grid_for(a.nitems, 65535)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_gemm(const OzGemmArgs &a, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL((gemm_i8_kernel<kOzBM, kOzBN>), dim3(a.ntiles, a.c.n), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_reconstruct(const OzRecArgs &a, hipStream_t s) {
    if (a.nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(reconstruct_kernel, dim3(grid_for(a.nitems, 65535)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace sipgpu
