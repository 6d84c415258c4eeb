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

// Zeroes the kOzZero bytes at the start of each plane (16 lanes x 16 bytes).
__global__ void zero_heads_kernel(int8_t *planes, long long plane_bytes, int nplanes) {
    static_assert(kOzZero == 256, "16 lanes of 16 bytes per plane");
    const int p = blockIdx.x * 4 + (threadIdx.x >> 4);
    if (p < nplanes)
        *reinterpret_cast<int4 *>(planes + p * plane_bytes + (threadIdx.x & 15) * 16) =
            int4{0, 0, 0, 0};
}

// Rows [x0, x0 + 32) x k [k0, k0 + 128) of block b: each thread loads 16
// values along the index that is adjacent in memory (32 lanes on 32
// consecutive rows, or 128 on consecutive k); entries outside the block read
// as 0.  Thread t holds (row, k) = (t % 32, t / 32 + 8 j) when rows are
// adjacent, else (t / 128 + 2 j, t % 128).
struct SubTile {
    double v[16];
};

__device__ __forceinline__ void fetch_subtile(const OzBlock &b, const int32_t *tab, int x0, int k0,
                                              int tid, SubTile &t) {
    if (b.xfast) {
        const int x = x0 + (tid & 31);
        const bool xin = x < b.X;
        const int xo = xin ? tab[b.xtab + x] : 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = k0 + (tid >> 5) + 8 * j;
            t.v[j] = (xin && k < b.K) ? b.data[tab[b.ktab + k] + xo] : 0.0;
        }
    } else {
        const int k = k0 + (tid & 127);
        const bool kin = k < b.K;
        const int ko = kin ? tab[b.ktab + k] : 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int x = x0 + (tid >> 7) + 2 * j;
            t.v[j] = (kin && x < b.X) ? b.data[ko + tab[b.xtab + x]] : 0.0;
        }
    }
}

__device__ __forceinline__ void store_subtile(const OzBlock &b, int tid, const SubTile &t,
                                              double (*S)[kOzTileK + 1]) {
    if (b.xfast) {
#pragma unroll
        for (int j = 0; j < 16; ++j) S[tid & 31][(tid >> 5) + 8 * j] = t.v[j];
    } else {
#pragma unroll
        for (int j = 0; j < 16; ++j) S[(tid >> 7) + 2 * j][tid & 127] = t.v[j];
    }
}

// One workgroup per 32-row tile of a block: max |value| of each row over
// the block's K, as a binary exponent folded into the row's component with
// atomicMax (order-independent, so the result is deterministic).  Maxima
// stay in registers across the K sweep (two sub-tiles in flight per step);
// one LDS reduction per tile folds them per row.
__global__ __launch_bounds__(256) void row_exponent_kernel(OzConvArgs a, const OzConvTile *items,
                                                           int nitems) {
    __shared__ double part[kOzTileX][129];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const OzConvTile w = items[it];
        const OzBlock b = a.blocks[w.block];
        double mx[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) mx[j] = 0.0;
        for (int k0 = 0; k0 < b.K; k0 += 2 * kOzTileK) {
            SubTile t0, t1;
            fetch_subtile(b, a.tab, w.x0, k0, tid, t0);
            fetch_subtile(b, a.tab, w.x0, k0 + kOzTileK, tid, t1);
#pragma unroll
            for (int j = 0; j < 16; ++j) mx[j] = fmax(mx[j], fmax(fabs(t0.v[j]), fabs(t1.v[j])));
        }
        // rows adjacent: thread holds row tid%32 for 16 k; else 16 rows
        __syncthreads();
        if (b.xfast) {
            double m = 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) m = fmax(m, mx[j]);
            part[tid & 31][tid >> 5] = m;
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) part[(tid >> 7) + 2 * j][tid & 127] = mx[j];
        }
        __syncthreads();
        if (tid < kOzTileX) {
            const int cols = b.xfast ? 8 : 128;
            double m = 0.0;
            for (int c = 0; c < cols; ++c) m = fmax(m, part[tid][c]);
            const int x = w.x0 + tid;
            if (x < b.X && m > 0.0) {
                int e;
                frexp(m, &e); // m = f * 2^e, f in [0.5, 1)
                atomicMax(&a.expo[b.exp_off + x], e);
            }
        }
    }
}

// t_i = (v mod p) u mod p in [0, p) of an INT32 accumulator |v| < 2^31:
// v = h 2^17 + l with l in [0, 2^17), so h c17 + l + p 2^17 (c17 = 2^17 mod
// p) is congruent to v, positive and below 2^26; two Barrett steps with
// bar = floor(2^32 / p) (each leaves r in [0, 2p)) and the multiplication
// by u (< p) finish it in 32-bit integer arithmetic.
__device__ __forceinline__ uint32_t crt_digit(int v, int p, int c17, int u, uint32_t bar) {
    const uint32_t t = (uint32_t)((v >> 17) * c17 + (v & 0x1ffff)) + ((uint32_t)p << 17);
    uint32_t c = t - __umulhi(t, bar) * (uint32_t)p;
    if (c >= (uint32_t)p) c -= p;
    uint32_t tt = c * (uint32_t)u;
    tt -= __umulhi(tt, bar) * (uint32_t)p;
    if (tt >= (uint32_t)p) tt -= p;
    return tt;
}

// Residue of an exact integer |v| < 2^52 held in a double, in [0, p):
// q = floor(v / p) from the rounded reciprocal is off by at most one, and
// v - p q is exact (fma of integers below 2^53).  (The residue panels use
// the cheaper limb / Barrett form below; this serves the GEMM epilogue.)
__device__ __forceinline__ int mod_exact(double v, double p, double rp) {
    const double q = floor(v * rp);
    double r = fma(-p, q, v);
    if (r < 0.0) r += p;
    if (r >= p) r -= p;
    return (int)r;
}

// Persistent workgroups over 32 x 128 conversion tiles: the next tile's
// values are loaded into registers while the current one (transposed
// through LDS) is converted to balanced residues (|r| <= 127) for every
// modulus and written as 16-byte row pieces of the block's panel in each
// modulus plane.  Panel bytes for k in [K, Kp) are written as zeros.
__global__ __launch_bounds__(256) void to_residues_kernel(OzConvArgs a) {
    __shared__ double S[kOzTileX][kOzTileK + 1];
    const int tid = threadIdx.x;
    int it = blockIdx.x;
    if (it >= a.nitems) return;
    SubTile next;
    OzConvTile w = a.items[it];
    OzBlock b = a.blocks[w.block];
    fetch_subtile(b, a.tab, w.x0, w.k0, tid, next);
    for (; it < a.nitems; it += gridDim.x) {
        __syncthreads(); // the previous tile's LDS reads are done
        store_subtile(b, tid, next, S);
        __syncthreads();
        const OzConvTile cw = w;
        const OzBlock cb = b;
        const int nit = it + gridDim.x;
        if (nit < a.nitems) {
            w = a.items[nit];
            b = a.blocks[w.block];
            fetch_subtile(b, a.tab, w.x0, w.k0, tid, next);
        }
        const int xx = tid >> 3, kc = (tid & 7) * 16; // 8 lanes: one 128-byte row piece
        const int x = cw.x0 + xx, k = cw.k0 + kc;
        if (x >= cb.X || k >= cb.Kp) continue;
        const int e = a.expo[cb.exp_off + x];
        const int shift = e == INT_MIN ? 0 : (a.c.P - 1) - e;
        // each scaled integer v (|v| <= 2^51, exact in FP64) split once into
        // v = h2 2^34 + h1 2^17 + h0 with h0, h1 in [0, 2^17), |h2| <= 2^17
        int h0[16], h1[16], h2[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double v = e == INT_MIN ? 0.0 : rint(ldexp(S[xx][kc + j], shift));
            const double d2 = floor(v * 0x1p-34);
            const double r1 = fma(-d2, 0x1p34, v); // exact, in [0, 2^34)
            const double d1 = floor(r1 * 0x1p-17);
            h2[j] = (int)d2;
            h1[j] = (int)d1;
            h0[j] = (int)fma(-d1, 0x1p17, r1);
        }
        int8_t *dst = a.planes + cb.panel_off + (long long)x * cb.Kp + k;
        for (int i = 0; i < a.c.n; ++i) {
            const int p = a.c.p[i], c17 = a.c.c17[i], c34 = a.c.c34[i], off = a.c.off[i];
            const uint32_t bar = a.c.bar[i];
            uint32_t word[4] = {0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                // t = v mod p up to a multiple of p, in (0, 2^28): 24-bit products
                const int t = __mul24(h2[j], c34) + __mul24(h1[j], c17) + h0[j] + off;
                const uint32_t q = __umulhi((uint32_t)t, bar); // floor(t/p) or one less
                int r = t - __mul24((int)q, p);                  // in [0, 2p)
                if (r >= p) r -= p;
                if (r > p / 2) r -= p; // balanced, |r| <= 127
                word[j >> 2] |= (uint32_t)(uint8_t)(int8_t)r << (8 * (j & 3));
            }
            *reinterpret_cast<int32x4 *>(dst + (long long)i * a.plane_bytes) =
                int32x4{(int)word[0], (int)word[1], (int)word[2], (int)word[3]};
        }
    }
}

// INT8 GEMM over residue panels, one modulus per blockIdx.y.  Workgroup tile
// BM x BN (4 waves, 2 x 2, each (BM/2) x (BN/2)); the task's virtual K runs
// in stages of 128 bytes = 8 chunks, double-buffered in LDS; loads run two
// stages ahead through two register sets (16 bytes per lane each): the
// loads of stage s + 2 are in flight while stage s is multiplied, and
// stage s + 1 is written to the other buffer after it.  LDS rows hold 8 chunk
// slots, chunk c of row r in slot c ^ (r & 7), so both the stores (one wave
// = 8 whole rows) and the fragment reads (16 rows x 4 chunks) spread over
// all banks.  Rows past the task's M / N read row M-1 /
// N-1 and are never stored; chunks past the pairs' K read the plane's zero
// bytes.  MFMA v_mfma_i32_16x16x64_i8: lane l holds row l%16, bytes
// 16*(l/16) .. +15 of a 64-byte k step; accumulator register r of lane l is
// C(4*(l/16) + r, l%16).
template <int BM, int BN, int WM, int WN, int MINB, int SETS>
__global__ __launch_bounds__(64 * WM * WN, MINB) void gemm_i8_kernel(OzGemmArgs a) {
    constexpr int NT = 64 * WM * WN; // threads
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int KS = kOzStage, CH = kOzStageChunks;
    // 8-row load instructions per wave per stage
    constexpr int IA = BM / (8 * WM * WN), IB = BN / (8 * WM * WN);
    static_assert(CH == 8 && KS == 128, "a stage is 8 chunks of 16 bytes");
    static_assert(IA * 8 * WM * WN == BM && IB * 8 * WM * WN == BN && TM % 16 == 0 &&
                      TN % 16 == 0, "tile shape");
    constexpr int kWin = 512; // chunk entries per LDS window (64 stages)
    static_assert(kWin % (2 * CH) == 0, "a window holds whole pairs of stages");
    __shared__ __attribute__((aligned(1024))) int8_t As[2][BM * KS];
    __shared__ __attribute__((aligned(1024))) int8_t Bs[2][BN * KS];
    __shared__ __attribute__((aligned(16))) OzChunkRef Es[kWin];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int g = (lane & 7) ^ (lane >> 3); // chunk of the stage this lane loads
    const int mod = blockIdx.y;
    const int p = a.c.p[mod], u = a.c.u[mod], c17 = a.c.c17[mod];
    const uint32_t bar = a.c.bar[mod];
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
        const int nst = task.nchunks / CH; // even: the planner pads to pairs of stages
        // SETS == 2: two register sets, the loads of stage s + 2 are issued
        // while stage s is multiplied and land in LDS after stage s + 1;
        // SETS == 1 (the 64 x 128 wave tiles): one set, one stage ahead
        int32x4 xa[IA], xb[IB], ya[SETS == 2 ? IA : 1], yb[SETS == 2 ? IB : 1];
        auto put = [&](int buf, const int32x4 *va, const int32x4 *vb) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                *reinterpret_cast<int32x4 *>(&As[buf][(wave * IA + i) * 1024 + lane * 16]) = va[i];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                *reinterpret_cast<int32x4 *>(&Bs[buf][(wave * IB + i) * 1024 + lane * 16]) = vb[i];
        };
        // B fragments in groups of FH (the 64 x 128 wave tiles read them in
        // two halves, which keeps the second prefetch register set in budget)
        constexpr int FH = FN > 4 ? 4 : FN;
        auto compute = [&](int buf) {
            const int8_t *sa = As[buf], *sb = Bs[buf];
#pragma unroll
            for (int ks = 0; ks < KS / 64; ++ks) {
                const int slot = ((ks * 4 + lk) ^ (li & 7)) * 16;
                int32x4 fa[FM];
#pragma unroll
                for (int i = 0; i < FM; ++i)
                    fa[i] = *reinterpret_cast<const int32x4 *>(sa + (wm * TM + i * 16 + li) * KS + slot);
#pragma unroll
                for (int j0 = 0; j0 < FN; j0 += FH) {
                    int32x4 fb[FH];
#pragma unroll
                    for (int j = 0; j < FH; ++j)
                        fb[j] = *reinterpret_cast<const int32x4 *>(
                            sb + (wn * TN + (j0 + j) * 16 + li) * KS + slot);
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FH; ++j)
                            acc[i][j0 + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                fa[i], fb[j], acc[i][j0 + j], 0, 0, 0);
                }
            }
        };

        // loads are unconditional (stage indices are clamped, so the last
        // stages are fetched again) so the staged registers need no merge
        // copies, which would wait on them
        // Chunk entries are staged in LDS, a window of kWin entries at a time,
        // and each lane reads its own with ds_read: waiting for an entry then
        // waits on the LDS counter only, never on the operand loads in flight.
        struct Sel {
            uint32_t a, b;
            int stride;
        };
        auto load_window = [&](int st0) {
            __syncthreads(); // nobody still reads the previous window
            const int c0 = st0 * CH, n = min(kWin, task.nchunks - c0);
            for (int i = tid; i < n; i += NT)
                *reinterpret_cast<int32x4 *>(&Es[i]) =
                    *reinterpret_cast<const int32x4 *>(&ch[c0 + i]);
            __syncthreads();
        };
        auto entry = [&](int st) {
            const int idx = (min(st, nst - 1) * CH + g) % kWin;
            const OzChunkRef e = Es[idx];
            return Sel{e.a, e.b, e.stride};
        };
        auto fetch2 = [&](const Sel &e, int32x4 *va, int32x4 *vb) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                va[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.a + ra[i] * e.stride);
#pragma unroll
            for (int i = 0; i < IB; ++i)
                vb[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.b + rb[i] * e.stride);
        };
        load_window(0);
        if constexpr (SETS == 1) {
            fetch2(entry(0), xa, xb);
            put(0, xa, xb);
            __syncthreads();
            int buf = 0;
            for (int s = 0; s < nst; ++s) {
                if (s + 1 < nst && ((s + 1) * CH) % kWin == 0) load_window(s + 1);
                fetch2(entry(s + 1), xa, xb);
                __builtin_amdgcn_sched_barrier(0); // keep the loads ahead of the MFMAs
                compute(buf);
                put(buf ^ 1, xa, xb); // after the last stage this writes a copy nothing reads
                __syncthreads();
                buf ^= 1;
            }
        } else {
        fetch2(entry(0), xa, xb);
        fetch2(entry(1), ya, yb);
        put(0, xa, xb);
        __syncthreads();
        for (int s = 0; s < nst; s += 2) {
            // entries of stages s + 2, s + 3 are in the next window (a window
            // holds an even number of stages, so they share one)
            if (s + 2 < nst && ((s + 2) * CH) % kWin == 0) load_window(s + 2);
            // stage s from buffer 0; stage s + 1 waits in (ya, yb)
            fetch2(entry(s + 2), xa, xb);
            __builtin_amdgcn_sched_barrier(0); // keep the loads ahead of the MFMAs
            compute(0);
            put(1, ya, yb);
            __syncthreads();
            // stage s + 1 from buffer 1; stage s + 2 waits in (xa, xb)
            fetch2(entry(s + 3), ya, yb);
            __builtin_amdgcn_sched_barrier(0);
            compute(1);
            put(0, xa, xb); // after the last pair of stages this writes a copy nothing reads
            __syncthreads();
        }
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
                        const uint32_t tt = crt_digit(acc[i][j][r], p, c17, u, bar);
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
        // four consecutive elements per thread: one 32-bit load per plane
        // (task regions start 16-byte aligned and planes end 16 bytes past
        // the last region, so the word never leaves the plane)
        for (int e0 = w.e0 + 4 * (int)threadIdx.x; e0 < e1; e0 += 4 * (int)blockDim.x) {
            U128 acc[4] = {{0ull, 0ull}, {0ull, 0ull}, {0ull, 0ull}, {0ull, 0ull}};
            double s[4] = {0.0, 0.0, 0.0, 0.0};
            // moduli four at a time: the four planes' words are loaded
            // before any is used, so their latencies overlap
            for (int i0 = 0; i0 < a.c.n; i0 += 4) {
                uint32_t w4[4];
#pragma unroll
                for (int d = 0; d < 4; ++d)
                    w4[d] = i0 + d < a.c.n
                                ? *reinterpret_cast<const uint32_t *>(
                                      res + (long long)(i0 + d) * a.res_plane_bytes + e0)
                                : 0u;
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    if (i0 + d >= a.c.n) break;
                    const int i = i0 + d;
                    const double rp = a.c.rp[i];
                    const uint64_t mi[2] = {a.c.mi[i][0], a.c.mi[i][1]};
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const unsigned t = (w4[d] >> (8 * q)) & 0xffu;
                        s[q] += (double)t * rp;
                        mad_small(acc[q], (unsigned long long)t, mi);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int e = e0 + q;
                if (e >= e1) break;
                int m, n;
                if (task.nfast) {
                    m = e / task.N;
                    n = e - m * task.N;
                } else {
                    n = e / task.M;
                    m = e - n * task.M;
                }
                U128 qm{0ull, 0ull};
                mad_small(qm, (unsigned long long)rint(s[q]), a.c.m);
                const U128 C{acc[q].lo - qm.lo,
                             acc[q].hi - qm.hi - (acc[q].lo < qm.lo ? 1ull : 0ull)};
                const int ea = a.expo[task.expA + m], eb = a.expo[task.expB + n];
                double v = 0.0;
                if (ea != INT_MIN && eb != INT_MIN) {
                    const int sa = (a.c.P - 1) - ea, sb = (a.c.P - 1) - eb;
                    v = ldexp(to_double(C), -(sa + sb));
                }
                int beta;
                double *p = oz_dest(task, a.subs, a.tab, m, n, beta);
                if (beta) v += *p;
                *p = v;
            }
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

hipError_t oz_zero_heads(int8_t *planes, long long plane_bytes, int nplanes, hipStream_t s) {
    if (nplanes <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_heads_kernel, dim3((nplanes + 3) / 4), dim3(64), 0, s, planes,
                       plane_bytes, nplanes);
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
    hipLaunchKernelGGL(to_residues_kernel, dim3(grid_for(a.nitems, 4 * 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_gemm(const OzGemmArgs &a, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    if (a.tile == kOzBigTile) return oz_gemm_big(a, s);
    if (a.tile == kOzWideTile) {
        hipLaunchKernelGGL((gemm_i8_kernel<kOzBigTile, kOzBigTile, 4, 2, 1, 1>),
                           dim3(a.ntiles, a.c.n), dim3(512), 0, s, a);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((gemm_i8_kernel<kOzBM, kOzBN, 2, 2, 2, 2>), dim3(a.ntiles, a.c.n), dim3(256),
                       0, s, a);
    return hipGetLastError();
}

hipError_t oz_reconstruct(const OzRecArgs &a, hipStream_t s) {
    if (a.nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(reconstruct_kernel, dim3(grid_for(a.nitems, 65535)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace sipgpu
