// packed.hip -- the FP64 contraction on packed operand panels.
//
// For large batches the operand permutations are applied once per block
// instead of once per tile: every operand instance (block + its k / x
// tables) is packed into X row-major rows of Kp doubles (K rounded up to
// an even count, zero filled), and the grouped GEMM runs over each task's
// virtual K -- its pairs' panels concatenated in 16-byte chunks through the
// same chunk map as the split-precision path -- with no table lookups in
// the K loop.  The product is scattered into the destination block through
// its tables, accumulating when the task does.
//
// MFMA v_mfma_f64_16x16x4_f64: lane l holds A(i = l%16, k = l/16); register
// r of lane l holds C(4r + l/16, l%16).  A 16-byte LDS read gives a lane two
// consecutive k of its row; the pair feeds two MFMAs (the first double to the
// first, the second to the second), so MFMA t of a 64-byte step sums
// k = 2 (4 s + l/16) + t over l/16 = 0..3: the same permutation of k for both
// operands, so the sum over all MFMAs is the full dot product.
#include "kernels.h"
#include "ozaki.h"

namespace sipgpu {

namespace {

typedef int int32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) int8_t gi8_t;

// Workgroups over 32 x 128 tiles of operand blocks, each done as two halves
// of 64 k (half the LDS of a whole tile, so twice the workgroups per CU keep
// their loads in flight): read along the index adjacent in memory,
// transposed through LDS, written as 128-byte row pieces of the panel.
__global__ __launch_bounds__(256) void pack_kernel(OzConvArgs a) {
    constexpr int HK = kOzTileK / 2;
    __shared__ __attribute__((aligned(16))) double S[kOzTileX][HK + 2];
    const int tid = threadIdx.x;
    for (int it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const OzConvTile w = a.items[it];
        const OzBlock b = a.blocks[w.block];
#pragma unroll 1
        for (int half = 0; half < 2; ++half) {
            const int kb = w.k0 + half * HK;
            double v[8];
            if (b.xfast) {
                const int x = w.x0 + (tid & 31);
                const bool xin = x < b.X;
                const int xo = xin ? a.tab[b.xtab + x] : 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = kb + (tid >> 5) + 8 * j;
                    v[j] = (xin && k < b.K) ? b.data[a.tab[b.ktab + k] + xo] : 0.0;
                }
            } else {
                const int k = kb + (tid & 63);
                const bool kin = k < b.K;
                const int ko = kin ? a.tab[b.ktab + k] : 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int x = w.x0 + (tid >> 6) + 4 * j;
                    v[j] = (kin && x < b.X) ? b.data[ko + a.tab[b.xtab + x]] : 0.0;
                }
            }
            __syncthreads(); // the previous LDS reads are done
            if (b.xfast) {
#pragma unroll
                for (int j = 0; j < 8; ++j) S[tid & 31][(tid >> 5) + 8 * j] = v[j];
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) S[(tid >> 6) + 4 * j][tid & 63] = v[j];
            }
            __syncthreads();
            // 8 lanes per row, 4 pairs each: a row's 8 lanes cover 128 bytes
            const int xx = tid >> 3, l = tid & 7;
            const int x = w.x0 + xx;
            if (x < b.X) {
                double *dst = reinterpret_cast<double *>(a.planes + b.panel_off) +
                              (long long)x * b.Kp + kb;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int k = 2 * (8 * j + l);
                    if (kb + k < b.Kp) // Kp is even: whole 16-byte pairs
                        *reinterpret_cast<f64x2 *>(dst + k) =
                            *reinterpret_cast<const f64x2 *>(&S[xx][k]);
                }
            }
        }
    }
}

// Grouped FP64 GEMM over packed panels.  Workgroup tile BM x BN = 128 x 128
// (4 waves, 2 x 2, each 64 x 64: 16 accumulator fragments) or 64 x 64 (each
// wave 32 x 32: for small batches, whose tiles are cut into K segments:
// a segment's partial sums are a quarter of the bytes); stages of
// 128 bytes (16 k) double-buffered in LDS with rows XOR-swizzled by 16-byte
// chunk; loads run one stage ahead through registers; chunk entries come
// from an LDS window (their plane offsets count 16-byte units, so panels may
// reach 64 GiB).  Rows past the task's M / N read row M-1 / N-1 and are
// never stored; chunks past the pairs' K read the plane's zero bytes.
template <int BM, int BN>
__global__ __launch_bounds__(256, 2) void gemm_f64_packed_kernel(OzGemmArgs a) {
    constexpr int WM = 2, WN = 2;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int KS = kOzStage, CH = kOzStageChunks;
    constexpr int IA = BM / 32, IB = BN / 32;
    constexpr int kWin = 512;
    static_assert(CH == 8 && KS == 128, "a stage is 8 chunks of 16 bytes");
    __shared__ __attribute__((aligned(1024))) int8_t As[2][BM * KS];
    __shared__ __attribute__((aligned(1024))) int8_t Bs[2][BN * KS];
    __shared__ __attribute__((aligned(16))) OzChunkRef Es[kWin];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int g = (lane & 7) ^ (lane >> 3);
    gi8_t *plane = (gi8_t *)a.planes;

    const int u0 = a.wg_first ? a.wg_first[blockIdx.x] : (int)blockIdx.x;
    const int u1 = a.wg_first ? a.wg_first[blockIdx.x + 1] : min((int)blockIdx.x + 1, a.ntiles);
    for (int t = u0; t < u1; ++t) {
        const OzTile tile = a.tiles[t];
        if (tile.task < 0) continue;
        const OzTask task = a.tasks[tile.task];
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        int ra[IA], rb[IB];
#pragma unroll
        for (int i = 0; i < IA; ++i) ra[i] = min(m0 + 8 * (wave * IA + i) + (lane >> 3), task.M - 1);
#pragma unroll
        for (int i = 0; i < IB; ++i) rb[i] = min(n0 + 8 * (wave * IB + i) + (lane >> 3), task.N - 1);

        f64x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

        const OzChunkRef *ch = a.chunks + task.chunk0;
        const int sb = tile.s0, se = tile.s1; // stages of this unit, se > sb
        // windows of kWin chunk entries start at stage sb + 64 j
        auto load_window = [&](int st0) {
            __syncthreads();
            const int c0 = st0 * CH, n = min(kWin, se * CH - c0);
            for (int i = tid; i < n; i += 256)
                *reinterpret_cast<int32x4 *>(&Es[i]) = *reinterpret_cast<const int32x4 *>(&ch[c0 + i]);
            __syncthreads();
        };
        struct Sel {
            uint32_t a, b;
            int stride;
        };
        auto entry = [&](int st) {
            const OzChunkRef e = Es[((min(st, se - 1) - sb) * CH + g) % kWin];
            return Sel{e.a, e.b, e.stride};
        };
        int32x4 xa[IA], xb[IB];
        auto fetch = [&](const Sel &e) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                xa[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + (long long)e.a * kOzChunk + (long long)ra[i] * e.stride);
#pragma unroll
            for (int i = 0; i < IB; ++i)
                xb[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + (long long)e.b * kOzChunk + (long long)rb[i] * e.stride);
        };
        auto put = [&](int buf) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                *reinterpret_cast<int32x4 *>(&As[buf][(wave * IA + i) * 1024 + lane * 16]) = xa[i];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                *reinterpret_cast<int32x4 *>(&Bs[buf][(wave * IB + i) * 1024 + lane * 16]) = xb[i];
        };
        auto compute = [&](int buf) {
            const int8_t *sa = As[buf], *sb = Bs[buf];
#pragma unroll
            for (int ks = 0; ks < KS / 64; ++ks) {
                const int slot = ((ks * 4 + lk) ^ (li & 7)) * 16;
                f64x2 fa[FM], fb[FN];
#pragma unroll
                for (int i = 0; i < FM; ++i)
                    fa[i] = *reinterpret_cast<const f64x2 *>(sa + (wm * TM + i * 16 + li) * KS + slot);
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    fb[j] = *reinterpret_cast<const f64x2 *>(sb + (wn * TN + j * 16 + li) * KS + slot);
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i][h], fb[j][h],
                                                                             acc[i][j], 0, 0, 0);
            }
        };

        // One stage of loads ahead is enough here: a stage is 64 MFMAs of
        // 64 cycles per wave.  Loads are unconditional (the last stage is
        // fetched twice) so the staged registers need no merge copies.
        load_window(sb);
        fetch(entry(sb));
        put(0);
        __syncthreads();
        int buf = 0;
        for (int s = sb; s < se; ++s) {
            if (s + 1 < se && ((s + 1 - sb) * CH) % kWin == 0) load_window(s + 1);
            fetch(entry(s + 1));
            __builtin_amdgcn_sched_barrier(0);
            compute(buf);
            put(buf ^ 1); // after the last stage this writes a copy nothing reads
            __syncthreads();
            buf ^= 1;
        }

        if (tile.part >= 0) {
            // one K segment of a split tile: raw accumulators, register order;
            // elements outside the task (which the fixup skips) are not written
            double *w = a.work + (long long)tile.part * kOzWorkSlot;
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    const bool nin = n0 + wn * TN + j * 16 + li < task.N;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (nin && m0 + wm * TM + i * 16 + 4 * r + lk < task.M)
                            w[(((wave * FM + i) * FN + j) * 4 + r) * 64 + lane] = acc[i][j][r];
                }
            continue;
        }
        // scatter into the destination block(s), 16 elements at a time:
        // their destinations first, then all loads (beta), then the stores,
        // so the loads overlap instead of waiting behind each other's stores
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * TN + j * 16 + li;
            const bool nin = n < task.N;
            const int no = (task.rmap < 0 && nin) ? a.tab[task.nD + n] : 0;
            double *p[FM * 4];
            int bt[FM * 4];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * TM + i * 16 + 4 * r + lk;
                    const int q = i * 4 + r;
                    p[q] = nullptr;
                    bt[q] = 0;
                    if (nin && m < task.M) {
                        if (task.rmap < 0) {
                            p[q] = task.D + a.tab[task.mD + m] + no;
                            bt[q] = task.beta;
                        } else {
                            p[q] = oz_dest(task, a.subs, a.tab, m, n, bt[q]);
                        }
                    }
                }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (p[i * 4 + r] && bt[i * 4 + r]) acc[i][j][r] += *p[i * 4 + r];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (p[i * 4 + r]) *p[i * 4 + r] = acc[i][j][r];
        }
    }
}

// Split tiles: one thread per element of a split tile (BM * BN / 256
// workgroups of 256 per tile) sums the tile's segments in segment order --
// the layout the GEMM wrote: element e = (((wave * FM + i) * FN + j) * 4 + r)
// * 64 + lane holds fragment (i, j, r) of that wave and lane -- and stores
// the sum like the GEMM epilogue.  One element per thread keeps thousands of
// the dependent destination lookups in flight.
template <int BM, int BN>
__global__ __launch_bounds__(256) void fixup_f64_kernel(OzGemmArgs a) {
    constexpr int TM = BM / 2, TN = BN / 2, FM = TM / 16, FN = TN / 16;
    constexpr int LFM = FM == 4 ? 2 : 1, LFN = FN == 4 ? 2 : 1; // log2
    static_assert((FM == 4 || FM == 2) && (FN == 4 || FN == 2), "tile shape");
    constexpr int kPer = BM * BN / 256; // workgroups per tile
    const long long nwork = (long long)a.nfix * kPer;
    for (long long w = blockIdx.x; w < nwork; w += gridDim.x) {
        const OzTile f = a.fix[w / kPer];
        const int e = (int)(w % kPer) * 256 + threadIdx.x;
        const int lane = e & 63, r = (e >> 6) & 3, j = (e >> 8) & (FN - 1);
        const int i = (e >> (8 + LFN)) & (FM - 1);
        const int wave = e >> (8 + LFN + LFM), wm = wave & 1, wn = wave >> 1;
        const OzTask task = a.tasks[f.task];
        const int m = f.mt * BM + wm * TM + i * 16 + 4 * r + (lane >> 4);
        const int n = f.nt * BN + wn * TN + j * 16 + (lane & 15);
        if (m >= task.M || n >= task.N) continue;
        const double *w0 = a.work + (long long)f.part * kOzWorkSlot + e;
        // segments summed in order; loaded eight at a time so the loads are
        // in flight together
        double v = 0.0;
        for (int q0 = 0; q0 < f.s0; q0 += 8) {
            double x[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                x[q] = q0 + q < f.s0 ? w0[(long long)(q0 + q) * kOzWorkSlot] : 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (q0 + q < f.s0) v = q0 + q == 0 ? x[q] : v + x[q];
        }
        int beta;
        double *p = oz_dest(task, a.subs, a.tab, m, n, beta);
        if (beta) v += *p;
        *p = v;
    }
}

} // namespace

hipError_t oz_fixup_f64(const OzGemmArgs &a, hipStream_t s) {
    if (a.nfix <= 0) return hipSuccess;
    const int edge = a.tile == kOzSmallTile ? kOzSmallTile : kOzBM;
    const long long work = (long long)a.nfix * (edge * edge / 256);
    const dim3 grid((unsigned)(work < 1048576 ? work : 1048576));
    if (a.tile == kOzSmallTile)
        hipLaunchKernelGGL((fixup_f64_kernel<kOzSmallTile, kOzSmallTile>), grid, dim3(256), 0, s,
                           a);
    else
        hipLaunchKernelGGL((fixup_f64_kernel<kOzBM, kOzBN>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_pack(const OzConvArgs &a, hipStream_t s) {
    if (a.nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3(a.nitems < 65535 ? a.nitems : 65535), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t oz_gemm_f64(const OzGemmArgs &a, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    const dim3 grid(a.wg_first ? a.nwg : a.ntiles);
    if (a.tile == kOzSmallTile)
        hipLaunchKernelGGL((gemm_f64_packed_kernel<kOzSmallTile, kOzSmallTile>), grid, dim3(256), 0,
                           s, a);
    else
        hipLaunchKernelGGL((gemm_f64_packed_kernel<kOzBM, kOzBN>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace sipgpu
