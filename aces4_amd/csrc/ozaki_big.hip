// ozaki_big.hip -- the split path's INT8 GEMM with 256 x 256 workgroup
// tiles, for batches of large result blocks.  Built without the VGPR-form
// MFMA option: its 256 accumulator registers per lane live in AGPRs.
#include "kernels.h"
#include "ozaki.h"

namespace sipgpu {

namespace {

typedef int int32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) int8_t gi8_t;

// The GEMM with a 256 x 256 workgroup tile: each of the 4 waves owns
// 128 x 128 (8 x 8 fragments, 256 accumulator registers, in AGPRs), which
// halves the LDS bytes read and written per MFMA against 64 x 64 wave tiles.
// Stages are 64 bytes of virtual K (4 chunks); LDS rows of 64 bytes hold
// chunk c of row r in slot c ^ ((r >> 1) & 3), so a wave's 16-row stores and
// 8-row fragment reads cover all banks.  Loads run two stages ahead through
// two register sets (16 bytes per lane per 16-row group); chunk entries come
// from an LDS window.
__global__ __launch_bounds__(256, 1) void gemm_i8_big_kernel(OzGemmArgs a) {
    constexpr int BM = 256, BN = 256, WM = 2, WN = 2;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int KS = 64, CH = KS / kOzChunk; // 4 chunks per stage
    constexpr int IA = BM / 64, IB = BN / 64;  // 16-row load groups per wave per stage
    constexpr int kWin = 512;
    static_assert(kWin % (2 * CH) == 0, "a window holds whole pairs of stages");
    // one LDS buffer: two stages of A and B during the K loop, then the
    // accumulators, 16 KB per wave per epilogue round
    constexpr int kStageBytes = (BM + BN) * KS;
    constexpr int kPitch = 20; // int32 per accumulator column in the epilogue (2-way banks)
    static_assert(2 * kStageBytes >= 4 * 128 * kPitch * 4, "epilogue rounds fit the stages");
    __shared__ __attribute__((aligned(1024))) int8_t Lds[2 * kStageBytes];
    __shared__ __attribute__((aligned(16))) OzChunkRef Es[kWin];
    int8_t *As[2] = {Lds, Lds + kStageBytes};
    int8_t *Bs[2] = {Lds + BM * KS, Lds + kStageBytes + BM * KS};

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int g = (lane & 3) ^ ((lane >> 3) & 3); // chunk of the stage this lane loads
    const int mod = blockIdx.y;
    const int p = a.c.p[mod], u = a.c.u[mod];
    const uint32_t bar = a.c.bar[mod];
    gi8_t *plane = (gi8_t *)(a.planes + (long long)mod * a.plane_bytes);

    for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        const OzTile tile = a.tiles[t];
        if (tile.task < 0) continue;
        const OzTask task = a.tasks[tile.task];
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        const OzChunkRef *ch = a.chunks + task.chunk0;
        const int nst = task.nchunks / CH; // even
        // rows this lane loads: group i of the wave covers 16 rows
        int ra[IA], rb[IB];
#pragma unroll
        for (int i = 0; i < IA; ++i) ra[i] = min(m0 + 64 * wave + 16 * i + (lane >> 2), task.M - 1);
#pragma unroll
        for (int i = 0; i < IB; ++i) rb[i] = min(n0 + 64 * wave + 16 * i + (lane >> 2), task.N - 1);

        int32x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = int32x4{0, 0, 0, 0};

        auto load_window = [&](int st0) {
            __syncthreads();
            const int c0 = st0 * CH, n = min(kWin, task.nchunks - c0);
            for (int i = tid; i < n; i += 256)
                *reinterpret_cast<int32x4 *>(&Es[i]) = *reinterpret_cast<const int32x4 *>(&ch[c0 + i]);
            __syncthreads();
        };
        struct Sel {
            uint32_t a, b;
            int stride;
        };
        auto entry = [&](int st) {
            const OzChunkRef e = Es[(min(st, nst - 1) * CH + g) % kWin];
            return Sel{e.a, e.b, e.stride};
        };
        int32x4 xa[IA], xb[IB], ya[IA], yb[IB];
        auto fetch = [&](const Sel &e, int32x4 (&va)[IA], int32x4 (&vb)[IB]) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                va[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.a + ra[i] * e.stride);
#pragma unroll
            for (int i = 0; i < IB; ++i)
                vb[i] = *reinterpret_cast<const __attribute__((address_space(1))) int32x4 *>(
                    plane + e.b + rb[i] * e.stride);
        };
        auto put = [&](int buf, const int32x4 (&va)[IA], const int32x4 (&vb)[IB]) {
#pragma unroll
            for (int i = 0; i < IA; ++i)
                *reinterpret_cast<int32x4 *>(&As[buf][(4 * wave + i) * 1024 + lane * 16]) = va[i];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                *reinterpret_cast<int32x4 *>(&Bs[buf][(4 * wave + i) * 1024 + lane * 16]) = vb[i];
        };
        auto compute = [&](int buf) {
            const int8_t *sa = As[buf], *sb = Bs[buf];
            int32x4 fb[FN];
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int r = wn * TN + j * 16 + li;
                fb[j] = *reinterpret_cast<const int32x4 *>(sb + r * KS + ((lk ^ ((r >> 1) & 3)) * 16));
            }
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int r = wm * TM + i * 16 + li;
                const int32x4 fa =
                    *reinterpret_cast<const int32x4 *>(sa + r * KS + ((lk ^ ((r >> 1) & 3)) * 16));
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa, fb[j], acc[i][j], 0, 0, 0);
            }
        };

        load_window(0);
        fetch(entry(0), xa, xb);
        fetch(entry(1), ya, yb);
        put(0, xa, xb);
        __syncthreads();
        for (int s = 0; s < nst; s += 2) {
            if (s + 2 < nst && ((s + 2) * CH) % kWin == 0) load_window(s + 2);
            fetch(entry(s + 2), xa, xb);
            __builtin_amdgcn_sched_barrier(0);
            compute(0);
            put(1, ya, yb);
            __syncthreads();
            fetch(entry(s + 3), ya, yb);
            __builtin_amdgcn_sched_barrier(0);
            compute(1);
            put(0, xa, xb);
            __syncthreads();
        }

        // Epilogue through LDS in eight rounds (fragment row q of every
        // wave): t = (C mod p) u mod p, written one byte per element in
        // the task's element order (|C| < 2^31, so C + p 2^23 is positive and
        // below 2^32 for the Barrett step).
        uint8_t *res = a.res_planes + (long long)mod * a.res_plane_bytes + task.res_off;
        const uint32_t cofs = (uint32_t)p << 17;
        const int c17 = a.c.c17[mod];
        int32_t *Cs = reinterpret_cast<int32_t *>(Lds);
#pragma unroll
        for (int q = 0; q < FM; ++q) {
            // wave region: 16 rows x 128 columns, column-major with pitch kPitch
            int32_t *R = Cs + wave * 128 * kPitch;
#pragma unroll
            for (int j = 0; j < FN; ++j)
                *reinterpret_cast<int32x4 *>(&R[(j * 16 + li) * kPitch + 4 * lk]) = acc[q][j];
            __syncthreads();
            for (int e = tid; e < 4 * 16 * 128; e += 256) {
                const int w = e >> 11, rem = e & 2047;
                int row, col;
                if (task.nfast) {
                    row = rem >> 7;
                    col = rem & 127;
                } else {
                    col = rem >> 4;
                    row = rem & 15;
                }
                const int m = m0 + (w % WM) * TM + q * 16 + row;
                const int n = n0 + (w / WM) * TN + col;
                if (m < task.M && n < task.N) {
                    // v = h 2^17 + l: h c17 + l + p 2^17 is congruent to v,
                    // positive and below 2^26 for every |v| < 2^31
                    const int v = Cs[w * 128 * kPitch + col * kPitch + row];
                    const uint32_t cv = (uint32_t)((v >> 17) * c17 + (v & 0x1ffff)) + cofs;
                    uint32_t c = cv - __umulhi(cv, bar) * (uint32_t)p;
                    if (c >= (uint32_t)p) c -= p;
                    uint32_t tt = c * (uint32_t)u;
                    tt -= __umulhi(tt, bar) * (uint32_t)p;
                    if (tt >= (uint32_t)p) tt -= p;
                    const long long idx =
                        task.nfast ? (long long)m * task.N + n : (long long)n * task.M + m;
                    res[idx] = (uint8_t)tt;
                }
            }
            __syncthreads();
        }
    }
}

} // namespace

hipError_t oz_gemm_big(const OzGemmArgs &a, hipStream_t s) {
    if (a.ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(gemm_i8_big_kernel, dim3(a.ntiles, a.c.n), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace sipgpu
