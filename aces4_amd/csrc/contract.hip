// contract.hip -- the contract super-instruction on gfx950: a grouped,
// persistent FP64-MFMA GEMM with the index permutations fused into the
// operand loads and the result scatter.
//
// Reference semantics: tensor_block_contract_ (tensor_dil_omp.F90:662-906)
// permutes L and R to contracted-first order, zeroes D, calls
// dgemm('T','N',M,N,K,1,L,K,R,K,1,D,M) and permutes the product into D.
// Here nothing is permuted in memory: a workgroup computes a BM x BN tile of
//   D(m, n) = sum_k L(k, m) R(k, n)
// reading L(k, m) = L[koffL[k] + moffL[m]] and R(k, n) = R[koffR[k] +
// noffR[n]] through the plan's offset tables, and writes the tile to
// D[moffD[m] + noffD[n]].  A task (one destination block) may sum several
// operand pairs in one accumulator -- the `T = A*B; R += T` idiom of SIAL
// programs done in registers -- and is written once, so results do not
// depend on how tasks are batched.
//
// MFMA: v_mfma_f64_16x16x4_f64.  Operand lane layout: A(i, k) in lane
// i + 16k, B(k, j) in lane j + 16k.  Accumulator: register r of lane l
// holds C(4r + l/16, l%16).
#include "kernels.h"

namespace sipgpu {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kBK = 16; // K per LDS stage (4 MFMA k-steps)

// Global -> register stage of a BK x BX operand tile.  `kfast` picks the
// lane order that keeps the global loads coalesced: consecutive lanes walk
// k when the operand's fastest dim is contracted, x otherwise.
template <int BX, int NT>
struct TileLoader {
    static constexpr int kItems = kBK * BX / NT;
    static_assert(kItems >= 1 && (kBK * BX) % NT == 0, "tile/threads mismatch");
    int xoff[kItems]; // per-pair x offsets (or -1 when out of range)
    int kk[kItems];
    int xx[kItems];

    __device__ void setup(bool kfast, int tid) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            const int e = tid + i * NT;
            if (kfast) { kk[i] = e % kBK; xx[i] = e / kBK; }
            else       { xx[i] = e % BX;  kk[i] = e / BX; }
        }
    }
    __device__ void pair(const int32_t *__restrict__ xtab, int x0, int X) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            const int x = x0 + xx[i];
            xoff[i] = x < X ? xtab[x] : -1;
        }
    }
    __device__ void load(double (&v)[kItems], const double *__restrict__ src,
                         const int32_t *__restrict__ ktab, int k0, int K) const {
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
            const int k = k0 + kk[i];
            v[i] = 0.0;
            if (k < K && xoff[i] >= 0) v[i] = src[ktab[k] + xoff[i]];
        }
    }
    __device__ void store(double *S, int pitch, const double (&v)[kItems]) const {
#pragma unroll
        for (int i = 0; i < kItems; ++i) S[kk[i] * pitch + xx[i]] = v[i];
    }
};

// BM x BN workgroup tile, WM x WN waves, each wave a (BM/WM) x (BN/WN) tile
// of 16x16 MFMA blocks.
template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(WM *WN * 64) void gemm_f64_kernel(GemmArgs args) {
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int PA = BM + 1, PB = BN + 1; // odd pitch: k-fast stores hit distinct banks
    static_assert(FM >= 1 && FN >= 1, "wave tile too small");
    __shared__ double As[kBK * PA];
    __shared__ double Bs[kBK * PB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;

    TileLoader<BM, NT> la;
    TileLoader<BN, NT> lb;
    double va[TileLoader<BM, NT>::kItems];
    double vb[TileLoader<BN, NT>::kItems];

    for (int t = blockIdx.x; t < args.ntiles; t += gridDim.x) {
        const GemmTile tile = args.tiles ? args.tiles[t]
                                         : GemmTile{0, t % args.tilesM, t / args.tilesM, 0};
        if (tile.task < 0) continue; // padding slot of the XCD schedule
        const GemmTask task = args.tiles ? args.tasks[tile.task] : args.task0;
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;

        f64x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

        for (int p = task.first; p < task.first + task.npairs; ++p) {
            const GemmPair pr = args.tiles ? args.pairs[p] : args.pair0;
            const int32_t *ktabL = args.tab + pr.kL;
            const int32_t *ktabR = args.tab + pr.kR;
            la.setup(pr.flags & 1, tid);
            lb.setup(pr.flags & 2, tid);
            la.pair(args.tab + pr.mL, m0, task.M);
            lb.pair(args.tab + pr.nR, n0, task.N);
            la.load(va, pr.L, ktabL, 0, pr.K);
            lb.load(vb, pr.R, ktabR, 0, pr.K);
            for (int k0 = 0; k0 < pr.K; k0 += kBK) {
                __syncthreads(); // previous stage fully consumed
                la.store(As, PA, va);
                lb.store(Bs, PB, vb);
                __syncthreads();
                if (k0 + kBK < pr.K) { // prefetch the next stage into registers
                    la.load(va, pr.L, ktabL, k0 + kBK, pr.K);
                    lb.load(vb, pr.R, ktabR, k0 + kBK, pr.K);
                }
#pragma unroll
                for (int ks = 0; ks < kBK / 4; ++ks) {
                    const int kr = ks * 4 + lk;
                    double a[FM], b[FN];
#pragma unroll
                    for (int i = 0; i < FM; ++i) a[i] = As[kr * PA + wm * TM + i * 16 + li];
#pragma unroll
                    for (int j = 0; j < FN; ++j) b[j] = Bs[kr * PB + wn * TN + j * 16 + li];
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j],
                                                                             0, 0, 0);
                }
            }
        }

        // Epilogue: scatter the tile into the destination block.
        const int32_t *mtab = args.tab + task.mD;
        const int32_t *ntab = args.tab + task.nD;
        double *__restrict__ D = task.D;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int n = n0 + wn * TN + j * 16 + li;
            if (n >= task.N) continue;
            const int no = ntab[n];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = m0 + wm * TM + i * 16 + 4 * r + lk;
                    if (m < task.M) {
                        const int off = mtab[m] + no;
                        double v = acc[i][j][r];
                        if (task.beta) v += D[off];
                        D[off] = v;
                    }
                }
            }
        }
        __syncthreads(); // LDS reuse across persistent iterations
    }
}

__global__ void scale_scatter_kernel(double *D, const int32_t *moffD, const double *X,
                                     const int32_t *moffX, const double *s, int M) {
    // tensor_block_add_ on a zeroed destination (tensor_dil_omp.F90:394-436):
    // a factor within 1e-13 of one is not applied.
    double f = *s;
    const double df = f - 1.0;
    if (!(fabs(df) > 1e-13)) f = 1.0;
    const bool scale = fabs(df) > 1e-13;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const double x = X[moffX[m]];
        D[moffD[m]] = 0.0 + (scale ? x * f : x);
    }
}

constexpr int kDotThreads = 256;

__global__ __launch_bounds__(kDotThreads) void dot_partial_kernel(
    double *part, const double *L, const int32_t *koffL, const double *R,
    const int32_t *koffR, int K) {
    __shared__ double red[kDotThreads];
    double v = 0.0;
    const int chunk = (K + gridDim.x - 1) / gridDim.x;
    const int kb = blockIdx.x * chunk, ke = min(K, kb + chunk);
    for (int k = kb + threadIdx.x; k < ke; k += kDotThreads) v += L[koffL[k]] * R[koffR[k]];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kDotThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void dot_final_kernel(double *D, const double *part, int nparts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double v = 0.0;
        for (int i = 0; i < nparts; ++i) v += part[i];
        D[0] = 0.0 + v;
    }
}

__global__ void scalar_mul_kernel(double *D, const double *L, const double *R) {
    if (threadIdx.x == 0 && blockIdx.x == 0) D[0] = L[0] * R[0];
}

template <int BM, int BN, int WM, int WN>
hipError_t launch_cfg(const GemmArgs &args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, WM, WN>), dim3(grid), dim3(WM * WN * 64), 0,
                       stream, args);
    return hipGetLastError();
}

} // namespace

int gemm_tile_m(int cfg) {
    switch (cfg) {
    case kGemm64x64: return 64;
    case kGemm128x128: return 128;
    case kGemm64x128: return 64;
    default: return 0;
    }
}

int gemm_tile_n(int cfg) {
    switch (cfg) {
    case kGemm64x64: return 64;
    case kGemm128x128: return 128;
    case kGemm64x128: return 128;
    default: return 0;
    }
}

hipError_t launch_gemm(int cfg, const GemmArgs &args, int grid, hipStream_t stream) {
    if (args.ntiles <= 0) return hipSuccess;
    switch (cfg) {
    case kGemm64x64: return launch_cfg<64, 64, 2, 2>(args, grid, stream);
    case kGemm128x128: return launch_cfg<128, 128, 2, 2>(args, grid, stream);
    case kGemm64x128: return launch_cfg<64, 128, 2, 2>(args, grid, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scale_scatter(double *D, const int32_t *moffD, const double *X,
                                const int32_t *moffX, const double *s, int M,
                                hipStream_t stream) {
    if (M <= 0) return hipSuccess;
    const int threads = 256;
    const int blocks = (M + threads - 1) / threads < 4096 ? (M + threads - 1) / threads : 4096;
    hipLaunchKernelGGL(scale_scatter_kernel, dim3(blocks), dim3(threads), 0, stream, D, moffD,
                       X, moffX, s, M);
    return hipGetLastError();
}

hipError_t launch_dot(double *D, const double *L, const int32_t *koffL, const double *R,
                      const int32_t *koffR, int K, double *scratch, hipStream_t stream) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(kDotThreads), 0, stream,
                       scratch, L, koffL, R, koffR, K);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(64), 0, stream, D, scratch, kDotBlocks);
    return hipGetLastError();
}

hipError_t launch_scalar_mul(double *D, const double *L, const double *R, hipStream_t stream) {
    hipLaunchKernelGGL(scalar_mul_kernel, dim3(1), dim3(64), 0, stream, D, L, R);
    return hipGetLastError();
}

} // namespace sipgpu
