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


typedef const __attribute__((address_space(1))) double *gptr_t;
typedef const __attribute__((address_space(1))) int32_t *gitab_t;

// Global -> register stage of a BK x BX operand tile.  `kfast` picks the
// lane order that keeps the global loads coalesced: consecutive lanes walk
// k when the operand's fastest dim is contracted, x otherwise.
//
// Software pipeline: the k-offset lookups of stage s+2 and the data loads of
// stage s+1 are in flight while stage s runs on the MFMA pipe.  All loads
// are unconditional (indices clamped into the block); out-of-range values
// are zeroed only when the stage is written to LDS, so no wait is forced
// before the MFMA loop.
template <int BX, int NT, int kBK>
struct TileLoader {
    static constexpr int kItems = kBK * BX / NT;
    static_assert(kItems >= 1 && (kBK * BX) % NT == 0 && NT % kBK == 0 && NT % BX == 0,
                  "tile/threads mismatch");
    static constexpr int kXStep = NT / kBK; // kfast: x advance per item
    static constexpr int kKStep = NT / BX;  // xfast: k advance per item
    bool kfast;
    int xoff[kItems]; // kfast: per item; xfast: xoff[0] only
    unsigned xmask;   // bit i: item i's x is inside the block
    int ko[kItems];   // k offsets of the next stage to issue (kfast: ko[0])
    gptr_t src;
    gitab_t ktab;
    int K;

    __device__ void pair(bool kf, const double *s, const int32_t *kt, int Kp,
                         const int32_t *xtab, int x0, int X, int tid) {
        kfast = kf;
        src = (gptr_t)s;
        ktab = (gitab_t)kt;
        K = Kp;
        gitab_t xt = (gitab_t)xtab;
        xmask = 0;
        if (kfast) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const int x = x0 + tid / kBK + i * kXStep;
                xoff[i] = xt[min(x, X - 1)];
                xmask |= (x < X ? 1u : 0u) << i;
            }
        } else {
            const int x = x0 + tid % BX;
            xoff[0] = xt[min(x, X - 1)];
            xmask = x < X ? ~0u : 0u;
        }
    }
    // Look up the k offsets of the stage starting at k0.
    __device__ void lookup(int k0, int tid) {
        if (kfast) {
            ko[0] = ktab[min(k0 + tid % kBK, K - 1)];
        } else {
#pragma unroll
            for (int i = 0; i < kItems; ++i) ko[i] = ktab[min(k0 + tid / BX + i * kKStep, K - 1)];
        }
    }
    // Issue the data loads of the looked-up stage.
    __device__ void issue(double (&v)[kItems]) const {
        if (kfast) {
#pragma unroll
            for (int i = 0; i < kItems; ++i) v[i] = src[ko[0] + xoff[i]];
        } else {
#pragma unroll
            for (int i = 0; i < kItems; ++i) v[i] = src[ko[i] + xoff[0]];
        }
    }
    // Write the stage starting at k0 to LDS, zeroing out-of-range elements.
    __device__ void store(double *S, int pitch, const double (&v)[kItems], int k0,
                          int tid) const {
        if (kfast) {
            const bool kin = k0 + tid % kBK < K;
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const bool ok = kin && ((xmask >> i) & 1u);
                S[(tid % kBK) * pitch + tid / kBK + i * kXStep] = ok ? v[i] : 0.0;
            }
        } else {
#pragma unroll
            for (int i = 0; i < kItems; ++i) {
                const bool ok = (k0 + tid / BX + i * kKStep < K) && xmask;
                S[(tid / BX + i * kKStep) * pitch + tid % BX] = ok ? v[i] : 0.0;
            }
        }
    }
};

// BM x BN workgroup tile, WM x WN waves, each wave a (BM/WM) x (BN/WN) tile
// of 16x16 MFMA blocks; K staged through LDS kBK at a time.
template <int BM, int BN, int WM, int WN, int kBK>
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

    TileLoader<BM, NT, kBK> la;
    TileLoader<BN, NT, kBK> lb;
    double va[TileLoader<BM, NT, kBK>::kItems];
    double vb[TileLoader<BN, NT, kBK>::kItems];

    for (int t = blockIdx.x; t < args.ntiles; t += gridDim.x) {
        const GemmTile tile = args.tiles ? args.tiles[t]
                                         : GemmTile{0, t % args.tilesM, t / args.tilesM, 0};
        if (tile.task < 0) continue; // padding slot of the XCD schedule
        const GemmTask task = args.tiles ? args.tasks[tile.task] : args.task0;
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        // rows / columns of this wave's sub-tile that lie inside the block
        const int mlive = __builtin_amdgcn_readfirstlane(task.M - (m0 + wm * TM));
        const int nlive = __builtin_amdgcn_readfirstlane(task.N - (n0 + wn * TN));

        f64x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

        for (int p = task.first; p < task.first + task.npairs; ++p) {
            const GemmPair pr = args.tiles ? args.pairs[p] : args.pair0;
            la.pair(pr.flags & 1, pr.L, args.tab + pr.kL, pr.K, args.tab + pr.mL, m0, task.M, tid);
            lb.pair(pr.flags & 2, pr.R, args.tab + pr.kR, pr.K, args.tab + pr.nR, n0, task.N, tid);
            la.lookup(0, tid);
            lb.lookup(0, tid);
            la.issue(va);
            lb.issue(vb);
            la.lookup(kBK, tid);
            lb.lookup(kBK, tid);
            for (int k0 = 0; k0 < pr.K; k0 += kBK) {
                __syncthreads(); // previous stage fully consumed
                la.store(As, PA, va, k0, tid);
                lb.store(Bs, PB, vb, k0, tid);
                __syncthreads();
                if (k0 + kBK < pr.K) {
                    la.issue(va);
                    lb.issue(vb);
                    la.lookup(k0 + 2 * kBK, tid);
                    lb.lookup(k0 + 2 * kBK, tid);
                }
                // Only the k-steps inside the pair's K and the 16x16 blocks
                // inside the destination are issued (wave-uniform tests): an
                // edge wave leaves its SIMD to the other resident workgroups.
                const int ksteps = min(kBK, pr.K - k0 + 3) / 4;
#pragma unroll
                for (int ks = 0; ks < kBK / 4; ++ks) {
                    if (ks >= ksteps) break;
                    const int kr = ks * 4 + lk;
                    double a[FM], b[FN];
#pragma unroll
                    for (int i = 0; i < FM; ++i) a[i] = As[kr * PA + wm * TM + i * 16 + li];
#pragma unroll
                    for (int j = 0; j < FN; ++j) b[j] = Bs[kr * PB + wn * TN + j * 16 + li];
#pragma unroll
                    for (int i = 0; i < FM; ++i) {
                        if (i * 16 >= mlive) break;
#pragma unroll
                        for (int j = 0; j < FN; ++j) {
                            if (j * 16 >= nlive) break;
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j],
                                                                             0, 0, 0);
                        }
                    }
                }
            }
        }

        // Epilogue: scatter the tile into the destination block.
        gitab_t mtab = (gitab_t)(args.tab + task.mD);
        gitab_t ntab = (gitab_t)(args.tab + task.nD);
        __attribute__((address_space(1))) double *D =
            (__attribute__((address_space(1))) double *)task.D;
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

template <int BM, int BN, int WM, int WN, int BK>
hipError_t launch_cfg(const GemmArgs &args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, WM, WN, BK>), dim3(grid), dim3(WM * WN * 64), 0,
                       stream, args);
    return hipGetLastError();
}

struct CfgShape {
    int bm, bn;
};
constexpr CfgShape kCfg[kGemmConfigs] = {
    {64, 64}, {128, 128}, {64, 128}, {64, 64}, {128, 128}, {128, 64}, {32, 64},
};

} // namespace

int gemm_tile_m(int cfg) { return cfg >= 0 && cfg < kGemmConfigs ? kCfg[cfg].bm : 0; }
int gemm_tile_n(int cfg) { return cfg >= 0 && cfg < kGemmConfigs ? kCfg[cfg].bn : 0; }

hipError_t launch_gemm(int cfg, const GemmArgs &args, int grid, hipStream_t stream) {
    if (args.ntiles <= 0) return hipSuccess;
    switch (cfg) {
    case 0: return launch_cfg<64, 64, 2, 2, 16>(args, grid, stream);
    case 1: return launch_cfg<128, 128, 2, 2, 16>(args, grid, stream);
    case 2: return launch_cfg<64, 128, 2, 2, 16>(args, grid, stream);
    case 3: return launch_cfg<64, 64, 2, 2, 32>(args, grid, stream);
    case 4: return launch_cfg<128, 128, 4, 2, 16>(args, grid, stream);
    case 5: return launch_cfg<128, 64, 2, 2, 16>(args, grid, stream);
    case 6: return launch_cfg<32, 64, 1, 2, 16>(args, grid, stream);
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
