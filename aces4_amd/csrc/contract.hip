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

// The MFMAs of one LDS stage.  Full stages (EDGE = false) carry no tests;
// edge stages issue only the k-steps inside K and the 16x16 blocks inside
// the destination (wave-uniform tests).
template <bool EDGE, int FM, int FN, int kBK, int PA, int PB, int TM, int TN>
__device__ __forceinline__ void stage_mma(const double *As, const double *Bs, f64x4 (&acc)[FM][FN],
                                          int wm, int wn, int li, int lk, int ksteps, int mlive,
                                          int nlive) {
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
        if (EDGE && ks >= ksteps) break;
        const int kr = ks * 4 + lk;
        double a[FM], b[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) a[i] = As[kr * PA + wm * TM + i * 16 + li];
#pragma unroll
        for (int j = 0; j < FN; ++j) b[j] = Bs[kr * PB + wn * TN + j * 16 + li];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            if (EDGE && i * 16 >= mlive) break;
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                if (EDGE && j * 16 >= nlive) break;
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
    }
}

// BM x BN workgroup tile, WM x WN waves, each wave a (BM/WM) x (BN/WN) tile
// of 16x16 MFMA blocks; K staged through LDS kBK at a time.
template <int BM, int BN, int WM, int WN, int kBK, int kWaves = 1>
__global__ __launch_bounds__(WM *WN * 64) __attribute__((amdgpu_waves_per_eu(kWaves)))
void gemm_f64_kernel(GemmArgs args) {
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
        // a single contraction split along K: segment ks of the tile
        const bool split = !args.tiles && args.ksplit > 1;
        const int ks = split ? t / (args.ntiles / args.ksplit) : 0;
        const int tt = t - ks * (split ? args.ntiles / args.ksplit : 0);
        const GemmTile tile = args.tiles ? args.tiles[t]
                                         : GemmTile{0, tt % args.tilesM, tt / args.tilesM, 0};
        if (tile.task < 0) continue; // padding slot of the XCD schedule
        GemmTask task = args.tiles ? args.tasks[tile.task] : args.task0;
        if (ks > 0) {
            task.D = args.ws + (ks - 1) * args.ws_stride;
            task.beta = 0;
        }
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
            GemmPair pr = args.tiles ? args.pairs[p] : args.pair0;
            if (split) { // this segment's k range: the tables from k0 on
                const int k0 = ks * args.kseg;
                pr.kL += k0;
                pr.kR += k0;
                pr.K = min(pr.K - k0, args.kseg);
            }
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
                // inside the destination are issued (wave-uniform tests, taken
                // only on edge stages): an edge wave leaves its SIMD to the
                // other resident workgroups.
                // (one code path with the tests inline keeps this kernel's
                // register budget at four waves per SIMD for 64x64 tiles)
                const int ksteps = min(kBK, pr.K - k0 + 3) / 4;
                stage_mma<true, FM, FN, kBK, PA, PB, TM, TN>(As, Bs, acc, wm, wn, li, lk, ksteps,
                                                             mlive, nlive);
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

// ---------------------------------------------------------------------------
// Double-buffered grouped kernel.
//
// A task's operand pairs are walked as one sequence of K stages (pair p,
// first k k0).  At the top of the iteration for stage s: LDS buffer s%2 holds
// stage s, registers hold the loads of stage s+1 (in flight) and the k-offset
// lookups of stage s+2 (in flight).  The wave runs the MFMAs of stage s, then
// moves stage s+1 into the other LDS buffer, issues the loads of stage s+2
// and the lookups of stage s+3, and one barrier ends the stage.  The M/N
// offset tables of kPairChunk pairs sit in LDS, so a pair switch costs no
// global round trip; full stages issue their MFMAs without edge tests.

constexpr int kPairChunk = 8;

struct Cursor {
    int p, k0, K; // absolute pair index, first k of the stage, the pair's K
};

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <class T> __device__ inline T *uni_ptr(T *p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<T *>(((uint64_t)hi << 32) | lo);
}

template <int BM, int BN, int WM, int WN, int BK>
struct DbKernel {
    static constexpr int NT = WM * WN * 64;
    static constexpr int TM = BM / WM, TN = BN / WN;
    static constexpr int FM = TM / 16, FN = TN / 16;
    static constexpr int PA = BM + 1, PB = BN + 1;
    static constexpr int IA = BK * BM / NT, IB = BK * BN / NT;
    static_assert(FM >= 1 && FN >= 1 && BK % 4 == 0, "bad tile");
    static_assert(IA >= 1 && IB >= 1 && (BK * BM) % NT == 0 && (BK * BN) % NT == 0, "bad split");
    static_assert(NT % BK == 0 && NT % BM == 0 && NT % BN == 0, "bad thread map");

    struct Smem {
        double A[2][BK * PA];
        double B[2][BK * PB];
        int xA[kPairChunk][BM];
        int xB[kPairChunk][BN];
    };

    // Per-thread loader state for one operand (X = BM or BN, I items).
    template <int X, int I>
    struct Side {
        unsigned xmask;  // item i inside the block (kfast), all items (xfast)
        int ko[I];       // k offsets of the looked-up stage
        bool kf_look;    // lane order of the looked-up stage
        double v[I];     // loads of the issued stage
        bool kf_issue;

        __device__ void init_mask(int x0, int Xn, int tid, bool any_kfast) {
            // x validity for both lane orders, resolved at use time
            xmask = 0;
#pragma unroll
            for (int i = 0; i < I; ++i)
                xmask |= ((x0 + tid / BK + i * (NT / BK)) < Xn ? 1u : 0u) << i;
            xmask |= ((x0 + tid % X) < Xn ? 1u : 0u) << 31;
            (void)any_kfast;
        }
        __device__ void lookup(bool kf, gitab_t ktab, int K, int k0, int tid) {
            kf_look = kf;
            if (kf) {
                ko[0] = ktab[min(k0 + tid % BK, K - 1)];
            } else {
#pragma unroll
                for (int i = 0; i < I; ++i) ko[i] = ktab[min(k0 + tid / X + i * (NT / X), K - 1)];
            }
        }
        __device__ void issue(gptr_t src, const int *xs, int tid) {
            kf_issue = kf_look;
            if (kf_look) {
#pragma unroll
                for (int i = 0; i < I; ++i) v[i] = src[ko[0] + xs[tid / BK + i * (NT / BK)]];
            } else {
                const int xo = xs[tid % X];
#pragma unroll
                for (int i = 0; i < I; ++i) v[i] = src[ko[i] + xo];
            }
        }
        __device__ void store(double *S, int pitch, int K, int k0, int tid) const {
            if (kf_issue) {
                const bool kin = k0 + tid % BK < K;
#pragma unroll
                for (int i = 0; i < I; ++i) {
                    const bool ok = kin && ((xmask >> i) & 1u);
                    S[(tid % BK) * pitch + tid / BK + i * (NT / BK)] = ok ? v[i] : 0.0;
                }
            } else {
                const bool xin = (xmask >> 31) & 1u;
#pragma unroll
                for (int i = 0; i < I; ++i) {
                    const int kk = tid / X + i * (NT / X);
                    S[kk * pitch + tid % X] = (xin && k0 + kk < K) ? v[i] : 0.0;
                }
            }
        }
    };

    template <bool EDGE>
    static __device__ void mma(const double *A, const double *B, f64x4 (&acc)[FM][FN], int wm,
                               int wn, int li, int lk, int ksteps, int mlive, int nlive) {
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            if (EDGE && ks >= ksteps) break;
            const int kr = ks * 4 + lk;
            double a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = A[kr * PA + wm * TM + i * 16 + li];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = B[kr * PB + wn * TN + j * 16 + li];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                if (EDGE && i * 16 >= mlive) break;
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if (EDGE && j * 16 >= nlive) break;
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                }
            }
        }
    }
};

template <int BM, int BN, int WM, int WN, int BK>
__global__ __launch_bounds__(WM *WN * 64) void gemm_f64_db_kernel(GemmArgs args) {
    using Kc = DbKernel<BM, BN, WM, WN, BK>;
    constexpr int NT = Kc::NT, TM = Kc::TM, TN = Kc::TN, FM = Kc::FM, FN = Kc::FN;
    __shared__ typename Kc::Smem sm;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    typename Kc::template Side<BM, Kc::IA> sa;
    typename Kc::template Side<BN, Kc::IB> sb;

    for (int t = blockIdx.x; t < args.ntiles; t += gridDim.x) {
        const GemmTile tile = args.tiles ? args.tiles[t]
                                         : GemmTile{0, t % args.tilesM, t / args.tilesM, 0};
        if (tile.task < 0) continue;
        const GemmTask task = args.tiles ? args.tasks[tile.task] : args.task0;
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        const int mlive = __builtin_amdgcn_readfirstlane(task.M - (m0 + wm * TM));
        const int nlive = __builtin_amdgcn_readfirstlane(task.N - (n0 + wn * TN));
        const int pend = task.first + task.npairs;
        // pair fields are wave-uniform: scalar loads, scalar branches
        auto pair_at = [&](int p) -> GemmPair {
            if (!args.tiles) return args.pair0;
            const GemmPair *pp = args.pairs + uni(p);
            GemmPair r;
            r.L = uni_ptr(pp->L);
            r.R = uni_ptr(pp->R);
            r.K = uni(pp->K);
            r.kL = uni(pp->kL);
            r.mL = uni(pp->mL);
            r.kR = uni(pp->kR);
            r.nR = uni(pp->nR);
            r.flags = uni(pp->flags);
            return r;
        };
        auto fill_chunk = [&](int c0) {
            const int cn = min(kPairChunk, pend - c0);
            // per-thread pair index: plain (vector) loads of the pair records
            for (int e = tid; e < cn * BM; e += NT) {
                const int mL = args.tiles ? args.pairs[c0 + e / BM].mL : args.pair0.mL;
                const int x = min(m0 + e % BM, task.M - 1);
                sm.xA[e / BM][e % BM] = args.tab[mL + x];
            }
            for (int e = tid; e < cn * BN; e += NT) {
                const int nR = args.tiles ? args.pairs[c0 + e / BN].nR : args.pair0.nR;
                const int x = min(n0 + e % BN, task.N - 1);
                sm.xB[e / BN][e % BN] = args.tab[nR + x];
            }
        };
        auto at_pair = [&](int p) -> Cursor { return Cursor{uni(p), 0, p < pend ? pair_at(p).K : 0}; };
        auto advance = [&](Cursor c) -> Cursor {
            c.k0 = uni(c.k0 + BK);
            if (c.k0 >= c.K) return at_pair(c.p + 1);
            return c;
        };
        auto lookup = [&](Cursor c) {
            const GemmPair pr = pair_at(c.p);
            sa.lookup(pr.flags & 1, (gitab_t)(args.tab + pr.kL), pr.K, c.k0, tid);
            sb.lookup(pr.flags & 2, (gitab_t)(args.tab + pr.kR), pr.K, c.k0, tid);
        };
        int chunk0 = task.first;
        auto issue = [&](Cursor c) {
            if (c.p >= chunk0 + kPairChunk) { // next chunk of pair tables
                __syncthreads();
                chunk0 = c.p;
                fill_chunk(chunk0);
                __syncthreads();
            }
            const GemmPair pr = pair_at(c.p);
            sa.issue((gptr_t)pr.L, sm.xA[c.p - chunk0], tid);
            sb.issue((gptr_t)pr.R, sm.xB[c.p - chunk0], tid);
        };
        auto store = [&](Cursor c, int buf) {
            sa.store(sm.A[buf], Kc::PA, c.K, c.k0, tid);
            sb.store(sm.B[buf], Kc::PB, c.K, c.k0, tid);
        };

        f64x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

        sa.init_mask(m0, task.M, tid, true);
        sb.init_mask(n0, task.N, tid, true);
        __syncthreads(); // previous tile done with the pair tables
        fill_chunk(chunk0);
        __syncthreads();

        // prologue: stage 0 to LDS, stage 1 loads and stage 2 lookups in flight
        Cursor C = at_pair(task.first);
        lookup(C);
        issue(C);
        store(C, 0);
        Cursor L = advance(C);
        if (L.p < pend) { lookup(L); issue(L); }
        Cursor Kn = L.p < pend ? advance(L) : L;
        if (Kn.p < pend) lookup(Kn);
        __syncthreads();

        int buf = 0;
        while (C.p < pend) {
            const int ksteps = uni(min(BK, C.K - C.k0 + 3) / 4);
            if (ksteps == BK / 4 && mlive >= TM && nlive >= TN)
                Kc::template mma<false>(sm.A[buf], sm.B[buf], acc, wm, wn, li, lk, ksteps, mlive,
                                        nlive);
            else
                Kc::template mma<true>(sm.A[buf], sm.B[buf], acc, wm, wn, li, lk, ksteps, mlive,
                                       nlive);
            if (L.p < pend) store(L, buf ^ 1);
            if (Kn.p < pend) issue(Kn);
            const Cursor K2 = Kn.p < pend ? advance(Kn) : Kn;
            if (K2.p < pend) lookup(K2);
            __syncthreads();
            C = L;
            L = Kn;
            Kn = K2;
            buf ^= 1;
        }

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
    }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS kernel.
//
// Operand tiles go from HBM straight into LDS (global_load_lds_dword): no
// staging registers, no LDS store instructions.  One wave instruction gathers
// 32 doubles of one k-row of the tile (lane 2j / 2j+1 move the low / high
// word of double j), through the offset tables: the row's k offset is
// wave-uniform (a scalar load), the lane's x offset is per-pair state.  Lanes
// whose x is outside the block are masked off and their LDS slots, zeroed once
// per tile, stay zero; rows past the pair's K are zeroed with plain LDS stores
// on the pair's last stage.  Two LDS buffers: the gathers of stage s+1 are in
// flight while stage s runs on the MFMA pipe; one barrier per stage.

template <int BM, int BN, int WM, int WN, int BK>
__global__ __launch_bounds__(WM *WN * 64) void gemm_f64_lds_kernel(GemmArgs args) {
    constexpr int NT = WM * WN * 64, NW = WM * WN;
    constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
    constexpr int PA = BM + 1, PB = BN + 1;
    constexpr int CA = BM / 32, CB = BN / 32;    // 32-double chunks per row
    constexpr int RW = BK / NW;                  // rows per wave per operand
    static_assert(BM % 32 == 0 && BN % 32 == 0 && BK % NW == 0 && BK % 4 == 0, "bad tile");
    __shared__ double As[2][BK * PA];
    __shared__ double Bs[2][BK * PB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM, wn = wave / WM;
    const int li = lane & 15, lk = lane >> 4;
    const int half = (lane & 1) * 4, xl = lane >> 1;

    for (int t = blockIdx.x; t < args.ntiles; t += gridDim.x) {
        const GemmTile tile = args.tiles ? args.tiles[t]
                                         : GemmTile{0, t % args.tilesM, t / args.tilesM, 0};
        if (tile.task < 0) continue;
        const GemmTask task = args.tiles ? args.tasks[tile.task] : args.task0;
        const int m0 = tile.mt * BM, n0 = tile.nt * BN;
        const int mlive = __builtin_amdgcn_readfirstlane(task.M - (m0 + wm * TM));
        const int nlive = __builtin_amdgcn_readfirstlane(task.N - (n0 + wn * TN));

        // zero both buffers once: masked-off columns then stay zero
        __syncthreads();
        for (int e = tid; e < 2 * BK * PA; e += NT) (&As[0][0])[e] = 0.0;
        for (int e = tid; e < 2 * BK * PB; e += NT) (&Bs[0][0])[e] = 0.0;

        f64x4 acc[FM][FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

        for (int p = task.first; p < task.first + task.npairs; ++p) {
            const GemmPair pr = args.tiles ? args.pairs[p] : args.pair0;
            // lane byte offsets of the tile's x columns (or -1 outside the block)
            int xa[CA], xb[CB];
#pragma unroll
            for (int c = 0; c < CA; ++c) {
                const int x = m0 + c * 32 + xl;
                xa[c] = x < task.M ? args.tab[pr.mL + x] * 8 + half : -1;
            }
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                const int x = n0 + c * 32 + xl;
                xb[c] = x < task.N ? args.tab[pr.nR + x] * 8 + half : -1;
            }
            const char *srcA = reinterpret_cast<const char *>(pr.L);
            const char *srcB = reinterpret_cast<const char *>(pr.R);
            // the tables are never written by the kernel: constant address
            // space, so wave-uniform lookups become scalar loads
            const __attribute__((address_space(4))) int32_t *ktA =
                (const __attribute__((address_space(4))) int32_t *)(args.tab + pr.kL);
            const __attribute__((address_space(4))) int32_t *ktB =
                (const __attribute__((address_space(4))) int32_t *)(args.tab + pr.kR);
            const int K = pr.K;

            // gather stage k0 into buffer b
            auto gather = [&](int k0, int b) {
#pragma unroll
                for (int q = 0; q < RW; ++q) {
                    const int r = wave + q * NW; // tile row (wave-uniform)
                    const int k = k0 + r;
                    if (k < K) {
                        const char *ra = srcA + (size_t)ktA[k] * 8;
                        const char *rb = srcB + (size_t)ktB[k] * 8;
#pragma unroll
                        for (int c = 0; c < CA; ++c)
                            if (xa[c] >= 0)
                                __builtin_amdgcn_global_load_lds(
                                    (const __attribute__((address_space(1))) void *)(ra + (unsigned)xa[c]),
                                    (__attribute__((address_space(3))) void *)(&As[b][r * PA + c * 32]),
                                    4, 0, 0);
#pragma unroll
                        for (int c = 0; c < CB; ++c)
                            if (xb[c] >= 0)
                                __builtin_amdgcn_global_load_lds(
                                    (const __attribute__((address_space(1))) void *)(rb + (unsigned)xb[c]),
                                    (__attribute__((address_space(3))) void *)(&Bs[b][r * PB + c * 32]),
                                    4, 0, 0);
                    } else {
                        // past the pair's K: this row contributes zero
                        for (int e = lane; e < BM; e += 64) As[b][r * PA + e] = 0.0;
                        for (int e = lane; e < BN; e += 64) Bs[b][r * PB + e] = 0.0;
                    }
                }
            };

            __syncthreads(); // previous pair's last stage consumed, zeroing done
            gather(0, 0);
            __builtin_amdgcn_s_waitcnt(0); // vmcnt(0) lgkmcnt(0) expcnt(0)
            __syncthreads();
            int b = 0;
            for (int k0 = 0; k0 < K; k0 += BK) {
                if (k0 + BK < K) gather(k0 + BK, b ^ 1);
                const int ksteps = min(BK, K - k0 + 3) / 4;
                if (ksteps == BK / 4 && mlive >= TM && nlive >= TN)
                    stage_mma<false, FM, FN, BK, PA, PB, TM, TN>(As[b], Bs[b], acc, wm, wn, li, lk,
                                                                 ksteps, mlive, nlive);
                else
                    stage_mma<true, FM, FN, BK, PA, PB, TM, TN>(As[b], Bs[b], acc, wm, wn, li, lk,
                                                                ksteps, mlive, nlive);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
                b ^= 1;
            }
        }

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

template <int BM, int BN, int WM, int WN, int BK, int W = 1>
hipError_t launch_cfg(const GemmArgs &args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, WM, WN, BK, W>), dim3(grid), dim3(WM * WN * 64), 0,
                       stream, args);
    return hipGetLastError();
}

template <int BM, int BN, int WM, int WN, int BK>
hipError_t launch_lds(const GemmArgs &args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((gemm_f64_lds_kernel<BM, BN, WM, WN, BK>), dim3(grid), dim3(WM * WN * 64),
                       0, stream, args);
    return hipGetLastError();
}

template <int BM, int BN, int WM, int WN, int BK>
hipError_t launch_db(const GemmArgs &args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL((gemm_f64_db_kernel<BM, BN, WM, WN, BK>), dim3(grid), dim3(WM * WN * 64),
                       0, stream, args);
    return hipGetLastError();
}

struct CfgShape {
    int bm, bn;
};
constexpr CfgShape kCfg[kGemmConfigs] = {
    {64, 64}, {128, 128}, {64, 128}, {64, 64}, {128, 128}, {128, 64}, {32, 64},
    {128, 128}, {128, 128}, {128, 64}, {64, 64},
    {64, 64}, {128, 64}, {128, 128},
    {64, 64}, {128, 64}, {128, 128}, {64, 128},
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
    case 7: return launch_db<128, 128, 2, 2, 8>(args, grid, stream);
    case 8: return launch_db<128, 128, 2, 2, 16>(args, grid, stream);
    case 9: return launch_db<128, 64, 2, 2, 16>(args, grid, stream);
    case 10: return launch_db<64, 64, 2, 2, 16>(args, grid, stream);
    case 11: return launch_cfg<64, 64, 2, 2, 16, 4>(args, grid, stream);
    case 12: return launch_cfg<128, 64, 2, 2, 16, 2>(args, grid, stream);
    case 13: return launch_cfg<128, 128, 2, 2, 16, 2>(args, grid, stream);
    case 14: return launch_lds<64, 64, 2, 2, 16>(args, grid, stream);
    case 15: return launch_lds<128, 64, 2, 2, 16>(args, grid, stream);
    case 16: return launch_lds<128, 128, 2, 2, 16>(args, grid, stream);
    case 17: return launch_lds<64, 128, 2, 2, 16>(args, grid, stream);
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
