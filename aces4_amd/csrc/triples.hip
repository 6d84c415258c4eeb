// triples.hip -- the (T) triples energy with W kept on chip.
//
// W(a,b,c,i,j,k) = sum over the 6 permutations P of the pairs (a,i), (b,j),
// (c,k) -- written (x1,y1), (x2,y2), (x3,y3) after permuting -- of
//     sum_d V1(x2,x3,d,y3) T(x1,d,y1,y2)  +  sum_l T(x3,x2,y3,l) V2(l,x1,y2,y1)
// (V2 holds -V), and E += factor * sum (W / D) * W with
// D = e_i + e_j + e_k - e_a - e_b - e_c summed in that order
// (rccsdpt_aaa.sialx:2164-2174, energy_denominator_rhf.F:399-444).
//
// For every P both terms split W's six indices into the same row set
// {x1, y1, y2} and column set {x2, x3, y3}, so a P is one GEMM with
// K = v + o (d, then l).  A workgroup owns a box of W of extent 5 in every
// index (15625 doubles, held in LDS): it runs the six GEMMs of the box
// (125 x 125 tiles, padded to 128) on FP64 MFMA, adds each product into
// the box in P order (the first assigns), then divides, squares and sums
// the box into one partial energy.  W is never written to HBM.
//
// The operands come from two packed panel arrays built once from T, V1 and
// V2: PA over row sets (x, y1, y2) of 5-blocks, PB over column sets
// (x2, x3, y3).  Within a set, rows are x + 5 (y1 + 5 y2) (padded to 128,
// zero rows), and each 16-row block holds its K in MFMA operand order:
// element (row 16 rb + l % 16, k = 8 kp + 4 h + l / 16) sits at
// ((set * 8 + rb) * NKP + kp) * 128 + 2 l + h, so a wave reads the operands
// of two MFMA k-steps with one 16-byte load per lane, 1 KiB contiguous per
// wave.  Every P reads the same two arrays: P only changes which box
// coordinates select the row and column sets.
#include "kernels.h"

namespace sipgpu {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kBoxE = kTriplesBox;                 // box extent per index
constexpr int kBoxRows = kBoxE * kBoxE * kBoxE;     // 125 rows / columns per P
constexpr int kBoxElems = kBoxRows * kBoxRows;      // 15625
constexpr int kThreads = 512;

// the permutations P of (0, 1, 2), in the order W sums them
__constant__ int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// box digit strides in LDS: a fastest, then b, c, i, j, k
__device__ __forceinline__ int box_stride(int d) {
    return d == 0 ? 1 : d == 1 ? 5 : d == 2 ? 25 : d == 3 ? 125 : d == 4 ? 625 : 3125;
}

// ---------------------------------------------------------------- pack --

// One thread per panel double; rows and K positions outside the blocks'
// extents are zeros.  Reads gather from the blocks (a one-time pass).
__global__ __launch_bounds__(256) void triples_pack_kernel(TriplesPackArgs a) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long per_set = (long long)8 * a.nkp * 128;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < a.total;
         e += stride) {
        const long long set = e / per_set;
        const int rem = (int)(e - set * per_set);
        const int rb = rem / (a.nkp * 128);
        const int r2 = rem - rb * (a.nkp * 128);
        const int kp = r2 >> 7;
        const int lh = r2 & 127;
        const int lane = lh >> 1, h = lh & 1;
        const int row = rb * 16 + (lane & 15);
        const int k = kp * 8 + h * 4 + (lane >> 4);
        double v = 0.0;
        if (row < kBoxRows && k < a.kd + a.kl) {
            const int r0 = row % 5, r1 = (row / 5) % 5, r2d = row / 25;
            // set = (s0 * n1 + s1) * n2 + s2 over the set's three 5-block lists
            const int s2 = (int)(set % a.n2);
            const int s1 = (int)((set / a.n2) % a.n1);
            const int s0 = (int)(set / ((long long)a.n2 * a.n1));
            const TriplesBlk b0 = a.blk0[s0], b1 = a.blk1[s1], b2 = a.blk2[s2];
            if (r0 < b0.ext && r1 < b1.ext && r2d < b2.ext) {
                // global indices (segment, offset within segment) of the row
                const int g0s = b0.seg, g0o = b0.off + r0;
                const int g1s = b1.seg, g1o = b1.off + r1;
                const int g2s = b2.seg, g2o = b2.off + r2d;
                if (a.side == 0) {
                    // PA row (x, y1, y2)
                    if (k < a.kd) { // T(x, d, y1, y2)
                        const int ds = a.vseg_of[k], doff = k - a.vseg_start[ds];
                        const double *blk = a.T[((g0s * a.nv + ds) * a.no + g1s) * a.no + g2s];
                        const int e0 = a.vext[g0s], e1 = a.vext[ds], e2 = a.oext[g1s];
                        v = blk[g0o + e0 * (doff + e1 * (g1o + e2 * g2o))];
                    } else { // V2(l, x, y2, y1)
                        const int l = k - a.kd;
                        const int ls = a.oseg_of[l], loff = l - a.oseg_start[ls];
                        const double *blk = a.V2[((ls * a.nv + g0s) * a.no + g2s) * a.no + g1s];
                        const int e0 = a.oext[ls], e1 = a.vext[g0s], e2 = a.oext[g2s];
                        v = blk[loff + e0 * (g0o + e1 * (g2o + e2 * g1o))];
                    }
                } else {
                    // PB column (x2, x3, y3)
                    if (k < a.kd) { // V1(x2, x3, d, y3)
                        const int ds = a.vseg_of[k], doff = k - a.vseg_start[ds];
                        const double *blk = a.V1[((g0s * a.nv + g1s) * a.nv + ds) * a.no + g2s];
                        const int e0 = a.vext[g0s], e1 = a.vext[g1s], e2 = a.vext[ds];
                        v = blk[g0o + e0 * (g1o + e1 * (doff + e2 * g2o))];
                    } else { // T(x3, x2, y3, l)
                        const int l = k - a.kd;
                        const int ls = a.oseg_of[l], loff = l - a.oseg_start[ls];
                        const double *blk = a.T[((g1s * a.nv + g0s) * a.no + g2s) * a.no + ls];
                        const int e0 = a.vext[g1s], e1 = a.vext[g0s], e2 = a.oext[g2s];
                        v = blk[g1o + e0 * (g0o + e1 * (g2o + e2 * loff))];
                    }
                }
            }
        }
        a.out[e] = v;
    }
}

// ---------------------------------------------------------------- boxes --

// Box coordinates (5-block indices of a, b, c, i, j, k) of this workgroup;
// false for a workgroup without a box.  Global mode: linear ids in
// 2x2x2x2x2x2 super-boxes (consecutive ids are neighbours in all six
// indices), workgroup w on XCD w % 8 taking ids in order within its XCD's
// share, so the boxes running together on one XCD share operand sets in its
// L2.  Tuple mode: the boxes of one tuple, first index fastest.
// Returns 0 (no slot: past the launch's range), 1 (a slot whose box lies in
// the padding of the super-box grid: its partial is zero) or 2 (a box).
__device__ __forceinline__ int box_of(const TriplesArgs &a, int (&bc)[6], long long &slot) {
    if (a.tuple_mode) {
        long long q = blockIdx.x;
        slot = q;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            bc[d] = a.first[d] + (int)(q % a.count[d]);
            q /= a.count[d];
        }
        return 2;
    }
    const long long w = blockIdx.x;
    const long long local = (w & 7) * a.per_xcd + (w >> 3);
    if (local >= a.nbox) return 0;
    const long long L = a.box0 + local;
    slot = L;
    const int q = (int)(L & 63);
    long long r = L >> 6;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int sd = (int)(r % a.super[d]);
        r /= a.super[d];
        bc[d] = 2 * sd + ((q >> d) & 1);
        ok = ok && bc[d] < a.nblk[d];
    }
    return ok ? 2 : 1;
}

// One workgroup: one 5^6 box of W, all six GEMMs, the energy partial.
// Eight waves in 2 x 4, each a 64 x 32 tile (4 x 2 MFMA blocks): two waves
// per SIMD, so one issues MFMAs while the other waits for its operands.
__global__ __launch_bounds__(kThreads) void triples_box_kernel(TriplesArgs a) {
    __shared__ double W[kBoxElems];
    __shared__ double eps[6][kBoxE];
    __shared__ int ext[6];
    __shared__ double red[kThreads / 64];
    __shared__ double occ[kBoxRows];
    // box offsets of product row / column x (< 128) under permutation g
    __shared__ int rofs[6][128], cofs[6][128];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int bc[6];
    long long slot = 0;
    const int kind = box_of(a, bc, slot);
    if (kind < 2) {
        if (kind == 1 && tid == 0) a.partial[slot] = 0.0;
        return;
    }
    if (tid < 6) {
        const TriplesBlk b = tid < 3 ? a.vblk[bc[tid]] : a.oblk[bc[tid]];
        ext[tid] = b.ext;
    }
    if (tid < 30) {
        const int d = tid / kBoxE, q = tid % kBoxE;
        const TriplesBlk b = d < 3 ? a.vblk[bc[d]] : a.oblk[bc[d]];
        const double *e = d < 3 ? a.e_vir : a.e_occ;
        eps[d][q] = q < b.ext ? e[b.gstart + q] : 0.0;
    }

    for (int t = tid; t < 6 * 128; t += kThreads) {
        const int g = t >> 7, x = t & 127;
        const int p0 = kPerm[g][0], p1 = kPerm[g][1], p2 = kPerm[g][2];
        const int d0 = x % kBoxE, d1 = (x / kBoxE) % kBoxE, d2 = x / (kBoxE * kBoxE);
        // row (x1, y1, y2), column (x2, x3, y3)
        rofs[g][x] = d0 * box_stride(p0) + d1 * box_stride(3 + p0) + d2 * box_stride(3 + p1);
        cofs[g][x] = d0 * box_stride(p1) + d1 * box_stride(p2) + d2 * box_stride(3 + p2);
    }

    const int nkp = a.nkp;
    // operand pointers of permutation g: its row set (x1, y1, y2) in PA and
    // column set (x2, x3, y3) in PB, at this wave's first 16-row block
    auto panels = [&](int g, const f64x2 *&pa, const f64x2 *&pb) {
        const int p0 = kPerm[g][0], p1 = kPerm[g][1], p2 = kPerm[g][2];
        const long long rowset = ((long long)bc[p0] * a.nob + bc[3 + p0]) * a.nob + bc[3 + p1];
        const long long colset = ((long long)bc[p1] * a.nvb + bc[p2]) * a.nob + bc[3 + p2];
        pa = reinterpret_cast<const f64x2 *>(a.PA) + ((rowset * 8 + wm * 4) * nkp) * 64 + lane;
        pb = reinterpret_cast<const f64x2 *>(a.PB) + ((colset * 8 + wn * 2) * nkp) * 64 + lane;
    };
    f64x2 fa[3][4], fb[3][2];
    auto load = [&](int st, const f64x2 *pa, const f64x2 *pb, int kp) {
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[st][i] = pa[((long long)i * nkp + kp) * 64];
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[st][j] = pb[((long long)j * nkp + kp) * 64];
    };
    // The k-pairs of the six permutations form one stream; loads run two
    // k-pairs ahead of the MFMAs through three register sets, across the
    // permutation boundaries.  Load cursor: permutation lg, k-pair lkp
    // (past the last k-pair of the last permutation it reloads that one,
    // which nothing reads).
    int lg = 0, lkp = 0;
    const f64x2 *lpa, *lpb;
    panels(0, lpa, lpb);
    auto load_next = [&](int st) {
        load(st, lpa, lpb, lkp);
        if (++lkp == nkp) {
            if (lg < 5) {
                lkp = 0;
                ++lg;
                panels(lg, lpa, lpb);
            } else {
                lkp = nkp - 1;
            }
        }
    };
    f64x4 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    auto mma = [&](int st) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[st][i][h], fb[st][j][h],
                                                                     acc[i][j], 0, 0, 0);
    };
    // add permutation g's product into the box: row (x1, y1, y2), column
    // (x2, x3, y3); the first permutation assigns
    auto scatter = [&](int g) {
        // the previous P's adds into the box are complete
        __syncthreads();
        int coff[2];
        bool cok[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int c = (wn * 2 + j) * 16 + (lane & 15);
            cok[j] = c < kBoxRows;
            coff[j] = cofs[g][c];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (wm * 4 + i) * 16 + 4 * r + (lane >> 4);
                if (row >= kBoxRows) continue;
                const int roff = rofs[g][row];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (!cok[j]) continue;
                    double &w = W[roff + coff[j]];
                    w = g == 0 ? acc[i][j][r] : w + acc[i][j][r];
                }
            }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    };
    // compute cursor: permutation cg, k-pairs done ckp
    int cg = 0, ckp = 0;
    auto step = [&](int st) {
        mma(st);
        if (++ckp == nkp) {
            scatter(cg);
            ckp = 0;
            ++cg;
        }
    };
    load_next(0);
    load_next(1);
    const int total = 6 * nkp;
    for (int t = 0; t < total; t += 3) {
        load_next(2);
        step(0);
        if (t + 1 < total) {
            load_next(0);
            step(1);
        }
        if (t + 2 < total) {
            load_next(1);
            step(2);
        }
    }
    // energy of the box: sum over its elements inside the blocks' extents.
    // Element e = abc + 125 ijk (abc = a + 5 b + 25 c, ijk = i + 5 j + 25 k).
    // D = ((((e_i + e_j) + e_k) - e_a) - e_b) - e_c: the occupied sums of the
    // box's 125 (i, j, k) are formed once (NaN marks an (i, j, k) outside the
    // extents); a thread keeps one abc, its virtual energies in registers,
    // and runs over ijk.
    double sum = 0.0;
    if (!a.wout) {
        if (tid < kBoxRows) {
            const int i = tid % kBoxE, j = (tid / kBoxE) % kBoxE, k = tid / (kBoxE * kBoxE);
            const bool in = i < ext[3] && j < ext[4] && k < ext[5];
            occ[tid] = in ? eps[3][i] + eps[4][j] + eps[5][k] : __builtin_nan("");
        }
        __syncthreads(); // the box and the occupied sums are complete
        const int abc = tid & 127;
        const int ea_i = abc % kBoxE, eb_i = (abc / kBoxE) % kBoxE, ec_i = abc / (kBoxE * kBoxE);
        if (abc < kBoxRows && ea_i < ext[0] && eb_i < ext[1] && ec_i < ext[2]) {
            const double ea = eps[0][ea_i], eb = eps[1][eb_i], ec = eps[2][ec_i];
            for (int ijk = tid >> 7; ijk < kBoxRows; ijk += kThreads / 128) {
                const double o = occ[ijk];
                if (o != o) continue; // (i, j, k) outside the extents
                const double w = W[abc + kBoxRows * ijk];
                const double D = ((o - ea) - eb) - ec;
                sum += (w / D) * w;
            }
        }
    }
    __syncthreads();

    // tuple mode with W written out: every element's global position
    for (int e = a.wout ? tid : kBoxElems; e < kBoxElems; e += kThreads) {
        int q = e, dg[6];
        bool in = true;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            dg[d] = q % kBoxE;
            q /= kBoxE;
            in = in && dg[d] < ext[d];
        }
        if (!in) continue;
        const double w = W[e];
        // occupied energies in index order, then the virtual ones subtracted
        const double D = eps[3][dg[3]] + eps[4][dg[4]] + eps[5][dg[5]] - eps[0][dg[0]] -
                         eps[1][dg[1]] - eps[2][dg[2]];
        sum += (w / D) * w;
        if (a.wout) {
            long long o = 0;
#pragma unroll
            for (int d = 5; d >= 0; --d)
                o = o * a.wdims[d] + ((bc[d] - a.first[d]) * kBoxE + dg[d]);
            a.wout[o] = w;
        }
    }
    // fixed-order reduction: lanes by xor butterfly, then waves in order
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) sum += __shfl_xor(sum, m);
    if (lane == 0) red[wave] = sum;
    __syncthreads();
    if (tid == 0) {
        double t = red[0];
        for (int w = 1; w < kThreads / 64; ++w) t += red[w];
        a.partial[slot] = t;
    }
}

// out[0] += factor * sum of partial[0 .. n) in index order: per-workgroup
// sums of contiguous ranges, then one thread over them.
__global__ __launch_bounds__(256) void triples_reduce_kernel(const double *partial, long long n,
                                                             long long chunk, double *sums) {
    __shared__ double red[4];
    const long long b0 = (long long)blockIdx.x * chunk;
    const long long b1 = b0 + chunk < n ? b0 + chunk : n;
    double s = 0.0;
    for (long long e = b0 + threadIdx.x; e < b1; e += blockDim.x) s += partial[e];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void triples_final_kernel(const double *sums, int n, double factor, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += sums[i];
    out[0] += factor * t;
}

} // namespace

hipError_t launch_triples_pack(const TriplesPackArgs &a, hipStream_t s) {
    if (a.total <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((a.total + 255) / 256, 65536);
    triples_pack_kernel<<<(int)blocks, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_triples_boxes(const TriplesArgs &a, long long grid, hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    // a dispatch holds at most 2^32 - 1 work-items per dimension
    if (grid * kThreads >= (1LL << 32)) return hipErrorInvalidConfiguration;
    triples_box_kernel<<<(unsigned)grid, kThreads, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_triples_reduce(const double *partial, long long n, double *sums, double factor,
                                 double *out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long chunk = (n + kTriplesSums - 1) / kTriplesSums;
    const int blocks = (int)((n + chunk - 1) / chunk);
    triples_reduce_kernel<<<blocks, 256, 0, s>>>(partial, n, chunk, sums);
    triples_final_kernel<<<1, 64, 0, s>>>(sums, blocks, factor, out);
    return hipGetLastError();
}

} // namespace sipgpu
