// energy.hip -- orbital-energy denominators and energy reductions of
// coupled-cluster blocks.
//
// For a block x of rank r (column-major, extents dims) with one orbital
// energy vector per index (eps_d, the energies of that index's segment) and
// a sign per index (+1 occupied, -1 virtual), the denominator of element
// (i_0 .. i_{r-1}) is the sum of the occupied energies in index order minus
// the virtual ones in index order (the order energy_denominator_rhf.F
// do_rhf_den2/4/6 add them in).  Kernels: x /= D in place, and the fixed-
// order reduction E += factor * sum (x / D) * y, whose per-workgroup
// partials are summed by one thread in workgroup order (deterministic).
#include "kernels.h"

namespace sipgpu {

namespace {

constexpr int kEnergyThreads = 256;

// Mixed-radix digits (index values) of element e, first index fastest.
__device__ __forceinline__ void decode(const EnergyArgs &a, long long e, int (&idx)[kEnergyMaxRank]) {
#pragma unroll
    for (int d = 0; d < kEnergyMaxRank; ++d) {
        idx[d] = 0;
        if (d < a.rank) {
            idx[d] = (int)(e % a.dims[d]);
            e /= a.dims[d];
        }
    }
}

// idx += step in mixed radix (both digit vectors are in range, so one
// subtraction per digit is enough); a carry out of the last digit only
// happens past the end of the block, where the loops stop.
__device__ __forceinline__ void advance(const EnergyArgs &a, int (&idx)[kEnergyMaxRank],
                                        const int (&step)[kEnergyMaxRank]) {
    int carry = 0;
#pragma unroll
    for (int d = 0; d < kEnergyMaxRank; ++d)
        if (d < a.rank) {
            int v = idx[d] + step[d] + carry;
            carry = v >= a.dims[d];
            idx[d] = carry ? v - a.dims[d] : v;
        }
}

__device__ __forceinline__ double denominator(const EnergyArgs &a,
                                              const int (&idx)[kEnergyMaxRank]) {
    double v[kEnergyMaxRank];
#pragma unroll
    for (int d = 0; d < kEnergyMaxRank; ++d) v[d] = d < a.rank ? a.eps[d][idx[d]] : 0.0;
    // occupied energies added in index order, then each virtual one
    // subtracted in index order: eps = e_b + e_d + e_f - e_a - e_c - e_e
    double D = 0.0;
    bool first = true;
#pragma unroll
    for (int d = 0; d < kEnergyMaxRank; ++d)
        if (d < a.rank && a.sign[d] > 0) {
            D = first ? v[d] : D + v[d];
            first = false;
        }
#pragma unroll
    for (int d = 0; d < kEnergyMaxRank; ++d)
        if (d < a.rank && a.sign[d] < 0) {
            D = first ? -v[d] : D - v[d];
            first = false;
        }
    return D;
}

// Grid-stride loops keep each thread's element index as digits: decoded
// once, then advanced by the grid stride (no divisions per element).
__global__ __launch_bounds__(kEnergyThreads) void denominator_divide_kernel(EnergyArgs a,
                                                                            double *x) {
    const long long e0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long S = (long long)gridDim.x * blockDim.x;
    int idx[kEnergyMaxRank], step[kEnergyMaxRank];
    decode(a, e0, idx);
    decode(a, S, step);
    for (long long e = e0; e < a.n; e += S) {
        x[e] = x[e] / denominator(a, idx);
        advance(a, idx, step);
    }
}

// Each thread keeps the loads of its next kEnergyAhead elements in flight (a
// queue rotated every iteration), so the latency of one element's load hides
// behind the arithmetic of the previous ones.  With y == x each element is
// loaded once.
constexpr int kEnergyAhead = 4;

__global__ __launch_bounds__(kEnergyThreads) void energy_partial_kernel(EnergyArgs a,
                                                                        const double *x,
                                                                        const double *y,
                                                                        double *part) {
    __shared__ double red[kEnergyThreads];
    const long long e0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long S = (long long)gridDim.x * blockDim.x;
    const bool same = x == y;
    int idx[kEnergyMaxRank], step[kEnergyMaxRank];
    decode(a, e0, idx);
    decode(a, S, step);
    double qx[kEnergyAhead], qy[kEnergyAhead];
#pragma unroll
    for (int u = 0; u < kEnergyAhead; ++u) {
        const long long eu = e0 + u * S;
        qx[u] = eu < a.n ? x[eu] : 0.0;
        qy[u] = (eu < a.n && !same) ? y[eu] : 0.0;
    }
    double v = 0.0;
    for (long long e = e0; e < a.n; e += S) {
        const double xv = qx[0], yv = same ? qx[0] : qy[0];
#pragma unroll
        for (int u = 0; u + 1 < kEnergyAhead; ++u) {
            qx[u] = qx[u + 1];
            qy[u] = qy[u + 1];
        }
        const long long ea = e + kEnergyAhead * S;
        qx[kEnergyAhead - 1] = ea < a.n ? x[ea] : 0.0;
        qy[kEnergyAhead - 1] = (ea < a.n && !same) ? y[ea] : 0.0;
        v += (xv / denominator(a, idx)) * yv;
        advance(a, idx, step);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The partial blocks' sum and its energy in one pass: t = ((t + p0) + p1) +
// ... is stored, and the energy of the summed element is accumulated exactly
// as energy_partial_kernel does with x = y = t (same elements per thread, same
// order), so the result is the one of the two separate passes.
__global__ __launch_bounds__(kEnergyThreads) void energy_parts_kernel(EnergyArgs a, double *t,
                                                                      AddManyArgs ps,
                                                                      double *part) {
    __shared__ double red[kEnergyThreads];
    const long long e0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long S = (long long)gridDim.x * blockDim.x;
    int idx[kEnergyMaxRank], step[kEnergyMaxRank];
    decode(a, e0, idx);
    decode(a, S, step);
    double v = 0.0;
    for (long long e = e0; e < a.n; e += S) {
        double x[kAddManyMax];
        const double t0 = t[e];
#pragma unroll
        for (int q = 0; q < kAddManyMax; ++q)
            if (q < ps.n) x[q] = ps.p[q][e];
        double w = t0;
#pragma unroll
        for (int q = 0; q < kAddManyMax; ++q)
            if (q < ps.n) w += x[q];
        t[e] = w;
        v += (w / denominator(a, idx)) * w;
        advance(a, idx, step);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// One workgroup: thread t sums partials t, t + 256, ... in that order, then
// a fixed tree over the threads (the same order on every run).
__global__ __launch_bounds__(kEnergyThreads) void energy_final_kernel(double *out,
                                                                      const double *part,
                                                                      int nparts, double factor) {
    __shared__ double red[kEnergyThreads];
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kEnergyThreads) v += part[i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] += factor * red[0];
}

} // namespace

hipError_t launch_denominator_divide(const EnergyArgs &a, double *x, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const long long blocks = (a.n + kEnergyThreads - 1) / kEnergyThreads;
    hipLaunchKernelGGL(denominator_divide_kernel, dim3((unsigned)(blocks < 16384 ? blocks : 16384)),
                       dim3(kEnergyThreads), 0, s, a, x);
    return hipGetLastError();
}

hipError_t launch_energy_sum_parts(const EnergyArgs &a, double *t, const AddManyArgs &ps,
                                   double factor, double *out, double *part, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(energy_parts_kernel, dim3(kEnergyBlocks), dim3(kEnergyThreads), 0, s, a, t,
                       ps, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(kEnergyThreads), 0, s, out, part,
                       kEnergyBlocks, factor);
    return hipGetLastError();
}

hipError_t launch_energy_sum(const EnergyArgs &a, const double *x, const double *y,
                             double factor, double *out, double *part, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(energy_partial_kernel, dim3(kEnergyBlocks), dim3(kEnergyThreads), 0, s, a,
                       x, y, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(kEnergyThreads), 0, s, out, part,
                       kEnergyBlocks, factor);
    return hipGetLastError();
}

} // namespace sipgpu
