// calib.hip -- measured ceilings of the device: FP64 MFMA throughput (and
// the shader clock it runs at), streaming-copy HBM bandwidth and LDS read
// bandwidth.  Used as
// roofline denominators beside the datasheet peaks, and as the byte-count
// calibration of HBM counters (the copy moves a known number of bytes).
// Built with the MFMA VGPR form so the accumulators stay in place across
// loop iterations.
#include "kernels.h"

namespace sipgpu {

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Every wave issues `iters` rounds of 8 independent 16x16x4 FP64 MFMAs.
// Thread 0 of each block records the shader-clock and constant-rate
// wall-clock ticks spent in the loop, giving the effective clock under load.
__global__ __launch_bounds__(256) void mfma_f64_peak_kernel(double *out, long long *ticks,
                                                            int iters) {
    f64x4 acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = f64x4{out[i], 0.0, 0.0, 0.0};
    const double a = 1.0 + 1e-3 * (threadIdx.x & 7), b = 1.0 - 1e-3 * (threadIdx.x & 3);
    const long long c0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    if (s == 12345.678) out[blockIdx.x * blockDim.x + threadIdx.x] = s; // keeps the loop live
    if (threadIdx.x == 0) {
        ticks[2 * blockIdx.x] = c1 - c0;
        ticks[2 * blockIdx.x + 1] = w1 - w0;
    }
}

// Grid-stride copy with 16-byte accesses.
__global__ __launch_bounds__(256) void copy_kernel(const double2 *__restrict__ in,
                                                   double2 *__restrict__ out, size_t n2) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        out[i] = in[i];
}

// LDS read bandwidth: every wave reads 16 bytes per lane, eight reads per
// round, from a 64 KiB array (indices masked into it), `iters` rounds.
// Thread 0 records shader-clock and wall-clock ticks like the MFMA loop.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void lds_read_kernel(int *out, long long *ticks,
                                                              int iters) {
    typedef int int4v __attribute__((ext_vector_type(4)));
    __shared__ int4v buf[4096];
    const int tid = threadIdx.x;
    for (int i = tid; i < 4096; i += 64 * WAVES) buf[i] = int4v{i, i + 1, i + 2, i + 3};
    __syncthreads();
    int4v acc = {0, 0, 0, 0};
    int idx = tid & 4095;
    const long long c0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += buf[(idx + u * 512) & 4095];
        idx = (idx + 64) & 4095;
    }
    const long long c1 = clock64(), w1 = wall_clock64();
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x7fffffff) out[blockIdx.x] = acc.x; // keeps the loop live
    if (tid == 0) {
        ticks[2 * blockIdx.x] = c1 - c0;
        ticks[2 * blockIdx.x + 1] = w1 - w0;
    }
}

} // namespace

hipError_t launch_lds_read(int waves, int blocks, int *out, long long *ticks, int iters,
                           hipStream_t s) {
    if (waves == 8)
        hipLaunchKernelGGL(lds_read_kernel<8>, dim3(blocks), dim3(512), 0, s, out, ticks, iters);
    else
        hipLaunchKernelGGL(lds_read_kernel<16>, dim3(blocks), dim3(1024), 0, s, out, ticks, iters);
    return hipGetLastError();
}

hipError_t launch_mfma_peak(double *out, long long *ticks, int blocks, int iters, hipStream_t s) {
    hipLaunchKernelGGL(mfma_f64_peak_kernel, dim3(blocks), dim3(256), 0, s, out, ticks, iters);
    return hipGetLastError();
}

hipError_t launch_copy(const double *in, double *out, size_t n, hipStream_t s) {
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s,
                       reinterpret_cast<const double2 *>(in), reinterpret_cast<double2 *>(out),
                       n2);
    return hipGetLastError();
}

} // namespace sipgpu
