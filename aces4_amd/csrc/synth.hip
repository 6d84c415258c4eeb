// synth.hip -- seeded synthetic block data generated in HBM.
//
// Counter-based SplitMix64: output j (1-based) of (seed, stream) is
// mix(seed * G + stream * H + j * G), G = 0x9E3779B97F4A7C15,
// H = 0xD1B54A32D192ED03; uniform draws take the top 53 bits.  The host-side
// generator in aces4_amd/workloads.py computes the same integers, so uniform
// blocks are bit-identical on host and device; normal draws (Box-Muller over
// streams 2s+1 and 2s+2, cosines first then sines) agree to the last bits of
// the device's log / sincos.
#include "kernels.h"

namespace sipgpu {

namespace {

constexpr unsigned long long kGolden = 0x9E3779B97F4A7C15ull;
constexpr unsigned long long kStream = 0xD1B54A32D192ED03ull;

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(unsigned long long seed, unsigned long long stream,
                                       unsigned long long j) {
    const unsigned long long z = seed * kGolden + stream * kStream + (j + 1) * kGolden;
    return (double)(mix64(z) >> 11) * 0x1p-53;
}

__global__ void uniform_kernel(double *out, long long n, unsigned long long seed,
                               unsigned long long stream, double lo, double hi) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x)
        out[j] = lo + (hi - lo) * unit(seed, stream, (unsigned long long)j);
}

__global__ void normal_kernel(double *out, long long n, unsigned long long seed,
                              unsigned long long stream, double sigma) {
    const long long m = (n + 1) / 2;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < m;
         j += (long long)gridDim.x * blockDim.x) {
        const double u1 = 1.0 - unit(seed, 2 * stream + 1, (unsigned long long)j);
        const double u2 = unit(seed, 2 * stream + 2, (unsigned long long)j);
        const double r = sqrt(-2.0 * log(u1));
        double s, c;
        sincos(2.0 * M_PI * u2, &s, &c);
        out[j] = sigma * (r * c);
        if (m + j < n) out[m + j] = sigma * (r * s);
    }
}

} // namespace

hipError_t launch_fill_random(double *out, long long n, unsigned long long seed,
                              unsigned long long stream, int dist, double a, double b,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long work = dist == 0 ? n : (n + 1) / 2;
    const long long blocks = (work + 255) / 256;
    const dim3 grid((unsigned)(blocks < 16384 ? blocks : 16384));
    if (dist == 0)
        hipLaunchKernelGGL(uniform_kernel, grid, dim3(256), 0, s, out, n, seed, stream, a, b);
    else
        hipLaunchKernelGGL(normal_kernel, grid, dim3(256), 0, s, out, n, seed, stream, a);
    return hipGetLastError();
}

} // namespace sipgpu
