// internal.h -- runtime pieces shared between the library's host sources.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sipgpu.h"
#include "kernels.h"

// A library stream (the HIP stream and its contract-to-scalar scratch).
struct sipgpu_stream_s {
    hipStream_t s = nullptr;
    double *scratch = nullptr;
};

namespace sipgpu {
namespace rt {

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// The library is initialised on its device (first use initialises device 0).
int ensure_ready();
// The library's device, and the HIP stream of `s` (null: the library's own).
int device_id();
hipStream_t stream_of(sipgpu_stream_t s);

// Host image of an FP64 batch: tiles | tasks | pairs | tables, each 256-byte
// aligned, with device pointers resolved against wherever it is uploaded.
struct BatchImage {
    std::vector<unsigned char> host;
    size_t o_tiles = 0, o_tasks = 0, o_pairs = 0, o_tab = 0;
    int cfg = 0, ntiles = 0, ntasks = 0, npairs = 0, real_tiles = 0;
    double flops = 0.0;
    GemmArgs args_at(const void *dev) const {
        const unsigned char *base = static_cast<const unsigned char *>(dev);
        GemmArgs a{};
        a.tiles = reinterpret_cast<const GemmTile *>(base + o_tiles);
        a.tasks = reinterpret_cast<const GemmTask *>(base + o_tasks);
        a.pairs = reinterpret_cast<const GemmPair *>(base + o_pairs);
        a.tab = reinterpret_cast<const int32_t *>(base + o_tab);
        a.ntiles = ntiles;
        return a;
    }
};

// Plans a list of block contractions (same rules as sipgpu_batch_plan).
int build_batch_image(const sipgpu_contraction *list, int n, BatchImage &img);

} // namespace rt
} // namespace sipgpu

#define SIPGPU_HIP_TRY(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return ::sipgpu::rt::hip_fail(e_, #expr);                   \
    } while (0)
