// internal.h -- runtime pieces shared between the library's host sources.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sipgpu.h"
#include "kernels.h"
#include "ozaki.h"

// A library stream (the HIP stream and its contract-to-scalar scratch).
struct sipgpu_stream_s {
    hipStream_t s = nullptr;
    double *scratch = nullptr;
    double *ws = nullptr; // K-segment partial results of split single contractions
    size_t ws_bytes = 0;  // (grow-only)
    // Earlier, smaller workspaces: a graph captured on this stream may still
    // address one, so they are freed only with the stream.
    std::vector<double *> retired;
};

namespace sipgpu {
namespace rt {

// Record `msg` as the calling thread's last error and return `code`.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
// The library is initialised on its device (first use initialises device 0).
int ensure_ready();
// True while `s` is being captured into a graph.
bool capturing(hipStream_t s);
// Streams the calling thread is capturing (planning, allocation, host copies
// and synchronisation would invalidate their capture).
int &thread_captures();
// SIPGPU_ECAPTURE when the calling thread is capturing, else 0.
int refuse_in_capture(const char *what);
// Elements of a cached block (*elems = 0 when the key is unknown).
int cache_block_elems(sipgpu_cache_t c, unsigned long long key, size_t *elems);
// The library's device, and the HIP stream of `s` (null: the library's own).
int device_id();
hipStream_t stream_of(sipgpu_stream_t s);

// Host image of an FP64 batch: tiles | tasks | pairs | tables, each 256-byte
// aligned, with device pointers resolved against wherever it is uploaded.
struct BatchImage {
    std::vector<unsigned char> host;
    size_t o_tiles = 0, o_tasks = 0, o_pairs = 0, o_tab = 0, o_parts = 0;
    int cfg = 0, ntiles = 0, ntasks = 0, npairs = 0, real_tiles = 0;
    // Result blocks cut into pair groups (a small batch): the GEMM tasks of
    // groups after the first write workspace copies of their block, at
    // ws_off[i] doubles for task ws_task[i]; nparts entries at o_parts add
    // them into the blocks after the GEMM (part offsets in parts_off until
    // bind_workspace).
    std::vector<int> ws_task;
    std::vector<long long> ws_off, parts_off;
    int nparts = 0;
    long long parts_max_n = 0, ws_doubles = 0;
    double flops = 0.0;
    void bind_workspace(double *ws) {
        GemmTask *t = reinterpret_cast<GemmTask *>(host.data() + o_tasks);
        for (size_t i = 0; i < ws_task.size(); ++i) t[ws_task[i]].D = ws + ws_off[i];
        AddPartsEntry *e = reinterpret_cast<AddPartsEntry *>(host.data() + o_parts);
        for (int i = 0; i < nparts; ++i) e[i].p = ws + parts_off[i];
    }
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

// Plans a list of block contractions (same rules as sipgpu_batch_plan) for
// the table kernel; cfg >= 0 pins its tile configuration.  split_cus > 0:
// when the tiles fill at most half of that many CUs, each result block's
// pairs are cut into groups that run side by side (workspace copies summed
// into the block afterwards, in group order).
int build_batch_image(const sipgpu_contraction *list, int n, BatchImage &img, int cfg = -1,
                      int split_cus = 0);

// Choices of the panel planner; -1 takes the default (environment, then the
// built-in rule).
struct PanelOptions {
    int merge = -1; // 1: merge result blocks sharing a virtual K into grids
    int tail = -1;  // 1: persistent schedule with a water-filled last round (FP64)
    int tile = -1;  // split path GEMM tile: 128 or 256
};

// Device state of a panel batch (packed FP64 or split precision).
struct SplitData {
    OzConst c{};
    int *expo = nullptr;
    int nexpo = 0;
    int8_t *planes = nullptr;
    long long plane_bytes = 0;
    uint8_t *res_planes = nullptr;
    long long res_plane_bytes = 0;
    double *work = nullptr; // split-tile partial products
    const OzConvTile *row_items = nullptr;
    int nrow_items = 0;
    OzConvArgs cargs{};
    OzGemmArgs gargs{};
    OzRecArgs rargs{};
};

// Host plan of a panel batch: its descriptor image (uploaded as is) and the
// sizes of the device memory it runs in.  nmod > 0: split precision over
// nmod moduli; 0: packed FP64.
struct PanelPlan {
    int nmod = 0;
    std::vector<unsigned char> host;
    size_t o_blocks = 0, o_tasks = 0, o_tiles = 0, o_chunks = 0, o_tab = 0, o_conv = 0,
           o_rows = 0, o_rec = 0, o_subs = 0, o_fix = 0, o_wg = 0, o_expo = 0;
    long long plane_bytes = 0, res_plane_bytes = 0;
    int nslots = 0; // workspace slots of kOzBM * kOzBN doubles
    OzConst c{};
    int tile = 0, ntiles = 0, nconv = 0, nrows = 0, nrec = 0, nfix = 0, nwg = 0, nexpo = 0;
    int ndest = 0, ngroups = 0, npairs = 0, real_tiles = 0;
    double flops = 0.0;
    size_t plane_total() const { return (size_t)plane_bytes * (nmod ? nmod : 1); }
    size_t res_total() const { return (size_t)res_plane_bytes * nmod; }
    size_t work_total() const { return (size_t)nslots * kOzBM * kOzBN * sizeof(double); }
};

int build_panel_plan(const sipgpu_contraction *list, int n, int nmod, PanelPlan &p,
                     const PanelOptions &opt = PanelOptions());
// Points the kernels' arguments at an uploaded image and its memory.
void bind_panel_plan(const PanelPlan &p, void *desc, int8_t *planes, uint8_t *res_planes,
                     double *work, SplitData &sd);
// Enqueues the stages (SIPGPU_STAGE_* bits) of a bound panel batch.
int execute_panels(SplitData &sd, bool split, int stages, hipStream_t s);

} // namespace rt
} // namespace sipgpu

#define SIPGPU_HIP_TRY(expr)                                                              \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return ::sipgpu::rt::hip_fail(e_, #expr);                   \
    } while (0)
