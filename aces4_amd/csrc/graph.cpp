// graph.cpp -- CC iterations captured as HIP graphs and replayed.
//
// A CC iteration enqueues the same launches every time (the same batches
// over the same resident blocks, then the energy), so the whole sequence is
// captured once from its stream and replayed with one hipGraphLaunch: the
// host enqueue cost of dozens of launches becomes one, and the device runs
// the launches back to back.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/sipgpu.h"
#include "internal.h"

struct sipgpu_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
};

using sipgpu::rt::fail;
using sipgpu::rt::hip_fail;

int &sipgpu::rt::thread_captures() {
    thread_local int n = 0;
    return n;
}

int sipgpu::rt::refuse_in_capture(const char *what) {
    if (thread_captures() == 0) return 0;
    return fail(SIPGPU_ECAPTURE, std::string(what) + " is not allowed during graph capture");
}

bool sipgpu::rt::capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st == hipStreamCaptureStatusActive;
}

extern "C" {

int sipgpu_graph_begin(sipgpu_stream_t s) {
    if (int rc = sipgpu::rt::ensure_ready()) return rc;
    const hipStream_t hs = sipgpu::rt::stream_of(s);
    if (sipgpu::rt::capturing(hs)) return fail(SIPGPU_ECAPTURE, "stream is already capturing");
    SIPGPU_HIP_TRY(hipStreamBeginCapture(hs, hipStreamCaptureModeThreadLocal));
    ++sipgpu::rt::thread_captures();
    return 0;
}

int sipgpu_graph_end(sipgpu_stream_t s, sipgpu_graph_t *graph) {
    if (!graph) return fail(SIPGPU_EARG, "null graph");
    *graph = nullptr;
    if (int rc = sipgpu::rt::ensure_ready()) return rc;
    const hipStream_t hs = sipgpu::rt::stream_of(s);
    if (!sipgpu::rt::capturing(hs)) return fail(SIPGPU_ECAPTURE, "stream is not capturing");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(hs, &g);
    if (sipgpu::rt::thread_captures() > 0) --sipgpu::rt::thread_captures();
    if (e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        return e != hipSuccess ? hip_fail(e, "hipStreamEndCapture")
                               : fail(SIPGPU_ECAPTURE, "capture produced no graph");
    }
    auto G = new sipgpu_graph_s;
    G->graph = g;
    hipError_t e2 = hipGraphGetNodes(g, nullptr, &G->nodes);
    if (e2 == hipSuccess) e2 = hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0);
    if (e2 != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete G;
        return hip_fail(e2, "hipGraphInstantiate");
    }
    *graph = G;
    return 0;
}

int sipgpu_graph_launch(sipgpu_graph_t graph, sipgpu_stream_t s) {
    if (!graph || !graph->exec) return fail(SIPGPU_EARG, "null graph");
    SIPGPU_HIP_TRY(hipGraphLaunch(graph->exec, sipgpu::rt::stream_of(s)));
    return 0;
}

int sipgpu_graph_nodes(sipgpu_graph_t graph, int *nodes) {
    if (!graph || !nodes) return fail(SIPGPU_EARG, "null graph");
    *nodes = (int)graph->nodes;
    return 0;
}

int sipgpu_graph_destroy(sipgpu_graph_t graph) {
    if (!graph) return fail(SIPGPU_EARG, "null graph");
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    delete graph;
    return 0;
}

} // extern "C"
