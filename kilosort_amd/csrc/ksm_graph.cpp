// Graph recording of the library's launches (ksm_graph.h) and its C entry
// points: ksm_graph_begin / _end / _launch / _destroy and ksm_copy_async.
#include <memory>
#include <new>
#include <vector>

#include "ksm_common.h"
#include "ksm_graph.h"

namespace ksm {

struct GraphRecording {
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
    const char *where = nullptr;
    std::vector<std::shared_ptr<void>> keep;     // kernel argument blocks
};

// The handle ksm_graph_end returns: the executable graph and the argument
// blocks its kernel nodes were made from.
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    std::vector<std::shared_ptr<void>> keep;
};

namespace {
thread_local GraphRecording *t_rec = nullptr;

void chain(GraphRecording *r, hipGraphNode_t n) { r->last = n; }
}  // namespace

GraphRecording *recording(hipStream_t s) {
    GraphRecording *r = t_rec;
    return (r && r->stream == s) ? r : nullptr;
}

void record_keep(GraphRecording *r, std::shared_ptr<void> blob) {
    r->keep.push_back(std::move(blob));
}

void record_kernel(GraphRecording *r, const void *func, dim3 grid, dim3 block, unsigned shm,
                   void **args) {
    if (r->err != hipSuccess) return;
    hipKernelNodeParams p = {};
    p.func = const_cast<void *>(func);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = shm;
    p.kernelParams = args;
    p.extra = nullptr;
    hipGraphNode_t n = nullptr;
    const hipError_t e =
        hipGraphAddKernelNode(&n, r->graph, r->last ? &r->last : nullptr, r->last ? 1 : 0, &p);
    if (e != hipSuccess) {
        r->err = e;
        r->where = "hipGraphAddKernelNode";
        return;
    }
    chain(r, n);
}

void record_copy(GraphRecording *r, void *dst, const void *src, size_t bytes,
                 hipMemcpyKind kind) {
    if (r->err != hipSuccess) return;
    hipGraphNode_t n = nullptr;
    const hipError_t e = hipGraphAddMemcpyNode1D(&n, r->graph, r->last ? &r->last : nullptr,
                                                 r->last ? 1 : 0, dst, src, bytes, kind);
    if (e != hipSuccess) {
        r->err = e;
        r->where = "hipGraphAddMemcpyNode1D";
        return;
    }
    chain(r, n);
}

}  // namespace ksm

extern "C" {

int ksm_graph_begin(void *stream) {
    if (ksm::t_rec)
        return ksm::fail(KSM_ERR_INVALID, "ksm_graph_begin: this thread is already recording");
    auto *r = new (std::nothrow) ksm::GraphRecording;
    if (!r) return ksm::fail(KSM_ERR_HIP, "ksm_graph_begin: out of host memory");
    const hipError_t e = hipGraphCreate(&r->graph, 0);
    if (e != hipSuccess) {
        delete r;
        return ksm::hip_fail(e, "ksm_graph_begin: hipGraphCreate");
    }
    r->stream = ksm::as_stream(stream);
    ksm::t_rec = r;
    return KSM_OK;
}

int ksm_graph_end(void *stream, void **graph_exec) {
    ksm::GraphRecording *r = ksm::t_rec;
    if (graph_exec) *graph_exec = nullptr;
    if (!r || r->stream != ksm::as_stream(stream))
        return ksm::fail(KSM_ERR_INVALID, "ksm_graph_end: no recording open on this stream");
    ksm::t_rec = nullptr;
    int rc = KSM_OK;
    if (r->err != hipSuccess) {
        rc = ksm::hip_fail(r->err, r->where);
    } else if (!r->last) {
        rc = ksm::fail(KSM_ERR_INVALID, "ksm_graph_end: nothing was recorded");
    } else if (graph_exec) {
        auto *g = new (std::nothrow) ksm::GraphExec;
        hipError_t e = g ? hipGraphInstantiate(&g->exec, r->graph, nullptr, nullptr, 0)
                         : hipErrorOutOfMemory;
        if (e != hipSuccess) {
            delete g;
            rc = ksm::hip_fail(e, "ksm_graph_end: hipGraphInstantiate");
        } else {
            g->keep = std::move(r->keep);
            *graph_exec = g;
        }
    }
    hipGraphDestroy(r->graph);
    delete r;
    return rc;
}

int ksm_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return ksm::fail(KSM_ERR_INVALID, "ksm_graph_launch: null graph");
    if (ksm::recording(ksm::as_stream(stream)))
        return ksm::fail(KSM_ERR_INVALID, "ksm_graph_launch: the stream is being recorded");
    const hipError_t e =
        hipGraphLaunch(static_cast<ksm::GraphExec *>(graph_exec)->exec, ksm::as_stream(stream));
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_graph_launch");
    return KSM_OK;
}

int ksm_graph_destroy(void *graph_exec) {
    if (!graph_exec) return KSM_OK;
    auto *g = static_cast<ksm::GraphExec *>(graph_exec);
    const hipError_t e = hipGraphExecDestroy(g->exec);
    delete g;
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_graph_destroy");
    return KSM_OK;
}

int ksm_copy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    if (kind != KSM_COPY_DEVICE_TO_DEVICE && kind != KSM_COPY_DEVICE_TO_HOST)
        return ksm::fail(KSM_ERR_INVALID, "ksm_copy_async: unknown kind %d", kind);
    if (bytes == 0) return KSM_OK;
    if (!dst || !src) return ksm::fail(KSM_ERR_INVALID, "ksm_copy_async: null pointer");
    const hipMemcpyKind k =
        kind == KSM_COPY_DEVICE_TO_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const hipStream_t s = ksm::as_stream(stream);
    if (ksm::GraphRecording *r = ksm::recording(s)) {
        ksm::record_copy(r, dst, src, bytes, k);
        return KSM_OK;
    }
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, k, s);
    if (e != hipSuccess) return ksm::hip_fail(e, "ksm_copy_async");
    return KSM_OK;
}

}  // extern "C"
