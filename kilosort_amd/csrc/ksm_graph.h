// Recording of the library's own launches into a HIP graph, without stream
// capture.
//
// Between ksm_graph_begin(stream) and ksm_graph_end on one host thread, every
// kernel the matching entry points would enqueue on `stream` (and every
// ksm_copy_async on it) becomes the next node of a linear chain in a new
// graph instead; nothing runs.  Stream capture would record the same chain,
// but a capture in progress makes other threads' runtime calls (allocation,
// synchronisation) fail or invalidate it, while building nodes directly
// involves no stream state at all, so host threads driving other streams are
// unaffected.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <tuple>
#include <type_traits>
#include <utility>

namespace ksm {

struct GraphRecording;

// The recording this host thread holds open on stream s, or nullptr.
GraphRecording *recording(hipStream_t s);

// Keeps `blob` alive as long as the graph recorded into r.
void record_keep(GraphRecording *r, std::shared_ptr<void> blob);

// Appends a kernel node (arguments as an array of pointers to their values)
// after the recording's last node.  A failure is kept and reported by
// ksm_graph_end.
void record_kernel(GraphRecording *r, const void *func, dim3 grid, dim3 block, unsigned shm,
                   void **args);

// Appends a 1-D copy node.
void record_copy(GraphRecording *r, void *dst, const void *src, size_t bytes,
                 hipMemcpyKind kind);

// A kernel launch on stream s, or its node when s is being recorded on this
// thread.  The recorded argument values are converted to the kernel's own
// parameter types first, exactly as the launch would pass them.
template <typename... KArgs, typename... Args>
inline void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shm, hipStream_t s,
                   Args &&...args) {
    static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
    if (GraphRecording *r = recording(s)) {
        // the values and the pointer array live with the graph, whether or
        // not the runtime copies them when the node is made
        struct Blob {
            std::tuple<std::remove_cv_t<KArgs>...> vals;
            void *ptrs[sizeof...(KArgs) + 1] = {};
        };
        auto blob = std::make_shared<Blob>(Blob{{std::forward<Args>(args)...}});
        std::apply(
            [&blob](auto &...v) {
                int i = 0;
                ((blob->ptrs[i++] = const_cast<void *>(static_cast<const void *>(&v))), ...);
            },
            blob->vals);
        record_kernel(r, reinterpret_cast<const void *>(kernel), grid, block, shm, blob->ptrs);
        record_keep(r, std::move(blob));
        return;
    }
    kernel<<<grid, block, shm, s>>>(std::forward<Args>(args)...);
}

}  // namespace ksm
