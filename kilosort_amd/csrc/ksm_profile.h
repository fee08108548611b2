// Kernel timing scope used around the library's main kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <utility>

#include "ksm_graph.h"

namespace ksm {

// Profiling groups (bit mask of ksm_profile_enable).
constexpr unsigned kProfProject = 1u;   // conv_pc, gemm_*
constexpr unsigned kProfPeel = 2u;      // colmax_all, peel_*

// One profiled kernel launch.  When the scope's group is enabled, launch()
// goes through hipExtLaunchKernel with a pair of events that the runtime
// stamps with the kernel's own start and end (no marker packets around it);
// otherwise it is a plain launch and nothing is recorded.  On a stream whose
// launches are being recorded into a graph it adds the kernel's node instead
// (and times nothing).
class ProfileScope {
  public:
    ProfileScope(const char *name, unsigned group, hipStream_t s);
    ~ProfileScope();
    ProfileScope(const ProfileScope &) = delete;
    ProfileScope &operator=(const ProfileScope &) = delete;

    template <typename... KArgs, typename... Args>
    void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shm,
                Args &&...args) {
        if (recording(stream_)) {
            ksm::launch(kernel, grid, block, shm, stream_, std::forward<Args>(args)...);
            return;
        }
        if (start_ && stop_) {
            hipExtLaunchKernelGGL(kernel, grid, block, shm, stream_, start_, stop_, 0,
                                  std::forward<Args>(args)...);
            launched_ = true;
        } else {
            kernel<<<grid, block, shm, stream_>>>(std::forward<Args>(args)...);
        }
    }

  private:
    const char *name_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
    hipEvent_t stop_ = nullptr;
    bool launched_ = false;
};

}  // namespace ksm
