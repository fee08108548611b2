// Event-pair timing scope used around the library's main kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ksm {

// Profiling groups (bit mask of ksm_profile_enable).
constexpr unsigned kProfProject = 1u;   // conv_pc, gemm_*
constexpr unsigned kProfPeel = 2u;      // colmax_all, peel_*

class ProfileScope {
  public:
    ProfileScope(const char *name, unsigned group, hipStream_t s);
    ~ProfileScope();
    ProfileScope(const ProfileScope &) = delete;
    ProfileScope &operator=(const ProfileScope &) = delete;

  private:
    const char *name_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};

}  // namespace ksm
