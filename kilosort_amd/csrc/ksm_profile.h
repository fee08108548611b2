// Event-pair timing scope used around the library's main kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace ksm {

class ProfileScope {
  public:
    ProfileScope(const char *name, hipStream_t s);
    ~ProfileScope();
    ProfileScope(const ProfileScope &) = delete;
    ProfileScope &operator=(const ProfileScope &) = delete;

  private:
    const char *name_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};

}  // namespace ksm
