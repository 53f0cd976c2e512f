// fs_hostutil.h -- host-side helpers shared by the library's translation
// units: status propagation and device allocation.
#ifndef FS_HOSTUTIL_H
#define FS_HOSTUTIL_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "fs_internal.h"

#define FS_HIP_TRY(expr)                                                           \
  do {                                                                             \
    const hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                        \
      fs::set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
      return FS_ERR_HIP;                                                           \
    }                                                                              \
  } while (0)

#define FS_TRY(expr)              \
  do {                            \
    const int s_ = (expr);        \
    if (s_ != FS_OK) return s_;   \
  } while (0)

namespace fs {

template <class T>
int dev_alloc(T **ptr, size_t count) {
  const hipError_t e = hipMalloc(reinterpret_cast<void **>(ptr), std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) {
    *ptr = nullptr;
    fs::set_last_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return FS_ERR_OUT_OF_MEMORY;
  }
  return FS_OK;
}

}  // namespace fs

#endif
