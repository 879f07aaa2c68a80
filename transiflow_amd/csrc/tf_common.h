// Shared host-side helpers for the TransiFlow HIP backend's C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

namespace tf {

// Last error message of the calling thread; returned by tf_last_error().
inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline int fail(const char* what) {
  last_error() = what;
  return -1;
}

inline int fail_hip(hipError_t e, const char* where) {
  char buf[256];
  std::snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
  last_error() = buf;
  return -2;
}

}  // namespace tf

#define TF_HIP(call)                                      \
  do {                                                    \
    hipError_t tf_err_ = (call);                          \
    if (tf_err_ != hipSuccess) return tf::fail_hip(tf_err_, #call); \
  } while (0)

#define TF_LAUNCH_CHECK(where)                            \
  do {                                                    \
    hipError_t tf_err_ = hipGetLastError();               \
    if (tf_err_ != hipSuccess) return tf::fail_hip(tf_err_, where); \
  } while (0)
