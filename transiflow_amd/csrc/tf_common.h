// Shared host-side helpers for the TransiFlow HIP backend's C ABI.
#pragma once

#include <mutex>

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

// ---- opt-in kernel timing -------------------------------------------------
// When enabled (tf_prof_enable), launches wrapped in TF_PROF_BEGIN/END record
// a pair of HIP events on the launch stream; tf_prof_read() synchronises and
// returns the launch count and summed duration per kernel id.
enum ProfId : int {
  PROF_JAC_COUNT = 0,
  PROF_JAC_FILL = 1,
  PROF_RHS = 2,
  PROF_GJ_PANEL = 3,
  PROF_GJ_UPDATE = 4,
  PROF_GEMM = 5,
  PROF_GEMV = 6,
  PROF_SPMV = 7,
  PROF_CONV = 8,
  PROF_SOLVE = 9,
  PROF_SOLVE2 = 10,  // two-column line-factor apply
  PROF_IDS = 16
};
bool prof_on();
// Host-to-device copy through a process-wide non-blocking stream (created once
// per device), synchronised before returning.  Used instead of hipMemcpy, whose
// legacy-stream semantics make it fail while another thread captures a graph.
hipError_t upload(void* dst, const void* src, size_t bytes);
void prof_begin(int id, hipStream_t st);
void prof_end(int id, hipStream_t st);
void prof_work(int id, double w);
// A graph capture that failed (another thread's work can invalidate it) and whose
// call ran eagerly instead: counted for tf_capture_fallbacks().
void note_capture_fallback();
int capture_fallbacks();
// Held around every stream capture and graph instantiation of the library, so that
// captures of different threads (one per worker stream) never overlap.
std::mutex& capture_mutex();

}  // namespace tf

#define TF_PROF_BEGIN(id, st) \
  do {                        \
    if (tf::prof_on()) tf::prof_begin((id), (st)); \
  } while (0)
#define TF_PROF_END(id, st) \
  do {                      \
    if (tf::prof_on()) tf::prof_end((id), (st)); \
  } while (0)
#define TF_PROF_WORK(id, w) \
  do {                      \
    if (tf::prof_on()) tf::prof_work((id), (w)); \
  } while (0)

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
