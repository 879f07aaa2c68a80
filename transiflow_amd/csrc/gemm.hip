// Tile autotuner of the batched MFMA GEMM (see gemm.h).
#include "gemm.h"

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace tf {

namespace {

struct TuneState {
  std::mutex mu;
  int mode = 1;  // 1: time the candidates, 0: fixed rule
  std::map<std::tuple<int, int, int, int, int>, int> table;
};
TuneState& tune() {
  static TuneState s;
  return s;
}

int fixed_rule(int single, int M, int N, int K, int batch) {
  (void)K;
  const long big = (long)((M + 127) / 128) * ((N + 127) / 128) * batch;
  return (single && big >= 1024) ? GEMM_128x128 : GEMM_64x64;
}

template <typename T>
int time_candidates(int M, int N, int K, int batch, hipStream_t st) {
  T *a = nullptr, *b = nullptr, *c = nullptr;
  T** lists = nullptr;
  int best = -1;
  if (hipMalloc(&a, sizeof(T) * (size_t)M * K) == hipSuccess && hipMalloc(&b, sizeof(T) * (size_t)K * N) == hipSuccess &&
      hipMalloc(&c, sizeof(T) * (size_t)M * N) == hipSuccess &&
      hipMalloc(&lists, sizeof(T*) * 3 * (size_t)batch) == hipSuccess) {
    std::vector<T*> h(3 * (size_t)batch);
    for (int i = 0; i < batch; ++i) {
      h[i] = a;
      h[batch + i] = b;
      h[2 * batch + i] = c;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (upload(lists, h.data(), sizeof(T*) * h.size()) == hipSuccess &&
        hipMemsetAsync(a, 0, sizeof(T) * (size_t)M * K, st) == hipSuccess &&
        hipMemsetAsync(b, 0, sizeof(T) * (size_t)K * N, st) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
      float best_ms = 0.f;
      for (int tile = 0; tile < GEMM_TILES; ++tile) {
        gemm_launch<T>(tile, M, N, K, lists, K, lists + batch, N, lists + 2 * batch, N, 1.0, 0.0, batch, st);
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < 2; ++r)
          gemm_launch<T>(tile, M, N, K, lists, K, lists + batch, N, lists + 2 * batch, N, 1.0, 0.0, batch, st);
        (void)hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            (best < 0 || ms < best_ms)) {
          best = tile;
          best_ms = ms;
        }
      }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(c);
  (void)hipFree(lists);
  (void)hipGetLastError();
  return best;
}

}  // namespace

int gemm_tuned_tile(int single, int M, int N, int K, int batch, hipStream_t st) {
  TuneState& s = tune();
  const auto key = std::make_tuple(single, M, N, K, batch);
  {
    std::lock_guard<std::mutex> g(s.mu);
    auto it = s.table.find(key);
    if (it != s.table.end()) return it->second;
    if (s.mode == 0) return fixed_rule(single, M, N, K, batch);
  }
  int tile = single ? time_candidates<float>(M, N, K, batch, st) : time_candidates<double>(M, N, K, batch, st);
  if (tile < 0) tile = fixed_rule(single, M, N, K, batch);
  std::lock_guard<std::mutex> g(s.mu);
  s.table[key] = tile;
  return tile;
}

}  // namespace tf

extern "C" {

// mode 1: time the tile candidates on first use of a shape (default); 0: fixed rule.
int tf_gemm_tuning(int mode) {
  std::lock_guard<std::mutex> g(tf::tune().mu);
  tf::tune().mode = mode ? 1 : 0;
  return 0;
}

// Rows of the tuned table: (single, M, N, K, batch, tile) into out_host[6*i..];
// returns the number of rows (all of them when max_rows is large enough).
int tf_gemm_table(int* out_host, int max_rows) {
  std::lock_guard<std::mutex> g(tf::tune().mu);
  int i = 0;
  for (const auto& kv : tf::tune().table) {
    if (i < max_rows && out_host) {
      int* r = out_host + 6 * i;
      r[0] = std::get<0>(kv.first);
      r[1] = std::get<1>(kv.first);
      r[2] = std::get<2>(kv.first);
      r[3] = std::get<3>(kv.first);
      r[4] = std::get<4>(kv.first);
      r[5] = kv.second;
    }
    ++i;
  }
  return i;
}

}  // extern "C"
