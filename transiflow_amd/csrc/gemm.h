// Batched dense GEMM on the CDNA matrix cores:  C_b = alpha A_b B_b + beta C_b
// (row major, one pointer per batch entry).
//
// Block tile BM x BN with 256 threads = 4 waves in a 2 x 2 arrangement; each
// wave owns a (BM/2) x (BN/2) sub-tile made of 16 x 16 MFMA tiles
// (v_mfma_f64_16x16x4f64 for double, v_mfma_f32_16x16x4f32 for float).  K is
// consumed in chunks of 16 staged through LDS (k-major, so the 16 lanes that
// share a k read 16 consecutive rows / columns); the next chunk is prefetched
// into registers while the current one is multiplied.
//
// MFMA 16x16x4 operand layout (one wave, lane l):
//   A fragment: A[l % 16][l / 16]   (16 x 4)
//   B fragment: B[l / 16][l % 16]   (4 x 16)
//   D: 4 values per lane, v = 0..3:
//      f32: D[4 (l / 16) + v][l % 16]
//      f64: D[(l / 16) + 4 v][l % 16]   (rows interleaved)
// (both measured on gfx950 with row-sum / column-sum inputs).
#pragma once

#include "tf_common.h"

namespace tf {

typedef double tf_double4 __attribute__((ext_vector_type(4)));
typedef float tf_float4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef tf_double4 acc_t;
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef tf_float4 acc_t;
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T, int BM, int BN, int MINW = 1, int BK = 16>
__global__ __launch_bounds__(256, MINW) void k_gemm_mfma(int M, int N, int K, T* const* __restrict__ Ap, int lda,
                                                   T* const* __restrict__ Bp, int ldb, T* const* __restrict__ Cp,
                                                   int ldc, double alpha, double beta) {
  constexpr int WM = BM / 2, WN = BN / 2;      // wave sub-tile
  constexpr int TM = WM / 16, TN = WN / 16;    // MFMA tiles per wave
  constexpr int PA = BM + 4, PB_ = BN + 4;     // padded LDS rows
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;  // global loads per thread
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[BK * PA];
  __shared__ T Bs[BK * PB_];

  const T* __restrict__ A = Ap[blockIdx.z];
  const T* __restrict__ B = Bp[blockIdx.z];
  T* __restrict__ C = Cp[blockIdx.z];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;

  acc_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  T ra[LA], rb[LB];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = t + q * 256;
      const int r = e / BK, kk = e % BK;  // BK consecutive threads read one row's BK k
      const int gm = m0 + r, gk = k0 + kk;
      ra[q] = (gm < M && gk < K) ? A[(long)gm * lda + gk] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int e = t + q * 256;
      const int kk = e / BN, c = e % BN;
      const int gk = k0 + kk, gn = n0 + c;
      rb[q] = (gk < K && gn < N) ? B[(long)gk * ldb + gn] : (T)0;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < LA; ++q) {
      const int e = t + q * 256;
      As[(e % BK) * PA + e / BK] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
      const int e = t + q * 256;
      Bs[(e / BN) * PB_ + e % BN] = rb[q];
    }
  };

  load(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();  // previous chunk fully consumed
    store();
    __syncthreads();
    if (k0 + BK < K) load(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[(ks + lk) * PA + wm0 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[(ks + lk) * PB_ + wn0 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(a[i], b[j], acc[i][j]);
    }
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int gn = n0 + wn0 + j * 16 + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm0 + i * 16 + Mfma<T>::row(lk, v);
        if (gm < M && gn < N) {
          T* c = C + (long)gm * ldc + gn;
          const double prev = beta != 0.0 ? (double)*c : 0.0;
          *c = (T)(alpha * (double)acc[i][j][v] + beta * prev);
        }
      }
    }
}

}  // namespace tf

// ---- launch + tile autotuner ------------------------------------------------
//
// Candidate tiles: 128x128 and 64x64 (both hinted to two waves per SIMD) and
// 128x64.  The first call with a given (type, M, N, K, batch) times every
// candidate on scratch operands of that shape (every batch entry aliases one
// scratch matrix, so the scratch stays at one A, B and C), and the fastest is
// cached for the life of the process.  With tuning off a fixed rule is used.
namespace tf {

enum GemmTile : int { GEMM_128x128 = 0, GEMM_64x64 = 1, GEMM_128x64 = 2, GEMM_TILES = 3 };

template <typename T>
inline void gemm_launch(int tile, int M, int N, int K, T* const* A, int lda, T* const* B, int ldb, T* const* C,
                        int ldc, double alpha, double beta, int batch, hipStream_t st) {
  switch (tile) {
    case GEMM_128x128:
      hipLaunchKernelGGL((k_gemm_mfma<T, 128, 128, 2>), dim3((N + 127) / 128, (M + 127) / 128, batch), dim3(256), 0,
                         st, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
      break;
    case GEMM_128x64:
      hipLaunchKernelGGL((k_gemm_mfma<T, 128, 64, 1>), dim3((N + 63) / 64, (M + 127) / 128, batch), dim3(256), 0, st,
                         M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_mfma<T, 64, 64, 2>), dim3((N + 63) / 64, (M + 63) / 64, batch), dim3(256), 0, st, M,
                         N, K, A, lda, B, ldb, C, ldc, alpha, beta);
  }
}

int gemm_tuned_tile(int single, int M, int N, int K, int batch, hipStream_t st);

// C_b = alpha A_b B_b + beta C_b for b < batch (device pointer lists).
template <typename T>
inline void gemm_batched(int M, int N, int K, T* const* A, int lda, T* const* B, int ldb, T* const* C, int ldc,
                         double alpha, double beta, int batch, hipStream_t st) {
  if (batch <= 0 || M <= 0 || N <= 0) return;
  const int tile = gemm_tuned_tile(sizeof(T) == 4, M, N, K, batch, st);
  gemm_launch<T>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, batch, st);
}

}  // namespace tf
