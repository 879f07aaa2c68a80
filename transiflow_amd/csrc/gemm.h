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

// FP32 operands as sums of BF16 terms for v_mfma_f32_16x16x16bf16_1k (lane l):
//   a[i] = A[l % 16][4 (l / 16) + i], b[i] = B[4 (l / 16) + i][l % 16],
//   D[4 (l / 16) + v][l % 16] (the f32 layout), measured on gfx950.
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to nearest even): the
// three-term product hi*hi + hi*lo + lo*hi drops lo*lo and the rounding of lo,
// about 2^-16 relative per product.
typedef short tf_bf16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned short tf_bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void tf_split4(const float v[4], tf_bf16x4& hi, tf_bf16x4& lo) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const unsigned short h = tf_bf16_rne(v[c]);
    hi[c] = (short)h;
    lo[c] = (short)tf_bf16_rne(v[c] - __uint_as_float((unsigned)h << 16));
  }
}

// SPLIT 0: the type's own MFMA; SPLIT 1 or 3 (float only): BF16 terms of the
// FP32 operands, one (hi*hi) or three products, FP32 accumulation.
template <typename T, int BM, int BN, int MINW = 1, int BK = 16, int SPLIT = 0>
__global__ __launch_bounds__(256, MINW) void k_gemm_mfma(int M, int N, int K, T* const* __restrict__ Ap, int lda,
                                                   T* const* __restrict__ Bp, int ldb, T* const* __restrict__ Cp,
                                                   int ldc, double alpha, double beta) {
  constexpr int WM = BM / 2, WN = BN / 2;      // wave sub-tile
  constexpr int TM = WM / 16, TN = WN / 16;    // MFMA tiles per wave
  constexpr int PA = BM + 4, PB_ = BN + 4;     // padded LDS rows
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;  // global loads per thread
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[BK * PA];
  __shared__ __align__(16) T Bs[BK * PB_];

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
  // FP32 operands with 16-byte aligned rows: four consecutive k (A) or n (B) per
  // thread in one load; B goes to LDS with one 16-byte store (uniform per block)
  constexpr bool V4 = sizeof(T) == 4 && LA % 4 == 0 && LB % 4 == 0 && BK % 4 == 0;
  const bool v4 = V4 && ((((uintptr_t)A | (uintptr_t)B) & 15) == 0) && lda % 4 == 0 && ldb % 4 == 0 && K % 4 == 0 &&
                  N % 4 == 0;
  typedef T T4 __attribute__((ext_vector_type(4)));
  auto load = [&](int k0) {
    if constexpr (V4) {
      if (v4) {
#pragma unroll
        for (int q = 0; q < LA / 4; ++q) {
          const int e = t + q * 256;
          const int r = e / (BK / 4), kq = e % (BK / 4);
          const int gm = m0 + r, gk = k0 + 4 * kq;
          T4 v = T4{0, 0, 0, 0};
          if (gm < M && gk < K) v = *(const T4*)(A + (long)gm * lda + gk);
#pragma unroll
          for (int j = 0; j < 4; ++j) ra[4 * q + j] = v[j];
        }
#pragma unroll
        for (int q = 0; q < LB / 4; ++q) {
          const int e = t + q * 256;
          const int kk = e / (BN / 4), c4 = e % (BN / 4);
          const int gk = k0 + kk, gn = n0 + 4 * c4;
          T4 v = T4{0, 0, 0, 0};
          if (gk < K && gn < N) v = *(const T4*)(B + (long)gk * ldb + gn);
#pragma unroll
          for (int j = 0; j < 4; ++j) rb[4 * q + j] = v[j];
        }
        return;
      }
    }
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
    if constexpr (V4) {
      if (v4) {
#pragma unroll
        for (int q = 0; q < LA / 4; ++q) {
          const int e = t + q * 256;
          const int r = e / (BK / 4), kq = e % (BK / 4);
#pragma unroll
          for (int j = 0; j < 4; ++j) As[(4 * kq + j) * PA + r] = ra[4 * q + j];
        }
#pragma unroll
        for (int q = 0; q < LB / 4; ++q) {
          const int e = t + q * 256;
          const int kk = e / (BN / 4), c4 = e % (BN / 4);
          *(T4*)(Bs + kk * PB_ + 4 * c4) = T4{rb[4 * q], rb[4 * q + 1], rb[4 * q + 2], rb[4 * q + 3]};
        }
        return;
      }
    }
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
    if constexpr (SPLIT == 0) {
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
    } else {
      static_assert(sizeof(T) == 4 && BK % 16 == 0, "BF16 terms need FP32 operands and 16-wide k chunks");
#pragma unroll
      for (int ks = 0; ks < BK; ks += 16) {
        tf_bf16x4 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          float v[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) v[c] = As[(ks + 4 * lk + c) * PA + wm0 + i * 16 + lr];
          tf_split4(v, ah[i], al[i]);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          float v[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) v[c] = Bs[(ks + 4 * lk + c) * PB_ + wn0 + j * 16 + lr];
          tf_split4(v, bh[j], bl[j]);
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if constexpr (SPLIT == 3) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[i], bh[j], acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bl[j], acc[i][j], 0, 0, 0);
            }
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[i], bh[j], acc[i][j], 0, 0, 0);
          }
      }
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

template <typename T, int SPLIT = 0>
inline void gemm_launch(int tile, int M, int N, int K, T* const* A, int lda, T* const* B, int ldb, T* const* C,
                        int ldc, double alpha, double beta, int batch, hipStream_t st) {
  switch (tile) {
    case GEMM_128x128:
      hipLaunchKernelGGL((k_gemm_mfma<T, 128, 128, 2, 16, SPLIT>), dim3((N + 127) / 128, (M + 127) / 128, batch),
                         dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
      break;
    case GEMM_128x64:
      hipLaunchKernelGGL((k_gemm_mfma<T, 128, 64, 1, 16, SPLIT>), dim3((N + 63) / 64, (M + 127) / 128, batch),
                         dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_mfma<T, 64, 64, 2, 16, SPLIT>), dim3((N + 63) / 64, (M + 63) / 64, batch),
                         dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
  }
}

int gemm_tuned_tile(int single, int M, int N, int K, int batch, hipStream_t st);

// FP32 GEMM on BF16 terms of the operands (split 1: hi*hi; 3: hi*hi + hi*lo + lo*hi),
// with the tile the tuner picked for the FP32 kernel of the same shape.
inline void gemm_batched_split(int split, int M, int N, int K, float* const* A, int lda, float* const* B, int ldb,
                               float* const* C, int ldc, double alpha, double beta, int batch, hipStream_t st) {
  if (batch <= 0 || M <= 0 || N <= 0) return;
  const int tile = gemm_tuned_tile(1, M, N, K, batch, st);
  if (split == 1)
    gemm_launch<float, 1>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, batch, st);
  else
    gemm_launch<float, 3>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, batch, st);
}

// C_b = alpha A_b B_b + beta C_b for b < batch (device pointer lists).
template <typename T>
inline void gemm_batched(int M, int N, int K, T* const* A, int lda, T* const* B, int ldb, T* const* C, int ldc,
                         double alpha, double beta, int batch, hipStream_t st) {
  if (batch <= 0 || M <= 0 || N <= 0) return;
  const int tile = gemm_tuned_tile(sizeof(T) == 4, M, N, K, batch, st);
  gemm_launch<T>(tile, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, batch, st);
}

}  // namespace tf
