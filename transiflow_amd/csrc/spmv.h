// CSR SpMV with the pressure gauge of the reference's direct solve
// (transiflow/interface/SciPy.py:95-103): column ``gauge`` is left out of every
// row and row ``gauge`` becomes -x[gauge].
//
// Two forms: spmv_rows (G lanes per row, shuffle reduction in a fixed order)
// and spmv_stream (a workgroup streams a row block's entries); spmv() picks one
// per value type.  Both are deterministic and accumulate in the vector type's
// precision (FP64 for FP64 vectors); both launch 256-thread workgroups.
#pragma once

#include "tf_common.h"

#include <type_traits>

namespace tf {

// XCD-aware block order: the hardware deals workgroups to the eight XCDs (each with its
// own L2) round-robin, so block b runs as logical block (b % 8) (nb / 8) + b / 8 and each
// XCD streams contiguous row ranges, fetching into its L2 only the part of x those rows
// gather.  A bijection on [0, nb) (the identity when nb is not a multiple of 8); the
// rows' arithmetic does not depend on which block runs them.  Measured against the plain
// order on repeated products with the same operands: 256^2 FP64 6.1 against 7.7 us, FP32
// 5.3 against 6.3 (there each XCD's L2 also keeps its rows between repetitions); 64^3
// 49.9 against 52.6 and 32.8 against 33.6; 1024^2 within 2% either way.
__device__ __forceinline__ int xcd_block(int b, int nb) { return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3); }

template <int G, typename TV, typename TX>
__device__ __forceinline__ void spmv_rows(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                          const TV* __restrict__ v, const TX* __restrict__ x, TX* __restrict__ y,
                                          int gauge) {
  // Grid-stride over row groups: a capped grid keeps each wavefront streaming
  // several rows instead of retiring after one.  The G lanes of a row take the
  // same trip count, so the shuffles always see a whole group.
  using Acc = typename std::conditional<sizeof(TX) == 8, double, float>::type;
  const int lane = threadIdx.x & (G - 1);
  const long stride = (long)gridDim.x * blockDim.x / G;
  const long b = xcd_block(blockIdx.x, gridDim.x);
  for (long row = (b * blockDim.x + threadIdx.x) / G; row < n; row += stride) {
    Acc acc = 0;
    if (row != gauge) {
      const int e = rp[row + 1];
      for (int k = rp[row] + lane; k < e; k += G) {
        const int c = ci[k];
        if (c != gauge) acc += (Acc)v[k] * (Acc)x[c];
      }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (lane == 0) y[row] = row == gauge ? -x[row] : (TX)acc;
  }
}

// CSR-stream: a workgroup of 256 threads takes SPMV_ROWS consecutive rows,
// whose entries are one contiguous range of the CSR arrays.  All threads
// stream that range with unit stride (every load coalesced, no idle lanes on
// short rows), store the products in LDS, and then one thread per row sums its
// products in column order.  A workgroup whose range exceeds SPMV_CAP entries
// (rows much longer than the stencils here) sums each row from global memory.
constexpr int SPMV_ROWS = 128;
constexpr int SPMV_CAP = 4096;

template <typename TV, typename TX>
__device__ __forceinline__ void spmv_stream(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                            const TV* __restrict__ v, const TX* __restrict__ x, TX* __restrict__ y,
                                            int gauge) {
  using Acc = typename std::conditional<sizeof(TX) == 8, double, float>::type;
  __shared__ Acc prod[SPMV_CAP];
  __shared__ int rps[SPMV_ROWS + 1];
  const int r0 = xcd_block(blockIdx.x, gridDim.x) * SPMV_ROWS;
  const int nr = min(SPMV_ROWS, n - r0);
  for (int t = threadIdx.x; t <= nr; t += blockDim.x) rps[t] = rp[r0 + t];
  __syncthreads();
  const int e0 = rps[0], e1 = rps[nr];
  const int t = threadIdx.x;
  if (e1 - e0 <= SPMV_CAP) {
    for (int e = e0 + t; e < e1; e += blockDim.x) {
      const int c = ci[e];
      prod[e - e0] = c != gauge ? (Acc)v[e] * (Acc)x[c] : (Acc)0;
    }
    __syncthreads();
    if (t < nr) {
      const int row = r0 + t;
      Acc s = 0;
      for (int e = rps[t]; e < rps[t + 1]; ++e) s += prod[e - e0];
      y[row] = row == gauge ? -x[row] : (TX)s;
    }
  } else if (t < nr) {
    const int row = r0 + t;
    Acc s = 0;
    for (int e = rps[t]; e < rps[t + 1]; ++e) {
      const int c = ci[e];
      s += c != gauge ? (Acc)v[e] * (Acc)x[c] : (Acc)0;
    }
    y[row] = row == gauge ? -x[row] : (TX)s;
  }
}

// The launched form per value type, as measured on the cavity Jacobians:
// FP32 values (the inner solves) stream (1024^2: 3.9 TB/s against 2.2 with 8
// lanes per row); FP64 values take 4 lanes per row (algorithmic bytes / time:
// 2D 1024^2 3.4 TB/s, 256^2 3.2 TB/s; 8 lanes 2.7 and 2.6, 16 lanes 2.0 and 1.9,
// streaming 2.4; 3D 64^3, 14 entries per row, 3.7 TB/s against 3.9 with 8
// lanes).  A larger grid cap (32768 blocks) was slower at 1024^2 (2.5 TB/s).
constexpr int SPMV_LANES = 4;
constexpr int SPMV_MAX_BLOCKS = 8192;

template <typename TV, typename TX>
__device__ __forceinline__ void spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                     const TV* __restrict__ v, const TX* __restrict__ x, TX* __restrict__ y, int gauge) {
  if constexpr (sizeof(TV) == 8)
    spmv_rows<SPMV_LANES>(n, rp, ci, v, x, y, gauge);
  else
    spmv_stream(n, rp, ci, v, x, y, gauge);
}

template <typename TV>
inline int spmv_blocks(int n) {
  if constexpr (sizeof(TV) == 8) {
    const long b = ((long)n * SPMV_LANES + 255) / 256;
    return (int)(b < SPMV_MAX_BLOCKS ? (b < 1 ? 1 : b) : SPMV_MAX_BLOCKS);
  } else {
    return (n + SPMV_ROWS - 1) / SPMV_ROWS;
  }
}

}  // namespace tf
