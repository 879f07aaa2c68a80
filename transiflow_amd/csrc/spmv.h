// CSR SpMV with the pressure gauge of the reference's direct solve
// (transiflow/interface/SciPy.py:95-103): column ``gauge`` is left out of every
// row and row ``gauge`` becomes -x[gauge].
//
// G lanes per row (a power of two <= 64): the rows of one wavefront are
// consecutive, so their entries -- contiguous in the CSR arrays -- are read
// coalesced; each lane sums a strided part of its row and the G partial sums
// are reduced with shuffles in a fixed order (deterministic).  Accumulation in
// the vector type's precision (FP64 for FP64 vectors).
#pragma once

#include "tf_common.h"

#include <type_traits>

namespace tf {

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
  for (long row = (blockIdx.x * (long)blockDim.x + threadIdx.x) / G; row < n; row += stride) {
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

constexpr int SPMV_LANES = 8;          // ~9 entries per row on the cavity Jacobians
constexpr int SPMV_MAX_BLOCKS = 8192;  // 32 workgroups of 256 per CU

inline int spmv_blocks(int n) {
  const long b = ((long)n * SPMV_LANES + 255) / 256;
  return (int)(b < SPMV_MAX_BLOCKS ? (b < 1 ? 1 : b) : SPMV_MAX_BLOCKS);
}

}  // namespace tf
