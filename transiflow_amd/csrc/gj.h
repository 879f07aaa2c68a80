// Batched in-place inversion of dense m x m matrices by blocked Gauss-Jordan:
// 64-wide panels, each inverted by one workgroup with implicit partial pivoting
// (k_panel), then a rank-64 update of the whole matrix by two batched GEMMs on
// the matrix cores.  Shared by the line factorisations (bcr.hip, lineilu.hip).
#pragma once

#include "tf_common.h"

namespace tf {
namespace gj {

constexpr int PB = 64;  // Gauss-Jordan panel width

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane);
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Inverse of the pb x pb diagonal panel block of X_b at k0 (padded with the
// identity to 64): Gauss-Jordan on [A | I], implicit partial pivoting; 8 waves,
// lane = row, 16 augmented columns per wave in registers.  One block per matrix.
template <typename T>
__global__ __launch_bounds__(512) void k_panel(int m, int k0, int pb, T* const* __restrict__ Xp,
                                               T* __restrict__ Pinv, int* __restrict__ nperturb, double repl = 0.0) {
  __shared__ T colk[2][64];
  __shared__ T wmax[8];
  const T* X = Xp[blockIdx.x];
  T* P = Pinv + (long)blockIdx.x * PB * PB;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  T a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = wave * 16 + q;
    T v;
    if (c < 64)
      v = (lane < pb && c < pb) ? X[(long)(k0 + lane) * m + k0 + c] : (lane == c ? (T)1 : (T)0);
    else
      v = (lane == c - 64) ? (T)1 : (T)0;
    a[q] = v;
  }
  T mx = 0;
  if (wave < 4) {
#pragma unroll
    for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(a[q]));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) wmax[wave] = mx;
  __syncthreads();
  T amax = 0;
  for (int w = 0; w < 4; ++w) amax = fmax(amax, wmax[w]);
  if (amax == (T)0) amax = 1;
  const T tol = (sizeof(T) == 8 ? (T)1e-13 : (T)1e-6) * amax;
  bool used = false;
  int mystep = -1;
  for (int k = 0; k < 64; ++k) {
    if (wave == (k >> 4)) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q == (k & 15)) colk[k & 1][lane] = a[q];
    }
    __syncthreads();
    const T f = colk[k & 1][lane];
    // pivot: largest |f| among unused rows; 32-bit key (magnitude bits, lane) reduced with shuffles
    // (unused rows get keys >= 64, ties go to the lower lane; used rows key 0)
    unsigned key = 0;
    if (!used) {
      const unsigned bits = __float_as_uint((float)fabs(f));
      key = (bits & ~63u) + 64u + (unsigned)(63 - lane);
    }
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned o = __shfl_xor(key, off, 64);
      key = o > key ? o : key;
    }
    const int p = __builtin_amdgcn_readfirstlane(63 - (int)((key - 64u) & 63u));
    T piv = readlane_t<T>(f, p);
    if (fabs(piv) < tol) {
      // a (numerically) zero pivot becomes +-tol, or +-repl amax when the caller
      // asks for a regularising replacement (approximate factorisations)
      const T rep = repl > 0.0 ? (T)repl * amax : tol;
      piv = piv < (T)0 ? -rep : rep;
      if (threadIdx.x == 0) atomicAdd(nperturb, 1);
    }
    const T ip = (T)1 / piv;
    if (lane == p) {
      used = true;
      mystep = k;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const T pr = readlane_t<T>(a[q], p) * ip;
      a[q] = (lane == p) ? pr : a[q] - f * pr;
    }
  }
  if (wave >= 4 && mystep >= 0 && mystep < pb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = wave * 16 + q - 64;
      if (c < pb) P[mystep * pb + c] = a[q];
    }
  }
}

// Snapshots for panel K = [k0, k0+pb) of X_b:
//   Kr = X[K, :] with X[K, K] replaced by the identity   (pb x m)
//   Kc = X[:, K] with the rows K zeroed                    (m x pb)
// and X[:, K] is zeroed (the update rebuilds it).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_copy(int m, int k0, int pb, T* const* __restrict__ Xp,
                                                    T* __restrict__ Krb, T* __restrict__ Kcb) {
  T* X = Xp[blockIdx.y];
  T* Kr = Krb + (long)blockIdx.y * PB * m;
  T* Kc = Kcb + (long)blockIdx.y * m * PB;
  const long total = (long)pb * m;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int s = (int)(e / m), c = (int)(e % m);
    const bool inK = c >= k0 && c < k0 + pb;
    Kr[(long)s * m + c] = inK ? (c - k0 == s ? (T)1 : (T)0) : X[(long)(k0 + s) * m + c];
  }
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / pb), s = (int)(e % pb);
    const bool rowK = i >= k0 && i < k0 + pb;
    T* x = X + (long)i * m + k0 + s;
    Kc[(long)i * pb + s] = rowK ? (T)0 : *x;
    *x = (T)0;
  }
}

// X[K, :] = Tbuf  (after the rank-pb update)
template <typename T>
__global__ __launch_bounds__(256) void k_panel_store(int m, int k0, int pb, T* const* __restrict__ Xp,
                                                     const T* __restrict__ Tb) {
  T* X = Xp[blockIdx.y];
  const T* Tm = Tb + (long)blockIdx.y * PB * m;
  const long total = (long)pb * m;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
    X[(long)(k0 + e / m) * m + e % m] = Tm[e];
}

}  // namespace gj
}  // namespace tf
