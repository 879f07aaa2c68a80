// Krylov building blocks for restarted (F)GMRES in FP64 and FP32.
//
// The Arnoldi vectors live in one row-major basis V[k][n].  Classical
// Gram-Schmidt with one re-orthogonalisation pass (CGS2) needs, per pass, a
// fused multi-dot h = V[0..k)^T w and a fused multi-axpy w -= V[0..k)^T h;
// each reads w once and every basis vector once.  Dot products are reduced
// deterministically (fixed partition, fixed tree order) in FP64 whatever the
// vector precision.  Scalars produced on the device (norms, Givens inputs)
// stay on the device; the host reads the small Hessenberg column once per
// iteration.
//
// Replaces scipy.sparse.linalg.gmres as used by the reference
// (transiflow/interface/SciPy.py:343-370) and the LU solve of
// SciPy.direct_solve (SciPy.py:204-258) when preconditioned by a direct factor.
#include "tf_common.h"

#include <type_traits>

namespace {

constexpr int MD_THREADS = 256;
constexpr int MD_MAX_BLOCKS = 1024;
constexpr int MAX_BASIS = 256;

constexpr int MD_WAVES = MD_THREADS / 64;
constexpr int MD_KB = 8;   // basis vectors per pass over w

// 16-byte vectors: each lane moves 16 bytes per load, so a wavefront keeps
// 1 KiB per basis vector in flight (streaming from HBM needs tens of KiB in
// flight per CU).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = f32x4; static constexpr int N = 4; };
template <> struct Vec16<double> { using type = f64x2; static constexpr int N = 2; };

template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The block count of a multidot over n entries: a fixed function of n (so the
// partition, and with it every sum, is deterministic), about two 16-byte
// loads per lane and basis vector, at most MD_MAX_BLOCKS.
int md_blocks(long n) {
  const long per = (long)MD_THREADS * 2 * 4;
  const long b = (n + per - 1) / per;
  return (int)(b < 64 ? 64 : (b > MD_MAX_BLOCKS ? MD_MAX_BLOCKS : b));
}

// Block b's sums of V[i] . w into partial[i * nb + b].  Accumulation in the
// vector type (FP32 vectors: FP32 accumulation inside each block), MD_KB
// basis vectors per pass so w is read once per MD_KB vectors, reduced by
// wavefront shuffles and then across the block's waves in a fixed order.
// VEC: 16-byte loads over entries [b0, b1) counted in vectors (rows of V and w
// 16-byte aligned, n a multiple of the vector width); else one entry per lane.
// A grid with gridDim.y == 2 does the same for a second system: blockIdx.y 1
// takes V1, w1 and partial1 (the other kernels below follow the same rule).
template <typename T, bool VEC>
__global__ __launch_bounds__(MD_THREADS) void k_multidot_partial(int n, int k, const T* __restrict__ V, long ldv,
                                                                 const T* __restrict__ w,
                                                                 double* __restrict__ partial,
                                                                 const T* __restrict__ V1 = nullptr,
                                                                 const T* __restrict__ w1 = nullptr,
                                                                 double* __restrict__ partial1 = nullptr) {
  if (blockIdx.y) {
    V = V1;
    w = w1;
    partial = partial1;
  }
  using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
  using VT = typename Vec16<T>::type;
  constexpr int VN = VEC ? Vec16<T>::N : 1;
  __shared__ Acc sh[MD_WAVES][MD_KB];
  const int nb = gridDim.x;
  const int units = n / VN;
  const int chunk = (units + nb - 1) / nb;
  const int b0 = blockIdx.x * chunk, b1 = min(units, b0 + chunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i0 = 0; i0 < k; i0 += MD_KB) {
    const int kb = min(MD_KB, k - i0);
    Acc acc[MD_KB];
#pragma unroll
    for (int q = 0; q < MD_KB; ++q) acc[q] = 0;
    for (int e = b0 + threadIdx.x; e < b1; e += MD_THREADS) {
      if constexpr (VEC) {
        const VT we = reinterpret_cast<const VT*>(w)[e];
        VT vq[MD_KB];
#pragma unroll
        for (int q = 0; q < MD_KB; ++q)
          if (q < kb) vq[q] = reinterpret_cast<const VT*>(V + (long)(i0 + q) * ldv)[e];
#pragma unroll
        for (int q = 0; q < MD_KB; ++q)
          if (q < kb) {
#pragma unroll
            for (int l = 0; l < VN; ++l) acc[q] += (Acc)vq[q][l] * (Acc)we[l];
          }
      } else {
        const Acc we = (Acc)w[e];
#pragma unroll
        for (int q = 0; q < MD_KB; ++q)
          if (q < kb) acc[q] += (Acc)V[(long)(i0 + q) * ldv + e] * we;
      }
    }
#pragma unroll
    for (int q = 0; q < MD_KB; ++q) {
      const Acc s = wave_sum(acc[q]);
      if (lane == 0) sh[wave][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < kb) {
      Acc s = 0;
      for (int v = 0; v < MD_WAVES; ++v) s += sh[v][threadIdx.x];
      partial[(long)(i0 + threadIdx.x) * nb + blockIdx.x] = (double)s;
    }
    __syncthreads();
  }
}

// h[i] (+)= sum_b partial[i][b]; one workgroup per i, wavefront reductions.
__global__ __launch_bounds__(MD_THREADS) void k_multidot_final(int k, int nb, const double* __restrict__ partial,
                                                               double* __restrict__ h, int accumulate,
                                                               const double* __restrict__ partial1 = nullptr,
                                                               double* __restrict__ h1 = nullptr) {
  if (blockIdx.y) {
    partial = partial1;
    h = h1;
  }
  __shared__ double sh[MD_WAVES];
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += MD_THREADS) acc += partial[(long)i * nb + b];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < MD_WAVES; ++v) s += sh[v];
    h[i] = accumulate ? h[i] + s : s;
  }
}

// w += sum_i c[i] V[i] with the coefficients c (sign applied) in LDS: accumulation
// in FP64 in the order i = 0, 1, ...; VEC: 16-byte loads, four basis vectors' loads
// issued together.  Grid-stride over the entries.
template <typename T, bool VEC>
__device__ __forceinline__ void axpy_body(int n, int k, const T* __restrict__ V, long ldv, const double* c,
                                          T* __restrict__ w) {
  using VT = typename Vec16<T>::type;
  constexpr int VN = VEC ? Vec16<T>::N : 1;
  const int units = n / VN;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < units; e += gridDim.x * blockDim.x) {
    if constexpr (VEC) {
      VT wv = reinterpret_cast<const VT*>(w)[e];
      double acc[VN];
#pragma unroll
      for (int l = 0; l < VN; ++l) acc[l] = (double)wv[l];
      int i = 0;
      for (; i + 4 <= k; i += 4) {
        VT a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = reinterpret_cast<const VT*>(V + (long)(i + q) * ldv)[e];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
          for (int l = 0; l < VN; ++l) acc[l] += c[i + q] * (double)a[q][l];
      }
      for (; i < k; ++i) {
        const VT a = reinterpret_cast<const VT*>(V + (long)i * ldv)[e];
#pragma unroll
        for (int l = 0; l < VN; ++l) acc[l] += c[i] * (double)a[l];
      }
#pragma unroll
      for (int l = 0; l < VN; ++l) wv[l] = (T)acc[l];
      reinterpret_cast<VT*>(w)[e] = wv;
    } else {
      double acc = (double)w[e];
      for (int i = 0; i < k; ++i) acc += c[i] * (double)V[(long)i * ldv + e];
      w[e] = (T)acc;
    }
  }
}

// w += sign * sum_i coef[i] V[i]   (coef on the device, FP64).
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void k_multiaxpy(int n, int k, const T* __restrict__ V, long ldv,
                                                   const double* __restrict__ coef, T* __restrict__ w, double sign,
                                                   const T* __restrict__ V1 = nullptr,
                                                   const double* __restrict__ coef1 = nullptr,
                                                   T* __restrict__ w1 = nullptr) {
  if (blockIdx.y) {
    V = V1;
    coef = coef1;
    w = w1;
  }
  __shared__ double c[MAX_BASIS];
  for (int i = threadIdx.x; i < k; i += blockDim.x) c[i] = sign * coef[i];
  __syncthreads();
  axpy_body<T, VEC>(n, k, V, ldv, c, w);
}

// ---- CGS2 with the block sums reduced where they are consumed --------------------
//
// The Arnoldi step's reductions for small vectors: instead of a separate kernel
// that sums the multidot's per-block partials, every block of the consuming kernel
// stages the k x nb partials in LDS and sums them itself (the same values in every
// block: for each i, b ascending, in FP64); block 0 stores them.  Used when the
// partials fit FIN_LDS doubles (n up to about 0.4M at restart 30), where the
// launches, not the bytes, set the step's time.
constexpr int FIN_LDS = 6144;

__device__ __forceinline__ void fin_sums(int k, int nb, const double* __restrict__ partial, double* sp, double* c) {
  for (int e = threadIdx.x; e < k * nb; e += blockDim.x) sp[e] = partial[e];
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += sp[i * nb + b];
    c[i] = s;
  }
  __syncthreads();
}

// c = the sums of partial (k x nb); block 0: hout[i] = c[i] (add: hout[i] += c[i]);
// then w -= sum_i c[i] V[i].  blockIdx.y 1 takes the second system.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void k_multiaxpy_fin(int n, int k, int nb, const T* __restrict__ V, long ldv,
                                                       const double* __restrict__ partial, double* __restrict__ hout,
                                                       int add, T* __restrict__ w, const T* __restrict__ V1,
                                                       const double* __restrict__ partial1,
                                                       double* __restrict__ hout1, T* __restrict__ w1) {
  if (blockIdx.y) {
    V = V1;
    partial = partial1;
    hout = hout1;
    w = w1;
  }
  __shared__ double sp[FIN_LDS];
  __shared__ double c[MAX_BASIS];
  fin_sums(k, nb, partial, sp, c);
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < k; i += blockDim.x) hout[i] = add ? hout[i] + c[i] : c[i];
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) c[i] = -c[i];
  __syncthreads();
  axpy_body<T, VEC>(n, k, V, ldv, c, w);
}

// nrm = sqrt(sum of partial[0..nb)) (block 0 stores it in h[k]), out = w / nrm.
// hhost (pinned host memory, or null): block 0 also writes the column h[0..k] there,
// so the host reads it without a copy after the launch.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void k_scale_fin(int n, int nb, int k, const double* __restrict__ partial,
                                                   double* __restrict__ h, double* __restrict__ hhost,
                                                   const T* __restrict__ w, T* __restrict__ out,
                                                   const double* __restrict__ partial1, double* __restrict__ h1,
                                                   double* __restrict__ hhost1, const T* __restrict__ w1,
                                                   T* __restrict__ out1) {
  if (blockIdx.y) {
    partial = partial1;
    h = h1;
    hhost = hhost1;
    w = w1;
    out = out1;
  }
  __shared__ double sp[FIN_LDS];
  __shared__ double c[1];
  fin_sums(1, nb, partial, sp, c);
  const double nrm = sqrt(c[0]);
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) h[k] = nrm;
    if (hhost) {
      // h[0..k) is final since the previous launch
      for (int i = threadIdx.x; i < k; i += blockDim.x) hhost[i] = h[i];
      if (threadIdx.x == 0) hhost[k] = nrm;
    }
  }
  using VT = typename Vec16<T>::type;
  constexpr int VN = VEC ? Vec16<T>::N : 1;
  const double inv = 1.0 / nrm;
  const int units = n / VN;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < units; e += gridDim.x * blockDim.x) {
    if constexpr (VEC) {
      VT v = reinterpret_cast<const VT*>(w)[e];
#pragma unroll
      for (int l = 0; l < VN; ++l) v[l] = (T)((double)v[l] * inv);
      reinterpret_cast<VT*>(out)[e] = v;
    } else {
      out[e] = (T)((double)w[e] * inv);
    }
  }
}

// out = w / nrm where nrm is a device scalar (the next Arnoldi vector).
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void k_scale_by_inv(int n, const T* __restrict__ w, const double* __restrict__ nrm,
                                                      T* __restrict__ out, const T* __restrict__ w1 = nullptr,
                                                      const double* __restrict__ nrm1 = nullptr,
                                                      T* __restrict__ out1 = nullptr) {
  if (blockIdx.y) {
    w = w1;
    nrm = nrm1;
    out = out1;
  }
  using VT = typename Vec16<T>::type;
  constexpr int VN = VEC ? Vec16<T>::N : 1;
  const double inv = 1.0 / *nrm;
  const int units = n / VN;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < units; e += gridDim.x * blockDim.x) {
    if constexpr (VEC) {
      VT v = reinterpret_cast<const VT*>(w)[e];
#pragma unroll
      for (int l = 0; l < VN; ++l) v[l] = (T)((double)v[l] * inv);
      reinterpret_cast<VT*>(out)[e] = v;
    } else {
      out[e] = (T)((double)w[e] * inv);
    }
  }
}

// sqrt(sum partial[0..nb)) into out
__global__ __launch_bounds__(MD_THREADS) void k_norm_final(int nb, const double* __restrict__ partial,
                                                           double* __restrict__ out,
                                                           const double* __restrict__ partial1 = nullptr,
                                                           double* __restrict__ out1 = nullptr) {
  if (blockIdx.y) {
    partial = partial1;
    out = out1;
  }
  __shared__ double sh[MD_WAVES];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += MD_THREADS) acc += partial[b];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < MD_WAVES; ++v) s += sh[v];
    *out = sqrt(s);
  }
}

// 16-byte access is possible when n and the leading dimension are multiples of
// the vector width and the pointers are 16-byte aligned.
template <typename T>
bool vec_ok(int n, long ldv, const void* a, const void* b) {
  constexpr int VN = Vec16<T>::N;
  return n % VN == 0 && ldv % VN == 0 && ((uintptr_t)a % 16) == 0 && ((uintptr_t)b % 16) == 0;
}

// ---------------------------------------------------------------------------
// Cell-block Jacobi: invert each cell's dof x dof diagonal block of the
// gauged matrix (Gauss-Jordan with partial pivoting in registers).  A block
// that is singular (e.g. the lid-driven cavity's north-east corner cell,
// whose continuity row has no own-cell entries) gets a unit pivot for the
// dependent variable, i.e. that component is passed through unscaled.
// ---------------------------------------------------------------------------
template <int DOF>
__global__ __launch_bounds__(256) void k_cellblock_setup(int ncell, const int* __restrict__ rp,
                                                         const int* __restrict__ ci, const double* __restrict__ v,
                                                         int gauge, double* __restrict__ inv) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncell) return;
  double a[DOF][DOF], b[DOF][DOF];
  const int r0 = cell * DOF;
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int c = 0; c < DOF; ++c) {
      a[r][c] = 0.0;
      b[r][c] = r == c ? 1.0 : 0.0;
    }
  double amax = 0.0;
#pragma unroll
  for (int r = 0; r < DOF; ++r) {
    const int row = r0 + r;
    if (row == gauge) {
      a[r][r] = -1.0;
      continue;
    }
    for (int k = rp[row]; k < rp[row + 1]; ++k) {
      const int c = ci[k] - r0;
      if (c >= 0 && c < DOF && ci[k] != gauge) {
#pragma unroll
        for (int cc = 0; cc < DOF; ++cc)
          if (cc == c) a[r][cc] = v[k];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int c = 0; c < DOF; ++c) amax = fmax(amax, fabs(a[r][c]));
  const double tiny = 1e-14 * (amax > 0 ? amax : 1.0);
#pragma unroll
  for (int k = 0; k < DOF; ++k) {
    int p = k;
    double best = fabs(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < DOF; ++r)
      if (fabs(a[r][k]) > best) {
        best = fabs(a[r][k]);
        p = r;
      }
#pragma unroll
    for (int r = k + 1; r < DOF; ++r)
      if (r == p) {
#pragma unroll
        for (int c = 0; c < DOF; ++c) {
          double t = a[k][c]; a[k][c] = a[r][c]; a[r][c] = t;
          t = b[k][c]; b[k][c] = b[r][c]; b[r][c] = t;
        }
      }
    double piv = a[k][k];
    if (fabs(piv) <= tiny) {
      piv = 1.0;
      a[k][k] = 1.0;
    }
    const double ip = 1.0 / piv;
#pragma unroll
    for (int c = 0; c < DOF; ++c) {
      a[k][c] *= ip;
      b[k][c] *= ip;
    }
#pragma unroll
    for (int r = 0; r < DOF; ++r)
      if (r != k) {
        const double f = a[r][k];
#pragma unroll
        for (int c = 0; c < DOF; ++c) {
          a[r][c] -= f * a[k][c];
          b[r][c] -= f * b[k][c];
        }
      }
  }
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int c = 0; c < DOF; ++c) inv[((long)cell * DOF + r) * DOF + c] = b[r][c];
}

template <int DOF, typename T>
__global__ __launch_bounds__(256) void k_cellblock_apply(int ncell, const double* __restrict__ inv,
                                                         const T* __restrict__ r, T* __restrict__ z) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncell) return;
  double x[DOF];
#pragma unroll
  for (int c = 0; c < DOF; ++c) x[c] = (double)r[cell * DOF + c];
#pragma unroll
  for (int rr = 0; rr < DOF; ++rr) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < DOF; ++c) acc += inv[((long)cell * DOF + rr) * DOF + c] * x[c];
    z[cell * DOF + rr] = (T)acc;
  }
}

int grid_for(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 2048 ? (g < 1 ? 1 : g) : 2048);
}

template <typename T>
int multidot(int n, int k, const T* V, long ldv, const T* w, double* work, double* h, int accumulate,
             hipStream_t st) {
  if (k < 1 || k > MAX_BASIS) return tf::fail("multidot: bad basis size");
  const int nb = md_blocks(n);
  if (vec_ok<T>(n, ldv, V, w))
    hipLaunchKernelGGL((k_multidot_partial<T, true>), dim3(nb), dim3(MD_THREADS), 0, st, n, k, V, ldv, w, work);
  else
    hipLaunchKernelGGL((k_multidot_partial<T, false>), dim3(nb), dim3(MD_THREADS), 0, st, n, k, V, ldv, w, work);
  hipLaunchKernelGGL(k_multidot_final, dim3(k), dim3(MD_THREADS), 0, st, k, nb, work, h, accumulate);
  TF_LAUNCH_CHECK("multidot");
  return 0;
}

template <typename T>
int multiaxpy(int n, int k, const T* V, long ldv, const double* coef, T* w, double sign, hipStream_t st) {
  if (k < 0 || k > MAX_BASIS) return tf::fail("multiaxpy: bad basis size");
  if (k == 0) return 0;
  if (vec_ok<T>(n, ldv, V, w))
    hipLaunchKernelGGL((k_multiaxpy<T, true>), dim3(grid_for(n / Vec16<T>::N)), dim3(256), 0, st, n, k, V, ldv,
                       coef, w, sign);
  else
    hipLaunchKernelGGL((k_multiaxpy<T, false>), dim3(grid_for(n)), dim3(256), 0, st, n, k, V, ldv, coef, w, sign);
  TF_LAUNCH_CHECK("multiaxpy");
  return 0;
}

template <typename T>
int normalize(int n, const T* w, double* work, double* nrm_dev, T* out, hipStream_t st) {
  const int nb = md_blocks(n);
  if (vec_ok<T>(n, n, w, w))
    hipLaunchKernelGGL((k_multidot_partial<T, true>), dim3(nb), dim3(MD_THREADS), 0, st, n, 1, w, (long)n, w, work);
  else
    hipLaunchKernelGGL((k_multidot_partial<T, false>), dim3(nb), dim3(MD_THREADS), 0, st, n, 1, w, (long)n, w, work);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(MD_THREADS), 0, st, nb, work, nrm_dev);
  if (vec_ok<T>(n, n, w, out))
    hipLaunchKernelGGL((k_scale_by_inv<T, true>), dim3(grid_for(n / Vec16<T>::N)), dim3(256), 0, st, n, w, nrm_dev,
                       out);
  else
    hipLaunchKernelGGL((k_scale_by_inv<T, false>), dim3(grid_for(n)), dim3(256), 0, st, n, w, nrm_dev, out);
  TF_LAUNCH_CHECK("normalize");
  return 0;
}

// h[i] += y[i], i <= j, for two systems (blockIdx.x 0 and 1)
__global__ void k_add_pair(int k, double* h0, const double* y0, double* h1, const double* y1) {
  double* h = blockIdx.x ? h1 : h0;
  const double* y = blockIdx.x ? y1 : y0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) h[i] = h[i] + y[i];
}

// One CGS2 orthogonalisation and normalisation of step j for nsys (1 or 2) systems
// in the same launches (blockIdx.y picks the system; the second system's pointers
// are unused for one): h = V^T w, w -= V h, y = V^T w, w -= V y, h += y,
// h[j+1] = |w|, V[j+1] = w / |w|.  With the partials in LDS reach (k nb <= FIN_LDS)
// the sums are reduced by the consuming kernels (k_multiaxpy_fin, k_scale_fin:
// five launches); else by separate kernels (k_multidot_final, k_add_pair,
// k_norm_final: eight).  The choice depends on n and j only, so one system and
// two systems take the same arithmetic.  hh0, hh1 (pinned host memory, or null):
// on the five-launch path the last launch also writes each column h[0..j+1] there
// (*to_host = 1); otherwise *to_host = 0 and the caller copies it.
template <typename T>
int cgs2_sys(int nsys, int n, int j, const T* V0, const T* V1, T* w0, T* w1, double* work0, double* work1,
             double* h0, double* h1, double* y0, double* y1, T* v0, T* v1, double* hh0, double* hh1, int* to_host,
             hipStream_t st) {
  const int k = j + 1;
  *to_host = 0;
  if (k < 1 || k + 1 > MAX_BASIS) return tf::fail("cgs2: bad basis size");
  if (nsys == 1) {
    V1 = V0;
    w1 = w0;
    work1 = work0;
    h1 = h0;
    y1 = y0;
    v1 = v0;
    hh1 = hh0;
  }
  const int nb = md_blocks(n);
  const bool vec = vec_ok<T>(n, n, V0, w0) && vec_ok<T>(n, n, V1, w1) && vec_ok<T>(n, n, v0, v1);
  const bool fused = (long)k * nb <= FIN_LDS;
  const dim3 md(nb, nsys), blk(MD_THREADS), fin(k, nsys);
  const dim3 ew(grid_for(vec ? n / Vec16<T>::N : n), nsys);
  for (int pass = 0; pass < 2; ++pass) {
    double* c0 = pass ? y0 : h0;
    double* c1 = pass ? y1 : h1;
    if (vec)
      hipLaunchKernelGGL((k_multidot_partial<T, true>), md, blk, 0, st, n, k, V0, (long)n, (const T*)w0, work0, V1,
                         (const T*)w1, work1);
    else
      hipLaunchKernelGGL((k_multidot_partial<T, false>), md, blk, 0, st, n, k, V0, (long)n, (const T*)w0, work0, V1,
                         (const T*)w1, work1);
    if (fused) {
      // pass 0: h = sums; pass 1: h += sums (y is not needed further)
      if (vec)
        hipLaunchKernelGGL((k_multiaxpy_fin<T, true>), ew, dim3(256), 0, st, n, k, nb, V0, (long)n,
                           (const double*)work0, h0, pass, w0, V1, (const double*)work1, h1, w1);
      else
        hipLaunchKernelGGL((k_multiaxpy_fin<T, false>), ew, dim3(256), 0, st, n, k, nb, V0, (long)n,
                           (const double*)work0, h0, pass, w0, V1, (const double*)work1, h1, w1);
      continue;
    }
    hipLaunchKernelGGL(k_multidot_final, fin, blk, 0, st, k, nb, (const double*)work0, c0, 0, (const double*)work1,
                       c1);
    if (vec)
      hipLaunchKernelGGL((k_multiaxpy<T, true>), ew, dim3(256), 0, st, n, k, V0, (long)n, (const double*)c0, w0, -1.0,
                         V1, (const double*)c1, w1);
    else
      hipLaunchKernelGGL((k_multiaxpy<T, false>), ew, dim3(256), 0, st, n, k, V0, (long)n, (const double*)c0, w0,
                         -1.0, V1, (const double*)c1, w1);
  }
  if (!fused) hipLaunchKernelGGL(k_add_pair, dim3(nsys), dim3(64), 0, st, k, h0, (const double*)y0, h1, (const double*)y1);
  // |w| into h[j+1], V[j+1] = w / |w|
  if (vec)
    hipLaunchKernelGGL((k_multidot_partial<T, true>), md, blk, 0, st, n, 1, (const T*)w0, (long)n, (const T*)w0,
                       work0, (const T*)w1, (const T*)w1, work1);
  else
    hipLaunchKernelGGL((k_multidot_partial<T, false>), md, blk, 0, st, n, 1, (const T*)w0, (long)n, (const T*)w0,
                       work0, (const T*)w1, (const T*)w1, work1);
  if (fused) {
    if (vec)
      hipLaunchKernelGGL((k_scale_fin<T, true>), ew, dim3(256), 0, st, n, nb, k, (const double*)work0, h0, hh0,
                         (const T*)w0, v0, (const double*)work1, h1, hh1, (const T*)w1, v1);
    else
      hipLaunchKernelGGL((k_scale_fin<T, false>), ew, dim3(256), 0, st, n, nb, k, (const double*)work0, h0, hh0,
                         (const T*)w0, v0, (const double*)work1, h1, hh1, (const T*)w1, v1);
    TF_LAUNCH_CHECK("cgs2");
    *to_host = hh0 != nullptr && hh1 != nullptr;
    return 0;
  }
  hipLaunchKernelGGL(k_norm_final, dim3(1, nsys), blk, 0, st, nb, (const double*)work0, h0 + k, (const double*)work1,
                     h1 + k);
  if (vec)
    hipLaunchKernelGGL((k_scale_by_inv<T, true>), ew, dim3(256), 0, st, n, (const T*)w0, (const double*)(h0 + k), v0,
                       (const T*)w1, (const double*)(h1 + k), v1);
  else
    hipLaunchKernelGGL((k_scale_by_inv<T, false>), ew, dim3(256), 0, st, n, (const T*)w0, (const double*)(h0 + k),
                       v0, (const T*)w1, (const double*)(h1 + k), v1);
  TF_LAUNCH_CHECK("cgs2");
  return 0;
}

}  // namespace

namespace tf {
// One CGS2 Arnoldi orthogonalisation and normalisation (cgs2_sys) for two systems
// in the same launches, or for one; the same arithmetic either way.
int cgs2_pair_f64(int n, int j, const double* V0, const double* V1, double* w0, double* w1, double* work0,
                  double* work1, double* h0, double* h1, double* y0, double* y1, double* v0, double* v1,
                  double* hh0, double* hh1, int* to_host, hipStream_t st) {
  return cgs2_sys<double>(2, n, j, V0, V1, w0, w1, work0, work1, h0, h1, y0, y1, v0, v1, hh0, hh1, to_host, st);
}
int cgs2_pair_f32(int n, int j, const float* V0, const float* V1, float* w0, float* w1, double* work0,
                  double* work1, double* h0, double* h1, double* y0, double* y1, float* v0, float* v1,
                  double* hh0, double* hh1, int* to_host, hipStream_t st) {
  return cgs2_sys<float>(2, n, j, V0, V1, w0, w1, work0, work1, h0, h1, y0, y1, v0, v1, hh0, hh1, to_host, st);
}
// the same for one system
int cgs2_one_f64(int n, int j, const double* V, double* w, double* work, double* h, double* y, double* v,
                 double* hh, int* to_host, hipStream_t st) {
  return cgs2_sys<double>(1, n, j, V, nullptr, w, nullptr, work, nullptr, h, nullptr, y, nullptr, v, nullptr, hh,
                          nullptr, to_host, st);
}
int cgs2_one_f32(int n, int j, const float* V, float* w, double* work, double* h, double* y, float* v,
                 double* hh, int* to_host, hipStream_t st) {
  return cgs2_sys<float>(1, n, j, V, nullptr, w, nullptr, work, nullptr, h, nullptr, y, nullptr, v, nullptr, hh,
                         nullptr, to_host, st);
}
}  // namespace tf

extern "C" {

int tf_krylov_work_size(void) { return MD_MAX_BLOCKS * MAX_BASIS; }
int tf_krylov_max_basis(void) { return MAX_BASIS; }

// h[0..k) (+)= V[0..k)^T w     (V row-major with leading dimension ldv)
int tf_multidot_f64(int n, int k, const double* V, long ldv, const double* w, double* work, double* h,
                    int accumulate, void* stream) {
  return multidot<double>(n, k, V, ldv, w, work, h, accumulate, (hipStream_t)stream);
}
int tf_multidot_f32(int n, int k, const float* V, long ldv, const float* w, double* work, double* h, int accumulate,
                    void* stream) {
  return multidot<float>(n, k, V, ldv, w, work, h, accumulate, (hipStream_t)stream);
}

// w += sign * sum_i coef[i] V[i]
int tf_multiaxpy_f64(int n, int k, const double* V, long ldv, const double* coef, double* w, double sign,
                     void* stream) {
  return multiaxpy<double>(n, k, V, ldv, coef, w, sign, (hipStream_t)stream);
}
int tf_multiaxpy_f32(int n, int k, const float* V, long ldv, const double* coef, float* w, double sign,
                     void* stream) {
  return multiaxpy<float>(n, k, V, ldv, coef, w, sign, (hipStream_t)stream);
}

// nrm_dev = ||w||_2 (deterministic), then out = w / nrm_dev
int tf_normalize_f64(int n, const double* w, double* work, double* nrm_dev, double* out, void* stream) {
  return normalize<double>(n, w, work, nrm_dev, out, (hipStream_t)stream);
}
int tf_normalize_f32(int n, const float* w, double* work, double* nrm_dev, float* out, void* stream) {
  return normalize<float>(n, w, work, nrm_dev, out, (hipStream_t)stream);
}

int tf_cellblock_setup(int ncell, int dof, const int* rp, const int* ci, const double* v, int gauge, double* inv,
                       void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((ncell + 255) / 256), b(256);
  switch (dof) {
    case 3: hipLaunchKernelGGL(k_cellblock_setup<3>, g, b, 0, st, ncell, rp, ci, v, gauge, inv); break;
    case 4: hipLaunchKernelGGL(k_cellblock_setup<4>, g, b, 0, st, ncell, rp, ci, v, gauge, inv); break;
    case 5: hipLaunchKernelGGL(k_cellblock_setup<5>, g, b, 0, st, ncell, rp, ci, v, gauge, inv); break;
    default: return tf::fail("cellblock: unsupported dof");
  }
  TF_LAUNCH_CHECK("cellblock setup");
  return 0;
}

int tf_cellblock_apply_f64(int ncell, int dof, const double* inv, const double* r, double* z, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((ncell + 255) / 256), b(256);
  switch (dof) {
    case 3: hipLaunchKernelGGL((k_cellblock_apply<3, double>), g, b, 0, st, ncell, inv, r, z); break;
    case 4: hipLaunchKernelGGL((k_cellblock_apply<4, double>), g, b, 0, st, ncell, inv, r, z); break;
    case 5: hipLaunchKernelGGL((k_cellblock_apply<5, double>), g, b, 0, st, ncell, inv, r, z); break;
    default: return tf::fail("cellblock: unsupported dof");
  }
  TF_LAUNCH_CHECK("cellblock apply");
  return 0;
}

int tf_cellblock_apply_f32(int ncell, int dof, const double* inv, const float* r, float* z, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((ncell + 255) / 256), b(256);
  switch (dof) {
    case 3: hipLaunchKernelGGL((k_cellblock_apply<3, float>), g, b, 0, st, ncell, inv, r, z); break;
    case 4: hipLaunchKernelGGL((k_cellblock_apply<4, float>), g, b, 0, st, ncell, inv, r, z); break;
    case 5: hipLaunchKernelGGL((k_cellblock_apply<5, float>), g, b, 0, st, ncell, inv, r, z); break;
    default: return tf::fail("cellblock: unsupported dof");
  }
  TF_LAUNCH_CHECK("cellblock apply");
  return 0;
}

}  // extern "C"
