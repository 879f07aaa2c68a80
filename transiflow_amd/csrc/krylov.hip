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
constexpr int MD_BLOCKS = 256;   // fixed partition -> deterministic sums
constexpr int MAX_BASIS = 256;

constexpr int MD_WAVES = MD_THREADS / 64;
constexpr int MD_KB = 8;   // basis vectors per pass over w

template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(MD_THREADS) void k_multidot_partial(int n, int k, const T* __restrict__ V, long ldv,
                                                                 const T* __restrict__ w,
                                                                 double* __restrict__ partial) {
  // partial[i * MD_BLOCKS + b] = sum over block b's slice of V[i] . w, accumulated
  // in the vector type (FP32 vectors: FP32 accumulation inside each block).
  // MD_KB basis vectors per pass, so w is read once per MD_KB vectors; sums are
  // reduced by wavefront shuffles, then across the block's waves in a fixed order.
  using Acc = typename std::conditional<sizeof(T) == 8, double, float>::type;
  __shared__ Acc sh[MD_WAVES][MD_KB];
  const int chunk = (n + MD_BLOCKS - 1) / MD_BLOCKS;
  const int b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i0 = 0; i0 < k; i0 += MD_KB) {
    const int kb = min(MD_KB, k - i0);
    Acc acc[MD_KB];
#pragma unroll
    for (int q = 0; q < MD_KB; ++q) acc[q] = 0;
    for (int e = b0 + threadIdx.x; e < b1; e += MD_THREADS) {
      const Acc we = (Acc)w[e];
#pragma unroll
      for (int q = 0; q < MD_KB; ++q)
        if (q < kb) acc[q] += (Acc)V[(long)(i0 + q) * ldv + e] * we;
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
      partial[(long)(i0 + threadIdx.x) * MD_BLOCKS + blockIdx.x] = (double)s;
    }
    __syncthreads();
  }
}

// h[i] (+)= sum_b partial[i][b]; one workgroup per i, wavefront reductions.
__global__ __launch_bounds__(MD_THREADS) void k_multidot_final(int k, const double* __restrict__ partial,
                                                               double* __restrict__ h, int accumulate) {
  __shared__ double sh[MD_WAVES];
  const int i = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < MD_BLOCKS; b += MD_THREADS) acc += partial[(long)i * MD_BLOCKS + b];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < MD_WAVES; ++v) s += sh[v];
    h[i] = accumulate ? h[i] + s : s;
  }
}

// w -= sum_i coef[i] V[i]   (coef on the device, FP64)
template <typename T>
__global__ __launch_bounds__(256) void k_multiaxpy(int n, int k, const T* __restrict__ V, long ldv,
                                                   const double* __restrict__ coef, T* __restrict__ w, double sign) {
  __shared__ double c[MAX_BASIS];
  for (int i = threadIdx.x; i < k; i += blockDim.x) c[i] = coef[i];
  __syncthreads();
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double acc = (double)w[e];
    for (int i = 0; i < k; ++i) acc += sign * c[i] * (double)V[(long)i * ldv + e];
    w[e] = (T)acc;
  }
}

// out = w / nrm where nrm is a device scalar (the next Arnoldi vector).
template <typename T>
__global__ __launch_bounds__(256) void k_scale_by_inv(int n, const T* __restrict__ w, const double* __restrict__ nrm,
                                                      T* __restrict__ out) {
  const double inv = 1.0 / *nrm;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    out[e] = (T)((double)w[e] * inv);
}

// sqrt(sum partial) into out
__global__ __launch_bounds__(MD_THREADS) void k_norm_final(const double* __restrict__ partial,
                                                           double* __restrict__ out) {
  __shared__ double sh[MD_WAVES];
  double acc = 0.0;
  for (int b = threadIdx.x; b < MD_BLOCKS; b += MD_THREADS) acc += partial[b];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int v = 0; v < MD_WAVES; ++v) s += sh[v];
    *out = sqrt(s);
  }
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
  hipLaunchKernelGGL(k_multidot_partial<T>, dim3(MD_BLOCKS), dim3(MD_THREADS), 0, st, n, k, V, ldv, w, work);
  hipLaunchKernelGGL(k_multidot_final, dim3(k), dim3(MD_THREADS), 0, st, k, work, h, accumulate);
  TF_LAUNCH_CHECK("multidot");
  return 0;
}

}  // namespace

extern "C" {

int tf_krylov_work_size(void) { return MD_BLOCKS * MAX_BASIS; }
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
  if (k < 0 || k > MAX_BASIS) return tf::fail("multiaxpy: bad basis size");
  if (k == 0) return 0;
  hipLaunchKernelGGL(k_multiaxpy<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, k, V, ldv, coef,
                     w, sign);
  TF_LAUNCH_CHECK("multiaxpy");
  return 0;
}
int tf_multiaxpy_f32(int n, int k, const float* V, long ldv, const double* coef, float* w, double sign,
                     void* stream) {
  if (k < 0 || k > MAX_BASIS) return tf::fail("multiaxpy: bad basis size");
  if (k == 0) return 0;
  hipLaunchKernelGGL(k_multiaxpy<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, k, V, ldv, coef,
                     w, sign);
  TF_LAUNCH_CHECK("multiaxpy");
  return 0;
}

// nrm_dev = ||w||_2 (deterministic), then out = w / nrm_dev
int tf_normalize_f64(int n, const double* w, double* work, double* nrm_dev, double* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_multidot_partial<double>, dim3(MD_BLOCKS), dim3(MD_THREADS), 0, st, n, 1, w, (long)n, w, work);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(MD_THREADS), 0, st, work, nrm_dev);
  hipLaunchKernelGGL(k_scale_by_inv<double>, dim3(grid_for(n)), dim3(256), 0, st, n, w, nrm_dev, out);
  TF_LAUNCH_CHECK("normalize");
  return 0;
}
int tf_normalize_f32(int n, const float* w, double* work, double* nrm_dev, float* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_multidot_partial<float>, dim3(MD_BLOCKS), dim3(MD_THREADS), 0, st, n, 1, w, (long)n, w, work);
  hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(MD_THREADS), 0, st, work, nrm_dev);
  hipLaunchKernelGGL(k_scale_by_inv<float>, dim3(grid_for(n)), dim3(256), 0, st, n, w, nrm_dev, out);
  TF_LAUNCH_CHECK("normalize");
  return 0;
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
