// Exact block-tridiagonal factorisation of the gauged Jacobian over grid lines.
//
// With the state ordered x-fastest, the rows of grid line j (all x, fixed y,
// 2D) form block row j of size m = nx*dof, and the stencil couples line j only
// to lines j-1 and j+1: the gauged Jacobian is block tridiagonal with blocks
// L_j (to j-1), D_j, U_j (to j+1).  Block LU by the Thomas recursion with
// explicit inverses:
//
//   S_0 = D_0,  S_j = D_j - L_j W_{j-1},   Y_j = S_j^{-1},
//   W_j = Y_j U_j   (next line's Schur update),  G_j = L_j Y_{j-1}
//
// L_j and U_j have at most a handful of entries per row/column, so W_j, G_j and
// the Schur update cost O(m^2); the only O(m^3) work is the inversion, done by
// a blocked Gauss-Jordan sweep in FP64 (panel width PB, partial pivoting within
// each diagonal panel, tiny pivots statically perturbed -- the factor is used
// as the preconditioner of GMRES, which absorbs the perturbation).
//
// Solve A x = r:  forward  z_0 = r_0,  z_j = r_j - G_j z_{j-1}
//                 backward x_{N-1} = Y_{N-1} z_{N-1},  x_j = Y_j z_j - W_j x_{j+1}
// one GEMV kernel per line and sweep.  Y/G/W are stored in FP64 or FP32
// (mixed-precision variant; the factorisation arithmetic stays FP64).
//
// Replaces the sparse LU of the reference's direct solve
// (scipy.sparse.linalg.splu in transiflow/interface/SciPy.py:131-162,194-258)
// for problems whose grid lines fit a dense block.
#include "tf_common.h"
#include "transiflow_hip.h"

#include <vector>

namespace {

constexpr int PB = 64;   // panel width of the Gauss-Jordan sweep
constexpr int KC = 8;    // max couplings of one row (column) to the previous (next) line

struct Coupling {
  int idx;
  int pad;
  double val;
};

// ---- coupling extraction --------------------------------------------------

// For every row r of line j: entries with columns in line j-1 -> Lc[r][*] (col offset);
// entries with columns in line j+1 -> appended to Ut[c][*] (row offset within line j).
// Errors (coupling beyond the neighbouring lines, more than KC couplings) set *err.
__global__ __launch_bounds__(256) void k_couplings(int n, int m, const int* __restrict__ rp,
                                                   const int* __restrict__ ci, const double* __restrict__ v, int gauge,
                                                   Coupling* __restrict__ Lc, int* __restrict__ lcount,
                                                   Coupling* __restrict__ Ut, int* __restrict__ ucount,
                                                   int* __restrict__ err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int j = r / m;
  int nl = 0;
  if (r != gauge) {
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int c = ci[k];
      if (c == gauge) continue;
      const int jc = c / m;
      if (jc == j) continue;
      if (jc == j - 1) {
        if (nl < KC) Lc[(long)r * KC + nl] = Coupling{c - (j - 1) * m, 0, v[k]};
        else atomicOr(err, 1);
        ++nl;
      } else if (jc == j + 1) {
        const int slot = atomicAdd(&ucount[c], 1);
        if (slot < KC) Ut[(long)c * KC + slot] = Coupling{r - j * m, 0, v[k]};
        else atomicOr(err, 1);
      } else {
        atomicOr(err, 2);
      }
    }
  }
  lcount[r] = nl < KC ? nl : KC;
}

// Sort each column's upper couplings by row (atomic append order is arbitrary).
__global__ __launch_bounds__(256) void k_sort_ut(int n, Coupling* __restrict__ Ut, int* __restrict__ ucount) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int cnt = ucount[c];
  if (cnt > KC) cnt = KC;
  ucount[c] = cnt;
  Coupling* u = Ut + (long)c * KC;
  for (int a = 1; a < cnt; ++a) {
    Coupling key = u[a];
    int b = a - 1;
    while (b >= 0 && u[b].idx > key.idx) {
      u[b + 1] = u[b];
      --b;
    }
    u[b + 1] = key;
  }
}

__global__ __launch_bounds__(256) void k_fill_zero(long n, int* __restrict__ p) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

// ---- dense line block -----------------------------------------------------

// X = D_j (gauged): zero, then scatter the in-line entries.  One workgroup per row.
__global__ __launch_bounds__(256) void k_diag_block(int m, int j, const int* __restrict__ rp,
                                                    const int* __restrict__ ci, const double* __restrict__ v,
                                                    int gauge, double* __restrict__ X) {
  const int rl = blockIdx.x;
  const int r = j * m + rl;
  double* row = X + (long)rl * m;
  for (int c = threadIdx.x; c < m; c += blockDim.x) row[c] = 0.0;
  __syncthreads();
  if (r == gauge) {
    if (threadIdx.x == 0) row[rl] = -1.0;
    return;
  }
  for (int k = rp[r] + threadIdx.x; k < rp[r + 1]; k += blockDim.x) {
    const int c = ci[k];
    if (c == gauge) continue;
    const int cl = c - j * m;
    if (cl >= 0 && cl < m) row[cl] = v[k];
  }
}

// X[r][:] -= sum_k Lc[r][k].val * W_prev[Lc.idx][:]  (Schur update of line j)
template <typename T>
__global__ __launch_bounds__(256) void k_schur(int m, int j, const Coupling* __restrict__ Lc,
                                               const int* __restrict__ lcount, const T* __restrict__ Wp,
                                               double* __restrict__ X) {
  const int rl = blockIdx.x;
  const long r = (long)j * m + rl;
  const int cnt = lcount[r];
  double* row = X + (long)rl * m;
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    double acc = row[c];
    for (int k = 0; k < cnt; ++k) {
      const Coupling cp = Lc[r * KC + k];
      acc -= cp.val * (double)Wp[(long)cp.idx * m + c];
    }
    row[c] = acc;
  }
}

// ---- blocked Gauss-Jordan inversion of X (m x m, FP64, in place) -------------

// Invert the pb x pb diagonal panel block (pb <= 64; rows/columns beyond pb are
// padded with the identity).  Gauss-Jordan on [A | I] with partial pivoting,
// implicit (no row swaps): 8 waves, lane = row, each wave holds 16 of the 128
// augmented columns in registers.  Per column: the owner wave publishes the
// column through LDS, every wave finds the same pivot row with shuffles, reads
// the pivot row with readlane and eliminates its own columns.  The row that
// pivoted column k ends holding row k of the inverse in the right half.
// Pivots below tol * max|block| are replaced by +-tol*max (counted).
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(512) void k_gj_panel(int m, int k0, int pb, const double* __restrict__ X,
                                                  double* __restrict__ Pinv, int* __restrict__ nperturb) {
  __shared__ double colk[2][64];
  __shared__ double wmax[8];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  double a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int c = wave * 16 + q;
    double v;
    if (c < 64)
      v = (lane < pb && c < pb) ? X[(long)(k0 + lane) * m + k0 + c] : (lane == c ? 1.0 : 0.0);
    else
      v = (lane == c - 64) ? 1.0 : 0.0;
    a[q] = v;
  }
  double mx = 0.0;
  if (wave < 4) {
#pragma unroll
    for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(a[q]));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) wmax[wave] = mx;
  __syncthreads();
  double amax = 0.0;
  for (int w = 0; w < 4; ++w) amax = fmax(amax, wmax[w]);
  if (amax == 0.0) amax = 1.0;
  const double tol = 1e-13 * amax;
  bool used = false;
  int mystep = -1;
  for (int k = 0; k < 64; ++k) {
    const int owner = k >> 4;
    if (wave == owner) {
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q == (k & 15)) colk[k & 1][lane] = a[q];
    }
    __syncthreads();
    const double f = colk[k & 1][lane];
    double best = used ? -1.0 : fabs(f);
    int bi = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    const int p = __builtin_amdgcn_readfirstlane(bi);
    double piv = readlane_f64(f, p);
    if (fabs(piv) < tol) {
      piv = piv < 0.0 ? -tol : tol;
      if (threadIdx.x == 0) atomicAdd(nperturb, 1);
    }
    const double ip = 1.0 / piv;
    if (lane == p) {
      used = true;
      mystep = k;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const double pr = readlane_f64(a[q], p) * ip;
      a[q] = (lane == p) ? pr : a[q] - f * pr;
    }
  }
  if (wave >= 4 && mystep >= 0 && mystep < pb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = wave * 16 + q - 64;
      if (c < pb) Pinv[mystep * pb + c] = a[q];
    }
  }
}

// Snapshot of the panel rows (Kr = X[k0:k0+pb, :]) and columns (Kc = X[:, k0:k0+pb])
// before the in-place update overwrites them.
__global__ __launch_bounds__(256) void k_gj_copy(int m, int k0, int pb, const double* __restrict__ X,
                                                 double* __restrict__ Kr, double* __restrict__ Kc) {
  const long total = (long)pb * m;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int s = (int)(e / m), c = (int)(e % m);
    Kr[(long)s * m + c] = X[(long)(k0 + s) * m + c];
    const int i = (int)(e / pb), t = (int)(e % pb);
    Kc[(long)i * PB + t] = X[(long)i * m + k0 + t];
  }
}

// One 64x64 output tile of the block Gauss-Jordan step for panel K:
//   K rows:     X[K, K] = Pinv,  X[K, R] = T = Pinv X[K, R]
//   other rows: X[O, K] = -X[O, K] Pinv,  X[O, R] -= X[O, K] T
// T's tile is recomputed by every row tile from the snapshots (cheap, no extra pass).
__global__ __launch_bounds__(256) void k_gj_update(int m, int k0, int pb, double* __restrict__ X,
                                                   const double* __restrict__ Pinv, const double* __restrict__ Kr,
                                                   const double* __restrict__ Kc) {
  __shared__ double P[PB][PB + 1];
  __shared__ double B[PB][64 + 1];   // Kr tile, then T tile (or Pinv columns)
  __shared__ double U[64][PB + 1];   // Kc tile
  const int i0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  const int t = threadIdx.x;
  const bool rowsK = (i0 == k0);
  const bool colsK = (c0 == k0);
  const int tr = (t >> 4) * 4, tc = (t & 15) * 4;
  for (int e = t; e < PB * PB; e += 256) {
    const int r = e >> 6, c = e & 63;
    P[r][c] = (r < pb && c < pb) ? Pinv[r * pb + c] : 0.0;
  }
  if (colsK) {
    for (int e = t; e < PB * 64; e += 256) {
      const int r = e >> 6, c = e & 63;
      B[r][c] = (r < pb && c < pb) ? Pinv[r * pb + c] : 0.0;
    }
    __syncthreads();
  } else {
    for (int e = t; e < PB * 64; e += 256) {
      const int s = e >> 6, b = e & 63;
      B[s][b] = (s < pb && c0 + b < m) ? Kr[(long)s * m + c0 + b] : 0.0;
    }
    __syncthreads();
    // T tile = P * B, 4x4 per thread, then written back into B
    double acc[4][4] = {};
    for (int s = 0; s < PB; ++s) {
      double u[4], w[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) u[a] = P[tr + a][s];
#pragma unroll
      for (int b = 0; b < 4; ++b) w[b] = B[s][tc + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += u[a] * w[b];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) B[tr + a][tc + b] = acc[a][b];
    __syncthreads();
  }
  if (rowsK) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = i0 + tr + a;
      if (tr + a >= pb || i >= m) continue;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int c = c0 + tc + b;
        if (c < m) X[(long)i * m + c] = B[tr + a][tc + b];
      }
    }
    return;
  }
  for (int e = t; e < 64 * PB; e += 256) {
    const int a = e >> 6, s = e & 63;
    U[a][s] = (i0 + a < m && s < pb) ? Kc[(long)(i0 + a) * PB + s] : 0.0;
  }
  __syncthreads();
  double acc[4][4] = {};
  for (int s = 0; s < PB; ++s) {
    double u[4], w[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) u[a] = U[tr + a][s];
#pragma unroll
    for (int b = 0; b < 4; ++b) w[b] = B[s][tc + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += u[a] * w[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + tr + a;
    if (i >= m) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int c = c0 + tc + b;
      if (c >= m) continue;
      double* x = X + (long)i * m + c;
      *x = colsK ? -acc[a][b] : *x - acc[a][b];
    }
  }
}

// Y = X cast to the storage type.
template <typename T>
__global__ __launch_bounds__(256) void k_store_inverse(int m, const double* __restrict__ X, T* __restrict__ Y) {
  const int r = blockIdx.x;
  for (int i = threadIdx.x; i < m; i += blockDim.x) Y[(long)r * m + i] = (T)X[(long)r * m + i];
}

// W_j[a][c] = sum_k Y_j[a][Ut[c].idx] * Ut[c].val   (c over line j+1)
template <typename T>
__global__ __launch_bounds__(256) void k_make_W(int m, int j, const T* __restrict__ Y, const Coupling* __restrict__ Ut,
                                                const int* __restrict__ ucount, T* __restrict__ W) {
  const int a = blockIdx.x;
  const long cbase = (long)(j + 1) * m;
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    const int cnt = ucount[cbase + c];
    double acc = 0.0;
    for (int k = 0; k < cnt; ++k) {
      const Coupling cp = Ut[(cbase + c) * KC + k];
      acc += (double)Y[(long)a * m + cp.idx] * cp.val;
    }
    W[(long)a * m + c] = (T)acc;
  }
}

// G_j[r][:] = sum_k Lc[r].val * Y_{j-1}[Lc.idx][:]
template <typename T>
__global__ __launch_bounds__(256) void k_make_G(int m, int j, const T* __restrict__ Yp, const Coupling* __restrict__ Lc,
                                                const int* __restrict__ lcount, T* __restrict__ G) {
  const int rl = blockIdx.x;
  const long r = (long)j * m + rl;
  const int cnt = lcount[r];
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < cnt; ++k) {
      const Coupling cp = Lc[r * KC + k];
      acc += cp.val * (double)Yp[(long)cp.idx * m + c];
    }
    G[(long)rl * m + c] = (T)acc;
  }
}

// ---- solve ------------------------------------------------------------------

// out[r] = base[r] + sa * (A[r,:] . a) + sb * (B[r,:] . b)   (one wave per row)
template <typename T, typename V>
__global__ __launch_bounds__(256) void k_gemv2(int m, const T* __restrict__ A, const V* __restrict__ a, double sa,
                                               const T* __restrict__ B, const V* __restrict__ b, double sb,
                                               const V* __restrict__ base, V* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= m) return;
  double acc = 0.0;
  const T* ar = A + (long)r * m;
  for (int c = lane; c < m; c += 64) acc += (double)ar[c] * (double)a[c];
  acc *= sa;
  if (B) {
    double acc2 = 0.0;
    const T* br = B + (long)r * m;
    for (int c = lane; c < m; c += 64) acc2 += (double)br[c] * (double)b[c];
    acc += sb * acc2;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) out[r] = (V)((base ? (double)base[r] : 0.0) + acc);
}

template <typename V>
__global__ __launch_bounds__(256) void k_copy_vec(long n, const V* __restrict__ x, V* __restrict__ y) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) y[e] = x[e];
}

}  // namespace

struct tf_linelu {
  int n, m, N;
  int single;  // storage: 1 -> FP32, 0 -> FP64
  double* X;
  void* Y;
  void* G;
  void* W;
  Coupling* Lc;
  Coupling* Ut;
  int* lcount;
  int* ucount;
  int* err;
  int* nperturb;
  double* Pinv;
  double* Tm;
  double* Uc;
  void* ztmp;
  int* host_flags;
};

namespace {

template <typename T>
int factor_impl(tf_linelu* h, const int* rp, const int* ci, const double* v, int gauge, hipStream_t st) {
  const int m = h->m, N = h->N, n = h->n;
  T* Y = (T*)h->Y;
  T* G = (T*)h->G;
  T* W = (T*)h->W;
  hipLaunchKernelGGL(k_fill_zero, dim3(64), dim3(256), 0, st, (long)n, h->ucount);
  hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, st, 2L, h->err);
  hipLaunchKernelGGL(k_couplings, dim3((n + 255) / 256), dim3(256), 0, st, n, m, rp, ci, v, gauge, h->Lc, h->lcount,
                     h->Ut, h->ucount, h->err);
  hipLaunchKernelGGL(k_sort_ut, dim3((n + 255) / 256), dim3(256), 0, st, n, h->Ut, h->ucount);
  const long mm = (long)m * m;
  const int tiles = (m + 63) / 64;
  for (int j = 0; j < N; ++j) {
    hipLaunchKernelGGL(k_diag_block, dim3(m), dim3(256), 0, st, m, j, rp, ci, v, gauge, h->X);
    if (j > 0) hipLaunchKernelGGL((k_schur<T>), dim3(m), dim3(256), 0, st, m, j, h->Lc, h->lcount, W + (j - 1) * mm, h->X);
    for (int k0 = 0; k0 < m; k0 += PB) {
      const int pb = m - k0 < PB ? m - k0 : PB;
      TF_PROF_BEGIN(tf::PROF_GJ_PANEL, st);
      hipLaunchKernelGGL(k_gj_panel, dim3(1), dim3(512), 0, st, m, k0, pb, h->X, h->Pinv, h->nperturb);
      TF_PROF_END(tf::PROF_GJ_PANEL, st);
      hipLaunchKernelGGL(k_gj_copy, dim3(((long)pb * m + 255) / 256), dim3(256), 0, st, m, k0, pb, h->X, h->Tm, h->Uc);
      TF_PROF_BEGIN(tf::PROF_GJ_UPDATE, st);
      hipLaunchKernelGGL(k_gj_update, dim3(tiles, tiles), dim3(256), 0, st, m, k0, pb, h->X, h->Pinv, h->Tm, h->Uc);
      TF_PROF_END(tf::PROF_GJ_UPDATE, st);
    }
    hipLaunchKernelGGL((k_store_inverse<T>), dim3(m), dim3(256), 0, st, m, h->X, Y + j * mm);
    if (j + 1 < N) hipLaunchKernelGGL((k_make_W<T>), dim3(m), dim3(256), 0, st, m, j, Y + j * mm, h->Ut, h->ucount, W + j * mm);
    if (j > 0) hipLaunchKernelGGL((k_make_G<T>), dim3(m), dim3(256), 0, st, m, j, Y + (j - 1) * mm, h->Lc, h->lcount, G + j * mm);
  }
  TF_LAUNCH_CHECK("linelu factor");
  return 0;
}

template <typename T, typename V>
int solve_impl(tf_linelu* h, const V* r, V* x, hipStream_t st) {
  const int m = h->m, N = h->N;
  const long mm = (long)m * m;
  const T* Y = (const T*)h->Y;
  const T* G = (const T*)h->G;
  const T* W = (const T*)h->W;
  V* z = (V*)h->ztmp;
  const dim3 gg((m + 3) / 4), bb(256);
  // forward: z_0 = r_0 ; z_j = r_j - G_j z_{j-1}
  hipLaunchKernelGGL(k_copy_vec<V>, dim3((m + 255) / 256), dim3(256), 0, st, (long)m, r, z);
  for (int j = 1; j < N; ++j)
    hipLaunchKernelGGL((k_gemv2<T, V>), gg, bb, 0, st, m, G + j * mm, z + (long)(j - 1) * m, -1.0, (const T*)nullptr,
                       (const V*)nullptr, 0.0, r + (long)j * m, z + (long)j * m);
  // backward: x_{N-1} = Y z ; x_j = Y_j z_j - W_j x_{j+1}
  hipLaunchKernelGGL((k_gemv2<T, V>), gg, bb, 0, st, m, Y + (N - 1) * mm, z + (long)(N - 1) * m, 1.0,
                     (const T*)nullptr, (const V*)nullptr, 0.0, (const V*)nullptr, x + (long)(N - 1) * m);
  for (int j = N - 2; j >= 0; --j)
    hipLaunchKernelGGL((k_gemv2<T, V>), gg, bb, 0, st, m, Y + j * mm, z + (long)j * m, 1.0, W + j * mm,
                       x + (long)(j + 1) * m, -1.0, (const V*)nullptr, x + (long)j * m);
  TF_LAUNCH_CHECK("linelu solve");
  return 0;
}

}  // namespace

extern "C" {

// n = N * m rows; single != 0 stores the factors in FP32.
int tf_linelu_create(tf_linelu** out, int n, int m, int single) {
  if (!out) return tf::fail("null handle");
  *out = nullptr;
  if (m < 1 || n % m != 0) return tf::fail("line block size must divide n");
  tf_linelu* h = new tf_linelu();
  h->n = n;
  h->m = m;
  h->N = n / m;
  h->single = single;
  const size_t es = single ? sizeof(float) : sizeof(double);
  const size_t mm = (size_t)m * m;
  auto alloc = [&]() -> int {
    TF_HIP(hipMalloc(&h->X, sizeof(double) * mm));
    TF_HIP(hipMalloc(&h->Y, es * mm * h->N));
    TF_HIP(hipMalloc(&h->G, es * mm * h->N));
    TF_HIP(hipMalloc(&h->W, es * mm * h->N));
    TF_HIP(hipMalloc(&h->Lc, sizeof(Coupling) * KC * (size_t)n));
    TF_HIP(hipMalloc(&h->Ut, sizeof(Coupling) * KC * (size_t)n));
    TF_HIP(hipMalloc(&h->lcount, sizeof(int) * n));
    TF_HIP(hipMalloc(&h->ucount, sizeof(int) * n));
    TF_HIP(hipMalloc(&h->err, sizeof(int) * 2));
    h->nperturb = h->err + 1;
    TF_HIP(hipMalloc(&h->Pinv, sizeof(double) * PB * PB));
    TF_HIP(hipMalloc(&h->Tm, sizeof(double) * PB * (size_t)m));
    TF_HIP(hipMalloc(&h->Uc, sizeof(double) * PB * (size_t)m));
    TF_HIP(hipMalloc(&h->ztmp, sizeof(double) * (size_t)n));
    TF_HIP(hipHostMalloc(&h->host_flags, sizeof(int) * 2, hipHostMallocDefault));
    return 0;
  };
  if (int e = alloc()) {
    tf_linelu_destroy(h);  // frees the partial set (pointers start null)
    return e;
  }
  *out = h;
  return 0;
}

int tf_linelu_destroy(tf_linelu* h) {
  if (!h) return 0;
  (void)hipFree(h->X);
  (void)hipFree(h->Y);
  (void)hipFree(h->G);
  (void)hipFree(h->W);
  (void)hipFree(h->Lc);
  (void)hipFree(h->Ut);
  (void)hipFree(h->lcount);
  (void)hipFree(h->ucount);
  (void)hipFree(h->err);
  (void)hipFree(h->Pinv);
  (void)hipFree(h->Tm);
  (void)hipFree(h->Uc);
  (void)hipFree(h->ztmp);
  (void)hipHostFree(h->host_flags);
  delete h;
  return 0;
}

// Factor the gauged matrix (CSR rp/ci/v, gauge row, -1 for none).  Asynchronous.
int tf_linelu_factor(tf_linelu* h, const int* rp, const int* ci, const double* v, int gauge, void* stream) {
  if (!h) return tf::fail("null handle");
  hipStream_t st = (hipStream_t)stream;
  return h->single ? factor_impl<float>(h, rp, ci, v, gauge, st) : factor_impl<double>(h, rp, ci, v, gauge, st);
}

// Synchronises and reports: flags[0] = structure error bits (1: too many
// couplings, 2: coupling beyond neighbouring lines), flags[1] = perturbed pivots.
int tf_linelu_status(tf_linelu* h, int* flags_host, void* stream) {
  if (!h || !flags_host) return tf::fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  TF_HIP(hipMemcpyAsync(h->host_flags, h->err, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
  TF_HIP(hipStreamSynchronize(st));
  flags_host[0] = h->host_flags[0];
  flags_host[1] = h->host_flags[1];
  return 0;
}

int tf_linelu_solve_f64(tf_linelu* h, const double* r, double* x, void* stream) {
  if (!h) return tf::fail("null handle");
  hipStream_t st = (hipStream_t)stream;
  return h->single ? solve_impl<float, double>(h, r, x, st) : solve_impl<double, double>(h, r, x, st);
}

int tf_linelu_solve_f32(tf_linelu* h, const float* r, float* x, void* stream) {
  if (!h) return tf::fail("null handle");
  hipStream_t st = (hipStream_t)stream;
  return h->single ? solve_impl<float, float>(h, r, x, st) : solve_impl<double, float>(h, r, x, st);
}

}  // extern "C"
