// Block cyclic reduction (BCR) of the gauged Jacobian over grid lines.
//
// The Jacobian is block tridiagonal over grid lines (see linelu.hip): line j
// couples to lines j-1 (L_j) and j+1 (U_j).  Odd-even reduction eliminates the
// lines at even positions of the active list, level by level, until one line
// is left.  For an eliminated line e with kept neighbours a (before) and b
// (after):
//
//   x_e = Y_e (r_e - L_e x_a - U_e x_b),   Y_e = D_e^{-1}
//
// and a kept line k with eliminated neighbours e1 (before) and e2 (after) is
// updated to
//
//   P1 = L_k Y_e1,  P2 = U_k Y_e2
//   D_k <- D_k - P1 U_e1 - P2 L_e2,   L_k <- -P1 L_e1,   U_k <- -P2 U_e2
//   r_k <- r_k - P1 r_e1 - P2 r_e2
//
// Every level is a batch of independent dense operations: the inversions of
// its eliminated lines (blocked Gauss-Jordan, 64-wide panels) and six GEMMs
// per kept line -- so the sequential depth is (levels x panels), about 100 for
// 256 lines, instead of (lines x panels) for the Thomas recursion.  At level
// 0 the couplings are the sparse CSR entries (a few per row), so the level-0
// products are sparse-dense and cost O(m^2).  The factor (Y, P1, P2, L, U) is
// kept in the storage type T (double, or float for the mixed-precision
// variant, whose factorisation also runs in FP32).  For the FP32 apply of that
// variant the blocks can also be copied to FP16 with one FP32 scale per row
// (tf_bcr_set_half), which halves the bytes the apply reads.
//
// Solve: forward  r_k -= P1 r_e1 + P2 r_e2  (level by level; at level 0, where
//                 L_k and U_k are sparse, as r_k -= L_k (Y_e1 r_e1) + U_k (Y_e2 r_e2)),
//        backward x_e = Y_e (r_e - L_e x_a - U_e x_b)  (levels in reverse).
//
// Replaces the sparse LU of the reference's direct solve
// (scipy.sparse.linalg.splu, transiflow/interface/SciPy.py:131-162,194-258).
#include "gemm.h"
#include "gj.h"
#include "tf_common.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <type_traits>
#include <vector>

namespace {

using tf::gj::PB;
using tf::gj::k_panel;
using tf::gj::k_panel_copy;
using tf::gj::k_panel_store;
using tf::gj::readlane_t;
constexpr int KC = 8;   // couplings per row/column to a neighbouring line

struct Cpl {
  int idx;
  int pad;
  double val;
};

// Line ordering of the unknowns.  Rows of one line are contiguous in the
// factor: natural order when lines are x-lines (2D) or planes (3D); for 2D
// y-lines (line = column i, m = ny dof) the natural index d + dof (i + nx j)
// maps to i m + j dof + d.
struct Order {
  int ylines, nx, dof, m;
  __host__ __device__ __forceinline__ int to_line(int r) const {
    if (!ylines) return r;
    const int d = r % dof, cell = r / dof, i = cell % nx, j = cell / nx;
    return i * m + j * dof + d;
  }
  __host__ __device__ __forceinline__ int to_natural(int q) const {
    if (!ylines) return q;
    const int i = q / m, rest = q - i * m, j = rest / dof, d = rest - j * dof;
    return d + dof * (i + nx * j);
  }
};

__global__ __launch_bounds__(256) void k_zero_int(long n, int* __restrict__ p) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

// zero n elements of each of a batch of matrices (grid.y = batch)
template <typename T>
__global__ __launch_bounds__(256) void k_zero_many(long n, T* const* __restrict__ ptrs) {
  T* p = ptrs[blockIdx.y];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = (T)0;
}

// Row-wise couplings: Lr[r] (columns in line(r)-1, as offsets), Ur[r] (line(r)+1).
// Column-wise couplings: Ut[c] = rows of line(c)-1 coupling into column c,
//                        Lt[c] = rows of line(c)+1 coupling into column c.
__global__ __launch_bounds__(256) void k_couplings4(int n, int m, Order o, const int* __restrict__ rp,
                                                    const int* __restrict__ ci, const double* __restrict__ v,
                                                    int gauge, Cpl* __restrict__ Lr, int* __restrict__ nLr,
                                                    Cpl* __restrict__ Ur, int* __restrict__ nUr,
                                                    Cpl* __restrict__ Ut, int* __restrict__ nUt,
                                                    Cpl* __restrict__ Lt, int* __restrict__ nLt,
                                                    int* __restrict__ err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;  // natural row
  if (r >= n) return;
  const int q = o.to_line(r);                              // line-major row
  const int j = q / m;
  int nl = 0, nu = 0;
  if (r != gauge) {
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int c = ci[k];
      if (c == gauge) continue;
      const int qc = o.to_line(c);
      const int jc = qc / m;
      if (jc == j) continue;
      if (jc == j - 1) {
        if (nl < KC) Lr[(long)q * KC + nl] = Cpl{qc - jc * m, 0, v[k]};
        else atomicOr(err, 1);
        ++nl;
        const int s = atomicAdd(&nLt[qc], 1);
        if (s < KC) Lt[(long)qc * KC + s] = Cpl{q - j * m, 0, v[k]};
        else atomicOr(err, 1);
      } else if (jc == j + 1) {
        if (nu < KC) Ur[(long)q * KC + nu] = Cpl{qc - jc * m, 0, v[k]};
        else atomicOr(err, 1);
        ++nu;
        const int s = atomicAdd(&nUt[qc], 1);
        if (s < KC) Ut[(long)qc * KC + s] = Cpl{q - j * m, 0, v[k]};
        else atomicOr(err, 1);
      } else {
        atomicOr(err, 2);
      }
    }
  }
  nLr[q] = nl < KC ? nl : KC;
  nUr[q] = nu < KC ? nu : KC;
}

// Sort the column-wise lists by row (the atomic append order is arbitrary).
__global__ __launch_bounds__(256) void k_sort_cols(int n, Cpl* __restrict__ A, int* __restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int k = cnt[c] < KC ? cnt[c] : KC;
  cnt[c] = k;
  Cpl* u = A + (long)c * KC;
  for (int a = 1; a < k; ++a) {
    Cpl key = u[a];
    int b = a - 1;
    while (b >= 0 && u[b].idx > key.idx) {
      u[b + 1] = u[b];
      --b;
    }
    u[b + 1] = key;
  }
}

// D[j] = dense diagonal block of line j (gauged).  grid (m, N).
template <typename T>
__global__ __launch_bounds__(256) void k_diag_blocks(int m, Order o, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const double* __restrict__ v,
                                                     int gauge, T* __restrict__ D) {
  const int rl = blockIdx.x, j = blockIdx.y;
  const long r = o.to_natural(j * m + rl);
  T* row = D + ((long)j * m + rl) * m;
  for (int c = threadIdx.x; c < m; c += blockDim.x) row[c] = (T)0;
  __syncthreads();
  if (r == gauge) {
    if (threadIdx.x == 0) row[rl] = (T)-1;
    return;
  }
  for (int k = rp[r] + threadIdx.x; k < rp[r + 1]; k += blockDim.x) {
    const int c = ci[k];
    if (c == gauge) continue;
    const long cl = (long)o.to_line(c) - (long)j * m;
    if (cl >= 0 && cl < m) row[cl] = (T)v[k];
  }
}

// ---- level-0 sparse products -----------------------------------------------

// O_b[r][:] = sum_q cpl[rows of line_b][r][q].val * D_b[idx][:]      (sparse rows x dense)
template <typename T>
__global__ __launch_bounds__(256) void k_sp_rows(int m, const Cpl* __restrict__ cpl, const int* __restrict__ ncpl,
                                                 const int* __restrict__ line, T* const* __restrict__ Dp,
                                                 T* const* __restrict__ Op) {
  const int rl = blockIdx.x, b = blockIdx.y;
  const long r = (long)line[b] * m + rl;
  const T* D = Dp[b];
  T* O = Op[b] + (long)rl * m;
  const int cnt = ncpl[r];
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    double acc = 0.0;
    for (int q = 0; q < cnt; ++q) {
      const Cpl cp = cpl[r * KC + q];
      acc += cp.val * (double)D[(long)cp.idx * m + c];
    }
    O[c] = (T)acc;
  }
}

// O_b[r][c] = (acc ? O_b[r][c] : 0) + sign * sum_q P_b[r][cpl[col c of line_b].idx] * val   (dense x sparse cols)
// A thread owns column c and keeps its (at most KC) couplings in registers while
// it sweeps SPC_ROWS rows; consecutive threads write consecutive columns.
constexpr int SPC_ROWS = 32;

inline dim3 spc_grid(int m, int batch) {
  return dim3((unsigned)((m + 255) / 256), (unsigned)((m + SPC_ROWS - 1) / SPC_ROWS), (unsigned)batch);
}

template <typename T>
__global__ __launch_bounds__(256) void k_sp_cols(int m, const Cpl* __restrict__ cpl, const int* __restrict__ ncpl,
                                                 const int* __restrict__ line, T* const* __restrict__ Pp,
                                                 T* const* __restrict__ Op, double sign, int accumulate) {
  const int b = blockIdx.z;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= m) return;
  const int r0 = blockIdx.y * SPC_ROWS;
  const int r1 = r0 + SPC_ROWS < m ? r0 + SPC_ROWS : m;
  const T* P = Pp[b];
  T* O = Op[b];
  const long col = (long)line[b] * m + c;
  const int cnt = ncpl[col];
  int idx[KC];
  double val[KC];
#pragma unroll
  for (int q = 0; q < KC; ++q) {
    idx[q] = 0;
    val[q] = 0.0;
    if (q < cnt) {
      const Cpl cp = cpl[col * KC + q];
      idx[q] = cp.idx;
      val[q] = cp.val;
    }
  }
  for (int r = r0; r < r1; ++r) {
    const T* prow = P + (long)r * m;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < KC; ++q)
      if (q < cnt) acc += (double)prow[idx[q]] * val[q];
    T* o = O + (long)r * m + c;
    *o = (T)((accumulate ? (double)*o : 0.0) + sign * acc);
  }
}

// ---- solve ----------------------------------------------------------------------
//
// Every solve kernel is a batch of row-times-vector products over m-long rows.
// A group of lanes owns one row and reads it with 16-byte loads when the row
// length and the vectors allow it (group_for sets the group size).  The rows
// come from the factor's own blocks (Blocks<T>), or from FP16 copies of the
// FP32 blocks with one power-of-two FP32 scale per row (HalfBlocks): half the
// bytes for the FP32 apply of the mixed-precision solver.

typedef _Float16 f16;
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int W = 4;
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int W = 2;
};

// Lanes per row G (16, 32 or 64; a launch argument): 16 when a level has many
// lines (measured: 8-lane groups are slower); wider groups for the few lines of
// the upper levels, where each group's chain of dependent loads sets the time.
__host__ __device__ __forceinline__ int rows_per_block(int G) { return 256 / G; }
inline int group_for(int batch) { return batch >= 32 ? 16 : (batch >= 8 ? 32 : 64); }
// NC columns, R rows: sum over c of a_i[c] x_j[c] (x_j = x + j cs) by the G-lane
// group g (every lane of the group gets the results).  Each row is read once for
// all columns and each x chunk once for all rows; every (row, column) sum has the
// order of the one-row, one-column case, so the result does not depend on NC or R
// (an NC-column apply equals NC single-column applies bit for bit).
template <int NC, int R, typename T, typename V>
__device__ __forceinline__ void group_dot(const T* const (&a)[R], const V* __restrict__ x, long cs, int m, int g,
                                          int G, bool vec, double (&out)[R][NC]) {
  double acc[R][NC];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = 0.0;
  if constexpr (sizeof(T) == sizeof(V)) {
    if (vec) {
      typedef typename Vec16<T>::type VT;
      constexpr int W = Vec16<T>::W;
#pragma unroll 4
      for (int c = g; c < m / W; c += G) {
        VT p[R];
#pragma unroll
        for (int i = 0; i < R; ++i) p[i] = ((const VT*)a[i])[c];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const VT q = ((const VT*)(x + j * cs))[c];
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int w = 0; w < W; ++w) acc[i][j] += (double)p[i][w] * (double)q[w];
        }
      }
    } else {
      for (int c = g; c < m; c += G)
#pragma unroll
        for (int j = 0; j < NC; ++j)
#pragma unroll
          for (int i = 0; i < R; ++i) acc[i][j] += (double)a[i][c] * (double)x[j * cs + c];
    }
  } else {
    for (int c = g; c < m; c += G)
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i][j] += (double)a[i][c] * (double)x[j * cs + c];
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      for (int off = G / 2; off > 0; off >>= 1) acc[i][j] += __shfl_xor(acc[i][j], off, G);
      out[i][j] = acc[i][j];
    }
}

// FP16 rows against FP32 vectors: 8 halves (16 bytes) per row and two float4 per
// column per lane per step
template <int NC, int R>
__device__ __forceinline__ void group_dot(const f16* const (&a)[R], const float* __restrict__ x, long cs, int m,
                                          int g, int G, bool vec, double (&out)[R][NC]) {
  double acc[R][NC];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = 0.0;
  if (vec) {
#pragma unroll 2
    for (int c = g; c < m / 8; c += G) {
      f16x8 p[R];
#pragma unroll
      for (int i = 0; i < R; ++i) p[i] = ((const f16x8*)a[i])[c];
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        const f32x4* xv = (const f32x4*)(x + j * cs);
        const f32x4 q0 = xv[2 * c], q1 = xv[2 * c + 1];
#pragma unroll
        for (int i = 0; i < R; ++i) {
          float sv = 0.0f;
#pragma unroll
          for (int w = 0; w < 4; ++w) sv = fmaf((float)p[i][w], q0[w], sv);
#pragma unroll
          for (int w = 0; w < 4; ++w) sv = fmaf((float)p[i][4 + w], q1[w], sv);
          acc[i][j] += (double)sv;
        }
      }
    }
  } else {
    for (int c = g; c < m; c += G)
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i][j] += (double)(float)a[i][c] * (double)x[j * cs + c];
  }
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      for (int off = G / 2; off > 0; off >>= 1) acc[i][j] += __shfl_xor(acc[i][j], off, G);
      out[i][j] = acc[i][j];
    }
}

// Rows of a batch of m x m blocks in the factor's storage type.
template <typename T>
struct Blocks {
  T* const* p;
  __device__ __forceinline__ const T* row(int b, int r, int m) const { return p[b] + (long)r * m; }
  __device__ __forceinline__ double scale(int, int) const { return 1.0; }
};

// FP16 copies: row r of block b is s[b][r] * p[b][r m ... r m + m).
struct HalfBlocks {
  f16* const* p;
  float* const* s;
  __device__ __forceinline__ const f16* row(int b, int r, int m) const { return p[b] + (long)r * m; }
  __device__ __forceinline__ double scale(int b, int r) const { return (double)s[b][r]; }
};

// One block's rows with its base pointers already resolved (the few-line levels
// pass them by value, so no pointer list is read before the rows).
template <typename B>
struct RowOf;
template <typename T>
struct RowOf<Blocks<T>> {
  struct type {
    const T* p;
    __device__ __forceinline__ const T* row(int r, int m) const { return p + (long)r * m; }
    __device__ __forceinline__ double scale(int) const { return 1.0; }
  };
};
template <>
struct RowOf<HalfBlocks> {
  struct type {
    const f16* p;
    const float* s;
    __device__ __forceinline__ const f16* row(int r, int m) const { return p + (long)r * m; }
    __device__ __forceinline__ double scale(int r) const { return (double)s[r]; }
  };
};

// Rows per lane group R: with two, the rows share each load of the vector chunk.
// Two for FP16 rows once a launch has 12288 rows (batch x m) or more, so that the
// halved group count still fills the device; one for rows in the factor's own
// type.  Measured on the line-factor apply (FP32 from FP16 blocks, one and two
// columns): 256 x 256 cavity 294 / 388 -> 295 / 355 us, 32^3 cavity 649 / 914 ->
// 627 / 735 us; two rows of FP64 blocks were slower at 32^3 (2296 -> 2438 us).
template <typename B>
int rows_for(int batch, int m) {
  return (std::is_same<B, HalfBlocks>::value && (long)batch * m >= 12288) ? 2 : 1;
}

// f(std::integral_constant<int, R>) for the run-time row count R (1 or 2)
template <typename F>
void with_rows(int R, F&& f) {
  if (R == 2)
    f(std::integral_constant<int, 2>());
  else
    f(std::integral_constant<int, 1>());
}

template <int NC, int R, typename B, typename V>
__device__ __forceinline__ void row_dot(const B& blk, int b, const int (&r)[R], const V* __restrict__ x, long cs,
                                        int m, int g, int G, bool vec, double (&out)[R][NC]) {
  decltype(blk.row(b, 0, m)) a[R];
#pragma unroll
  for (int i = 0; i < R; ++i) a[i] = blk.row(b, r[i], m);
  group_dot<NC, R>(a, x, cs, m, g, G, vec, out);
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const double sc = blk.scale(b, r[i]);
#pragma unroll
    for (int j = 0; j < NC; ++j) out[i][j] *= sc;
  }
}

// The R rows of this lane group: a block covers R * rows_per_block(G) rows; rows
// past m are clamped to row 0 (read, never written).
template <int R>
__device__ __forceinline__ void group_rows(int m, int G, int (&rr)[R], bool (&live)[R]) {
  const int rpb = rows_per_block(G);
  const int r0 = blockIdx.x * rpb * R + (int)threadIdx.x / G;
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int r = r0 + i * rpb;
    live[i] = r < m;
    rr[i] = live[i] ? r : 0;
  }
}

// Every vector argument below holds NC columns at stride cs.

// rw[k] -= P1 rw[e1] + P2 rw[e2] (terms with a missing neighbour dropped) for a batch of kept lines
template <typename B, typename V, int NC, int R>
__global__ __launch_bounds__(256) void k_fwd(int m, int G, int vec, const int* __restrict__ k_line,
                                             const int* __restrict__ e1, const int* __restrict__ e2, B P1, B P2,
                                             V* __restrict__ rw, long cs) {
  const int b = blockIdx.y;
  const int g = threadIdx.x % G;
  int rr[R];
  bool live[R];
  group_rows<R>(m, G, rr, live);
  double acc[R][NC], d[R][NC];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = 0.0;
  if (e1[b] >= 0) {
    row_dot<NC, R>(P1, b, rr, rw + (long)e1[b] * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] += d[i][j];
  }
  if (e2[b] >= 0) {
    row_dot<NC, R>(P2, b, rr, rw + (long)e2[b] * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] += d[i][j];
  }
  if (g == 0)
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (live[i])
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          V* out = rw + j * cs + (long)k_line[b] * m + rr[i];
          *out = (V)((double)*out - acc[i][j]);
        }
}

// t[e] = rw[e] - L_e x[a] - U_e x[b]  (dense L/U, levels >= 1)
template <typename B, typename V, int NC, int R>
__global__ __launch_bounds__(256) void k_bwd_t(int m, int G, int vec, const int* __restrict__ e_line,
                                               const int* __restrict__ a_line, const int* __restrict__ b_line,
                                               B Lb, B Ub, const V* __restrict__ rw, const V* __restrict__ x,
                                               V* __restrict__ tbuf, long cs) {
  const int bi = blockIdx.y;
  const int g = threadIdx.x % G;
  int rr[R];
  bool live[R];
  group_rows<R>(m, G, rr, live);
  const long base = (long)e_line[bi] * m;
  const int a = a_line[bi], bb = b_line[bi];
  double acc[R][NC], d[R][NC];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[i][j] = 0.0;
  if (a >= 0) {
    row_dot<NC, R>(Lb, bi, rr, x + (long)a * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] += d[i][j];
  }
  if (bb >= 0) {
    row_dot<NC, R>(Ub, bi, rr, x + (long)bb * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] += d[i][j];
  }
  if (g == 0)
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (live[i])
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          const long row = base + rr[i];
          tbuf[j * cs + row] = (V)((double)rw[j * cs + row] - acc[i][j]);
        }
}

// Level 0: t[e] = rw[e] - L_e x[a] - U_e x[b] with the sparse couplings.  Thread per
// row; the row's couplings are loaded once for all columns, each column's sum keeps
// the coupling order.
template <typename V, int NC>
__global__ __launch_bounds__(256) void k_bwd_t_sparse(int m, const int* __restrict__ e_line,
                                                      const int* __restrict__ a_line, const int* __restrict__ b_line,
                                                      const Cpl* __restrict__ Lr, const int* __restrict__ nLr,
                                                      const Cpl* __restrict__ Ur, const int* __restrict__ nUr,
                                                      const V* __restrict__ rw, const V* __restrict__ x,
                                                      V* __restrict__ tbuf, long cs) {
  const int bi = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const long row = (long)e_line[bi] * m + r;
  const int a = a_line[bi], bb = b_line[bi];
  const int nl = a >= 0 ? nLr[row] : 0, nu = bb >= 0 ? nUr[row] : 0;
  Cpl lc[KC], uc[KC];
#pragma unroll
  for (int q = 0; q < KC; ++q) {
    if (q < nl) lc[q] = Lr[row * KC + q];
    if (q < nu) uc[q] = Ur[row * KC + q];
  }
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const V* xj = x + j * cs;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < KC; ++q)
      if (q < nl) acc += lc[q].val * (double)xj[(long)a * m + lc[q].idx];
#pragma unroll
    for (int q = 0; q < KC; ++q)
      if (q < nu) acc += uc[q].val * (double)xj[(long)bb * m + uc[q].idx];
    tbuf[j * cs + row] = (V)((double)rw[j * cs + row] - acc);
  }
}

// x[e] = Y_e t[e]
template <typename B, typename V, int NC, int R>
__global__ __launch_bounds__(256) void k_bwd_x(int m, int G, int vec, const int* __restrict__ e_line, B Yb,
                                               const V* __restrict__ tbuf, V* __restrict__ x, long cs) {
  const int bi = blockIdx.y;
  const int g = threadIdx.x % G;
  int rr[R];
  bool live[R];
  group_rows<R>(m, G, rr, live);
  const long base = (long)e_line[bi] * m;
  double d[R][NC];
  row_dot<NC, R>(Yb, bi, rr, tbuf + base, cs, m, g, G, vec, d);
  if (g == 0)
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (live[i])
#pragma unroll
        for (int j = 0; j < NC; ++j) x[j * cs + base + rr[i]] = (V)d[i][j];
}

// Levels >= 1 of a factor made with the products Y L and Y U in place of L and U
// (tf_bcr_set_fused): x[e] = Y_e rw[e] - (Y_e L_e) x[a] - (Y_e U_e) x[b] in one
// pass, the same three blocks read and no intermediate vector.
template <typename B, typename V, int NC, int R>
__global__ __launch_bounds__(256) void k_bwd_fused(int m, int G, int vec, const int* __restrict__ e_line,
                                                   const int* __restrict__ a_line, const int* __restrict__ b_line,
                                                   B Yb, B YLb, B YUb, const V* __restrict__ rw, V* __restrict__ x,
                                                   long cs) {
  const int bi = blockIdx.y;
  const int g = threadIdx.x % G;
  int rr[R];
  bool live[R];
  group_rows<R>(m, G, rr, live);
  const long base = (long)e_line[bi] * m;
  const int a = a_line[bi], bb = b_line[bi];
  double acc[R][NC], d[R][NC];
  row_dot<NC, R>(Yb, bi, rr, rw + base, cs, m, g, G, vec, acc);
  if (a >= 0) {
    row_dot<NC, R>(YLb, bi, rr, x + (long)a * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] -= d[i][j];
  }
  if (bb >= 0) {
    row_dot<NC, R>(YUb, bi, rr, x + (long)bb * m, cs, m, g, G, vec, d);
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[i][j] -= d[i][j];
  }
  if (g == 0)
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (live[i])
#pragma unroll
        for (int j = 0; j < NC; ++j) x[j * cs + base + rr[i]] = (V)acc[i][j];
}

// ---- the few-line levels with their operands by value ----------------------------
//
// A level with at most SMALL_MAX lines is latency-bound: its launch reads little,
// and every load of a line index or a block pointer from the schedule's lists
// delays the row loads behind it.  k_small takes the line indices and the blocks'
// row bases as kernel arguments (resolved on the host from a copy of the lists),
// so the row and vector loads issue at once.  dir 0: the forward step of k_fwd;
// dir 1: the fused backward step of k_bwd_fused; the same arithmetic (R = 1).
constexpr int SMALL_MAX = 8;

template <typename B>
struct SmallLevel {
  int line[SMALL_MAX], n1[SMALL_MAX], n2[SMALL_MAX];
  typename RowOf<B>::type blk[3][SMALL_MAX];  // forward: P1, P2; backward: Y, Y L, Y U
};

template <int NC, typename Row, typename V>
__device__ __forceinline__ void small_dot(const Row& row, int r, const V* __restrict__ x, long cs, int m, int g,
                                          int G, bool vec, double (&out)[NC]) {
  decltype(row.row(0, m)) a[1] = {row.row(r, m)};
  double o[1][NC];
  group_dot<NC, 1>(a, x, cs, m, g, G, vec, o);
  const double sc = row.scale(r);
#pragma unroll
  for (int j = 0; j < NC; ++j) out[j] = o[0][j] * sc;
}

template <typename B, typename V, int NC>
__global__ __launch_bounds__(256) void k_small(int m, int G, int vec, int dir, SmallLevel<B> a, V* __restrict__ rw,
                                               V* __restrict__ xo, long cs) {
  const int b = blockIdx.y;
  const int g = threadIdx.x % G;
  const int r = blockIdx.x * rows_per_block(G) + (int)threadIdx.x / G;
  const bool live = r < m;
  const int rr = live ? r : 0;
  const int n1 = a.n1[b], n2 = a.n2[b];
  double d[NC], acc[NC];
  if (dir == 0) {
#pragma unroll
    for (int j = 0; j < NC; ++j) acc[j] = 0.0;
    if (n1 >= 0) {
      small_dot<NC>(a.blk[0][b], rr, rw + (long)n1 * m, cs, m, g, G, vec, d);
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += d[j];
    }
    if (n2 >= 0) {
      small_dot<NC>(a.blk[1][b], rr, rw + (long)n2 * m, cs, m, g, G, vec, d);
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += d[j];
    }
    if (live && g == 0)
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        V* out = rw + j * cs + (long)a.line[b] * m + r;
        *out = (V)((double)*out - acc[j]);
      }
  } else {
    const long base = (long)a.line[b] * m;
    small_dot<NC>(a.blk[0][b], rr, rw + base, cs, m, g, G, vec, acc);
    if (n1 >= 0) {
      small_dot<NC>(a.blk[1][b], rr, xo + (long)n1 * m, cs, m, g, G, vec, d);
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] -= d[j];
    }
    if (n2 >= 0) {
      small_dot<NC>(a.blk[2][b], rr, xo + (long)n2 * m, cs, m, g, G, vec, d);
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] -= d[j];
    }
    if (live && g == 0)
#pragma unroll
      for (int j = 0; j < NC; ++j) xo[j * cs + base + r] = (V)acc[j];
  }
}

// ---- the top of the reduction as one dense map ------------------------------------
//
// Once the reduction is down to at most TOPD_MAX lines (the lines active at level
// zl), their solutions are a linear map of their right-hand sides after the forward
// levels below zl: x[line i] = sum_j Z[i][j] rw[line j], with Z the inverse of that
// reduced system, formed at factor time from the factor's own blocks
// (tf_bcr_set_dense_top).  One launch replaces the forward levels from zl up and the
// backward levels down to zl (five launches at four lines), for four more blocks
// read at four lines (16 instead of 12).
constexpr int TOPD_MAX = 4;

template <typename B>
struct DenseTop {
  int n;
  int line[TOPD_MAX];
  typename RowOf<B>::type blk[TOPD_MAX][TOPD_MAX];  // [output line][input line]
};

template <typename B, typename V, int NC>
__global__ __launch_bounds__(256) void k_dense_top(int m, int G, int vec, DenseTop<B> a, const V* __restrict__ rw,
                                                   V* __restrict__ xo, long cs) {
  const int b = blockIdx.y;
  const int g = threadIdx.x % G;
  const int r = blockIdx.x * rows_per_block(G) + (int)threadIdx.x / G;
  const bool live = r < m;
  const int rr = live ? r : 0;
  double acc[NC], d[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) acc[j] = 0.0;
#pragma unroll
  for (int t = 0; t < TOPD_MAX; ++t) {
    if (t < a.n) {
      small_dot<NC>(a.blk[b][t], rr, rw + (long)a.line[t] * m, cs, m, g, G, vec, d);
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] += d[j];
    }
  }
  if (live && g == 0)
#pragma unroll
    for (int j = 0; j < NC; ++j) xo[j * cs + (long)a.line[b] * m + r] = (V)acc[j];
}

// ---- the upper levels in one launch ---------------------------------------------
//
// The levels with few rows (batch x m < TOP_ROWS: the last forward levels and the
// first backward levels, one group row each, R = 1) take a few microseconds of
// launch and dependent-load latency each while reading little.  k_top runs them
// all in one launch as an ordered list of tasks: a task is one block of rows of
// one line at one level step (forward levels upward, then the fused backward
// levels downward), handed out by an atomic ticket in that order.  A task waits
// until every task of the previous step has finished (a per-step completion
// counter), so a block only ever waits for tasks with smaller tickets, which
// blocks that are already running hold: the launch makes progress whatever
// number of its blocks is resident.  Each row's arithmetic is that of k_fwd /
// k_bwd_fused, so the results are those of the per-level launches bit for bit.
// Waits are bounded (TOP_SPIN polls); a wait that runs out sets the error word
// and the task goes on.  The last block to leave resets the counters for the
// next launch.  Applies of one handle must not overlap (as for its work vectors).
constexpr int TOP_ROWS = 12288;
constexpr int MAX_TOP = 16;          // steps of one launch
constexpr long TOP_SPIN = 1L << 16;  // polls (about 1 us each) before a wait gives up
// sync words: [0] ticket, [1] blocks finished, [2 .. 2+MAX_TOP) tasks done per step, [2+MAX_TOP] error
constexpr int TOP_SYNC_WORDS = 3 + MAX_TOP;
constexpr int TOP_BLOCKS = 256;        // blocks of a k_top launch

template <typename B>
struct TopStep {
  int dir;    // 0: forward (kept lines, rw[k] -= P1 rw[e1] + P2 rw[e2]); 1: fused backward
  int batch;  // lines
  int G;      // lanes per row
  int nrb;    // row blocks per line
  int first;  // first ticket of the step
  const int* line;
  const int* n1;
  const int* n2;
  B blk[3];   // forward: P1, P2; backward: Y, Y L, Y U
};

template <typename B>
struct TopPlan {
  int nsteps, total;
  TopStep<B> s[MAX_TOP];
};

template <typename B, typename V, int NC>
__global__ __launch_bounds__(256) void k_top(int m, int vec, TopPlan<B> tp, V* __restrict__ rw, V* __restrict__ xo,
                                             long cs, unsigned* __restrict__ sync) {
  __shared__ int s_t;
  for (;;) {
    if (threadIdx.x == 0) s_t = (int)atomicAdd(&sync[0], 1u);
    __syncthreads();
    const int t = s_t;
    __syncthreads();
    if (t >= tp.total) break;
    int si = 0;
    while (si + 1 < tp.nsteps && t >= tp.s[si + 1].first) ++si;
    const TopStep<B>& S = tp.s[si];
    if (si > 0) {
      if (threadIdx.x == 0) {
        // relaxed polls (an acquire load per poll would invalidate the CU's and the
        // XCD's caches on every iteration, under the feet of the working tasks), one
        // acquire fence once the count is reached
        const unsigned need = (unsigned)(S.first - tp.s[si - 1].first);
        long spins = 0;
        while (__hip_atomic_load(&sync[2 + si - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
          if (++spins > TOP_SPIN) {
            atomicOr(&sync[2 + MAX_TOP], 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      __syncthreads();
    }
    const int local = t - S.first;
    const int b = local / S.nrb, rb = local - b * S.nrb;
    const int G = S.G;
    const int g = threadIdx.x % G;
    const int r = rb * rows_per_block(G) + (int)threadIdx.x / G;
    const bool live = r < m;
    const int rr[1] = {live ? r : 0};
    double d[1][NC];
    if (S.dir == 0) {
      double acc[NC];
#pragma unroll
      for (int j = 0; j < NC; ++j) acc[j] = 0.0;
      const int e1 = S.n1[b], e2 = S.n2[b];
      if (e1 >= 0) {
        row_dot<NC, 1>(S.blk[0], b, rr, rw + (long)e1 * m, cs, m, g, G, vec, d);
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[j] += d[0][j];
      }
      if (e2 >= 0) {
        row_dot<NC, 1>(S.blk[1], b, rr, rw + (long)e2 * m, cs, m, g, G, vec, d);
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[j] += d[0][j];
      }
      if (live && g == 0)
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          V* out = rw + j * cs + (long)S.line[b] * m + r;
          *out = (V)((double)*out - acc[j]);
        }
    } else {
      const long base = (long)S.line[b] * m;
      const int a = S.n1[b], bb = S.n2[b];
      double acc[1][NC];
      row_dot<NC, 1>(S.blk[0], b, rr, rw + base, cs, m, g, G, vec, acc);
      if (a >= 0) {
        row_dot<NC, 1>(S.blk[1], b, rr, xo + (long)a * m, cs, m, g, G, vec, d);
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[0][j] -= d[0][j];
      }
      if (bb >= 0) {
        row_dot<NC, 1>(S.blk[2], b, rr, xo + (long)bb * m, cs, m, g, G, vec, d);
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[0][j] -= d[0][j];
      }
      if (live && g == 0)
#pragma unroll
        for (int j = 0; j < NC; ++j) xo[j * cs + base + r] = (V)acc[0][j];
    }
    __syncthreads();  // the block's writes of this task are done
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      atomicAdd(&sync[2 + si], 1u);
    }
  }
  if (threadIdx.x == 0) {
    if (atomicAdd(&sync[1], 1u) == gridDim.x - 1) {
      // every block has left: no task is waiting; reset for the next launch
      for (int q = 0; q < 2 + MAX_TOP; ++q) __hip_atomic_store(&sync[q], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// copy a batch of m x m blocks: dst[b] = src[b]   (grid.y = batch)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_blocks(long mm, T* const* __restrict__ src, T* const* __restrict__ dst) {
  const T* s = src[blockIdx.y];
  T* d = dst[blockIdx.y];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < mm; e += (long)gridDim.x * blockDim.x) d[e] = s[e];
}

// Level 0, kept lines: rw[k] -= L_k y[e1] + U_k y[e2] with the sparse couplings and
// y[e] = Y_e rw[e] (so P1 rw[e1] = L_k Y_e1 rw[e1] without reading P1).  Thread per
// row; the row's couplings are loaded once for all columns.
template <typename V, int NC>
__global__ __launch_bounds__(256) void k_fwd_sparse(int m, const int* __restrict__ k_line,
                                                    const int* __restrict__ e1, const int* __restrict__ e2,
                                                    const Cpl* __restrict__ Lr, const int* __restrict__ nLr,
                                                    const Cpl* __restrict__ Ur, const int* __restrict__ nUr,
                                                    const V* __restrict__ y, V* __restrict__ rw, long cs) {
  const int b = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const long row = (long)k_line[b] * m + r;
  const int a = e1[b], c = e2[b];
  const int nl = a >= 0 ? nLr[row] : 0, nu = c >= 0 ? nUr[row] : 0;
  Cpl lc[KC], uc[KC];
#pragma unroll
  for (int q = 0; q < KC; ++q) {
    if (q < nl) lc[q] = Lr[row * KC + q];
    if (q < nu) uc[q] = Ur[row * KC + q];
  }
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    const V* yj = y + j * cs;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < KC; ++q)
      if (q < nl) acc += lc[q].val * (double)yj[(long)a * m + lc[q].idx];
#pragma unroll
    for (int q = 0; q < KC; ++q)
      if (q < nu) acc += uc[q].val * (double)yj[(long)c * m + uc[q].idx];
    rw[j * cs + row] = (V)((double)rw[j * cs + row] - acc);
  }
}

// FP16 copy of a batch of m x m FP32 blocks with one power-of-two scale per row:
// the row maximum maps below 2^15 (no overflow), the scaling itself is exact, and
// only entries more than 2^-29 below the row maximum fall into FP16 subnormals.
// One wave per row; grid (ceil(m / 4), batch).
__global__ __launch_bounds__(256) void k_to_half(int m, float* const* __restrict__ src, f16* const* __restrict__ dst,
                                                 float* const* __restrict__ scl) {
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= m) return;  // the whole wave leaves together
  const float* a = src[b] + (long)r * m;
  f16* o = dst[b] + (long)r * m;
  // rows up to 64 * TH_REG entries are read once, into registers; longer rows twice
  constexpr int TH_REG = 16;
  const bool held = m <= 64 * TH_REG;
  float v[TH_REG];
  float mx = 0.0f;
  if (held) {
#pragma unroll
    for (int t = 0; t < TH_REG; ++t) {
      const int c = lane + 64 * t;
      v[t] = c < m ? a[c] : 0.0f;
      mx = fmaxf(mx, fabsf(v[t]));
    }
  } else {
    for (int c = lane; c < m; c += 64) mx = fmaxf(mx, fabsf(a[c]));
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  int e = 0;
  if (mx > 0.0f && mx <= 3.0e38f) (void)frexpf(mx, &e);  // mx < 2^e
  if (e < -100) e = -100;
  const float inv = ldexpf(1.0f, 15 - e);
  if (held) {
#pragma unroll
    for (int t = 0; t < TH_REG; ++t) {
      const int c = lane + 64 * t;
      if (c < m) o[c] = (f16)(v[t] * inv);
    }
  } else {
    for (int c = lane; c < m; c += 64) o[c] = (f16)(a[c] * inv);
  }
  if (lane == 0) scl[b][r] = ldexpf(1.0f, e - 15);
}

// NC column pointers passed by value
template <typename P, int NC>
struct Cols {
  P p[NC];
};

// rw (line-major, NC columns at stride cs) <- the right-hand sides r[j] in natural order;
// xtab (or null) <- the output pointers x[j] for the replayed graph's k_stage_out_tab
template <typename V, int NC>
__global__ __launch_bounds__(256) void k_stage_in(int n, Order o, Cols<const V*, NC> r, V* __restrict__ rw, long cs,
                                                  Cols<V*, NC> x, V** __restrict__ xtab) {
  if (xtab && blockIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (threadIdx.x == j) xtab[j] = x.p[j];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int s = o.to_natural(q);
#pragma unroll
    for (int j = 0; j < NC; ++j) rw[j * cs + q] = r.p[j][s];
  }
}

// the solutions x[j] in natural order <- xo (line-major, NC columns at stride cs)
template <typename V, int NC>
__global__ __launch_bounds__(256) void k_stage_out(int n, Order o, const V* __restrict__ xo, long cs, Cols<V*, NC> x) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int s = o.to_natural(q);
#pragma unroll
    for (int j = 0; j < NC; ++j) x.p[j][s] = xo[j * cs + q];
  }
}

// the same with the output pointers read from xtab (written by this apply's k_stage_in),
// so that the captured graph ends with the scatter whatever the caller's x
template <typename V, int NC>
__global__ __launch_bounds__(256) void k_stage_out_tab(int n, Order o, const V* __restrict__ xo, long cs,
                                                       V* const* __restrict__ xtab) {
  V* x[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) x[j] = xtab[j];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int s = o.to_natural(q);
#pragma unroll
    for (int j = 0; j < NC; ++j) x[j][s] = xo[j * cs + q];
  }
}

// ---- host-side schedule ----------------------------------------------------------

struct LevelPlan {
  // eliminated lines
  std::vector<int> e, ea, eb;
  // kept lines
  std::vector<int> k, k1, k2, kfa, kfb, kslot;
};

std::vector<LevelPlan> make_plan(int N) {
  // The last active line is always kept (it is eliminated alone at the final
  // level): with a no-slip lid its raw diagonal block carries the 1D pressure
  // null space and is singular until the lines below it have been folded in.
  std::vector<LevelPlan> plan;
  std::vector<int> A(N);
  for (int j = 0; j < N; ++j) A[j] = j;
  int slot = 0;
  while (!A.empty()) {
    LevelPlan L;
    const int na = (int)A.size();
    const int kp = na > 1 ? (na - 1) % 2 : 1;  // parity of the kept positions
    for (int p = 0; p < na; ++p) {
      if (p % 2 != kp) {
        L.e.push_back(A[p]);
        L.ea.push_back(p >= 1 ? A[p - 1] : -1);
        L.eb.push_back(p + 1 < na ? A[p + 1] : -1);
      } else {
        L.k.push_back(A[p]);
        L.k1.push_back(p >= 1 ? A[p - 1] : -1);
        L.k2.push_back(p + 1 < na ? A[p + 1] : -1);
        L.kfa.push_back(p >= 2 ? A[p - 2] : -1);
        L.kfb.push_back(p + 2 < na ? A[p + 2] : -1);
        L.kslot.push_back(slot++);
      }
    }
    plan.push_back(L);
    A = L.k;
  }
  return plan;
}

}  // namespace

// Device-side pointer/index lists of one level, all in one pool.
struct BcrLevelDev {
  int ne, nk;
  int *e, *ea, *eb;        // eliminated lines and their kept neighbours (-1: none)
  void **Ye, **Le, **Ue;   // their D (inverted in place), L, U
  int *k, *k1, *k2;        // kept lines and their eliminated neighbours (-1: none)
  void **P1, **P2;         // per kept line
  // operand lists of the sub-batches (kept lines with e1 / with e2 / with a line before e1 / after e2)
  int n_p1, n_p2, n_fa, n_fb, n_nofa, n_nofb;
  int *k_hase1_line, *k_hase2_line, *k_fa_line, *k_fb_line;
  void **g_p1_A, **g_p1_B, **g_p1_C;  // P1 = L_k Y_e1
  void **g_d1_A, **g_d1_B;            // D_k -= P1 U_e1
  void **g_p2_A, **g_p2_B, **g_p2_C;  // P2 = U_k Y_e2
  void **g_d2_A, **g_d2_B;            // D_k -= P2 L_e2
  void **g_fa_A, **g_fa_B, **g_fa_C;  // L_k = -P1 L_e1
  void **g_fb_A, **g_fb_B, **g_fb_C;  // U_k = -P2 U_e2
  void **z_fa, **z_fb;                // L_k (U_k) that become zero
  // fused factors: L_e <- Y_e L_e, U_e <- Y_e U_e (through the scratch W) at levels >= 1
  int n_yl, n_yu;
  void **g_yl_A, **g_yl_B, **g_yl_C;  // Y_e, L_e, W
  void **g_yu_A, **g_yu_B, **g_yu_C;  // Y_e, U_e, W
  // FP16 copies read by the FP32 apply (tf_bcr_set_half): blocks and row scales
  void **Yh, **Lh, **Uh, **P1h, **P2h;
  void **Ys, **Ls, **Us, **P1s, **P2s;
};

struct BcrGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// One step of forming the dense top Z at factor time: kind 0 zeroes the blocks of
// list a, 1 copies a[i] -> b[i], 2 is the batched product c = alpha a b + beta c.
struct BcrZStep {
  int kind, count;
  void **a, **b, **c;
  double alpha, beta;
};

struct tf_bcr {
  int n, m, N, single;
  Order order;
  void* xw;    // 2n (line-major solution when the order is not natural; two columns)
  void* xtab;  // 2 pointers: the outputs of the current graph apply (k_stage_in -> k_stage_out_tab)
  // captured apply: the sweeps rw -> xw replayed as one graph launch (per vector type and column count)
  int use_graph;
  hipStream_t cap;
  BcrGraph g64[2], g32[2];   // [columns - 1]
  size_t es;
  void* D;    // N m^2: diagonal blocks, inverted in place (Y)
  void* Lm;   // N m^2: dense lower couplings (levels >= 1)
  void* Um;   // N m^2
  void* P1;   // slots m^2 (one per kept line per level)
  void* P2;   // slots m^2
  void* Pinv; // batch x 64 x 64
  void* Kr;   // batch x 64 x m
  void* Kc;   // batch x m x 64
  void* Tb;   // batch x 64 x m
  Cpl *Lr, *Ur, *Ut, *Lt;
  int *nLr, *nUr, *nUt, *nLt;
  int* flags;  // [err, nperturb]
  int* host_flags;
  void* rw;    // 2n (solve workspace, up to two columns at stride n)
  void* tb;    // 2n
  void* yb;    // 2n (level-0 Y_e r_e)
  void* pool;  // device pointer/index pool
  std::vector<LevelPlan> plan;
  std::vector<BcrLevelDev> lev;
  size_t solve_bytes;  // algorithmic bytes of one solve (dense blocks read once, vectors)
  // FP16 copies of the blocks the apply reads (single-precision factors only)
  int half;            // the FP32 apply reads the FP16 copies
  int factored;        // a factorisation has run (the copies can be made)
  void* H;             // FP16 blocks: D, L, U (N m^2 each), P1, P2 (slots m^2 each)
  void* S;             // FP32 row scales, same layout with m per block
  void* hpool;         // per-level FP16 pointer lists + the conversion list
  void **cv_src, **cv_dst, **cv_scl;  // blocks converted after each factorisation
  int ncv;
  size_t slots;
  size_t solve_bytes_half;
  int split;           // level GEMMs of an FP32 factor on BF16 terms: 0 off, 1 or 3 terms
  int fused;           // L, U of the levels >= 1 hold Y L, Y U (one backward kernel per level)
  int use_top;         // the upper levels of the fused backward and their forward levels in one k_top launch
  int use_small;       // levels with at most SMALL_MAX lines through k_small (operands by value)
  std::vector<char> pool_host, hpool_host;  // host copies of pool and hpool (k_small's operands)
  unsigned* topsync;   // k_top's counters and error word (TOP_SYNC_WORDS)
  void* W;             // scratch of the fused products (largest level, m^2 each)
  // dense top (tf_bcr_set_dense_top): Z, the inverse of the reduced system of the
  // zn <= TOPD_MAX lines active at level zl, block (i, j) at (i zn + j) m^2
  int zmax;                     // requested line limit (0: off)
  int zl, zn;                   // zn == 0: no dense top
  std::vector<int> zline;       // the lines of level zl in position order
  void* Z;                      // factor precision (also read by the FP16 apply)
  void* zpool;                  // device lists of the steps
  std::vector<BcrZStep> zsteps;
  long zreplaced;               // blocks of the levels the dense top replaces in an apply
  int z_ok;                     // Z holds the inverse for the current factor
};

namespace {

template <typename T>
void gemm(int M, int N, int K, void** A, int lda, void** B, int ldb, void** C, int ldc, double alpha, double beta,
          int batch, hipStream_t st) {
  if (batch <= 0) return;
  TF_PROF_WORK(tf::PROF_GEMM, 2.0 * M * N * K * batch);
  TF_PROF_BEGIN(tf::PROF_GEMM, st);
  tf::gemm_batched<T>(M, N, K, (T* const*)A, lda, (T* const*)B, ldb, (T* const*)C, ldc, alpha, beta, batch, st);
  TF_PROF_END(tf::PROF_GEMM, st);
}

// The level GEMMs of an FP32 factor, optionally on BF16 terms (tf_bcr_set_gemm_split).
template <typename T>
void level_gemm(tf_bcr* h, int M, int N, int K, void** A, int lda, void** B, int ldb, void** C, int ldc, double alpha,
                double beta, int batch, hipStream_t st) {
  if constexpr (sizeof(T) == 4) {
    if (h->split && batch > 0) {
      TF_PROF_WORK(tf::PROF_GEMM, 2.0 * M * N * K * batch);
      TF_PROF_BEGIN(tf::PROF_GEMM, st);
      tf::gemm_batched_split(h->split, M, N, K, (float* const*)A, lda, (float* const*)B, ldb, (float* const*)C, ldc,
                             alpha, beta, batch, st);
      TF_PROF_END(tf::PROF_GEMM, st);
      return;
    }
  }
  gemm<T>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, batch, st);
}

template <typename T>
void invert_batch(tf_bcr* h, void** Xp, int batch, hipStream_t st) {
  const int m = h->m;
  T* Pinv = (T*)h->Pinv;
  T* Kr = (T*)h->Kr;
  T* Kc = (T*)h->Kc;
  T* Tb = (T*)h->Tb;
  // per-batch pointers to the snapshot buffers are laid out contiguously: the
  // GEMMs need pointer lists, which live right after Xp in the pool
  void** pinvp = Xp + batch;
  void** krp = pinvp + batch;
  void** kcp = krp + batch;
  void** tbp = kcp + batch;
  const long sz = (long)PB * m;
  const int gcopy = (int)((sz + 255) / 256) < 1024 ? (int)((sz + 255) / 256) : 1024;
  for (int k0 = 0; k0 < m; k0 += PB) {
    const int pb = m - k0 < PB ? m - k0 : PB;
    TF_PROF_WORK(tf::PROF_GJ_PANEL, 2.0 * 64 * 64 * 128 * batch);
    TF_PROF_BEGIN(tf::PROF_GJ_PANEL, st);
    hipLaunchKernelGGL(k_panel<T>, dim3(batch), dim3(512), 0, st, m, k0, pb, (T* const*)Xp, Pinv, h->flags + 1);
    TF_PROF_END(tf::PROF_GJ_PANEL, st);
    hipLaunchKernelGGL(k_panel_copy<T>, dim3(gcopy, batch), dim3(256), 0, st, m, k0, pb, (T* const*)Xp, Kr, Kc);
    // T = Pinv * Kr   (pb x pb) x (pb x m)
    gemm<T>(pb, m, pb, pinvp, pb, krp, m, tbp, m, 1.0, 0.0, batch, st);
    // X -= Kc * T     (m x pb) x (pb x m)
    gemm<T>(m, m, pb, kcp, pb, tbp, m, Xp, m, -1.0, 1.0, batch, st);
    hipLaunchKernelGGL(k_panel_store<T>, dim3(gcopy, batch), dim3(256), 0, st, m, k0, pb, (T* const*)Xp,
                       (const T*)Tb);
  }
  (void)Pinv; (void)Kr; (void)Kc; (void)Tb;
}

// the captured applies (every vector type and column count) are made again on next use
void drop_graphs(tf_bcr* h) {
  for (BcrGraph* g : {&h->g64[0], &h->g64[1], &h->g32[0], &h->g32[1]}) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    *g = BcrGraph{};
  }
}

template <typename T>
int factor_impl(tf_bcr* h, const int* rp, const int* ci, const double* v, int gauge, hipStream_t st) {
  const int m = h->m, N = h->N, n = h->n;
  hipLaunchKernelGGL(k_zero_int, dim3(4), dim3(256), 0, st, 2L, h->flags);
  hipLaunchKernelGGL(k_zero_int, dim3(256), dim3(256), 0, st, (long)n, h->nUt);
  hipLaunchKernelGGL(k_zero_int, dim3(256), dim3(256), 0, st, (long)n, h->nLt);
  hipLaunchKernelGGL(k_couplings4, dim3((n + 255) / 256), dim3(256), 0, st, n, m, h->order, rp, ci, v, gauge, h->Lr,
                     h->nLr,
                     h->Ur, h->nUr, h->Ut, h->nUt, h->Lt, h->nLt, h->flags);
  hipLaunchKernelGGL(k_sort_cols, dim3((n + 255) / 256), dim3(256), 0, st, n, h->Ut, h->nUt);
  hipLaunchKernelGGL(k_sort_cols, dim3((n + 255) / 256), dim3(256), 0, st, n, h->Lt, h->nLt);
  hipLaunchKernelGGL(k_diag_blocks<T>, dim3(m, N), dim3(256), 0, st, m, h->order, rp, ci, v, gauge, (T*)h->D);
  for (size_t l = 0; l < h->lev.size(); ++l) {
    BcrLevelDev& L = h->lev[l];
    // 1. invert the eliminated lines' diagonal blocks in place
    invert_batch<T>(h, L.Ye, L.ne, st);
    if (L.nk == 0) continue;
    if (l == 0) {
      // sparse couplings at level 0 (e1 = k-1, e2 = k+1):
      // P1 = L_k Y_e1, P2 = U_k Y_e2 (sparse rows of line k)
      if (L.n_p1)
        hipLaunchKernelGGL(k_sp_rows<T>, dim3(m, L.n_p1), dim3(256), 0, st, m, h->Lr, h->nLr, L.k_hase1_line,
                           (T* const*)L.g_p1_B, (T* const*)L.g_p1_C);
      if (L.n_p2)
        hipLaunchKernelGGL(k_sp_rows<T>, dim3(m, L.n_p2), dim3(256), 0, st, m, h->Ur, h->nUr, L.k_hase2_line,
                           (T* const*)L.g_p2_B, (T* const*)L.g_p2_C);
      // D_k -= P1 U_e1 (rows of line k-1 into the columns of line k: Ut)
      // D_k -= P2 L_e2 (rows of line k+1 into the columns of line k: Lt)
      if (L.n_p1)
        hipLaunchKernelGGL(k_sp_cols<T>, spc_grid(m, L.n_p1), dim3(256), 0, st, m, h->Ut, h->nUt,
                           L.k_hase1_line,
                           (T* const*)L.g_p1_C, (T* const*)L.g_d1_A, -1.0, 1);
      if (L.n_p2)
        hipLaunchKernelGGL(k_sp_cols<T>, spc_grid(m, L.n_p2), dim3(256), 0, st, m, h->Lt, h->nLt,
                           L.k_hase2_line,
                           (T* const*)L.g_p2_C, (T* const*)L.g_d2_A, -1.0, 1);
      // L_k = -P1 L_e1 (columns of line k-2: Lt) ; U_k = -P2 U_e2 (columns of line k+2: Ut)
      if (L.n_fa)
        hipLaunchKernelGGL(k_sp_cols<T>, spc_grid(m, L.n_fa), dim3(256), 0, st, m, h->Lt, h->nLt,
                           L.k_fa_line,
                           (T* const*)L.g_fa_A, (T* const*)L.g_fa_C, -1.0, 0);
      if (L.n_fb)
        hipLaunchKernelGGL(k_sp_cols<T>, spc_grid(m, L.n_fb), dim3(256), 0, st, m, h->Ut, h->nUt,
                           L.k_fb_line,
                           (T* const*)L.g_fb_A, (T* const*)L.g_fb_C, -1.0, 0);
    } else {
      level_gemm<T>(h, m, m, m, L.g_p1_A, m, L.g_p1_B, m, L.g_p1_C, m, 1.0, 0.0, L.n_p1, st);
      level_gemm<T>(h, m, m, m, L.g_p2_A, m, L.g_p2_B, m, L.g_p2_C, m, 1.0, 0.0, L.n_p2, st);
      level_gemm<T>(h, m, m, m, L.g_p1_C, m, L.g_d1_B, m, L.g_d1_A, m, -1.0, 1.0, L.n_p1, st);
      level_gemm<T>(h, m, m, m, L.g_p2_C, m, L.g_d2_B, m, L.g_d2_A, m, -1.0, 1.0, L.n_p2, st);
      level_gemm<T>(h, m, m, m, L.g_fa_A, m, L.g_fa_B, m, L.g_fa_C, m, -1.0, 0.0, L.n_fa, st);
      level_gemm<T>(h, m, m, m, L.g_fb_A, m, L.g_fb_B, m, L.g_fb_C, m, -1.0, 0.0, L.n_fb, st);
      if (h->fused) {
        // the eliminated lines' L_e, U_e are not used by the factorisation after
        // this level: replace them by Y_e L_e and Y_e U_e for the one-pass backward
        const long mm2 = (long)m * m;
        gemm<T>(m, m, m, L.g_yl_A, m, L.g_yl_B, m, L.g_yl_C, m, 1.0, 0.0, L.n_yl, st);
        if (L.n_yl)
          hipLaunchKernelGGL(k_copy_blocks<T>, dim3(256, L.n_yl), dim3(256), 0, st, mm2, (T* const*)L.g_yl_C,
                             (T* const*)L.g_yl_B);
        gemm<T>(m, m, m, L.g_yu_A, m, L.g_yu_B, m, L.g_yu_C, m, 1.0, 0.0, L.n_yu, st);
        if (L.n_yu)
          hipLaunchKernelGGL(k_copy_blocks<T>, dim3(256, L.n_yu), dim3(256), 0, st, mm2, (T* const*)L.g_yu_C,
                             (T* const*)L.g_yu_B);
      }
    }
    // couplings that no longer exist
    const long mm = (long)m * m;
    if (L.n_nofa)
      hipLaunchKernelGGL(k_zero_many<T>, dim3(256, L.n_nofa), dim3(256), 0, st, mm, (T* const*)L.z_fa);
    if (L.n_nofb)
      hipLaunchKernelGGL(k_zero_many<T>, dim3(256, L.n_nofb), dim3(256), 0, st, mm, (T* const*)L.z_fb);
  }
  h->factored = 1;
  if constexpr (sizeof(T) == 4) {
    if (h->half && h->ncv > 0)
      hipLaunchKernelGGL(k_to_half, dim3((m + 3) / 4, h->ncv), dim3(256), 0, st, m, (float* const*)h->cv_src,
                         (f16* const*)h->cv_dst, (float* const*)h->cv_scl);
  }
  // the dense top reads Y L and Y U of the levels from zl up: fused factors only
  const int zok = (h->zn > 0 && h->fused) ? 1 : 0;
  if (zok) {
    const long mm = (long)m * m;
    for (const BcrZStep& s : h->zsteps) {
      if (s.kind == 0)
        hipLaunchKernelGGL(k_zero_many<T>, dim3(256, s.count), dim3(256), 0, st, mm, (T* const*)s.a);
      else if (s.kind == 1)
        hipLaunchKernelGGL(k_copy_blocks<T>, dim3(256, s.count), dim3(256), 0, st, mm, (T* const*)s.a,
                           (T* const*)s.b);
      else
        gemm<T>(m, m, m, s.a, m, s.b, m, s.c, m, s.alpha, s.beta, s.count, st);
    }
  }
  if (zok != h->z_ok) drop_graphs(h);  // the captured sweeps take the other path
  h->z_ok = zok;
  TF_LAUNCH_CHECK("bcr factor");
  return 0;
}

// Host value of entry i of a device list that lives in the handle's pool or hpool
// (false when the entry is in neither).
template <typename P>
bool host_entry(const tf_bcr* h, const void* dev_list, int i, P& v) {
  const char* d = (const char*)dev_list;
  const size_t at = sizeof(P) * (size_t)i;
  for (int q = 0; q < 2; ++q) {
    const char* d0 = (const char*)(q == 0 ? h->pool : h->hpool);
    const std::vector<char>& mirror = q == 0 ? h->pool_host : h->hpool_host;
    if (d0 && d >= d0 && (size_t)(d - d0) + at + sizeof(P) <= mirror.size()) {
      memcpy(&v, mirror.data() + (d - d0) + at, sizeof(P));
      return true;
    }
  }
  return false;
}

template <typename T>
bool host_row(const tf_bcr* h, const Blocks<T>& b, int i, typename RowOf<Blocks<T>>::type& r) {
  return host_entry(h, b.p, i, r.p);
}
inline bool host_row(const tf_bcr* h, const HalfBlocks& b, int i, RowOf<HalfBlocks>::type& r) {
  return host_entry(h, b.p, i, r.p) && host_entry(h, b.s, i, r.s);
}

// k_small's operands of one level step (batch <= SMALL_MAX lines); false when a list
// is not in the handle's pools (the caller then uses the list-reading kernels)
template <typename B>
bool small_level(const tf_bcr* h, int batch, const int* line, const int* n1, const int* n2, const B& b0,
                 const B& b1, const B* b2, SmallLevel<B>& a) {
  a = SmallLevel<B>{};
  if (batch < 1 || batch > SMALL_MAX) return false;
  for (int i = 0; i < batch; ++i) {
    if (!host_entry(h, line, i, a.line[i]) || !host_entry(h, n1, i, a.n1[i]) || !host_entry(h, n2, i, a.n2[i]) ||
        !host_row(h, b0, i, a.blk[0][i]) || !host_row(h, b1, i, a.blk[1][i]) ||
        (b2 && !host_row(h, *b2, i, a.blk[2][i])))
      return false;
  }
  return true;
}

// the dense top's block (i, j) as rows of the factor's storage
template <typename T>
bool dense_row(const tf_bcr* h, const Blocks<T>*, int i, int j, typename RowOf<Blocks<T>>::type& r) {
  if (sizeof(T) != h->es || !h->Z) return false;
  r.p = (const T*)h->Z + (size_t)h->m * h->m * (size_t)(i * h->zn + j);
  return true;
}

// the rows the dense top is read through: the factor's storage, also under the FP16
// apply (FP16 copies of Z made a weaker preconditioner: about 13% more inner
// iterations in the 256 x 256 cavity continuation)
template <typename B>
struct ZRows {
  typedef B type;
};
template <>
struct ZRows<HalfBlocks> {
  typedef Blocks<float> type;
};

// k_dense_top's operands; false when the handle holds no dense top for this factor
template <typename B>
bool dense_top(const tf_bcr* h, DenseTop<B>& a) {
  a = DenseTop<B>{};
  if (!h->z_ok || h->zn < 2 || h->zn > TOPD_MAX || (int)h->zline.size() != h->zn) return false;
  a.n = h->zn;
  for (int i = 0; i < h->zn; ++i) {
    a.line[i] = h->zline[i];
    for (int j = 0; j < h->zn; ++j)
      if (!dense_row(h, (const B*)nullptr, i, j, a.blk[i][j])) return false;
  }
  return true;
}

// forward and backward sweeps over the levels, rows read through the block
// source B, for NC columns at stride cs in rw, tb, xo (and the handle's yb)
template <typename B, typename V, int NC>
void sweep(tf_bcr* h, int vec, B P1(const BcrLevelDev&), B P2(const BcrLevelDev&), B Lb(const BcrLevelDev&),
           B Ub(const BcrLevelDev&), B Yb(const BcrLevelDev&), V* rw, V* tb, V* xo, long cs, hipStream_t st) {
  const int m = h->m;
  const dim3 blk(256);
  auto gx = [m](int G, int R) { return (unsigned)((m + R * rows_per_block(G) - 1) / (R * rows_per_block(G))); };
  const unsigned gs = (unsigned)((m + 255) / 256);
  V* yb = (V*)h->yb;
  const int nlev = (int)h->lev.size();
  // the upper levels (every level from lt up has fewer than TOP_ROWS rows per
  // launch) go to one k_top launch when the backward is fused
  int lt = nlev;
  while (lt > 1 && (long)std::max(h->lev[lt - 1].nk, h->lev[lt - 1].ne) * m < TOP_ROWS) --lt;
  auto steps_from = [&](int l0) {
    int c = 0;
    for (int l = l0; l < nlev; ++l) c += (h->lev[l].nk > 0) + (h->lev[l].ne > 0);
    return c;
  };
  while (lt < nlev && steps_from(lt) > MAX_TOP) ++lt;
  typedef typename ZRows<B>::type ZB;
  DenseTop<ZB> zt;
  const bool dense = dense_top<ZB>(h, zt);
  TopPlan<B> tp{};
  if (!dense && h->fused && h->use_top && lt < nlev) {
    auto add = [&](int dir, int batch, const int* line, const int* n1, const int* n2, B b0, B b1, B b2) {
      TopStep<B>& S = tp.s[tp.nsteps++];
      S.dir = dir;
      S.batch = batch;
      S.G = group_for(batch);
      S.nrb = (m + rows_per_block(S.G) - 1) / rows_per_block(S.G);
      S.first = tp.total;
      S.line = line;
      S.n1 = n1;
      S.n2 = n2;
      S.blk[0] = b0;
      S.blk[1] = b1;
      S.blk[2] = b2;
      tp.total += batch * S.nrb;
    };
    for (int l = lt; l < nlev; ++l) {
      const BcrLevelDev& L = h->lev[l];
      if (L.nk) add(0, L.nk, L.k, L.k1, L.k2, P1(L), P2(L), P2(L));
    }
    for (int l = nlev - 1; l >= lt; --l) {
      const BcrLevelDev& L = h->lev[l];
      if (L.ne) add(1, L.ne, L.e, L.ea, L.eb, Yb(L), Lb(L), Ub(L));
    }
    if (tp.nsteps < 2) tp = TopPlan<B>{};
  }
  const bool top = tp.nsteps >= 2;
  const int fwd_end = dense ? h->zl : (top ? lt : nlev);
  for (int l = 0; l < fwd_end; ++l) {
    const BcrLevelDev& L = h->lev[l];
    if (!L.nk) continue;
    if (l == 0) {
      // the level-0 couplings are sparse: y_e = Y_e r_e, then r_k -= L_k y_e1 + U_k y_e2
      // (one Y block per eliminated line instead of two multipliers per kept line)
      const int G = group_for(L.ne);
      with_rows(rows_for<B>(L.ne, m), [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        hipLaunchKernelGGL((k_bwd_x<B, V, NC, R>), dim3(gx(G, R), L.ne), blk, 0, st, m, G, vec, L.e, Yb(L),
                           (const V*)rw, yb, cs);
      });
      hipLaunchKernelGGL((k_fwd_sparse<V, NC>), dim3(gs, L.nk), blk, 0, st, m, L.k, L.k1, L.k2, h->Lr, h->nLr, h->Ur,
                         h->nUr, (const V*)yb, rw, cs);
    } else if (SmallLevel<B> sa; h->use_small && L.nk <= SMALL_MAX && rows_for<B>(L.nk, m) == 1 &&
                                 small_level<B>(h, L.nk, L.k, L.k1, L.k2, P1(L), P2(L), nullptr, sa)) {
      hipLaunchKernelGGL((k_small<B, V, NC>), dim3(gx(group_for(L.nk), 1), L.nk), blk, 0, st, m, group_for(L.nk),
                         vec, 0, sa, rw, xo, cs);
    } else {
      const int G = group_for(L.nk);
      with_rows(rows_for<B>(L.nk, m), [&](auto Rc) {
        constexpr int R = decltype(Rc)::value;
        hipLaunchKernelGGL((k_fwd<B, V, NC, R>), dim3(gx(G, R), L.nk), blk, 0, st, m, G, vec, L.k, L.k1, L.k2, P1(L),
                           P2(L), rw, cs);
      });
    }
  }
  if (top)
    hipLaunchKernelGGL((k_top<B, V, NC>), dim3(TOP_BLOCKS), blk, 0, st, m, vec, tp, rw, xo, cs, h->topsync);
  if (dense) {
    const int G = group_for(zt.n);
    hipLaunchKernelGGL((k_dense_top<ZB, V, NC>), dim3(gx(G, 1), zt.n), blk, 0, st, m, G, vec, zt, (const V*)rw, xo,
                       cs);
  }
  for (int l = dense ? h->zl - 1 : (top ? lt - 1 : nlev - 1); l >= 0; --l) {
    const BcrLevelDev& L = h->lev[l];
    const int G = group_for(L.ne);
    if (SmallLevel<B> sa; l > 0 && h->fused && h->use_small && L.ne <= SMALL_MAX && rows_for<B>(L.ne, m) == 1) {
      const B ub = Ub(L);
      if (small_level<B>(h, L.ne, L.e, L.ea, L.eb, Yb(L), Lb(L), &ub, sa)) {
        hipLaunchKernelGGL((k_small<B, V, NC>), dim3(gx(G, 1), L.ne), blk, 0, st, m, G, vec, 1, sa, rw, xo, cs);
        continue;
      }
    }
    with_rows(rows_for<B>(L.ne, m), [&](auto Rc) {
      constexpr int R = decltype(Rc)::value;
      if (l > 0 && h->fused) {
        hipLaunchKernelGGL((k_bwd_fused<B, V, NC, R>), dim3(gx(G, R), L.ne), blk, 0, st, m, G, vec, L.e, L.ea, L.eb,
                           Yb(L), Lb(L), Ub(L), (const V*)rw, xo, cs);
        return;
      }
      if (l == 0)
        hipLaunchKernelGGL((k_bwd_t_sparse<V, NC>), dim3(gs, L.ne), blk, 0, st, m, L.e, L.ea, L.eb, h->Lr, h->nLr,
                           h->Ur, h->nUr, (const V*)rw, (const V*)xo, tb, cs);
      else
        hipLaunchKernelGGL((k_bwd_t<B, V, NC, R>), dim3(gx(G, R), L.ne), blk, 0, st, m, G, vec, L.e, L.ea, L.eb, Lb(L),
                           Ub(L), (const V*)rw, (const V*)xo, tb, cs);
      hipLaunchKernelGGL((k_bwd_x<B, V, NC, R>), dim3(gx(G, R), L.ne), blk, 0, st, m, G, vec, L.e, Yb(L),
                         (const V*)tb, xo, cs);
    });
  }
}

// 16-byte row loads need m a multiple of the vector width and every column of
// the vectors 16-byte aligned
inline bool aligned16(const void* p, long cs, int nc, size_t es) {
  return ((uintptr_t)p & 15) == 0 && (nc == 1 || (cs * (long)es) % 16 == 0);
}

template <typename T, typename V, int NC>
void levels(tf_bcr* h, V* rw, V* tb, V* xo, long cs, hipStream_t st) {
  constexpr int W = 16 / (int)sizeof(T);
  const int vec = (sizeof(T) == sizeof(V) && h->m % W == 0 && aligned16(xo, cs, NC, sizeof(V))) ? 1 : 0;
  typedef Blocks<T> B;
  sweep<B, V, NC>(
      h, vec, [](const BcrLevelDev& L) { return B{(T* const*)L.P1}; },
      [](const BcrLevelDev& L) { return B{(T* const*)L.P2}; }, [](const BcrLevelDev& L) { return B{(T* const*)L.Le}; },
      [](const BcrLevelDev& L) { return B{(T* const*)L.Ue}; }, [](const BcrLevelDev& L) { return B{(T* const*)L.Ye}; },
      rw, tb, xo, cs, st);
}

template <int NC>
void levels_half(tf_bcr* h, float* rw, float* tb, float* xo, long cs, hipStream_t st) {
  const int vec = (h->m % 8 == 0 && aligned16(xo, cs, NC, sizeof(float))) ? 1 : 0;
  typedef HalfBlocks B;
  sweep<B, float, NC>(
      h, vec, [](const BcrLevelDev& L) { return B{(f16* const*)L.P1h, (float* const*)L.P1s}; },
      [](const BcrLevelDev& L) { return B{(f16* const*)L.P2h, (float* const*)L.P2s}; },
      [](const BcrLevelDev& L) { return B{(f16* const*)L.Lh, (float* const*)L.Ls}; },
      [](const BcrLevelDev& L) { return B{(f16* const*)L.Uh, (float* const*)L.Us}; },
      [](const BcrLevelDev& L) { return B{(f16* const*)L.Yh, (float* const*)L.Ys}; }, rw, tb, xo, cs, st);
}

// algorithmic bytes of an NC-column apply: the factor blocks once (the dense top's in
// place of the levels it replaces), six vectors per column
template <typename T, typename V, int NC>
double solve_work(const tf_bcr* h) {
  const bool hb = sizeof(T) == 4 && sizeof(V) == 4 && h->half;
  double one = (double)(hb ? h->solve_bytes_half : h->solve_bytes);
  if (h->z_ok) {
    // the dense top's blocks (factor precision) in place of those of the levels it replaces
    const double mm = (double)h->m * h->m;
    one += (double)h->zn * h->zn * h->es * mm - (double)h->zreplaced * (hb ? 2.0 * mm + 4.0 * h->m : h->es * mm);
  }
  return one + (NC - 1) * 6.0 * sizeof(V) * (double)h->n;
}

constexpr int solve_prof_id(int nc) { return nc == 1 ? tf::PROF_SOLVE : tf::PROF_SOLVE2; }

// The sweeps of an apply on the work vectors: rw (right-hand sides, line-major,
// overwritten) -> xo (solutions, line-major), NC columns at stride n.
template <typename T, typename V, int NC>
void run_levels(tf_bcr* h, V* rw, V* xo, hipStream_t st) {
  const long n = h->n;
  V* tb = (V*)h->tb;
  bool half = false;
  if constexpr (sizeof(T) == 4 && sizeof(V) == 4) half = h->half != 0;
  if constexpr (sizeof(T) == 4 && sizeof(V) == 4) {
    if (half) levels_half<NC>(h, (float*)rw, (float*)tb, (float*)xo, n, st);
  }
  if (!half) levels<T, V, NC>(h, rw, tb, xo, n, st);
}

// r[j] (natural order) -> the work vector rw, one launch for all columns; x (or null):
// the output pointers into the handle's table for the graph's scatter
template <typename V, int NC>
void stage_in(const tf_bcr* h, const V* const (&r)[NC], V* rw, V* const (*x)[NC], hipStream_t st) {
  Cols<const V*, NC> c;
  Cols<V*, NC> xc{};
  for (int j = 0; j < NC; ++j) {
    c.p[j] = r[j];
    if (x) xc.p[j] = (*x)[j];
  }
  hipLaunchKernelGGL((k_stage_in<V, NC>), dim3(1024), dim3(256), 0, st, h->n, h->order, c, rw, (long)h->n, xc,
                     x ? (V**)h->xtab : (V**)nullptr);
}

// xo (line-major) -> x[j] (natural order), one launch for all columns
template <typename V, int NC>
void stage_out(const tf_bcr* h, const V* xo, V* const (&x)[NC], hipStream_t st) {
  Cols<V*, NC> c;
  for (int j = 0; j < NC; ++j) c.p[j] = x[j];
  hipLaunchKernelGGL((k_stage_out<V, NC>), dim3(1024), dim3(256), 0, st, h->n, h->order, xo, (long)h->n, c);
}

// NC right-hand sides r[j] -> solutions x[j] (host arrays of device pointers), eager.
template <typename T, typename V, int NC>
int solve_impl(tf_bcr* h, const V* const (&r)[NC], V* const (&x)[NC], hipStream_t st) {
  V* rw = (V*)h->rw;
  TF_PROF_WORK(solve_prof_id(NC), (solve_work<T, V, NC>(h)));
  TF_PROF_BEGIN(solve_prof_id(NC), st);
  // one column in natural order: the kernels write the solution in place
  const bool direct = !h->order.ylines && NC == 1;
  V* xo = direct ? x[0] : (V*)h->xw;
  stage_in<V, NC>(h, r, rw, nullptr, st);
  run_levels<T, V, NC>(h, rw, xo, st);
  if (!direct) stage_out<V, NC>(h, xo, x, st);
  TF_PROF_END(solve_prof_id(NC), st);
  TF_LAUNCH_CHECK("bcr solve");
  return 0;
}

// The sweeps (rw -> xw) and the scatter to the caller's x captured once per (vector
// type, column count) on a private stream and replayed with one graph launch after
// the gather kernel, which also writes this call's x pointers into the table the
// scatter reads (a scatter launched after the replay started about 8 us after its
// end).  The factor refactors in place, so the captured pointers stay valid.  The
// graph holds kernel nodes only.
template <typename T, typename V, int NC>
int solve_graph(tf_bcr* h, const V* const (&r)[NC], V* const (&x)[NC], hipStream_t st) {
  BcrGraph& gs = (sizeof(V) == 8 ? h->g64 : h->g32)[NC - 1];
  V* rw = (V*)h->rw;
  V* xw = (V*)h->xw;
  if (!gs.exec) {
    // non-blocking: a capture must not order against the legacy stream that other
    // threads use for synchronous copies
    if (!h->cap) TF_HIP(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    hipError_t le = hipSuccess, ce = hipSuccess, ie = hipSuccess;
    {
      std::lock_guard<std::mutex> lock(tf::capture_mutex());
      TF_HIP(hipStreamBeginCapture(h->cap, hipStreamCaptureModeThreadLocal));
      run_levels<T, V, NC>(h, rw, xw, h->cap);
      hipLaunchKernelGGL((k_stage_out_tab<V, NC>), dim3(1024), dim3(256), 0, h->cap, h->n, h->order,
                         (const V*)xw, (long)h->n, (V* const*)h->xtab);
      le = hipGetLastError();
      ce = hipStreamEndCapture(h->cap, &g);
      if (le == hipSuccess && ce == hipSuccess && g) ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    }
    if (le != hipSuccess || ce != hipSuccess || !g || ie != hipSuccess) {
      // a failed capture or instantiation: this apply runs eagerly, with the same
      // kernels, and is captured again next time
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      tf::note_capture_fallback();
      return solve_impl<T, V, NC>(h, r, x, st);
    }
    gs.graph = g;
    gs.exec = ex;
  }
  // the profiler's events bracket what an apply costs on this path: the staging kernels and the replay
  TF_PROF_WORK(solve_prof_id(NC), (solve_work<T, V, NC>(h)));
  TF_PROF_BEGIN(solve_prof_id(NC), st);
  stage_in<V, NC>(h, r, rw, &x, st);
  TF_HIP(hipGraphLaunch(gs.exec, st));
  TF_PROF_END(solve_prof_id(NC), st);
  TF_LAUNCH_CHECK("bcr graph apply");
  return 0;
}

template <typename V, int NC>
int solve_any(tf_bcr* h, const V* const (&r)[NC], V* const (&x)[NC], hipStream_t st) {
  if (h->use_graph)
    return h->single ? solve_graph<float, V, NC>(h, r, x, st) : solve_graph<double, V, NC>(h, r, x, st);
  return h->single ? solve_impl<float, V, NC>(h, r, x, st) : solve_impl<double, V, NC>(h, r, x, st);
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------------

namespace {

struct PoolBuilder {
  std::vector<char> bytes;
  // returns the byte offset of n pointer-sized (or int) slots
  size_t add_ptrs(const std::vector<void*>& v) {
    size_t off = align();
    bytes.resize(off + v.size() * sizeof(void*));
    if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(void*));
    return off;
  }
  size_t add_ints(const std::vector<int>& v) {
    size_t off = align();
    bytes.resize(off + v.size() * sizeof(int));
    if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(int));
    return off;
  }
  size_t align() {
    size_t o = (bytes.size() + 63) & ~size_t(63);
    bytes.resize(o);
    return o;
  }
};

}  // namespace

extern "C" {

// ylines != 0 (2D only): lines are the grid columns i (m = ny dof rows each, nx
// lines) instead of the rows j; n = nx ny dof.
int tf_bcr_destroy(tf_bcr* h);
static int bcr_create_impl(tf_bcr* h, int n, int m, int single, int nx, int dof, int ylines);

int tf_bcr_create(tf_bcr** out, int n, int m, int single, int nx, int dof, int ylines) {
  if (!out) return tf::fail("null handle");
  *out = nullptr;
  if (m < 1 || n % m != 0) return tf::fail("line block size must divide n");
  if (ylines && (nx < 1 || dof < 1 || m % dof != 0 || (long)nx * m != n))
    return tf::fail("y-line order needs n = nx * m with m = ny * dof");
  tf_bcr* h = new tf_bcr();
  const int rc = bcr_create_impl(h, n, m, single, nx, dof, ylines);
  if (rc) {
    // frees whatever was allocated before the failure (unset members are null)
    tf_bcr_destroy(h);
    return rc;
  }
  *out = h;
  return 0;
}

static int bcr_create_impl(tf_bcr* h, int n, int m, int single, int nx, int dof, int ylines) {
  h->order = Order{ylines ? 1 : 0, nx, dof, m};
  h->n = n;
  h->m = m;
  h->N = n / m;
  h->single = single;
  h->es = single ? sizeof(float) : sizeof(double);
  const size_t mm = (size_t)m * m, N = h->N;
  const size_t maxb = (N + 1) / 2;
  TF_HIP(hipMalloc(&h->D, h->es * mm * N));
  TF_HIP(hipMalloc(&h->Lm, h->es * mm * N));
  TF_HIP(hipMalloc(&h->Um, h->es * mm * N));
  // one multiplier pair per kept line per level (more than N - 1 when the
  // active count is odd at some level)
  h->plan = make_plan(h->N);
  size_t slots = 0;
  for (const LevelPlan& L : h->plan) slots += L.k.size();
  if (slots < 1) slots = 1;
  h->slots = slots;
  TF_HIP(hipMalloc(&h->P1, h->es * mm * slots));
  TF_HIP(hipMalloc(&h->P2, h->es * mm * slots));
  TF_HIP(hipMalloc(&h->Pinv, h->es * PB * PB * maxb));
  TF_HIP(hipMalloc(&h->Kr, h->es * PB * (size_t)m * maxb));
  TF_HIP(hipMalloc(&h->Kc, h->es * PB * (size_t)m * maxb));
  TF_HIP(hipMalloc(&h->Tb, h->es * PB * (size_t)m * maxb));
  for (Cpl** c : {&h->Lr, &h->Ur, &h->Ut, &h->Lt}) TF_HIP(hipMalloc(c, sizeof(Cpl) * KC * (size_t)n));
  for (int** c : {&h->nLr, &h->nUr, &h->nUt, &h->nLt}) TF_HIP(hipMalloc(c, sizeof(int) * (size_t)n));
  TF_HIP(hipMalloc(&h->flags, sizeof(int) * 2));
  TF_HIP(hipHostMalloc(&h->host_flags, sizeof(int) * 3, hipHostMallocDefault));  // [err, nperturb, top waits]
  TF_HIP(hipMalloc(&h->rw, 2 * sizeof(double) * (size_t)n));   // two columns
  TF_HIP(hipMalloc(&h->tb, 2 * sizeof(double) * (size_t)n));   // two columns
  TF_HIP(hipMalloc(&h->yb, 2 * sizeof(double) * (size_t)n));   // two columns
  TF_HIP(hipMalloc(&h->xw, 2 * sizeof(double) * (size_t)n));   // two columns
  TF_HIP(hipMalloc(&h->xtab, 2 * sizeof(void*)));
  TF_HIP(hipMalloc(&h->topsync, sizeof(unsigned) * TOP_SYNC_WORDS));
  {
    // zeroed through the library's non-blocking upload stream: a hipMemset on the legacy
    // stream fails, or invalidates the capture, while another thread captures a graph
    const unsigned zeros[TOP_SYNC_WORDS] = {};
    TF_HIP(tf::upload(h->topsync, zeros, sizeof(zeros)));
  }
  h->use_top = 0;  // off: see tf_bcr_set_top
  h->use_small = 1;

  // device pointer lists of the schedule
  char* D = (char*)h->D;
  char* Lm = (char*)h->Lm;
  char* Um = (char*)h->Um;
  char* P1 = (char*)h->P1;
  char* P2 = (char*)h->P2;
  const size_t bm = h->es * mm;
  auto blk = [&](char* base, int j) -> void* { return (void*)(base + bm * (size_t)j); };
  PoolBuilder pb;
  struct Off {
    size_t e, ea, eb, Ye, Le, Ue, k, k1, k2, P1, P2, p1A, p1B, p1C, d1A, d1B, p2A, p2B, p2C, d2A, d2B, faA, faB,
        faC, fbA, fbB, fbC, zfa, zfb, hase1, hase2, faline, fbline, ylA, ylB, ylC, yuA, yuB, yuC;
    int ne, nk, np1, np2, nfa, nfb, nnofa, nnofb, nyl, nyu;
  };
  std::vector<Off> offs;
  // scratch of the fused products: the largest batch of one level (levels >= 1)
  size_t wslots = 0;
  for (size_t l = 1; l < h->plan.size(); ++l) {
    size_t na = 0, nb = 0;
    for (size_t b = 0; b < h->plan[l].e.size(); ++b) {
      na += h->plan[l].ea[b] >= 0;
      nb += h->plan[l].eb[b] >= 0;
    }
    wslots = std::max(wslots, std::max(na, nb));
  }
  if (wslots) TF_HIP(hipMalloc(&h->W, bm * wslots));
  int l_index = 0;
  for (const LevelPlan& L : h->plan) {
    Off o{};
    o.ne = (int)L.e.size();
    o.nk = (int)L.k.size();
    o.e = pb.add_ints(L.e);
    o.ea = pb.add_ints(L.ea);
    o.eb = pb.add_ints(L.eb);
    std::vector<void*> ye, le, ue, pinv, kr, kc, tbv;
    for (int b = 0; b < o.ne; ++b) {
      ye.push_back(blk(D, L.e[b]));
      le.push_back(blk(Lm, L.e[b]));
      ue.push_back(blk(Um, L.e[b]));
      pinv.push_back((char*)h->Pinv + h->es * PB * PB * b);
      kr.push_back((char*)h->Kr + h->es * PB * (size_t)m * b);
      kc.push_back((char*)h->Kc + h->es * PB * (size_t)m * b);
      tbv.push_back((char*)h->Tb + h->es * PB * (size_t)m * b);
    }
    // Ye must be followed by the snapshot pointer lists (see invert_batch)
    std::vector<void*> ychain = ye;
    for (auto* v : {&pinv, &kr, &kc, &tbv}) ychain.insert(ychain.end(), v->begin(), v->end());
    o.Ye = pb.add_ptrs(ychain);
    o.Le = pb.add_ptrs(le);
    o.Ue = pb.add_ptrs(ue);
    o.k = pb.add_ints(L.k);
    o.k1 = pb.add_ints(L.k1);
    o.k2 = pb.add_ints(L.k2);
    std::vector<void*> p1, p2;
    std::vector<void*> p1A, p1B, p1C, d1A, d1B, p2A, p2B, p2C, d2A, d2B, faA, faB, faC, fbA, fbB, fbC, zfa, zfb;
    std::vector<int> hase1, hase2, faline, fbline;
    for (int b = 0; b < o.nk; ++b) {
      const int k = L.k[b], e1 = L.k1[b], e2 = L.k2[b], s = L.kslot[b];
      p1.push_back(blk(P1, s));
      p2.push_back(blk(P2, s));
      if (e1 >= 0) {
        p1A.push_back(blk(Lm, k));
        p1B.push_back(blk(D, e1));
        p1C.push_back(blk(P1, s));
        d1A.push_back(blk(D, k));
        d1B.push_back(blk(Um, e1));
        hase1.push_back(k);
      }
      if (e2 >= 0) {
        p2A.push_back(blk(Um, k));
        p2B.push_back(blk(D, e2));
        p2C.push_back(blk(P2, s));
        d2A.push_back(blk(D, k));
        d2B.push_back(blk(Lm, e2));
        hase2.push_back(k);
      }
      if (e1 >= 0 && L.kfa[b] >= 0) {
        faA.push_back(blk(P1, s));
        faB.push_back(blk(Lm, e1));
        faC.push_back(blk(Lm, k));
        faline.push_back(L.kfa[b]);
      } else {
        zfa.push_back(blk(Lm, k));
      }
      if (e2 >= 0 && L.kfb[b] >= 0) {
        fbA.push_back(blk(P2, s));
        fbB.push_back(blk(Um, e2));
        fbC.push_back(blk(Um, k));
        fbline.push_back(L.kfb[b]);
      } else {
        zfb.push_back(blk(Um, k));
      }
    }
    o.P1 = pb.add_ptrs(p1);
    o.P2 = pb.add_ptrs(p2);
    o.np1 = (int)p1A.size();
    o.p1A = pb.add_ptrs(p1A);
    o.p1B = pb.add_ptrs(p1B);
    o.p1C = pb.add_ptrs(p1C);
    o.d1A = pb.add_ptrs(d1A);
    o.d1B = pb.add_ptrs(d1B);
    o.hase1 = pb.add_ints(hase1);
    o.np2 = (int)p2A.size();
    o.p2A = pb.add_ptrs(p2A);
    o.p2B = pb.add_ptrs(p2B);
    o.p2C = pb.add_ptrs(p2C);
    o.d2A = pb.add_ptrs(d2A);
    o.d2B = pb.add_ptrs(d2B);
    o.hase2 = pb.add_ints(hase2);
    o.nfa = (int)faA.size();
    o.faA = pb.add_ptrs(faA);
    o.faB = pb.add_ptrs(faB);
    o.faC = pb.add_ptrs(faC);
    o.faline = pb.add_ints(faline);
    o.nfb = (int)fbA.size();
    o.fbA = pb.add_ptrs(fbA);
    o.fbB = pb.add_ptrs(fbB);
    o.fbC = pb.add_ptrs(fbC);
    o.fbline = pb.add_ints(fbline);
    o.nnofa = (int)zfa.size();
    o.zfa = pb.add_ptrs(zfa);
    o.nnofb = (int)zfb.size();
    o.zfb = pb.add_ptrs(zfb);
    // fused products Y_e L_e, Y_e U_e of the eliminated lines at levels >= 1 (W: scratch slots)
    std::vector<void*> ylA, ylB, ylC, yuA, yuB, yuC;
    if (l_index > 0) {
      for (int b = 0; b < o.ne; ++b) {
        const int e = L.e[b];
        if (L.ea[b] >= 0) {
          ylC.push_back(blk((char*)h->W, (int)ylA.size()));
          ylA.push_back(blk(D, e));
          ylB.push_back(blk(Lm, e));
        }
        if (L.eb[b] >= 0) {
          yuC.push_back(blk((char*)h->W, (int)yuA.size()));
          yuA.push_back(blk(D, e));
          yuB.push_back(blk(Um, e));
        }
      }
    }
    o.nyl = (int)ylA.size();
    o.nyu = (int)yuA.size();
    o.ylA = pb.add_ptrs(ylA);
    o.ylB = pb.add_ptrs(ylB);
    o.ylC = pb.add_ptrs(ylC);
    o.yuA = pb.add_ptrs(yuA);
    o.yuB = pb.add_ptrs(yuB);
    o.yuC = pb.add_ptrs(yuC);
    offs.push_back(o);
    ++l_index;
  }
  // one solve reads every Y once, the dense L/U of the lines eliminated at
  // levels >= 1, the multipliers of the kept lines at levels >= 1, the level-0
  // Y blocks a second time (forward), and moves ~6 vectors
  {
    size_t blocks = 0;
    for (size_t l = 0; l < h->plan.size(); ++l) {
      const LevelPlan& L = h->plan[l];
      blocks += L.e.size();
      if (l > 0) {
        for (size_t b = 0; b < L.e.size(); ++b) blocks += (L.ea[b] >= 0) + (L.eb[b] >= 0);
        for (size_t b = 0; b < L.k.size(); ++b) blocks += (L.k1[b] >= 0) + (L.k2[b] >= 0);
      } else if (!L.k.empty()) {
        blocks += L.e.size();
      }
    }
    h->solve_bytes = blocks * bm + 6 * sizeof(double) * (size_t)n;
    // FP16 copies: 2 bytes per entry + one FP32 scale per row; FP32 vectors
    h->solve_bytes_half = blocks * (2 * mm + 4 * (size_t)m) + 6 * sizeof(float) * (size_t)n;
  }
  pb.align();
  TF_HIP(hipMalloc(&h->pool, pb.bytes.size() + 64));
  TF_HIP(tf::upload(h->pool, pb.bytes.data(), pb.bytes.size()));
  h->pool_host = pb.bytes;
  char* P = (char*)h->pool;
  for (const Off& o : offs) {
    BcrLevelDev L{};
    L.ne = o.ne;
    L.nk = o.nk;
    L.e = (int*)(P + o.e);
    L.ea = (int*)(P + o.ea);
    L.eb = (int*)(P + o.eb);
    L.Ye = (void**)(P + o.Ye);
    L.Le = (void**)(P + o.Le);
    L.Ue = (void**)(P + o.Ue);
    L.k = (int*)(P + o.k);
    L.k1 = (int*)(P + o.k1);
    L.k2 = (int*)(P + o.k2);
    L.P1 = (void**)(P + o.P1);
    L.P2 = (void**)(P + o.P2);
    L.n_p1 = o.np1;
    L.g_p1_A = (void**)(P + o.p1A);
    L.g_p1_B = (void**)(P + o.p1B);
    L.g_p1_C = (void**)(P + o.p1C);
    L.g_d1_A = (void**)(P + o.d1A);
    L.g_d1_B = (void**)(P + o.d1B);
    L.k_hase1_line = (int*)(P + o.hase1);
    L.n_p2 = o.np2;
    L.g_p2_A = (void**)(P + o.p2A);
    L.g_p2_B = (void**)(P + o.p2B);
    L.g_p2_C = (void**)(P + o.p2C);
    L.g_d2_A = (void**)(P + o.d2A);
    L.g_d2_B = (void**)(P + o.d2B);
    L.k_hase2_line = (int*)(P + o.hase2);
    L.n_fa = o.nfa;
    L.g_fa_A = (void**)(P + o.faA);
    L.g_fa_B = (void**)(P + o.faB);
    L.g_fa_C = (void**)(P + o.faC);
    L.k_fa_line = (int*)(P + o.faline);
    L.n_fb = o.nfb;
    L.g_fb_A = (void**)(P + o.fbA);
    L.g_fb_B = (void**)(P + o.fbB);
    L.g_fb_C = (void**)(P + o.fbC);
    L.k_fb_line = (int*)(P + o.fbline);
    L.n_nofa = o.nnofa;
    L.z_fa = (void**)(P + o.zfa);
    L.n_nofb = o.nnofb;
    L.z_fb = (void**)(P + o.zfb);
    L.n_yl = o.nyl;
    L.g_yl_A = (void**)(P + o.ylA);
    L.g_yl_B = (void**)(P + o.ylB);
    L.g_yl_C = (void**)(P + o.ylC);
    L.n_yu = o.nyu;
    L.g_yu_A = (void**)(P + o.yuA);
    L.g_yu_B = (void**)(P + o.yuB);
    L.g_yu_C = (void**)(P + o.yuC);
    h->lev.push_back(L);
  }
  return 0;
}

int tf_bcr_destroy(tf_bcr* h) {
  if (!h) return 0;
  for (BcrGraph* g : {&h->g64[0], &h->g64[1], &h->g32[0], &h->g32[1]}) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
  }
  if (h->cap) (void)hipStreamDestroy(h->cap);
  for (void* p : {h->D, h->Lm, h->Um, h->P1, h->P2, h->Pinv, h->Kr, h->Kc, h->Tb, (void*)h->Lr, (void*)h->Ur,
                  (void*)h->Ut, (void*)h->Lt, (void*)h->nLr, (void*)h->nUr, (void*)h->nUt, (void*)h->nLt,
                  (void*)h->flags, h->rw, h->tb, h->yb, h->xw, h->pool, h->H, h->S, h->hpool, h->W,
                  (void*)h->topsync, h->Z, h->zpool, h->xtab})
    (void)hipFree(p);
  (void)hipHostFree(h->host_flags);
  delete h;
  return 0;
}

// Factor the gauged matrix (CSR, gauge row or -1).  Asynchronous on ``stream``.
int tf_bcr_factor(tf_bcr* h, const int* rp, const int* ci, const double* v, int gauge, void* stream) {
  if (!h) return tf::fail("null handle");
  hipStream_t st = (hipStream_t)stream;
  return h->single ? factor_impl<float>(h, rp, ci, v, gauge, st) : factor_impl<double>(h, rp, ci, v, gauge, st);
}

// flags[0]: structure errors (1 too many couplings, 2 coupling beyond the next line);
// flags[1]: statically perturbed pivots.  Synchronises the stream.
int tf_bcr_status(tf_bcr* h, int* flags_host, void* stream) {
  if (!h || !flags_host) return tf::fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  TF_HIP(hipMemcpyAsync(h->host_flags, h->flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
  TF_HIP(hipStreamSynchronize(st));
  flags_host[0] = h->host_flags[0];
  flags_host[1] = h->host_flags[1];
  return 0;
}

// on != 0 would run the upper levels of an apply with the fused backward as one
// k_top launch; it is refused: on the GPU the launch did not finish within the
// test's time limit (a 48 x 48 FP32 factor), with acquire polls and again with
// relaxed polls and one acquire fence, and the cause is not established.  0 (the
// default) keeps one launch per level.
int tf_bcr_set_top(tf_bcr* h, int on) {
  if (!h) return tf::fail("null handle");
  if (on) return tf::fail("tf_bcr_set_top: the one-launch upper levels are disabled (they did not finish on the GPU)");
  if ((on != 0) != (h->use_top != 0)) {
    for (BcrGraph* g : {&h->g64[0], &h->g64[1], &h->g32[0], &h->g32[1]}) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      *g = BcrGraph{};
    }
  }
  h->use_top = on ? 1 : 0;
  return 0;
}

// on != 0 (default): levels with at most SMALL_MAX lines run k_small, with the line
// indices and block row bases passed by value; 0: k_fwd / k_bwd_fused as for the
// other levels.  The results are the same bit for bit.
int tf_bcr_set_small(tf_bcr* h, int on) {
  if (!h) return tf::fail("null handle");
  if ((on != 0) != (h->use_small != 0)) {
    for (BcrGraph* g : {&h->g64[0], &h->g64[1], &h->g32[0], &h->g32[1]}) {
      if (g->exec) (void)hipGraphExecDestroy(g->exec);
      if (g->graph) (void)hipGraphDestroy(g->graph);
      *g = BcrGraph{};
    }
  }
  h->use_small = on ? 1 : 0;
  return 0;
}

// The number of k_top waits that gave up since the last call (then cleared);
// synchronises the stream.  Non-zero means the upper levels of some apply read
// vectors before the level below them had finished.
int tf_bcr_top_errors(tf_bcr* h, int* count_host, void* stream) {
  if (!h || !count_host) return tf::fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  // through the handle's pinned host words (a copy to pageable memory can stall other
  // threads' stream captures) and a memset on the stream
  TF_HIP(hipMemcpyAsync(h->host_flags + 2, h->topsync + 2 + MAX_TOP, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  TF_HIP(hipStreamSynchronize(st));
  const int v = h->host_flags[2];
  if (v) {
    TF_HIP(hipMemsetAsync(h->topsync + 2 + MAX_TOP, 0, sizeof(unsigned), st));
    TF_HIP(hipStreamSynchronize(st));
  }
  *count_host = v;
  return 0;
}

// on != 0: applies replay a captured graph (one launch instead of one per
// kernel); off while the kernel profiler is on.
int tf_bcr_set_graph(tf_bcr* h, int on) {
  if (!h) return tf::fail("null handle");
  h->use_graph = on ? 1 : 0;
  return 0;
}

// on != 0 (single-precision factors only): the FP32 apply reads FP16 copies of
// the factor blocks, one power-of-two FP32 scale per row, remade after every
// factorisation -- half the bytes of the FP32 blocks.  The first call allocates
// the copies; a factor that already exists is converted on ``stream``.
int tf_bcr_set_half(tf_bcr* h, int on, void* stream) {
  if (!h) return tf::fail("null handle");
  if (on && !h->single) return tf::fail("FP16 block copies need a single-precision factor");
  hipStream_t st = (hipStream_t)stream;
  const int m = h->m;
  const size_t mm = (size_t)m * m, N = h->N, slots = h->slots;
  if (on && !h->H) {
    // all three allocations succeed before any is kept (a failed call leaves no copies)
    void *H_ = nullptr, *S_ = nullptr;
    hipError_t e = hipMalloc(&H_, sizeof(f16) * mm * (3 * N + 2 * slots));
    if (e == hipSuccess) e = hipMalloc(&S_, sizeof(float) * (size_t)m * (3 * N + 2 * slots));
    if (e != hipSuccess) {
      (void)hipFree(H_);
      (void)hipFree(S_);
      return tf::fail_hip(e, "hipMalloc (FP16 block copies)");
    }
    f16* H = (f16*)H_;
    float* S = (float*)S_;
    // block j of buffer q (0 D, 1 L, 2 U, 3 P1, 4 P2) in the copies and in the FP32 factor
    auto hb = [&](int q, int j) -> void* {
      const size_t base = q < 3 ? q * N : 3 * N + (q - 3) * slots;
      return (void*)(H + (base + j) * mm);
    };
    auto sb = [&](int q, int j) -> void* {
      const size_t base = q < 3 ? q * N : 3 * N + (q - 3) * slots;
      return (void*)(S + (base + j) * (size_t)m);
    };
    char* src[5] = {(char*)h->D, (char*)h->Lm, (char*)h->Um, (char*)h->P1, (char*)h->P2};
    auto fb = [&](int q, int j) -> void* { return (void*)(src[q] + sizeof(float) * mm * (size_t)j); };
    PoolBuilder pb;
    struct HOff {
      size_t Yh, Lh, Uh, P1h, P2h, Ys, Ls, Us, P1s, P2s;
    };
    std::vector<HOff> offs;
    std::vector<void*> cs, cd, cc;
    auto conv = [&](int q, int j) {
      cs.push_back(fb(q, j));
      cd.push_back(hb(q, j));
      cc.push_back(sb(q, j));
    };
    for (size_t l = 0; l < h->plan.size(); ++l) {
      const LevelPlan& L = h->plan[l];
      std::vector<void*> yh, lh, uh, ys, ls, us, p1h, p2h, p1s, p2s;
      for (size_t b = 0; b < L.e.size(); ++b) {
        const int e = L.e[b];
        yh.push_back(hb(0, e));
        ys.push_back(sb(0, e));
        lh.push_back(hb(1, e));
        ls.push_back(sb(1, e));
        uh.push_back(hb(2, e));
        us.push_back(sb(2, e));
        conv(0, e);
        if (l > 0 && L.ea[b] >= 0) conv(1, e);
        if (l > 0 && L.eb[b] >= 0) conv(2, e);
      }
      for (size_t b = 0; b < L.k.size(); ++b) {
        const int sl = L.kslot[b];
        p1h.push_back(hb(3, sl));
        p1s.push_back(sb(3, sl));
        p2h.push_back(hb(4, sl));
        p2s.push_back(sb(4, sl));
        // (level 0 applies its multipliers through Y and the sparse couplings)
        if (l > 0 && L.k1[b] >= 0) conv(3, sl);
        if (l > 0 && L.k2[b] >= 0) conv(4, sl);
      }
      HOff o;
      o.Yh = pb.add_ptrs(yh);
      o.Lh = pb.add_ptrs(lh);
      o.Uh = pb.add_ptrs(uh);
      o.Ys = pb.add_ptrs(ys);
      o.Ls = pb.add_ptrs(ls);
      o.Us = pb.add_ptrs(us);
      o.P1h = pb.add_ptrs(p1h);
      o.P2h = pb.add_ptrs(p2h);
      o.P1s = pb.add_ptrs(p1s);
      o.P2s = pb.add_ptrs(p2s);
      offs.push_back(o);
    }
    const size_t ocs = pb.add_ptrs(cs), ocd = pb.add_ptrs(cd), occ = pb.add_ptrs(cc);
    pb.align();
    void* pool = nullptr;
    e = hipMalloc(&pool, pb.bytes.size() + 64);
    if (e == hipSuccess) e = tf::upload(pool, pb.bytes.data(), pb.bytes.size());
    if (e != hipSuccess) {
      (void)hipFree(pool);
      (void)hipFree(H_);
      (void)hipFree(S_);
      return tf::fail_hip(e, "FP16 copy pointer lists");
    }
    h->H = H_;
    h->S = S_;
    h->hpool = pool;
    h->hpool_host = pb.bytes;
    char* P = (char*)h->hpool;
    for (size_t l = 0; l < offs.size(); ++l) {
      BcrLevelDev& L = h->lev[l];
      const HOff& o = offs[l];
      L.Yh = (void**)(P + o.Yh);
      L.Lh = (void**)(P + o.Lh);
      L.Uh = (void**)(P + o.Uh);
      L.Ys = (void**)(P + o.Ys);
      L.Ls = (void**)(P + o.Ls);
      L.Us = (void**)(P + o.Us);
      L.P1h = (void**)(P + o.P1h);
      L.P2h = (void**)(P + o.P2h);
      L.P1s = (void**)(P + o.P1s);
      L.P2s = (void**)(P + o.P2s);
    }
    h->cv_src = (void**)(P + ocs);
    h->cv_dst = (void**)(P + ocd);
    h->cv_scl = (void**)(P + occ);
    h->ncv = (int)cs.size();
  }
  if ((on != 0) != (h->half != 0)) {
    // the captured FP32 applies read the other storage: capture them again
    for (BcrGraph& g : h->g32) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
      g = BcrGraph{};
    }
  }
  h->half = on ? 1 : 0;
  if (h->half && h->factored && h->ncv > 0) {
    hipLaunchKernelGGL(k_to_half, dim3((m + 3) / 4, h->ncv), dim3(256), 0, st, m, (float* const*)h->cv_src,
                       (f16* const*)h->cv_dst, (float* const*)h->cv_scl);
    TF_LAUNCH_CHECK("bcr FP16 copies");
  }
  return 0;
}

// split 0 (default): the level GEMMs of an FP32 factor use FP32 MFMA; 3: BF16 MFMA on
// three-term splits of the FP32 operands (about 2^-16 relative per product); 1: one term.
int tf_bcr_set_gemm_split(tf_bcr* h, int split) {
  if (!h) return tf::fail("null handle");
  if (split != 0 && split != 1 && split != 3) return tf::fail("split must be 0, 1 or 3");
  if (split && !h->single) return tf::fail("BF16 GEMM terms need a single-precision factor");
  h->split = split;
  return 0;
}

// on != 0: the factorisation stores Y L and Y U in place of the levels' L and U
// and the backward sweep is one kernel per level (two more GEMMs per eliminated
// line in the factorisation, one launch and one vector pass less per level in
// every apply).  Only before the first factorisation.
int tf_bcr_set_fused(tf_bcr* h, int on) {
  if (!h) return tf::fail("null handle");
  if (h->factored && (on != 0) != (h->fused != 0))
    return tf::fail("tf_bcr_set_fused: the handle already holds a factor of the other form");
  h->fused = on ? 1 : 0;
  return 0;
}

// lines in 2..TOPD_MAX: once the reduction is down to at most that many lines (at a
// level >= 1), each factorisation also forms the inverse Z of their reduced system
// (about 20 m x m x m products at four lines), and every apply maps those lines'
// right-hand sides to their solutions with one launch instead of running the
// forward and backward levels above.  Fused factors only (tf_bcr_set_fused; an
// unfused factor keeps the per-level sweeps).  0 (default): off.  Allocates Z;
// from the next factorisation on.
int tf_bcr_set_dense_top(tf_bcr* h, int lines) {
  if (!h) return tf::fail("null handle");
  if (lines != 0 && (lines < 2 || lines > TOPD_MAX)) return tf::fail("tf_bcr_set_dense_top: lines must be 0 or 2..4");
  drop_graphs(h);
  for (void* p : {h->Z, h->zpool}) (void)hipFree(p);
  h->Z = h->zpool = nullptr;
  h->zsteps.clear();
  h->zline.clear();
  h->zn = 0;
  h->zl = 0;
  h->zreplaced = 0;
  h->z_ok = 0;
  h->zmax = lines;
  if (!lines) return 0;
  // the lines active at level l are the kept lines of level l - 1
  int zl = -1;
  for (size_t l = 1; l < h->plan.size(); ++l)
    if ((int)h->plan[l - 1].k.size() <= lines) {
      zl = (int)l;
      break;
    }
  if (zl < 0 || h->plan[zl - 1].k.size() < 2) return 0;  // only the last line: nothing to replace
  const std::vector<int> A = h->plan[zl - 1].k;
  const int zn = (int)A.size();
  const int m = h->m;
  const size_t mm = (size_t)m * m, bm = h->es * mm;
  std::vector<int> pos(h->N, -1);
  for (int i = 0; i < zn; ++i) pos[A[i]] = i;
  void *Z = nullptr, *pool = nullptr;
  auto release = [&](hipError_t e, const char* what) {
    for (void* p : {Z, pool}) (void)hipFree(p);
    return tf::fail_hip(e, what);
  };
  hipError_t e = hipMalloc(&Z, bm * zn * zn);
  if (e != hipSuccess) return release(e, "hipMalloc (dense top)");
  auto zb = [&](int li, int lj) -> void* { return (char*)Z + bm * (size_t)(pos[li] * zn + pos[lj]); };
  auto fb = [&](void* base, int j) -> void* { return (char*)base + bm * (size_t)j; };
  // the steps, from the last level down to zl: after level l, Z[i][j] holds the
  // inverse of the reduced system of the lines active at level l
  struct HStep {
    int kind;
    std::vector<void*> a, b, c;
    double alpha, beta;
  };
  std::vector<HStep> hs;
  long replaced = 0;
  for (int l = (int)h->plan.size() - 1; l >= zl; --l) {
    const LevelPlan& L = h->plan[l];
    std::vector<int> act;  // the lines active at level l
    act.insert(act.end(), L.e.begin(), L.e.end());
    act.insert(act.end(), L.k.begin(), L.k.end());
    // kept rows, eliminated columns: Z[k][e] = -sum Z[k][k'] P1_k' (e = e1 of k') - sum Z[k][k'] P2_k' (e = e2 of k')
    if (!L.k.empty() && !L.e.empty()) {
      HStep z{0, {}, {}, {}, 0.0, 0.0}, g1{2, {}, {}, {}, -1.0, 1.0}, g2{2, {}, {}, {}, -1.0, 1.0};
      for (int k : L.k)
        for (int ee : L.e) z.a.push_back(zb(k, ee));
      for (size_t b = 0; b < L.k.size(); ++b)
        for (int k : L.k) {
          if (L.k1[b] >= 0) {
            g1.a.push_back(zb(k, L.k[b]));
            g1.b.push_back(fb(h->P1, L.kslot[b]));
            g1.c.push_back(zb(k, L.k1[b]));
          }
          if (L.k2[b] >= 0) {
            g2.a.push_back(zb(k, L.k[b]));
            g2.b.push_back(fb(h->P2, L.kslot[b]));
            g2.c.push_back(zb(k, L.k2[b]));
          }
        }
      for (HStep* s : {&z, &g1, &g2})
        if (!s->a.empty()) hs.push_back(*s);
    }
    // eliminated rows: Z[e][j] = Y_e [e == j] - (Y_e L_e) Z[a][j] - (Y_e U_e) Z[b][j]
    HStep cp{1, {}, {}, {}, 0.0, 0.0}, z{0, {}, {}, {}, 0.0, 0.0}, gl{2, {}, {}, {}, -1.0, 1.0},
        gu{2, {}, {}, {}, -1.0, 1.0};
    for (size_t b = 0; b < L.e.size(); ++b) {
      const int ee = L.e[b];
      cp.a.push_back(fb(h->D, ee));
      cp.b.push_back(zb(ee, ee));
      for (int j : act) {
        if (j != ee) z.a.push_back(zb(ee, j));
        if (L.ea[b] >= 0) {
          gl.a.push_back(fb(h->Lm, ee));
          gl.b.push_back(zb(L.ea[b], j));
          gl.c.push_back(zb(ee, j));
        }
        if (L.eb[b] >= 0) {
          gu.a.push_back(fb(h->Um, ee));
          gu.b.push_back(zb(L.eb[b], j));
          gu.c.push_back(zb(ee, j));
        }
      }
      replaced += 1 + (L.ea[b] >= 0) + (L.eb[b] >= 0);
    }
    for (size_t b = 0; b < L.k.size(); ++b) replaced += (L.k1[b] >= 0) + (L.k2[b] >= 0);
    for (HStep* s : {&cp, &z, &gl, &gu})
      if (!s->a.empty()) hs.push_back(*s);
  }
  PoolBuilder pb;
  std::vector<std::array<size_t, 3>> off;
  for (const HStep& s : hs) off.push_back({pb.add_ptrs(s.a), pb.add_ptrs(s.b), pb.add_ptrs(s.c)});
  pb.align();
  e = hipMalloc(&pool, pb.bytes.size() + 64);
  if (e == hipSuccess) e = tf::upload(pool, pb.bytes.data(), pb.bytes.size());
  if (e != hipSuccess) return release(e, "dense top pointer lists");
  char* P = (char*)pool;
  for (size_t i = 0; i < hs.size(); ++i)
    h->zsteps.push_back(BcrZStep{hs[i].kind, (int)hs[i].a.size(), (void**)(P + off[i][0]), (void**)(P + off[i][1]),
                                 (void**)(P + off[i][2]), hs[i].alpha, hs[i].beta});
  h->Z = Z;
  h->zpool = pool;
  h->zl = zl;
  h->zn = zn;
  h->zline = A;
  h->zreplaced = replaced;
  return 0;
}

int tf_bcr_solve_f64(tf_bcr* h, const double* r, double* x, void* stream) {
  if (!h) return tf::fail("null handle");
  const double* rr[1] = {r};
  double* xx[1] = {x};
  return solve_any<double, 1>(h, rr, xx, (hipStream_t)stream);
}

int tf_bcr_solve_f32(tf_bcr* h, const float* r, float* x, void* stream) {
  if (!h) return tf::fail("null handle");
  const float* rr[1] = {r};
  float* xx[1] = {x};
  return solve_any<float, 1>(h, rr, xx, (hipStream_t)stream);
}

// Two right-hand sides in one pass over the factor blocks (x0 = M r0, x1 = M r1).
int tf_bcr_solve2_f64(tf_bcr* h, const double* r0, const double* r1, double* x0, double* x1, void* stream) {
  if (!h) return tf::fail("null handle");
  const double* rr[2] = {r0, r1};
  double* xx[2] = {x0, x1};
  return solve_any<double, 2>(h, rr, xx, (hipStream_t)stream);
}

int tf_bcr_solve2_f32(tf_bcr* h, const float* r0, const float* r1, float* x0, float* x1, void* stream) {
  if (!h) return tf::fail("null handle");
  const float* rr[2] = {r0, r1};
  float* xx[2] = {x0, x1};
  return solve_any<float, 2>(h, rr, xx, (hipStream_t)stream);
}

}  // extern "C"
