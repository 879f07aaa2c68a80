// Block ILU over grid x-lines with a four-colour line order (two colours in 2D).
//
// The unknowns of one x-line (cells i = 0..nx-1 at fixed j, k; m = nx dof
// rows) are contiguous in the natural order, so the gauged Jacobian is a
// sparse matrix of m x m blocks over the (j, k) grid of lines.  A line couples
// to the lines at (j +- 1, k), (j, k +- 1), (j - 1, k + 1) and (j + 1, k - 1)
// (the convective cross terms).  Colouring the lines by the parities of j and
// k leaves no two neighbours in one colour, so with the lines taken colour by
// colour the block D-ILU
//
//   M = (P + L) P^-1 (P + U),   P_a = D_a - sum_{b earlier neighbour} A_ab P_b^-1 A_ba
//
// (L, U the strictly earlier / later couplings of A, kept sparse) has every
// colour's pivots independent of each other: a factorisation is four batches
// of Schur updates and dense inversions, an apply is four batched forward and
// four batched backward steps.  The pivot inverses are stored dense (N m^2
// entries: 1 GiB in FP32 for the 64^3 cavity); the couplings stay in the
// Jacobian's sparsity.
//
// This is the "block ILU(0)" of a line-block preconditioner and serves as the
// smoother of the two-grid method (mg.hip).  It replaces no single reference
// function: the reference's iterative path preconditions GMRES with scipy's
// spilu (transiflow/interface/SciPy.py:164-192,260-309).
#include "gemm.h"
#include "gj.h"
#include "tf_common.h"
#include "transiflow_hip.h"

#include <vector>

namespace {

using tf::gj::PB;
constexpr int KL = 16;  // off-line couplings kept per row and per column
constexpr int NDIR = 6;
__constant__ int c_dj[NDIR] = {-1, 1, 0, 0, -1, 1};
__constant__ int c_dk[NDIR] = {0, 0, -1, 1, 1, -1};
const int h_dj[NDIR] = {-1, 1, 0, 0, -1, 1};
const int h_dk[NDIR] = {0, 0, -1, 1, 1, -1};

struct Cp {
  int line;  // line of the other end
  int loc;   // row (column lists) or column (row lists) within that line
  double val;
};

// Colours count from the far corner (j = ny-1, k = nz-1), which always gets the
// last colour: its raw block, walled on the north and top, carries the line's
// pressure null space, and the Schur updates from the neighbours that come
// before it remove that null space unless it has only one of them.
__host__ __device__ __forceinline__ int colour_of(int line, int ny, int nz, int ncolour) {
  const int cj = (ny - 1 - line % ny) & 1, ck = (nz - 1 - line / ny) & 1;
  return ncolour == 4 ? 3 - (cj + 2 * ck) : 1 - cj;
}

__global__ __launch_bounds__(256) void k_zero_i(long n, int* __restrict__ p) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero_t(long n, T* __restrict__ p) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = (T)0;
}

// Dense diagonal blocks and the off-line couplings of the gauged matrix (row
// ``gauge`` -> -e_gauge, column ``gauge`` dropped elsewhere).  Row lists keep
// the CSR order; column lists are appended atomically and sorted afterwards.
template <typename T>
__global__ __launch_bounds__(256) void k_extract(int n, int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const double* __restrict__ v, int gauge, T* __restrict__ D,
                                                 Cp* __restrict__ rc, int* __restrict__ nrc, Cp* __restrict__ cc,
                                                 int* __restrict__ ncc, int* __restrict__ err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int line = r / m, loc = r - line * m;
  T* Drow = D + ((size_t)line * m + loc) * m;
  if (r == gauge) {
    Drow[loc] = (T)-1;
    nrc[r] = 0;
    return;
  }
  int cnt = 0;
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    const int c = ci[k];
    if (c == gauge) continue;
    const int lc = c / m;
    if (lc == line) {
      Drow[c - lc * m] = (T)v[k];
      continue;
    }
    if (cnt < KL) rc[(size_t)r * KL + cnt] = Cp{lc, c - lc * m, v[k]};
    else atomicOr(err, 1);
    ++cnt;
    const int s = atomicAdd(&ncc[c], 1);
    if (s < KL) cc[(size_t)c * KL + s] = Cp{line, loc, v[k]};
    else atomicOr(err, 1);
  }
  nrc[r] = cnt < KL ? cnt : KL;
}

// Column lists in (line, row) order, so the factor does not depend on the
// atomic append order.
__global__ __launch_bounds__(256) void k_sort_columns(int n, Cp* __restrict__ cc, int* __restrict__ ncc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  int k = ncc[c];
  if (k > KL) {
    k = KL;
    ncc[c] = KL;
  }
  Cp* a = cc + (size_t)c * KL;
  for (int i = 1; i < k; ++i) {
    const Cp x = a[i];
    int j = i - 1;
    while (j >= 0 && (a[j].line > x.line || (a[j].line == x.line && a[j].loc > x.loc))) {
      a[j + 1] = a[j];
      --j;
    }
    a[j + 1] = x;
  }
}

// Neighbour of line a in direction dir that comes earlier in the colour order
// than ``colour``, or -1.
__device__ __forceinline__ int earlier_neighbour(int a, int dir, int ny, int nz, int ncolour, int colour) {
  const int j = a % ny + c_dj[dir], k = a / ny + c_dk[dir];
  if (j < 0 || j >= ny || k < 0 || k >= nz) return -1;
  const int b = j + ny * k;
  return colour_of(b, ny, nz, ncolour) < colour ? b : -1;
}

// Schur update of the pivots of one colour's lines from their neighbours in one
// direction, in two passes through a dense m x m buffer per line:
//   (a)  W = P_b^-1 A_ba     W[s][c] = sum over the couplings of column c of line a into line b
//   (b)  P_a -= A_ab W       P_a[i][c] -= sum over the couplings of row i of line a into line b
template <typename T>
__global__ __launch_bounds__(256) void k_schur_a(int m, int ny, int nz, int ncolour, int colour, int dir,
                                                 const int* __restrict__ lines, const T* __restrict__ P,
                                                 const Cp* __restrict__ cc, const int* __restrict__ ncc,
                                                 T* __restrict__ W) {
  const int a = lines[blockIdx.y];
  const int b = earlier_neighbour(a, dir, ny, nz, ncolour, colour);
  if (b < 0) return;
  const T* Pb = P + (size_t)b * m * m;
  T* Wl = W + (size_t)blockIdx.y * m * m;
  const long mm = (long)m * m;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < mm; e += (long)gridDim.x * blockDim.x) {
    const int s = (int)(e / m), c = (int)(e - (long)s * m);
    const size_t col = (size_t)a * m + c;
    const Cp* lst = cc + col * KL;
    const int cnt = ncc[col];
    T acc = (T)0;
    for (int q = 0; q < cnt; ++q) {
      const Cp p = lst[q];
      if (p.line == b) acc += Pb[(size_t)s * m + p.loc] * (T)p.val;
    }
    Wl[e] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_schur_b(int m, int ny, int nz, int ncolour, int colour, int dir,
                                                 const int* __restrict__ lines, T* __restrict__ P,
                                                 const Cp* __restrict__ rc, const int* __restrict__ nrc,
                                                 const T* __restrict__ W) {
  const int a = lines[blockIdx.y];
  const int b = earlier_neighbour(a, dir, ny, nz, ncolour, colour);
  if (b < 0) return;
  T* Pa = P + (size_t)a * m * m;
  const T* Wl = W + (size_t)blockIdx.y * m * m;
  const long mm = (long)m * m;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < mm; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / m), c = (int)(e - (long)i * m);
    const size_t row = (size_t)a * m + i;
    const Cp* lst = rc + row * KL;
    const int cnt = nrc[row];
    T acc = (T)0;
    for (int q = 0; q < cnt; ++q) {
      const Cp p = lst[q];
      if (p.line == b) acc += (T)p.val * Wl[(size_t)p.loc * m + c];
    }
    Pa[e] -= acc;
  }
}

// Apply, sparse part of one colour step:
//   forward   t = r - sum over couplings into earlier colours of a x
//   backward  t =     sum over couplings into later colours of a x
template <typename V>
__global__ __launch_bounds__(256) void k_sparse_step(int m, int ny, int nz, int ncolour, int colour, int forward,
                                                     int nlines,
                                                     const int* __restrict__ lines, const Cp* __restrict__ rc,
                                                     const int* __restrict__ nrc, const V* __restrict__ r,
                                                     const V* __restrict__ x, V* __restrict__ t) {
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= (long)nlines * m) return;
  const int a = lines[idx / m];
  const size_t row = (size_t)a * m + (idx % m);
  const Cp* lst = rc + row * KL;
  const int cnt = nrc[row];
  double acc = 0.0;
  for (int q = 0; q < cnt; ++q) {
    const Cp p = lst[q];
    const int cb = colour_of(p.line, ny, nz, ncolour);
    if (forward ? cb < colour : cb > colour) acc += p.val * (double)x[(size_t)p.line * m + p.loc];
  }
  t[row] = forward ? (V)((double)r[row] - acc) : (V)acc;
}

// Apply, dense part of one colour step: forward x_a = P_a^-1 t_a, backward
// x_a -= P_a^-1 t_a.  16 lanes per row (16 rows per workgroup), 16-byte loads
// when the rows allow them.
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef float4 type;
  static constexpr int w = 4;
};
template <>
struct Vec4<double> {
  typedef double2 type;
  static constexpr int w = 2;
};

template <typename T>
__global__ __launch_bounds__(256) void k_dense_step(int m, int forward, int vec, int nlines,
                                                    const int* __restrict__ lines, const T* __restrict__ P,
                                                    const T* __restrict__ t, T* __restrict__ x) {
  const long row_g = blockIdx.x * 16L + (threadIdx.x >> 4);
  const int lane = threadIdx.x & 15;
  const bool live = row_g < (long)nlines * m;
  T acc = (T)0;
  int a = 0, loc = 0;
  if (live) {
    a = lines[row_g / m];
    loc = (int)(row_g % m);
    const T* prow = P + ((size_t)a * m + loc) * m;
    const T* tv = t + (size_t)a * m;
    if (vec) {
      typedef typename Vec4<T>::type V4;
      constexpr int w = Vec4<T>::w;
      const V4* p4 = reinterpret_cast<const V4*>(prow);
      const V4* t4 = reinterpret_cast<const V4*>(tv);
      for (int q = lane; q < m / w; q += 16) {
        const V4 pv = p4[q], tq = t4[q];
        if constexpr (w == 4) acc += pv.x * tq.x + pv.y * tq.y + pv.z * tq.z + pv.w * tq.w;
        else acc += pv.x * tq.x + pv.y * tq.y;
      }
    } else {
      for (int c = lane; c < m; c += 16) acc += prow[c] * tv[c];
    }
  }
  for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
  if (live && lane == 0) {
    const size_t o = (size_t)a * m + loc;
    x[o] = forward ? acc : x[o] - acc;
  }
}

}  // namespace

struct tf_lineilu {
  int n, m, nx, ny, nz, dof, N, single, ncolour;
  size_t es;
  int count[4];
  int* lines_d;        // lines of every colour, colour by colour
  int offset[4];       // start of colour c in lines_d
  void* P;             // N m x m pivots, inverted in place
  void* W;             // max colour count x m x m (Schur pass buffer)
  Cp *rc, *cc;         // n x KL row / column coupling lists
  int *nrc, *ncc;
  void *Pinv, *Kr, *Kc, *Tb;  // Gauss-Jordan panel scratch (max colour count)
  void** lists;        // per colour: Xp | pinv | kr | kc | tb pointer lists (device)
  int lists_off[4];
  int* flags;          // [0] structure errors, [1] perturbed pivots
  int* flags_host;
  void* t;             // n (apply scratch, storage type)
};

namespace {

template <typename T>
void gj_gemm(int M, int N, int K, void** A, int lda, void** B, int ldb, void** C, int ldc, double alpha, double beta,
             int batch, hipStream_t st) {
  if (batch <= 0) return;
  TF_PROF_WORK(tf::PROF_GEMM, 2.0 * M * N * K * batch);
  TF_PROF_BEGIN(tf::PROF_GEMM, st);
  tf::gemm_batched<T>(M, N, K, (T* const*)A, lda, (T* const*)B, ldb, (T* const*)C, ldc, alpha, beta, batch, st);
  TF_PROF_END(tf::PROF_GEMM, st);
}

// In-place inversion of the pivots of colour c (blocked Gauss-Jordan).
template <typename T>
void invert_colour(tf_lineilu* h, int c, hipStream_t st) {
  const int m = h->m, batch = h->count[c];
  if (batch == 0) return;
  void** Xp = h->lists + h->lists_off[c];
  void** pinvp = Xp + batch;
  void** krp = pinvp + batch;
  void** kcp = krp + batch;
  void** tbp = kcp + batch;
  const long sz = (long)PB * m;
  const int gcopy = (int)((sz + 255) / 256) < 1024 ? (int)((sz + 255) / 256) : 1024;
  for (int k0 = 0; k0 < m; k0 += PB) {
    const int pb = m - k0 < PB ? m - k0 : PB;
    // a pivot that is still singular (a local pressure null space the Schur
    // updates did not remove) is regularised: its zero pivot becomes 1e-3 of the
    // panel's largest entry, the scale of the smallest regular pivots
    hipLaunchKernelGGL(tf::gj::k_panel<T>, dim3(batch), dim3(512), 0, st, m, k0, pb, (T* const*)Xp, (T*)h->Pinv,
                       h->flags + 1, 1e-3);
    hipLaunchKernelGGL(tf::gj::k_panel_copy<T>, dim3(gcopy, batch), dim3(256), 0, st, m, k0, pb, (T* const*)Xp,
                       (T*)h->Kr, (T*)h->Kc);
    gj_gemm<T>(pb, m, pb, pinvp, pb, krp, m, tbp, m, 1.0, 0.0, batch, st);
    gj_gemm<T>(m, m, pb, kcp, pb, tbp, m, Xp, m, -1.0, 1.0, batch, st);
    hipLaunchKernelGGL(tf::gj::k_panel_store<T>, dim3(gcopy, batch), dim3(256), 0, st, m, k0, pb, (T* const*)Xp,
                       (const T*)h->Tb);
  }
}

int grid_of(long work) {
  long g = (work + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

template <typename T>
int factor_impl(tf_lineilu* h, const int* rp, const int* ci, const double* v, int gauge, hipStream_t st) {
  const int n = h->n, m = h->m;
  const long mm = (long)m * m;
  hipLaunchKernelGGL(k_zero_t<T>, dim3(grid_of((long)h->N * mm)), dim3(256), 0, st, (long)h->N * mm, (T*)h->P);
  hipLaunchKernelGGL(k_zero_i, dim3(grid_of(n)), dim3(256), 0, st, (long)n, h->ncc);
  hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(256), 0, st, 2L, h->flags);
  hipLaunchKernelGGL(k_extract<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, m, rp, ci, v, gauge, (T*)h->P, h->rc,
                     h->nrc, h->cc, h->ncc, h->flags);
  hipLaunchKernelGGL(k_sort_columns, dim3((n + 255) / 256), dim3(256), 0, st, n, h->cc, h->ncc);
  TF_LAUNCH_CHECK("line ILU extraction");
  const int gx = (int)((mm + 255) / 256) < 256 ? (int)((mm + 255) / 256) : 256;
  for (int c = 0; c < h->ncolour; ++c) {
    const int batch = h->count[c];
    if (batch == 0) continue;
    const int* lines = h->lines_d + h->offset[c];
    if (c > 0) {
      for (int dir = 0; dir < NDIR; ++dir) {
        if (h->nz == 1 && h_dk[dir] != 0) continue;
        hipLaunchKernelGGL(k_schur_a<T>, dim3(gx, batch), dim3(256), 0, st, m, h->ny, h->nz, h->ncolour, c, dir,
                           lines, (const T*)h->P, h->cc, h->ncc, (T*)h->W);
        hipLaunchKernelGGL(k_schur_b<T>, dim3(gx, batch), dim3(256), 0, st, m, h->ny, h->nz, h->ncolour, c, dir,
                           lines, (T*)h->P, h->rc, h->nrc, (const T*)h->W);
      }
    }
    invert_colour<T>(h, c, st);
    TF_LAUNCH_CHECK("line ILU colour");
  }
  return 0;
}

template <typename T>
int solve_impl(tf_lineilu* h, const T* r, T* x, hipStream_t st) {
  const int m = h->m;
  T* t = (T*)h->t;
  const int vec = (m % Vec4<T>::w == 0) ? 1 : 0;
  for (int pass = 0; pass < 2; ++pass) {
    const int forward = pass == 0;
    for (int s = 0; s < h->ncolour; ++s) {
      const int c = forward ? s : h->ncolour - 1 - s;
      const int nl = h->count[c];
      if (nl == 0) continue;
      if (!forward && c == h->ncolour - 1) continue;  // no later colour: x_a = y_a
      const int* lines = h->lines_d + h->offset[c];
      const long rows = (long)nl * m;
      hipLaunchKernelGGL(k_sparse_step<T>, dim3((int)((rows + 255) / 256)), dim3(256), 0, st, m, h->ny, h->nz, h->ncolour, c,
                         forward, nl, lines, h->rc, h->nrc, r, (const T*)x, t);
      hipLaunchKernelGGL(k_dense_step<T>, dim3((int)((rows + 15) / 16)), dim3(256), 0, st, m, forward, vec, nl, lines,
                         (const T*)h->P, (const T*)t, x);
    }
  }
  TF_LAUNCH_CHECK("line ILU apply");
  return 0;
}

int create_impl(tf_lineilu* h) {
  const int m = h->m, N = h->N;
  const size_t mm = (size_t)m * m;
  int maxc = 0;
  for (int c = 0; c < h->ncolour; ++c) maxc = h->count[c] > maxc ? h->count[c] : maxc;
  TF_HIP(hipMalloc(&h->P, h->es * mm * N));
  TF_HIP(hipMalloc(&h->W, h->es * mm * maxc));
  TF_HIP(hipMalloc(&h->rc, sizeof(Cp) * KL * (size_t)h->n));
  TF_HIP(hipMalloc(&h->cc, sizeof(Cp) * KL * (size_t)h->n));
  TF_HIP(hipMalloc(&h->nrc, sizeof(int) * (size_t)h->n));
  TF_HIP(hipMalloc(&h->ncc, sizeof(int) * (size_t)h->n));
  TF_HIP(hipMalloc(&h->Pinv, h->es * PB * PB * maxc));
  TF_HIP(hipMalloc(&h->Kr, h->es * PB * (size_t)m * maxc));
  TF_HIP(hipMalloc(&h->Kc, h->es * PB * (size_t)m * maxc));
  TF_HIP(hipMalloc(&h->Tb, h->es * PB * (size_t)m * maxc));
  TF_HIP(hipMalloc(&h->flags, sizeof(int) * 2));
  TF_HIP(hipHostMalloc(&h->flags_host, sizeof(int) * 2, hipHostMallocDefault));
  TF_HIP(hipMalloc(&h->t, h->es * (size_t)h->n));
  // lines grouped by colour, and the Gauss-Jordan pointer lists of each colour
  std::vector<int> lines;
  std::vector<void*> ptrs;
  for (int c = 0; c < h->ncolour; ++c) {
    h->offset[c] = (int)lines.size();
    h->lists_off[c] = (int)ptrs.size();
    std::vector<int> mine;
    for (int l = 0; l < N; ++l)
      if (colour_of(l, h->ny, h->nz, h->ncolour) == c) mine.push_back(l);
    for (int l : mine) ptrs.push_back((char*)h->P + h->es * mm * l);
    for (size_t b = 0; b < mine.size(); ++b) ptrs.push_back((char*)h->Pinv + h->es * PB * PB * b);
    for (size_t b = 0; b < mine.size(); ++b) ptrs.push_back((char*)h->Kr + h->es * PB * (size_t)m * b);
    for (size_t b = 0; b < mine.size(); ++b) ptrs.push_back((char*)h->Kc + h->es * PB * (size_t)m * b);
    for (size_t b = 0; b < mine.size(); ++b) ptrs.push_back((char*)h->Tb + h->es * PB * (size_t)m * b);
    lines.insert(lines.end(), mine.begin(), mine.end());
  }
  TF_HIP(hipMalloc(&h->lines_d, sizeof(int) * lines.size()));
  TF_HIP(tf::upload(h->lines_d, lines.data(), sizeof(int) * lines.size()));
  TF_HIP(hipMalloc(&h->lists, sizeof(void*) * ptrs.size()));
  TF_HIP(tf::upload(h->lists, ptrs.data(), sizeof(void*) * ptrs.size()));
  return 0;
}

}  // namespace

extern "C" {

int tf_lineilu_create(tf_lineilu** out, int nx, int ny, int nz, int dof, int single) {
  if (!out) return tf::fail("null handle");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1 || dof < 1) return tf::fail("tf_lineilu_create: bad grid");
  if ((long long)nx * ny * nz * dof >= (1LL << 31)) return tf::fail("tf_lineilu_create: too many unknowns");
  tf_lineilu* h = new tf_lineilu();
  h->nx = nx;
  h->ny = ny;
  h->nz = nz;
  h->dof = dof;
  h->m = nx * dof;
  h->N = ny * nz;
  h->n = h->m * h->N;
  h->single = single;
  h->es = single ? sizeof(float) : sizeof(double);
  h->ncolour = nz > 1 ? 4 : 2;
  for (int c = 0; c < 4; ++c) h->count[c] = 0;
  for (int l = 0; l < h->N; ++l) h->count[colour_of(l, ny, nz, h->ncolour)]++;
  if (int e = create_impl(h)) {
    tf_lineilu_destroy(h);
    return e;
  }
  *out = h;
  return 0;
}

int tf_lineilu_destroy(tf_lineilu* h) {
  if (!h) return 0;
  for (void* p : {h->P, h->W, (void*)h->rc, (void*)h->cc, (void*)h->nrc, (void*)h->ncc, h->Pinv, h->Kr, h->Kc, h->Tb,
                  (void*)h->flags, h->t, (void*)h->lines_d, (void*)h->lists})
    (void)hipFree(p);
  (void)hipHostFree(h->flags_host);
  delete h;
  return 0;
}

int tf_lineilu_factor(tf_lineilu* h, const int* row_ptr, const int* col, const double* val, int gauge,
                      void* stream) {
  if (!h) return tf::fail("null handle");
  hipStream_t st = (hipStream_t)stream;
  return h->single ? factor_impl<float>(h, row_ptr, col, val, gauge, st)
                   : factor_impl<double>(h, row_ptr, col, val, gauge, st);
}

int tf_lineilu_status(tf_lineilu* h, int* flags_host, void* stream) {
  if (!h || !flags_host) return tf::fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  TF_HIP(hipMemcpyAsync(h->flags_host, h->flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
  TF_HIP(hipStreamSynchronize(st));
  flags_host[0] = h->flags_host[0];
  flags_host[1] = h->flags_host[1];
  return 0;
}

int tf_lineilu_solve_f64(tf_lineilu* h, const double* r, double* x, void* stream) {
  if (!h) return tf::fail("null handle");
  if (h->single) return tf::fail("tf_lineilu_solve_f64 on a single-precision factor");
  if (r == x) return tf::fail("tf_lineilu_solve_f64: r and x must differ");
  return solve_impl<double>(h, r, x, (hipStream_t)stream);
}

int tf_lineilu_solve_f32(tf_lineilu* h, const float* r, float* x, void* stream) {
  if (!h) return tf::fail("null handle");
  if (!h->single) return tf::fail("tf_lineilu_solve_f32 on a double-precision factor");
  if (r == x) return tf::fail("tf_lineilu_solve_f32: r and x must differ");
  return solve_impl<float>(h, r, x, (hipStream_t)stream);
}

}  // extern "C"
