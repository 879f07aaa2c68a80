// Restarted flexible GMRES on the host side of the library: the whole solve is
// one C call, so the Python caller (and the interpreter lock) is out of the
// Arnoldi loop, and concurrent solves on different streams proceed
// independently.  Right preconditioning with the line factorisation
// (tf_bcr / tf_linelu) or cell-block Jacobi, CGS2 orthogonalisation with the
// device kernels of krylov.hip, Givens rotations on the host (one (j+2)-entry
// read-back per iteration).  Same iteration and stopping rules as the
// reference's scipy.sparse.linalg.gmres use (SciPy.py:343-370) restated for
// FGMRES: stop when the residual estimate |g_{j+1}| <= tol ||b||, restart with
// the true residual b - A x.
//
// The device part of Arnoldi step j (SpMV of Z[j], two Gram-Schmidt passes
// against V[0..j], the normalisation into V[j+1] and the copy of the (j+2)-entry
// column to pinned host memory) is captured once per j into a hipGraph and
// replayed with one launch; the SpMV reads the matrix through a small device
// descriptor that each solve points at its own CSR arrays, so the graphs stay
// valid across Jacobians.  The host then waits for the column by polling (no
// blocking wait, whose wake-up would leave the GPU idle longer); while the
// residual estimate is far from the tolerance, the preconditioner on the new basis
// vector is queued before that wait, so the device is busy through the host's part.
#include "spmv.h"
#include "tf_common.h"
#include "transiflow_hip.h"

#include <atomic>
#include <cmath>
#include <vector>

namespace {
constexpr int MAX_STEPS = 256;   // Arnoldi steps per restart cycle (tf_krylov_max_basis)

struct MatrixDesc {
  const int* rp;
  const int* ci;
  const void* val;
  int gauge;
};
}  // namespace

struct tf_krylov {
  int n, m, single;
  size_t es;
  void* V;     // (m+1) x n
  void* Z;     // m x n
  void* w;     // n
  double* h;   // m+2 (device)
  double* y;   // m+1 (device)
  double* work;
  double* h_host;      // m+2 (pinned, coherent)
  double* h_host_dev;  // h_host as kernels address it
  MatrixDesc* desc;          // the matrix of the current solve (device)
  hipStream_t cap;           // private stream the step graphs are captured on
  hipGraphExec_t step[MAX_STEPS];
  hipGraph_t step_graph[MAX_STEPS];
  // two-system steps with the partner handle step2_partner (captured on this handle)
  hipGraphExec_t step2[MAX_STEPS];
  hipGraph_t step2_graph[MAX_STEPS];
  unsigned long long uid;            // unique per handle (a freed handle's address can come back)
  unsigned long long step2_partner;  // uid of the partner the step2 graphs were captured with
  hipEvent_t step_done;              // recorded after a step's Hessenberg column is queued to the host
  double last_sn0;                   // |sine| of the first rotation of the last cycle (0: none yet)
  int use_graph;
};

namespace tf {
// hh0, hh1: the device address of pinned host copies of h (or null); *to_host = 1
// when the launches wrote the column there
int cgs2_pair_f64(int n, int j, const double* V0, const double* V1, double* w0, double* w1, double* work0,
                  double* work1, double* h0, double* h1, double* y0, double* y1, double* v0, double* v1,
                  double* hh0, double* hh1, int* to_host, hipStream_t st);
int cgs2_pair_f32(int n, int j, const float* V0, const float* V1, float* w0, float* w1, double* work0,
                  double* work1, double* h0, double* h1, double* y0, double* y1, float* v0, float* v1,
                  double* hh0, double* hh1, int* to_host, hipStream_t st);
int cgs2_one_f64(int n, int j, const double* V, double* w, double* work, double* h, double* y, double* v,
                 double* hh, int* to_host, hipStream_t st);
int cgs2_one_f32(int n, int j, const float* V, float* w, double* work, double* h, double* y, float* v,
                 double* hh, int* to_host, hipStream_t st);
}  // namespace tf

namespace {

__global__ void k_set_desc(MatrixDesc* d, const int* rp, const int* ci, const void* val, int gauge) {
  d->rp = rp;
  d->ci = ci;
  d->val = val;
  d->gauge = gauge;
}

template <typename T>
__global__ __launch_bounds__(256) void k_spmv_desc(int n, const MatrixDesc* __restrict__ d, const T* __restrict__ x,
                                                   T* __restrict__ y) {
  tf::spmv(n, d->rp, d->ci, (const T*)d->val, x, y, d->gauge);
}

// the SpMV of two systems (blockIdx.y picks descriptor and vectors)
template <typename T>
__global__ __launch_bounds__(256) void k_spmv_desc2(int n, const MatrixDesc* __restrict__ d0,
                                                    const T* __restrict__ x0, T* __restrict__ y0,
                                                    const MatrixDesc* __restrict__ d1, const T* __restrict__ x1,
                                                    T* __restrict__ y1) {
  const MatrixDesc* d = blockIdx.y ? d1 : d0;
  tf::spmv(n, d->rp, d->ci, (const T*)d->val, blockIdx.y ? x1 : x0, blockIdx.y ? y1 : y0, d->gauge);
}

struct Ops {
  const tf_krylov* k;
  int n;
  hipStream_t st;
  char* V(int j) const { return (char*)k->V + k->es * (size_t)n * j; }
  char* Z(int j) const { return (char*)k->Z + k->es * (size_t)n * j; }
};

int spmv(const tf_krylov* k, int n, const int* rp, const int* ci, const void* val, int gauge, const void* x, void* y,
         hipStream_t st) {
  return k->single ? tf_spmv_f32(n, rp, ci, (const float*)val, (const float*)x, (float*)y, gauge, st)
                   : tf_spmv_f64(n, rp, ci, (const double*)val, (const double*)x, (double*)y, gauge, st);
}

int precond(const tf_krylov* k, int kind, void* handle, int ncell, int dof, const void* r, void* z, hipStream_t st) {
  switch (kind) {
    case 1:
      return k->single ? tf_bcr_solve_f32((tf_bcr*)handle, (const float*)r, (float*)z, st)
                       : tf_bcr_solve_f64((tf_bcr*)handle, (const double*)r, (double*)z, st);
    case 2:
      return k->single ? tf_linelu_solve_f32((tf_linelu*)handle, (const float*)r, (float*)z, st)
                       : tf_linelu_solve_f64((tf_linelu*)handle, (const double*)r, (double*)z, st);
    case 3:
      return k->single ? tf_cellblock_apply_f32(ncell, dof, (const double*)handle, (const float*)r, (float*)z, st)
                       : tf_cellblock_apply_f64(ncell, dof, (const double*)handle, (const double*)r, (double*)z, st);
    case 4:
      return k->single ? tf_mg_solve_f32((tf_mg*)handle, (const float*)r, (float*)z, st)
                       : tf_mg_solve_f64((tf_mg*)handle, (const double*)r, (double*)z, st);
    case 5:
      return k->single ? tf_lineilu_solve_f32((tf_lineilu*)handle, (const float*)r, (float*)z, st)
                       : tf_lineilu_solve_f64((tf_lineilu*)handle, (const double*)r, (double*)z, st);
    default:
      return hipMemcpyAsync(z, r, k->es * (size_t)k->n, hipMemcpyDeviceToDevice, st) == hipSuccess
                 ? 0
                 : tf::fail("copy preconditioner");
  }
}

// |x|^2 queued into the handle's pinned word h_host[m + 1]: the reduction's last
// kernel writes it there (no copy); read it after the stream is synchronised
int norm_queue(const tf_krylov* k, int n, const void* x, hipStream_t st) {
  double* out = k->h_host_dev + k->m + 1;
  return k->single ? tf_multidot_f32(n, 1, (const float*)x, n, (const float*)x, k->work, out, 0, st)
                   : tf_multidot_f64(n, 1, (const double*)x, n, (const double*)x, k->work, out, 0, st);
}

double norm_read(const tf_krylov* k) { return std::sqrt(k->h_host[k->m + 1]); }

// ||x|| (host double; synchronises the stream)
int norm(const tf_krylov* k, int n, const void* x, double* out, hipStream_t st) {
  if (int e = norm_queue(k, n, x, st)) return e;
  TF_HIP(hipStreamSynchronize(st));
  *out = norm_read(k);
  return 0;
}

// The device part of Arnoldi step j on stream st: w = A Z[j] (through the
// descriptor), CGS2 h = V^T w, w -= V h, y = V^T w, w -= V y, h += y, V[j+1] =
// w / |w| with |w| into h[j+1], and h[0..j+1] into h_host (written there by the
// step's last kernel, or copied).
int arnoldi_step(tf_krylov* k, int j, hipStream_t st) {
  const int n = k->n;
  Ops o{k, n, st};
  if (k->single)
    hipLaunchKernelGGL(k_spmv_desc<float>, dim3(tf::spmv_blocks<float>(n)), dim3(256), 0, st, n, k->desc,
                       (const float*)o.Z(j), (float*)k->w);
  else
    hipLaunchKernelGGL(k_spmv_desc<double>, dim3(tf::spmv_blocks<double>(n)), dim3(256), 0, st, n, k->desc,
                       (const double*)o.Z(j), (double*)k->w);
  TF_LAUNCH_CHECK("fgmres spmv");
  int to_host = 0;
  if (int e = k->single ? tf::cgs2_one_f32(n, j, (const float*)k->V, (float*)k->w, k->work, k->h, k->y,
                                           (float*)o.V(j + 1), k->h_host_dev, &to_host, st)
                        : tf::cgs2_one_f64(n, j, (const double*)k->V, (double*)k->w, k->work, k->h, k->y,
                                           (double*)o.V(j + 1), k->h_host_dev, &to_host, st))
    return e;
  if (!to_host) TF_HIP(hipMemcpyAsync(k->h_host, k->h, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, st));
  return 0;
}

// Step j replayed from its captured graph (captured on first use).
int arnoldi_step_graph(tf_krylov* k, int j, hipStream_t st) {
  if (!k->step[j]) {
    if (!k->cap) TF_HIP(hipStreamCreateWithFlags(&k->cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    int e = 0;
    hipError_t ce = hipSuccess, ie = hipSuccess;
    {
      std::lock_guard<std::mutex> lock(tf::capture_mutex());
      TF_HIP(hipStreamBeginCapture(k->cap, hipStreamCaptureModeThreadLocal));
      e = arnoldi_step(k, j, k->cap);
      ce = hipStreamEndCapture(k->cap, &g);
      if (!e && ce == hipSuccess && g) ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    }
    if (e || ce != hipSuccess || !g || ie != hipSuccess) {
      // a failed capture or instantiation: this step runs eagerly, with the same
      // kernels, and is captured again next time
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      tf::note_capture_fallback();
      return arnoldi_step(k, j, st);
    }
    k->step_graph[j] = g;
    k->step[j] = ex;
  }
  TF_HIP(hipGraphLaunch(k->step[j], st));
  return 0;
}

// Arnoldi step j of two systems with handles a and b (same n, restart and precision)
// in the same launches: both SpMVs, both CGS2 orthogonalisations and
// normalisations (tf::cgs2_pair_*), each Hessenberg column to its handle's
// h_host.  The arithmetic is that of arnoldi_step on each.
int arnoldi_step2(tf_krylov* a, tf_krylov* b, int j, hipStream_t st) {
  const int n = a->n;
  Ops oa{a, n, st}, ob{b, n, st};
  if (a->single)
    hipLaunchKernelGGL(k_spmv_desc2<float>, dim3(tf::spmv_blocks<float>(n), 2), dim3(256), 0, st, n, a->desc,
                       (const float*)oa.Z(j), (float*)a->w, b->desc, (const float*)ob.Z(j), (float*)b->w);
  else
    hipLaunchKernelGGL(k_spmv_desc2<double>, dim3(tf::spmv_blocks<double>(n), 2), dim3(256), 0, st, n, a->desc,
                       (const double*)oa.Z(j), (double*)a->w, b->desc, (const double*)ob.Z(j), (double*)b->w);
  TF_LAUNCH_CHECK("fgmres spmv pair");
  int to_host = 0;
  if (int e = a->single ? tf::cgs2_pair_f32(n, j, (const float*)a->V, (const float*)b->V, (float*)a->w, (float*)b->w,
                                            a->work, b->work, a->h, b->h, a->y, b->y, (float*)oa.V(j + 1),
                                            (float*)ob.V(j + 1), a->h_host_dev, b->h_host_dev, &to_host, st)
                        : tf::cgs2_pair_f64(n, j, (const double*)a->V, (const double*)b->V, (double*)a->w,
                                            (double*)b->w, a->work, b->work, a->h, b->h, a->y, b->y,
                                            (double*)oa.V(j + 1), (double*)ob.V(j + 1), a->h_host_dev, b->h_host_dev,
                                            &to_host, st))
    return e;
  if (!to_host) {
    TF_HIP(hipMemcpyAsync(a->h_host, a->h, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, st));
    TF_HIP(hipMemcpyAsync(b->h_host, b->h, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, st));
  }
  return 0;
}

// The two-system step j replayed from a graph captured on handle a (recaptured
// when a's partner changes).
int arnoldi_step2_graph(tf_krylov* a, tf_krylov* b, int j, hipStream_t st) {
  if (a->step2_partner != b->uid) {
    for (int q = 0; q < MAX_STEPS; ++q) {
      if (a->step2[q]) (void)hipGraphExecDestroy(a->step2[q]);
      if (a->step2_graph[q]) (void)hipGraphDestroy(a->step2_graph[q]);
      a->step2[q] = nullptr;
      a->step2_graph[q] = nullptr;
    }
    a->step2_partner = b->uid;
  }
  if (!a->step2[j]) {
    if (!a->cap) TF_HIP(hipStreamCreateWithFlags(&a->cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    int e = 0;
    hipError_t ce = hipSuccess, ie = hipSuccess;
    {
      std::lock_guard<std::mutex> lock(tf::capture_mutex());
      TF_HIP(hipStreamBeginCapture(a->cap, hipStreamCaptureModeThreadLocal));
      e = arnoldi_step2(a, b, j, a->cap);
      ce = hipStreamEndCapture(a->cap, &g);
      if (!e && ce == hipSuccess && g) ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    }
    if (e || ce != hipSuccess || !g || ie != hipSuccess) {
      // as in arnoldi_step_graph: this step runs eagerly and is captured again next time
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      tf::note_capture_fallback();
      return arnoldi_step2(a, b, j, st);
    }
    a->step2_graph[j] = g;
    a->step2[j] = ex;
  }
  TF_HIP(hipGraphLaunch(a->step2[j], st));
  return 0;
}

// Wait for event ev by polling.
int poll_event(hipEvent_t ev) {
  for (;;) {
    const hipError_t q = hipEventQuery(ev);
    if (q == hipSuccess) return 0;
    if (q != hipErrorNotReady) return tf::fail_hip(q, "hipEventQuery");
  }
}

// ---- restarted FGMRES, one or two columns ---------------------------------------

// The state of one system's FGMRES (its Krylov handle, right-hand side, iterate,
// Hessenberg matrix, Givens rotations and counters).
struct Column {
  tf_krylov* k = nullptr;
  const void* b = nullptr;
  void* x = nullptr;
  double bnrm = 0.0, rel = INFINITY, tol = 0.0;
  int its = 0, kk = 0;
  bool first = true, done = false, in_cycle = false, stepping = false;
  std::vector<double> H, cs, sn, g, yk;
  double& Hij(int i, int j) { return H[(size_t)i * k->m + j]; }
};

// bnrm = |b|; a zero right-hand side has the solution zero (done)
int col_zero_check(Column& c, double bnrm, hipStream_t st) {
  const int n = c.k->n;
  c.bnrm = bnrm;
  c.rel = 0.0;
  if (c.bnrm == 0.0) {
    c.done = true;
    return c.k->single ? tf_fill_f32(n, 0.0, (float*)c.x, st) : tf_fill_f64(n, 0.0, (double*)c.x, st);
  }
  c.rel = INFINITY;
  return 0;
}

// The columns' state and |b| (the norms of all columns queued, then one wait)
int col_start(Column* cols, int nc, hipStream_t st) {
  for (int q = 0; q < nc; ++q) {
    Column& c = cols[q];
    const int m = c.k->m;
    c.H.assign((size_t)(m + 1) * m, 0.0);
    c.cs.assign(m, 0.0);
    c.sn.assign(m, 0.0);
    c.g.assign(m + 1, 0.0);
    c.yk.assign(m, 0.0);
    if (int e = norm_queue(c.k, c.k->n, c.b, st)) return e;
  }
  TF_HIP(hipStreamSynchronize(st));
  for (int q = 0; q < nc; ++q)
    if (int e = col_zero_check(cols[q], norm_read(cols[q].k), st)) return e;
  return 0;
}

// A restart cycle's first vector V[0] = r / |r|, or done (converged, out of
// iterations, or a zero residual after the first cycle).
int col_begin_cycle(Column& c, const int* rp, const int* ci, const void* val, int gauge, int maxit, int x_is_zero,
                    hipStream_t st) {
  const double tol = c.tol;
  tf_krylov* k = c.k;
  const int n = k->n;
  Ops o{k, n, st};
  double beta;
  if (c.first && x_is_zero) {
    TF_HIP(hipMemcpyAsync(o.V(0), c.b, k->es * (size_t)n, hipMemcpyDeviceToDevice, st));
    beta = c.bnrm;
  } else {
    if (int e = spmv(k, n, rp, ci, val, gauge, c.x, k->w, st)) return e;
    if (int e = k->single
                    ? tf_axpby_f32(n, 1.0, (const float*)c.b, -1.0, (const float*)k->w, (float*)o.V(0), st)
                    : tf_axpby_f64(n, 1.0, (const double*)c.b, -1.0, (const double*)k->w, (double*)o.V(0), st))
      return e;
    if (int e = norm(k, n, o.V(0), &beta, st)) return e;
  }
  c.rel = beta / c.bnrm;
  if (c.rel <= tol || c.its >= maxit || (!c.first && beta == 0.0)) {
    c.done = true;
    return 0;
  }
  c.first = false;
  if (int e = k->single ? tf_scale_f32(n, 1.0 / beta, (const float*)o.V(0), (float*)o.V(0), st)
                        : tf_scale_f64(n, 1.0 / beta, (const double*)o.V(0), (double*)o.V(0), st))
    return e;
  std::fill(c.H.begin(), c.H.end(), 0.0);
  std::fill(c.g.begin(), c.g.end(), 0.0);
  c.g[0] = beta;
  c.kk = 0;
  c.in_cycle = true;
  c.stepping = true;
  return 0;
}

// Column j of the Hessenberg matrix (in k->h_host) through the previous Givens
// rotations and a new one; stepping stops on convergence of the residual
// estimate, the iteration limit, a breakdown or a non-finite estimate.
void col_givens(Column& c, int j, int maxit) {
  const double tol = c.tol;
  for (int i = 0; i <= j + 1; ++i) c.Hij(i, j) = c.k->h_host[i];
  for (int i = 0; i < j; ++i) {
    const double t = c.cs[i] * c.Hij(i, j) + c.sn[i] * c.Hij(i + 1, j);
    c.Hij(i + 1, j) = -c.sn[i] * c.Hij(i, j) + c.cs[i] * c.Hij(i + 1, j);
    c.Hij(i, j) = t;
  }
  const double d = std::hypot(c.Hij(j, j), c.Hij(j + 1, j));
  if (d == 0.0) {
    c.kk = j;
    c.stepping = false;
    return;
  }
  c.cs[j] = c.Hij(j, j) / d;
  c.sn[j] = c.Hij(j + 1, j) / d;
  c.Hij(j, j) = d;
  c.Hij(j + 1, j) = 0.0;
  c.g[j + 1] = -c.sn[j] * c.g[j];
  c.g[j] = c.cs[j] * c.g[j];
  if (j == 0) c.k->last_sn0 = std::fabs(c.sn[0]);
  ++c.its;
  c.kk = j + 1;
  if (std::fabs(c.g[j + 1]) / c.bnrm <= tol || c.its >= maxit || !std::isfinite(c.g[j + 1])) c.stepping = false;
}

// Whether a stepping column will most likely still be stepping after step j: the
// iteration limit is not reached and its residual estimate |g_j| times the last
// reduction factor |s_{j-1}| (the sine of the previous rotation; at j = 0 the first
// rotation's of the handle's previous cycle) stays above four times the tolerance.
bool col_continues(const Column& c, int j, int maxit) {
  if (c.its + 1 >= maxit) return false;
  const double s = j >= 1 ? std::fabs(c.sn[j - 1]) : c.k->last_sn0;
  const double pred = std::fabs(c.g[j]) * s;
  return s > 0.0 && std::isfinite(pred) && pred > 4.0 * c.tol * c.bnrm;
}

// x += Z y with H y = g (upper triangular); a single-precision solve that ran out
// of iterations ends on its residual estimate.
int col_end_cycle(Column& c, int maxit, hipStream_t st) {
  tf_krylov* k = c.k;
  const int n = k->n, kk = c.kk;
  c.in_cycle = false;
  c.stepping = false;
  if (kk == 0) {
    c.done = true;
    return 0;
  }
  for (int i = kk - 1; i >= 0; --i) {
    double sum = c.g[i];
    for (int q = i + 1; q < kk; ++q) sum -= c.Hij(i, q) * c.yk[q];
    c.yk[i] = sum / c.Hij(i, i);
  }
  TF_HIP(hipMemcpyAsync(k->y, c.yk.data(), sizeof(double) * kk, hipMemcpyHostToDevice, st));
  if (int e = k->single ? tf_multiaxpy_f32(n, kk, (const float*)k->Z, n, k->y, (float*)c.x, 1.0, st)
                        : tf_multiaxpy_f64(n, kk, (const double*)k->Z, n, k->y, (double*)c.x, 1.0, st))
    return e;
  if (c.its >= maxit && k->single) {
    c.rel = std::fabs(c.g[kk]) / c.bnrm;
    c.done = true;
  }
  return 0;
}

// The preconditioner on the stepping columns' V[j] -> Z[j]: one two-column apply of
// the line factor when two columns step together, else one apply each.
int precond_columns(Column* c, int nc, int j, int kind, void* prec, int ncell, int dof, hipStream_t st) {
  int act[2], na = 0;
  for (int q = 0; q < nc; ++q)
    if (c[q].stepping) act[na++] = q;
  if (na == 2 && kind == 1) {
    tf_krylov* k0 = c[act[0]].k;
    tf_krylov* k1 = c[act[1]].k;
    Ops o0{k0, k0->n, st}, o1{k1, k1->n, st};
    return k0->single ? tf_bcr_solve2_f32((tf_bcr*)prec, (const float*)o0.V(j), (const float*)o1.V(j),
                                          (float*)o0.Z(j), (float*)o1.Z(j), st)
                      : tf_bcr_solve2_f64((tf_bcr*)prec, (const double*)o0.V(j), (const double*)o1.V(j),
                                          (double*)o0.Z(j), (double*)o1.Z(j), st);
  }
  for (int a = 0; a < na; ++a) {
    tf_krylov* k = c[act[a]].k;
    Ops o{k, k->n, st};
    if (int e = precond(k, kind, prec, ncell, dof, o.V(j), o.Z(j), st)) return e;
  }
  return 0;
}

int fgmres_columns(Column* c, int nc, const int* rp, const int* ci, const void* val, int gauge, int prec_kind,
                   void* prec, int ncell, int dof, int maxit, int x_is_zero, hipStream_t st) {
  if (int e = col_start(c, nc, st)) return e;
  // the Arnoldi step graphs read the matrix through each handle's descriptor
  for (int q = 0; q < nc; ++q)
    hipLaunchKernelGGL(k_set_desc, dim3(1), dim3(1), 0, st, c[q].k->desc, rp, ci, val, gauge);
  TF_LAUNCH_CHECK("fgmres descriptor");
  const int m = c[0].k->m;
  for (;;) {
    bool any = false;
    for (int q = 0; q < nc; ++q) {
      if (c[q].done) continue;
      if (int e = col_begin_cycle(c[q], rp, ci, val, gauge, maxit, x_is_zero, st)) return e;
      any = any || c[q].in_cycle;
    }
    if (!any) break;
    // Z[j] of the stepping columns already queued by the previous step's look-ahead
    bool have_z = false;
    for (int j = 0; j < m; ++j) {
      bool stepping = false;
      for (int q = 0; q < nc; ++q) stepping = stepping || c[q].stepping;
      if (!stepping) break;
      if (!have_z)
        if (int e = precond_columns(c, nc, j, prec_kind, prec, ncell, dof, st)) return e;
      if (nc == 2 && c[0].stepping && c[1].stepping) {
        tf_krylov* a = c[0].k;
        tf_krylov* b = c[1].k;
        if (int e = (a->use_graph && !tf::prof_on()) ? arnoldi_step2_graph(a, b, j, st) : arnoldi_step2(a, b, j, st))
          return e;
      } else {
        for (int q = 0; q < nc; ++q) {
          if (!c[q].stepping) continue;
          tf_krylov* k = c[q].k;
          if (int e = (k->use_graph && !tf::prof_on()) ? arnoldi_step_graph(k, j, st) : arnoldi_step(k, j, st))
            return e;
        }
      }
      // Look-ahead: when every stepping column will most likely go on, the
      // preconditioner on V[j+1] (written by the step just queued) is queued before
      // the host waits for the Hessenberg column, so the device works through the
      // host's Givens update.  The arithmetic is unchanged; a wrong guess costs one
      // unused Z[j+1].
      bool ahead = j + 1 < m;
      for (int q = 0; q < nc; ++q)
        if (c[q].stepping && !col_continues(c[q], j, maxit)) ahead = false;
      have_z = ahead;
      // the event marks the step's Hessenberg columns as visible to the host
      hipEvent_t ev = c[0].k->step_done;
      TF_HIP(hipEventRecord(ev, st));
      if (ahead)
        if (int e = precond_columns(c, nc, j + 1, prec_kind, prec, ncell, dof, st)) return e;
      if (int e = poll_event(ev)) return e;
      for (int q = 0; q < nc; ++q)
        if (c[q].stepping) col_givens(c[q], j, maxit);
    }
    for (int q = 0; q < nc; ++q)
      if (c[q].in_cycle)
        if (int e = col_end_cycle(c[q], maxit, st)) return e;
  }
  // the host copies of y must stay valid until the multiaxpys have read them
  TF_HIP(hipStreamSynchronize(st));
  return 0;
}

}  // namespace

extern "C" {

int tf_krylov_create(tf_krylov** out, int n, int restart, int single) {
  if (!out) return tf::fail("null handle");
  *out = nullptr;
  if (n < 1 || restart < 1 || restart > tf_krylov_max_basis() - 1) return tf::fail("bad Krylov size");
  tf_krylov* k = new tf_krylov();
  k->n = n;
  k->m = restart;
  k->single = single;
  k->es = single ? sizeof(float) : sizeof(double);
  auto alloc = [&]() -> int {
    TF_HIP(hipMalloc(&k->V, k->es * (size_t)n * (restart + 1)));
    TF_HIP(hipMalloc(&k->Z, k->es * (size_t)n * restart));
    TF_HIP(hipMalloc(&k->w, k->es * (size_t)n));
    TF_HIP(hipMalloc(&k->h, sizeof(double) * (restart + 2)));
    TF_HIP(hipMalloc(&k->y, sizeof(double) * (restart + 1)));
    TF_HIP(hipMalloc(&k->work, sizeof(double) * (size_t)tf_krylov_work_size()));
    // coherent: the Arnoldi step's last kernel writes the Hessenberg column here
    // and the host reads it as soon as the step's completion is seen
    TF_HIP(hipHostMalloc(&k->h_host, sizeof(double) * (restart + 2), hipHostMallocCoherent));
    TF_HIP(hipHostGetDevicePointer((void**)&k->h_host_dev, k->h_host, 0));
    TF_HIP(hipMalloc(&k->desc, sizeof(MatrixDesc)));
    TF_HIP(hipEventCreateWithFlags(&k->step_done, hipEventDisableTiming));
    return 0;
  };
  k->use_graph = 1;
  static std::atomic<unsigned long long> next_uid{1};
  k->uid = next_uid++;
  if (int e = alloc()) {
    tf_krylov_destroy(k);  // frees the partial set (pointers start null)
    return e;
  }
  *out = k;
  return 0;
}

int tf_krylov_set_graph(tf_krylov* k, int on) {
  if (!k) return tf::fail("null handle");
  k->use_graph = on ? 1 : 0;
  return 0;
}

int tf_krylov_destroy(tf_krylov* k) {
  if (!k) return 0;
  for (int j = 0; j < MAX_STEPS; ++j) {
    if (k->step[j]) (void)hipGraphExecDestroy(k->step[j]);
    if (k->step_graph[j]) (void)hipGraphDestroy(k->step_graph[j]);
    if (k->step2[j]) (void)hipGraphExecDestroy(k->step2[j]);
    if (k->step2_graph[j]) (void)hipGraphDestroy(k->step2_graph[j]);
  }
  if (k->cap) (void)hipStreamDestroy(k->cap);
  if (k->step_done) (void)hipEventDestroy(k->step_done);
  for (void* p : {k->V, k->Z, k->w, (void*)k->h, (void*)k->y, (void*)k->work, (void*)k->desc}) (void)hipFree(p);
  (void)hipHostFree(k->h_host);
  delete k;
  return 0;
}

// Solve (gauged A) x = b by FGMRES with A in CSR (values in the handle's precision),
// preconditioner kind 0 none / 1 tf_bcr / 2 tf_linelu / 3 cell-block inverse (handle =
// the FP64 inverse blocks, ncell, dof) / 4 tf_mg / 5 tf_lineilu.  x holds the initial
// guess (x_is_zero: it is zero and the first residual is b).  its_out / rel_out:
// iterations and the final relative residual -- the true one, except for a
// single-precision solve that ends on the iteration limit (the residual estimate
// then; callers refine in FP64).
int tf_fgmres(tf_krylov* k, const int* rp, const int* ci, const void* val, int gauge, int prec_kind,
              void* prec, int ncell, int dof, const void* b, void* x, double tol, int maxit, int x_is_zero,
              int* its_out, double* rel_out, void* stream) {
  if (!k || !its_out || !rel_out) return tf::fail("null argument");
  Column c[1];
  c[0].k = k;
  c[0].b = b;
  c[0].x = x;
  c[0].tol = tol;
  const int e = fgmres_columns(c, 1, rp, ci, val, gauge, prec_kind, prec, ncell, dof, maxit, x_is_zero,
                               (hipStream_t)stream);
  *its_out = c[0].its;
  *rel_out = c[0].rel;
  return e;
}

// Two systems with the same matrix and preconditioner (A x0 = b0, A x1 = b1), each
// with its own Krylov handle (same size, restart and precision), solved in lockstep:
// while both are in a restart cycle each preconditioner application serves both
// (tf_bcr_solve2_*: one pass over the factor blocks).  Each column's iterates are
// those of tf_fgmres on its own system with its own tolerance.
int tf_fgmres2(tf_krylov* k0, tf_krylov* k1, const int* rp, const int* ci, const void* val, int gauge,
               int prec_kind, void* prec, int ncell, int dof, const void* b0, const void* b1, void* x0, void* x1,
               double tol0, double tol1, int maxit, int x_is_zero, int* its_out, double* rel_out, void* stream) {
  if (!k0 || !k1 || !its_out || !rel_out) return tf::fail("null argument");
  if (k0 == k1 || k0->n != k1->n || k0->m != k1->m || k0->single != k1->single)
    return tf::fail("tf_fgmres2: two distinct Krylov handles of one size, restart and precision needed");
  Column c[2];
  c[0].k = k0;
  c[0].b = b0;
  c[0].x = x0;
  c[1].k = k1;
  c[1].b = b1;
  c[1].x = x1;
  c[0].tol = tol0;
  c[1].tol = tol1;
  const int e = fgmres_columns(c, 2, rp, ci, val, gauge, prec_kind, prec, ncell, dof, maxit, x_is_zero,
                               (hipStream_t)stream);
  for (int j = 0; j < 2; ++j) {
    its_out[j] = c[j].its;
    rel_out[j] = c[j].rel;
  }
  return e;
}

}  // extern "C"
