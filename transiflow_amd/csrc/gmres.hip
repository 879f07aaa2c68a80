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
// valid across Jacobians.  The host then waits for the column by polling the
// stream (no blocking wait, whose wake-up would leave the GPU idle longer).
#include "spmv.h"
#include "tf_common.h"
#include "transiflow_hip.h"

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
  double* h_host;  // m+2 (pinned)
  MatrixDesc* desc;          // the matrix of the current solve (device)
  hipStream_t cap;           // private stream the step graphs are captured on
  hipGraphExec_t step[MAX_STEPS];
  hipGraph_t step_graph[MAX_STEPS];
  int use_graph;
};

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

// ||x|| (host double; synchronises the stream)
int norm(const tf_krylov* k, int n, const void* x, double* out, hipStream_t st) {
  if (k->single) {
    if (int e = tf_multidot_f32(n, 1, (const float*)x, n, (const float*)x, k->work, k->h + k->m + 1, 0, st)) return e;
  } else {
    if (int e = tf_multidot_f64(n, 1, (const double*)x, n, (const double*)x, k->work, k->h + k->m + 1, 0, st))
      return e;
  }
  TF_HIP(hipMemcpyAsync(k->h_host + k->m + 1, k->h + k->m + 1, sizeof(double), hipMemcpyDeviceToHost, st));
  TF_HIP(hipStreamSynchronize(st));
  *out = std::sqrt(k->h_host[k->m + 1]);
  return 0;
}

// The device part of Arnoldi step j on stream st: w = A Z[j] (through the
// descriptor), CGS2 h = V^T w, w -= V h, y = V^T w, w -= V y, h += y, V[j+1] =
// w / |w| with |w| into h[j+1], and h[0..j+1] copied to h_host.
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
  if (k->single) {
    const float* Vf = (const float*)k->V;
    float* wf = (float*)k->w;
    if (int e = tf_multidot_f32(n, j + 1, Vf, n, wf, k->work, k->h, 0, st)) return e;
    if (int e = tf_multiaxpy_f32(n, j + 1, Vf, n, k->h, wf, -1.0, st)) return e;
    if (int e = tf_multidot_f32(n, j + 1, Vf, n, wf, k->work, k->y, 0, st)) return e;
    if (int e = tf_multiaxpy_f32(n, j + 1, Vf, n, k->y, wf, -1.0, st)) return e;
  } else {
    const double* Vd = (const double*)k->V;
    double* wd = (double*)k->w;
    if (int e = tf_multidot_f64(n, j + 1, Vd, n, wd, k->work, k->h, 0, st)) return e;
    if (int e = tf_multiaxpy_f64(n, j + 1, Vd, n, k->h, wd, -1.0, st)) return e;
    if (int e = tf_multidot_f64(n, j + 1, Vd, n, wd, k->work, k->y, 0, st)) return e;
    if (int e = tf_multiaxpy_f64(n, j + 1, Vd, n, k->y, wd, -1.0, st)) return e;
  }
  if (int e = tf_axpby_f64(j + 1, 1.0, k->h, 1.0, k->y, k->h, st)) return e;
  if (int e = k->single ? tf_normalize_f32(n, (const float*)k->w, k->work, k->h + j + 1, (float*)o.V(j + 1), st)
                        : tf_normalize_f64(n, (const double*)k->w, k->work, k->h + j + 1, (double*)o.V(j + 1), st))
    return e;
  TF_HIP(hipMemcpyAsync(k->h_host, k->h, sizeof(double) * (j + 2), hipMemcpyDeviceToHost, st));
  return 0;
}

// Step j replayed from its captured graph (captured on first use).
int arnoldi_step_graph(tf_krylov* k, int j, hipStream_t st) {
  if (!k->step[j]) {
    if (!k->cap) TF_HIP(hipStreamCreateWithFlags(&k->cap, hipStreamNonBlocking));
    TF_HIP(hipStreamBeginCapture(k->cap, hipStreamCaptureModeThreadLocal));
    const int e = arnoldi_step(k, j, k->cap);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(k->cap, &g);
    if (e) {
      if (g) (void)hipGraphDestroy(g);
      return e;
    }
    TF_HIP(ce);
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(g);
      return tf::fail_hip(ie, "hipGraphInstantiate (Arnoldi step)");
    }
    k->step_graph[j] = g;
    k->step[j] = ex;
  }
  TF_HIP(hipGraphLaunch(k->step[j], st));
  return 0;
}

// Wait for stream st by polling: returns once everything queued has finished.
int poll_wait(hipStream_t st) {
  for (;;) {
    const hipError_t q = hipStreamQuery(st);
    if (q == hipSuccess) return 0;
    if (q != hipErrorNotReady) return tf::fail_hip(q, "hipStreamQuery");
  }
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
    TF_HIP(hipHostMalloc(&k->h_host, sizeof(double) * (restart + 2), hipHostMallocDefault));
    TF_HIP(hipMalloc(&k->desc, sizeof(MatrixDesc)));
    return 0;
  };
  k->use_graph = 1;
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
  }
  if (k->cap) (void)hipStreamDestroy(k->cap);
  for (void* p : {k->V, k->Z, k->w, (void*)k->h, (void*)k->y, (void*)k->work, (void*)k->desc}) (void)hipFree(p);
  (void)hipHostFree(k->h_host);
  delete k;
  return 0;
}

// Solve (gauged A) x = b by FGMRES with A in CSR (values in the handle's precision),
// preconditioner kind 0 none / 1 tf_bcr / 2 tf_linelu / 3 cell-block inverse (handle =
// the FP64 inverse blocks, ncell, dof).  x holds the initial guess (x_is_zero: it is
// zero and the first residual is b).  its_out / rel_out: iterations and the final
// relative residual -- the true one, except for a single-precision solve that ends on
// the iteration limit (the residual estimate then; callers refine in FP64).
int tf_fgmres(tf_krylov* k, const int* rp, const int* ci, const void* val, int gauge, int prec_kind,
              void* prec, int ncell, int dof, const void* b, void* x, double tol, int maxit, int x_is_zero,
              int* its_out, double* rel_out, void* stream) {
  if (!k || !its_out || !rel_out) return tf::fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  const int n = k->n, m = k->m;
  Ops o{k, n, st};
  double bnrm = 0.0;
  if (int e = norm(k, n, b, &bnrm, st)) return e;
  *its_out = 0;
  *rel_out = 0.0;
  if (bnrm == 0.0)
    return k->single ? tf_fill_f32(n, 0.0, (float*)x, st) : tf_fill_f64(n, 0.0, (double*)x, st);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), yk(m);
  // the Arnoldi step graphs read the matrix through the descriptor
  hipLaunchKernelGGL(k_set_desc, dim3(1), dim3(1), 0, st, k->desc, rp, ci, val, gauge);
  TF_LAUNCH_CHECK("fgmres descriptor");
  auto Hij = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
  int its = 0;
  double rel = INFINITY;
  bool first = true;
  for (;;) {
    double beta;
    if (first && x_is_zero) {
      TF_HIP(hipMemcpyAsync(o.V(0), b, k->es * (size_t)n, hipMemcpyDeviceToDevice, st));
      beta = bnrm;
    } else {
      if (int e = spmv(k, n, rp, ci, val, gauge, x, k->w, st)) return e;
      if (int e = k->single ? tf_axpby_f32(n, 1.0, (const float*)b, -1.0, (const float*)k->w, (float*)o.V(0), st)
                            : tf_axpby_f64(n, 1.0, (const double*)b, -1.0, (const double*)k->w, (double*)o.V(0), st))
        return e;
      if (int e = norm(k, n, o.V(0), &beta, st)) return e;
    }
    rel = beta / bnrm;
    if (rel <= tol || its >= maxit || (!first && beta == 0.0)) break;
    first = false;
    if (int e = k->single ? tf_scale_f32(n, 1.0 / beta, (const float*)o.V(0), (float*)o.V(0), st)
                          : tf_scale_f64(n, 1.0 / beta, (const double*)o.V(0), (double*)o.V(0), st))
      return e;
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int kk = 0;
    bool estimate_exit = false;
    for (int j = 0; j < m; ++j) {
      if (int e = precond(k, prec_kind, prec, ncell, dof, o.V(j), o.Z(j), st)) return e;
      if (int e = (k->use_graph && !tf::prof_on()) ? arnoldi_step_graph(k, j, st) : arnoldi_step(k, j, st)) return e;
      if (int e = poll_wait(st)) return e;
      for (int i = 0; i <= j + 1; ++i) Hij(i, j) = k->h_host[i];
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * Hij(i, j) + sn[i] * Hij(i + 1, j);
        Hij(i + 1, j) = -sn[i] * Hij(i, j) + cs[i] * Hij(i + 1, j);
        Hij(i, j) = t;
      }
      const double d = std::hypot(Hij(j, j), Hij(j + 1, j));
      if (d == 0.0) {
        kk = j;
        break;
      }
      cs[j] = Hij(j, j) / d;
      sn[j] = Hij(j + 1, j) / d;
      Hij(j, j) = d;
      Hij(j + 1, j) = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++its;
      kk = j + 1;
      if (std::fabs(g[j + 1]) / bnrm <= tol || its >= maxit || !std::isfinite(g[j + 1])) break;
    }
    if (kk == 0) break;
    for (int i = kk - 1; i >= 0; --i) {
      double s = g[i];
      for (int c = i + 1; c < kk; ++c) s -= Hij(i, c) * yk[c];
      yk[i] = s / Hij(i, i);
    }
    TF_HIP(hipMemcpyAsync(k->y, yk.data(), sizeof(double) * kk, hipMemcpyHostToDevice, st));
    if (int e = k->single ? tf_multiaxpy_f32(n, kk, (const float*)k->Z, n, k->y, (float*)x, 1.0, st)
                          : tf_multiaxpy_f64(n, kk, (const double*)k->Z, n, k->y, (double*)x, 1.0, st))
      return e;
    if (its >= maxit && k->single) {
      rel = std::fabs(g[kk]) / bnrm;
      estimate_exit = true;
    }
    if (estimate_exit) break;
  }
  // the host copy of y must stay valid until the multiaxpy has read it
  TF_HIP(hipStreamSynchronize(st));
  *its_out = its;
  *rel_out = rel;
  return 0;
}

}  // extern "C"
