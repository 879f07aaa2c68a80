// Vector kernels and CSR SpMV for the FP64 / FP32 solve path.
//
// Reductions are deterministic: a fixed grid of TF_RED_BLOCKS workgroups each
// reduces a fixed contiguous slice in a fixed tree order, then one workgroup
// reduces the partials; the same inputs give the same bits on every run.
#include "spmv.h"
#include "tf_common.h"

#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

namespace tf {

namespace {
struct ProfState {
  bool on = false;
  std::mutex mu;
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  hipEvent_t open[PROF_IDS] = {};
  long count[PROF_IDS] = {};
  double ms[PROF_IDS] = {};
  double work[PROF_IDS] = {};
};
ProfState& ps() {
  static ProfState s;
  return s;
}
hipEvent_t take_event(ProfState& s) {
  if (!s.pool.empty()) {
    hipEvent_t e = s.pool.back();
    s.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
}  // namespace

bool prof_on() { return ps().on; }

hipError_t upload(void* dst, const void* src, size_t bytes) {
  static std::mutex mu;
  static std::vector<hipStream_t> streams;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipStream_t s = nullptr;
  {
    std::lock_guard<std::mutex> g(mu);
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) {
      e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
      if (e != hipSuccess) return e;
    }
    s = streams[dev];
  }
  if (bytes == 0) return hipSuccess;
  e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

void prof_work(int id, double w) {
  ProfState& s = ps();
  std::lock_guard<std::mutex> g(s.mu);
  s.work[id] += w;
}

static std::atomic<int> g_capture_fallbacks{0};
void note_capture_fallback() { ++g_capture_fallbacks; }
int capture_fallbacks() { return g_capture_fallbacks.load(); }
std::mutex& capture_mutex() {
  static std::mutex m;
  return m;
}

void prof_begin(int id, hipStream_t st) {
  ProfState& s = ps();
  std::lock_guard<std::mutex> g(s.mu);
  hipEvent_t e = take_event(s);
  (void)hipEventRecord(e, st);
  s.open[id] = e;
}

void prof_end(int id, hipStream_t st) {
  ProfState& s = ps();
  std::lock_guard<std::mutex> g(s.mu);
  hipEvent_t e = take_event(s);
  (void)hipEventRecord(e, st);
  s.pending.push_back({id, {s.open[id], e}});
}

}  // namespace tf

namespace {

constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;

// out may alias x or y (element-wise read-then-write by the same thread)
template <typename T>
__global__ __launch_bounds__(256) void k_axpby(long n, double alpha, const T* x, double beta, const T* y, T* out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (T)(alpha * (double)x[i] + beta * (double)y[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale(long n, double alpha, const T* x, T* out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (T)(alpha * (double)x[i]);
}

// true division (numpy's x / a rounds differently from x * (1/a))
__global__ __launch_bounds__(256) void k_div(long n, const double* x, double a, double* out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = x[i] / a;
}

// Residual of one theta-method step from u to v, divided by theta dt:
// G = F(v) + cb F(u) + cm m .* (u - v), m the diagonal of the (diagonal) mass matrix.
__global__ __launch_bounds__(256) void k_theta_residual(long n, const double* __restrict__ Fv,
                                                        const double* __restrict__ Fu, const double* __restrict__ m,
                                                        const double* __restrict__ u, const double* __restrict__ v,
                                                        double cb, double cm, double* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = Fv[i] + cb * Fu[i] + cm * (m[i] * (u[i] - v[i]));
}

// Real form of the complex matrix A + iB, interleaved: unknown c becomes the pair
// (2c, 2c+1) = (re, im), row r the rows 2r (re) and 2r+1 (im):
//   row 2r:   a (2c) and -b (2c+1);   row 2r+1:  b (2c) and a (2c+1)
// over the union of the columns of row r in A and B.  Row and column ``gauge``
// become -e (both halves), as the pressure gauge of the real solve.
// count: cnt[r] = entries of each of the two output rows.
__global__ __launch_bounds__(256) void k_interleave_count(int n, const int* __restrict__ ra, const int* __restrict__ ca,
                                                          const int* __restrict__ rb, const int* __restrict__ cb,
                                                          int gauge, int* __restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if (r == gauge) {
    cnt[2 * r] = cnt[2 * r + 1] = 1;
    return;
  }
  int u = 0;
  for (int p = ra[r]; p < ra[r + 1]; ++p) u += ca[p] != gauge;
  for (int q = rb[r]; q < rb[r + 1]; ++q) {
    if (cb[q] == gauge) continue;
    bool found = false;
    for (int p = ra[r]; p < ra[r + 1] && !found; ++p) found = ca[p] == cb[q];
    u += !found;
  }
  cnt[2 * r] = cnt[2 * r + 1] = 2 * u;
}

__global__ __launch_bounds__(256) void k_interleave_fill(int n, const int* __restrict__ ra, const int* __restrict__ ca,
                                                         const double* __restrict__ va, const int* __restrict__ rb,
                                                         const int* __restrict__ cb, const double* __restrict__ vb,
                                                         int gauge, const int* __restrict__ rc, int* __restrict__ cc,
                                                         double* __restrict__ vc) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int o0 = rc[2 * r], o1 = rc[2 * r + 1];
  if (r == gauge) {
    cc[o0] = 2 * r;
    vc[o0] = -1.0;
    cc[o1] = 2 * r + 1;
    vc[o1] = -1.0;
    return;
  }
  // columns of A (with the matching B value, if any), then B's columns missing in A
  for (int p = ra[r]; p < ra[r + 1]; ++p) {
    const int c = ca[p];
    if (c == gauge) continue;
    double b = 0.0;
    for (int q = rb[r]; q < rb[r + 1]; ++q)
      if (cb[q] == c) b = vb[q];
    cc[o0] = 2 * c; vc[o0++] = va[p];
    cc[o0] = 2 * c + 1; vc[o0++] = -b;
    cc[o1] = 2 * c; vc[o1++] = b;
    cc[o1] = 2 * c + 1; vc[o1++] = va[p];
  }
  for (int q = rb[r]; q < rb[r + 1]; ++q) {
    const int c = cb[q];
    if (c == gauge) continue;
    bool found = false;
    for (int p = ra[r]; p < ra[r + 1] && !found; ++p) found = ca[p] == c;
    if (found) continue;
    cc[o0] = 2 * c; vc[o0++] = 0.0;
    cc[o0] = 2 * c + 1; vc[o0++] = -vb[q];
    cc[o1] = 2 * c; vc[o1++] = vb[q];
    cc[o1] = 2 * c + 1; vc[o1++] = 0.0;
  }
}

// diag[r] = A[r][r] (0 when the row holds no diagonal entry)
__global__ __launch_bounds__(256) void k_csr_diagonal(int n, const int* __restrict__ row_ptr,
                                                      const int* __restrict__ col, const double* __restrict__ val,
                                                      double* __restrict__ diag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
    if (col[p] == r) d += val[p];
  diag[r] = d;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void k_convert(long n, const Tin* __restrict__ x, Tout* __restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (Tout)x[i];
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_dot_partial(long n, const T* __restrict__ x, const T* __restrict__ y,
                                                             double* __restrict__ partial) {
  __shared__ double sh[RED_THREADS];
  const long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long b0 = blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  double acc = 0.0;
  for (long i = b0 + threadIdx.x; i < b1; i += RED_THREADS) acc += (double)x[i] * (double)y[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* __restrict__ partial, int np,
                                                              double* __restrict__ out, int sqrt_abs) {
  __shared__ double sh[RED_THREADS];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += RED_THREADS) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = sqrt_abs ? sqrt(fabs(s)) : s;
}

// y = A x for CSR A; with gauge >= 0 the operator is the pressure-gauged
// system of the SciPy backend (SciPy.py:95-103): column ``gauge`` is removed
// and row ``gauge`` becomes -x[gauge].
template <typename TV, typename TX>
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                              const TV* __restrict__ v, const TX* __restrict__ x, TX* __restrict__ y,
                                              int gauge) {
  tf::spmv(n, rp, ci, v, x, y, gauge);
}

__global__ void k_set_entry(double* x, long i, double v) { x[i] = v; }

template <typename T>
__global__ __launch_bounds__(256) void k_fill(long n, double v, T* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

// C = alpha A + beta B (CSR), row by row: A's entries in their order, B's
// entries added where A has the column, otherwise appended in B's order.
__global__ __launch_bounds__(256) void k_csr_add_count(int n, const int* __restrict__ ra, const int* __restrict__ ca,
                                                       const int* __restrict__ rb, const int* __restrict__ cb,
                                                       int* __restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int c = ra[r + 1] - ra[r];
  for (int q = rb[r]; q < rb[r + 1]; ++q) {
    bool found = false;
    for (int p = ra[r]; p < ra[r + 1] && !found; ++p) found = ca[p] == cb[q];
    if (!found) ++c;
  }
  cnt[r] = c;
}

__global__ void k_exclusive_scan_serial(int n, const int* __restrict__ cnt, int* __restrict__ rp) {
  // one thread: n is the row count of one matrix (sub-millisecond at these sizes)
  int s = 0;
  for (int r = 0; r < n; ++r) {
    rp[r] = s;
    s += cnt[r];
  }
  rp[n] = s;
}

__global__ __launch_bounds__(256) void k_csr_add_fill(int n, const int* __restrict__ ra, const int* __restrict__ ca,
                                                      const double* __restrict__ va, double alpha,
                                                      const int* __restrict__ rb, const int* __restrict__ cb,
                                                      const double* __restrict__ vb, double beta,
                                                      const int* __restrict__ rc, int* __restrict__ cc,
                                                      double* __restrict__ vc) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int o = rc[r];
  const int a0 = ra[r], a1 = ra[r + 1];
  for (int p = a0; p < a1; ++p) {
    cc[o + p - a0] = ca[p];
    vc[o + p - a0] = alpha * va[p];
  }
  int tail = o + (a1 - a0);
  for (int q = rb[r]; q < rb[r + 1]; ++q) {
    int hit = -1;
    for (int p = a0; p < a1 && hit < 0; ++p)
      if (ca[p] == cb[q]) hit = p;
    if (hit >= 0) {
      vc[o + hit - a0] += beta * vb[q];
    } else {
      cc[tail] = cb[q];
      vc[tail] = beta * vb[q];
      ++tail;
    }
  }
}

int grid_for(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

}  // namespace

extern "C" {

int tf_prof_enable(int on) {
  tf::ps().on = on != 0;
  return 0;
}

// Synchronises the recorded events; count/ms_total receive the totals of kernel id.
// Graph captures that failed and whose call ran eagerly instead, since the process started.
int tf_capture_fallbacks(void) { return tf::capture_fallbacks(); }

int tf_prof_read(int id, long* count, double* ms_total) {
  if (id < 0 || id >= tf::PROF_IDS) return tf::fail("bad profiling id");
  tf::ProfState& s = tf::ps();
  std::lock_guard<std::mutex> g(s.mu);
  for (auto& p : s.pending) {
    (void)hipEventSynchronize(p.second.second);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
    s.count[p.first] += 1;
    s.ms[p.first] += ms;
    s.pool.push_back(p.second.first);
    s.pool.push_back(p.second.second);
  }
  s.pending.clear();
  *count = s.count[id];
  *ms_total = s.ms[id];
  return 0;
}

int tf_prof_reset(void) {
  long c;
  double m;
  tf_prof_read(0, &c, &m);
  tf::ProfState& s = tf::ps();
  std::lock_guard<std::mutex> g(s.mu);
  for (int i = 0; i < tf::PROF_IDS; ++i) {
    s.count[i] = 0;
    s.ms[i] = 0.0;
    s.work[i] = 0.0;
  }
  return 0;
}

// Algorithmic work (flops or bytes, as the launch site states) summed over the
// profiled launches of kernel id.
int tf_prof_work(int id, double* work_total) {
  if (id < 0 || id >= tf::PROF_IDS || !work_total) return tf::fail("bad profiling id");
  tf::ProfState& s = tf::ps();
  std::lock_guard<std::mutex> g(s.mu);
  *work_total = s.work[id];
  return 0;
}

// C = alpha A + beta B.  count writes rc[0..n] (rc[n] = nnz of C) and nnz_host after a
// stream sync, using cnt (n ints of scratch); fill writes cc, vc.
int tf_csr_add_count(int n, const int* ra, const int* ca, const int* rb, const int* cb, int* cnt, int* rc,
                     int* nnz_host, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_csr_add_count, dim3((n + 255) / 256), dim3(256), 0, st, n, ra, ca, rb, cb, cnt);
  hipLaunchKernelGGL(k_exclusive_scan_serial, dim3(1), dim3(1), 0, st, n, (const int*)cnt, rc);
  TF_LAUNCH_CHECK("csr add count");
  if (nnz_host) {
    TF_HIP(hipMemcpyAsync(nnz_host, rc + n, sizeof(int), hipMemcpyDeviceToHost, st));
    TF_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

// ---- tf_csr_add_count_noalloc: the count pass with library-owned scratch ----
// The per-row counts live in a file-scope device array, one copy per device
// for the life of the process, so the call allocates nothing.  Every call on a
// device shares that device's array: calls that overlap on one device race.
// Each call writes the counts of rows 0..n-1 before reading them, so nothing
// has to be cleared between calls.
constexpr int CSR_ADD_SCRATCH_ROWS = 1 << 22;  // 4,194,304 rows, 16 MiB per device
static __device__ __align__(256) int scratch[CSR_ADD_SCRATCH_ROWS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int* g_scratch_dev[MAX_DEVICES] = {};  // device address of `scratch`, per device ordinal

// Device address of the current device's `scratch`, looked up once per device
// (the array's host-side name is not a device address).
int scratch_of_current_device(int** out) {
  int dev = 0;
  TF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return tf::fail("tf_csr_add_count_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    TF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return tf::fail("tf_csr_add_count_noalloc: no device address for the scratch array");
    g_scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}
}  // namespace

int tf_csr_add_count_noalloc(int n, const int* ra, const int* ca, const int* rb, const int* cb, int* rc,
                             int* nnz_host, void* stream) {
  if (n < 1 || n > CSR_ADD_SCRATCH_ROWS)
    return tf::fail("tf_csr_add_count_noalloc: n must be in [1, 4194304] (the size of its scratch)");
  int* cnt = nullptr;
  if (int e = scratch_of_current_device(&cnt)) return e;
  return tf_csr_add_count(n, ra, ca, rb, cb, cnt, rc, nnz_host, stream);
}

int tf_csr_add_fill(int n, const int* ra, const int* ca, const double* va, double alpha, const int* rb,
                    const int* cb, const double* vb, double beta, const int* rc, int* cc, double* vc, void* stream) {
  hipLaunchKernelGGL(k_csr_add_fill, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, ra, ca, va, alpha,
                     rb, cb, vb, beta, rc, cc, vc);
  TF_LAUNCH_CHECK("csr add fill");
  return 0;
}

int tf_theta_residual(long n, const double* Fv, const double* Fu, const double* mdiag, const double* u,
                      const double* v, double cb, double cm, double* out, void* stream) {
  if (n < 1) return tf::fail("tf_theta_residual: empty vector");
  hipLaunchKernelGGL(k_theta_residual, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, Fv, Fu, mdiag, u, v,
                     cb, cm, out);
  TF_LAUNCH_CHECK("k_theta_residual");
  return 0;
}

int tf_csr_interleave_count(int n, const int* ra, const int* ca, const int* rb, const int* cb, int gauge, int* cnt,
                            int* rc, int* nnz_host, void* stream) {
  if (n < 1) return tf::fail("tf_csr_interleave_count: empty matrix");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_interleave_count, dim3((n + 255) / 256), dim3(256), 0, st, n, ra, ca, rb, cb, gauge, cnt);
  hipLaunchKernelGGL(k_exclusive_scan_serial, dim3(1), dim3(1), 0, st, 2 * n, (const int*)cnt, rc);
  TF_LAUNCH_CHECK("csr interleave count");
  if (nnz_host) {
    TF_HIP(hipMemcpyAsync(nnz_host, rc + 2 * n, sizeof(int), hipMemcpyDeviceToHost, st));
    TF_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

int tf_csr_interleave_fill(int n, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
                           const double* vb, int gauge, const int* rc, int* cc, double* vc, void* stream) {
  if (n < 1) return tf::fail("tf_csr_interleave_fill: empty matrix");
  hipLaunchKernelGGL(k_interleave_fill, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, ra, ca, va, rb, cb,
                     vb, gauge, rc, cc, vc);
  TF_LAUNCH_CHECK("csr interleave fill");
  return 0;
}

int tf_csr_diagonal(int n, const int* row_ptr, const int* col, const double* val, double* diag, void* stream) {
  if (n < 1) return tf::fail("tf_csr_diagonal: empty matrix");
  hipLaunchKernelGGL(k_csr_diagonal, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, row_ptr, col, val,
                     diag);
  TF_LAUNCH_CHECK("k_csr_diagonal");
  return 0;
}

int tf_set_entry_f64(double* x, long n, long i, double v, void* stream) {
  if (i < 0 || i >= n) return tf::fail("tf_set_entry_f64: index outside the vector");
  hipLaunchKernelGGL(k_set_entry, dim3(1), dim3(1), 0, (hipStream_t)stream, x, i, v);
  TF_LAUNCH_CHECK("k_set_entry");
  return 0;
}

int tf_fill_f64(long n, double v, double* x, void* stream) {
  hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, v, x);
  TF_LAUNCH_CHECK("k_fill");
  return 0;
}

int tf_axpby_f64(long n, double alpha, const double* x, double beta, const double* y, double* out, void* stream) {
  hipLaunchKernelGGL(k_axpby<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, alpha, x, beta, y, out);
  TF_LAUNCH_CHECK("k_axpby");
  return 0;
}

int tf_fill_f32(long n, double v, float* x, void* stream) {
  hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, v, x);
  TF_LAUNCH_CHECK("k_fill");
  return 0;
}

int tf_axpby_f32(long n, double alpha, const float* x, double beta, const float* y, float* out, void* stream) {
  hipLaunchKernelGGL(k_axpby<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, alpha, x, beta, y, out);
  TF_LAUNCH_CHECK("k_axpby");
  return 0;
}

int tf_scale_f32(long n, double alpha, const float* x, float* out, void* stream) {
  hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, alpha, x, out);
  TF_LAUNCH_CHECK("k_scale");
  return 0;
}

int tf_scale_f64(long n, double alpha, const double* x, double* out, void* stream) {
  hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, alpha, x, out);
  TF_LAUNCH_CHECK("k_scale");
  return 0;
}

int tf_div_f64(long n, const double* x, double a, double* out, void* stream) {
  hipLaunchKernelGGL(k_div, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, x, a, out);
  TF_LAUNCH_CHECK("k_div");
  return 0;
}

int tf_f64_to_f32(long n, const double* x, float* out, void* stream) {
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, x, out);
  TF_LAUNCH_CHECK("k_convert");
  return 0;
}

int tf_f32_to_f64(long n, const float* x, double* out, void* stream) {
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, x, out);
  TF_LAUNCH_CHECK("k_convert");
  return 0;
}

// Deterministic dot product (or 2-norm when norm != 0) into out_dev[0];
// ``work`` holds at least tf_reduce_work_size() doubles.  When out_host is
// not null the stream is synchronised and the value copied there.
int tf_reduce_work_size(void) { return RED_BLOCKS; }

int tf_dot_f64(long n, const double* x, const double* y, double* work, double* out_dev, double* out_host, int norm,
               void* stream) {
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_dot_partial<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, n, x, y, work);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, st, work, RED_BLOCKS, out_dev, norm);
  TF_LAUNCH_CHECK("tf_dot_f64");
  if (out_host) {
    TF_HIP(hipMemcpyAsync(out_host, out_dev, sizeof(double), hipMemcpyDeviceToHost, st));
    TF_HIP(hipStreamSynchronize(st));
  }
  return 0;
}

int tf_spmv_f64(int n, const int* rp, const int* ci, const double* v, const double* x, double* y, int gauge,
                void* stream) {
  TF_PROF_BEGIN(tf::PROF_SPMV, (hipStream_t)stream);
  hipLaunchKernelGGL((k_spmv<double, double>), dim3(tf::spmv_blocks<double>(n)), dim3(256), 0, (hipStream_t)stream, n, rp,
                     ci, v, x, y, gauge);
  TF_PROF_END(tf::PROF_SPMV, (hipStream_t)stream);
  TF_LAUNCH_CHECK("k_spmv f64");
  return 0;
}

int tf_refine_residual(int n, const int* rp, const int* ci, const double* val, int gauge, const double* x,
                       const double* b, double* r, float* r32, float* d32, double* work, double* norm_dev,
                       void* stream) {
  if (n < 1 || !rp || !ci || !val || !x || !b || !r || !r32 || !d32 || !work || !norm_dev)
    return tf::fail("tf_refine_residual: bad arguments");
  if (int e = tf_spmv_f64(n, rp, ci, val, x, r, gauge, stream)) return e;
  if (int e = tf_axpby_f64(n, 1.0, b, -1.0, r, r, stream)) return e;
  if (int e = tf_dot_f64(n, r, r, work, norm_dev, nullptr, 1, stream)) return e;
  if (int e = tf_f64_to_f32(n, r, r32, stream)) return e;
  return tf_fill_f32(n, 0.0, d32, stream);
}

int tf_refine_update(long n, const float* d32, double* d, double* x, void* stream) {
  if (n < 1 || !d32 || !d || !x) return tf::fail("tf_refine_update: bad arguments");
  if (int e = tf_f32_to_f64(n, d32, d, stream)) return e;
  return tf_axpby_f64(n, 1.0, x, 1.0, d, x, stream);
}

int tf_read_f64(const double* dev, int count, double* host, void* stream) {
  if (count < 0 || (count > 0 && (!dev || !host))) return tf::fail("tf_read_f64: bad arguments");
  if (count == 0) return 0;
  TF_HIP(hipMemcpyAsync(host, dev, sizeof(double) * (size_t)count, hipMemcpyDeviceToHost, (hipStream_t)stream));
  TF_HIP(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}

int tf_spmv_f32(int n, const int* rp, const int* ci, const float* v, const float* x, float* y, int gauge,
                void* stream) {
  hipLaunchKernelGGL((k_spmv<float, float>), dim3(tf::spmv_blocks<float>(n)), dim3(256), 0, (hipStream_t)stream, n, rp, ci, v,
                     x, y, gauge);
  TF_LAUNCH_CHECK("k_spmv f32");
  return 0;
}

}  // extern "C"
