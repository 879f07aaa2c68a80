// Two-grid preconditioner for the 3D cavity: line block ILU smoothing on the
// grid, and an exact solve on the grid coarsened by two in every direction.
//
// The coarse operator is the Jacobian of the same problem re-assembled on the
// coarse grid at the restricted state (the assembly kernels, any grid size),
// factored exactly by the plane factorisation (bcr.hip).  Transfers follow the
// staggered (MAC) layout of the discretisation (Discretization.py:20-32,
// utils.py:58-59: u on the east face, v on the north face, w on the top face,
// p and the scalars at the cell centre), and the finite-volume scaling of its
// equations (every equation is integrated over its control volume):
//
//   restriction  scalar rows: the sum over the 8 (2D: 4) fine cells of a coarse
//                cell; the gauge row (coarse cell 0 pressure) takes the fine
//                gauge row's residual;
//                velocity rows: along the normal, fine faces 2I, 2I+1, 2I+2 with
//                weights 1/2, 1, 1/2; summed over the 2 x 2 (2D: 2) tangential
//                fine faces; the wall face (last face, a Dirichlet row) gets 0;
//   prolongation scalars: the coarse cell's value;
//                velocities: linear along the normal (fine face 2I+1 is coarse
//                face I, fine face 2I the mean of coarse faces I-1 and I, with 0
//                at the start wall), linear in the tangential directions with
//                weights 3/4, 1/4 and a reflection (-value) beyond the walls.
//   With these, R A_h P equals the coarse Jacobian up to O(h^2) on smooth
//   fields (the rediscretised coarse operator is the Galerkin one).
//   state restriction (for the coarse Jacobian): scalars the mean of the fine
//                cells, velocities the mean of the coincident fine faces.
//
// One application (V(1,1) on two levels):
//   z = S r;  z += P B (R (r - A z));  z += S (r - A z)
// with S the line ILU (lineilu.hip), B the exact coarse solve, A the gauged
// fine Jacobian.  Replaces, as the preconditioner of the GMRES solve, the
// reference's spilu (transiflow/interface/SciPy.py:164-192,260-309) for
// grids whose exact factor does not fit.
#include "tf_common.h"
#include "transiflow_hip.h"

namespace {

struct Grid3 {
  int n[3];   // fine cells per direction
  int N[3];   // coarse cells per direction
  int dim, dof;
  __device__ __forceinline__ long fine_index(int i, int j, int k, int d) const {
    return d + (long)dof * (i + (long)n[0] * (j + (long)n[1] * k));
  }
  __device__ __forceinline__ long coarse_index(int I, int J, int K, int d) const {
    return d + (long)dof * (I + (long)N[0] * (J + (long)N[1] * K));
  }
};

// fine cells of coarse index I along axis a: 2I and 2I+1 (one cell if the axis is not coarsened)
__device__ __forceinline__ int fine_count(const Grid3& g, int a) { return g.n[a] == g.N[a] ? 1 : 2; }

template <typename V>
__global__ __launch_bounds__(256) void k_restrict(Grid3 g, int gauge, const V* __restrict__ rf, V* __restrict__ rc) {
  const long cell = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long ncell = (long)g.N[0] * g.N[1] * g.N[2];
  if (cell >= ncell) return;
  int C[3];
  C[0] = (int)(cell % g.N[0]);
  C[1] = (int)((cell / g.N[0]) % g.N[1]);
  C[2] = (int)(cell / ((long)g.N[0] * g.N[1]));
  const int f[3] = {fine_count(g, 0), fine_count(g, 1), fine_count(g, 2)};
  for (int d = 0; d < g.dof; ++d) {
    double s = 0.0;
    if (d >= g.dim) {
      for (int c = 0; c < f[2]; ++c)
        for (int b = 0; b < f[1]; ++b)
          for (int a = 0; a < f[0]; ++a)
            s += (double)rf[g.fine_index(f[0] * C[0] + a, f[1] * C[1] + b, f[2] * C[2] + c, d)];
    } else {
      const int ax = d;
      if (C[ax] == g.N[ax] - 1) {
        s = 0.0;  // wall face: Dirichlet row
      } else {
        const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
        for (int u = 0; u < f[t2]; ++u)
          for (int w = 0; w < f[t1]; ++w)
            for (int q = 0; q < 3; ++q) {
              int F[3];
              F[t1] = f[t1] * C[t1] + w;
              F[t2] = f[t2] * C[t2] + u;
              F[ax] = 2 * C[ax] + q;
              if (F[ax] >= g.n[ax]) continue;
              const double wt = q == 1 ? 1.0 : 0.5;
              s += wt * (double)rf[g.fine_index(F[0], F[1], F[2], d)];
            }
      }
    }
    rc[g.coarse_index(C[0], C[1], C[2], d)] = (V)s;
  }
  if (cell == 0 && gauge >= 0) rc[gauge] = rf[gauge];  // the gauge row (cell 0 pressure)
}

template <typename V>
__device__ __forceinline__ double coarse_val(const Grid3& g, const V* ec, int I, int J, int K, int d) {
  return (double)ec[g.coarse_index(I, J, K, d)];
}

// z_f += P e_c
template <typename V>
__global__ __launch_bounds__(256) void k_prolong_add(Grid3 g, const V* __restrict__ ec, V* __restrict__ zf) {
  const long cell = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long ncell = (long)g.n[0] * g.n[1] * g.n[2];
  if (cell >= ncell) return;
  int F[3];
  F[0] = (int)(cell % g.n[0]);
  F[1] = (int)((cell / g.n[0]) % g.n[1]);
  F[2] = (int)(cell / ((long)g.n[0] * g.n[1]));
  const int f[3] = {fine_count(g, 0), fine_count(g, 1), fine_count(g, 2)};
  int C[3] = {F[0] / f[0], F[1] / f[1], F[2] / f[2]};
  for (int d = 0; d < g.dof; ++d) {
    double val;
    if (d >= g.dim) {
      val = coarse_val(g, ec, C[0], C[1], C[2], d);
    } else {
      const int ax = d;
      // normal direction: one or two coarse faces
      int nI[2];
      double nw[2];
      int nn;
      if (F[ax] & 1) {
        nI[0] = (F[ax] - 1) / 2;
        nw[0] = 1.0;
        nn = 1;
      } else {
        nI[0] = F[ax] / 2 - 1;
        nI[1] = F[ax] / 2;
        nw[0] = 0.5;
        nw[1] = 0.5;
        nn = 2;
      }
      // tangential directions: 3/4 own coarse cell, 1/4 the neighbour on the fine cell's side
      int tI[2][2];
      double tw[2][2];
      int tn[2];
      const int ts[2] = {ax == 0 ? 1 : 0, ax == 2 ? 1 : 2};
      for (int q = 0; q < 2; ++q) {
        const int t = ts[q];
        if (f[t] == 1) {
          tI[q][0] = C[t];
          tw[q][0] = 1.0;
          tn[q] = 1;
        } else {
          tI[q][0] = C[t];
          tw[q][0] = 0.75;
          tI[q][1] = C[t] + ((F[t] & 1) ? 1 : -1);
          tw[q][1] = 0.25;
          tn[q] = 2;
        }
      }
      val = 0.0;
      for (int a = 0; a < nn; ++a) {
        if (nI[a] < 0) continue;  // start wall: zero correction
        for (int b = 0; b < tn[0]; ++b)
          for (int c = 0; c < tn[1]; ++c) {
            int I[3];
            I[ax] = nI[a];
            I[ts[0]] = tI[0][b];
            I[ts[1]] = tI[1][c];
            double sign = 1.0;
            for (int q = 0; q < 2; ++q) {
              const int t = ts[q];
              if (I[t] < 0) {
                I[t] = 0;
                sign = -sign;
              } else if (I[t] >= g.N[t]) {
                I[t] = g.N[t] - 1;
                sign = -sign;
              }
            }
            val += nw[a] * tw[0][b] * tw[1][c] * sign * coarse_val(g, ec, I[0], I[1], I[2], d);
          }
      }
    }
    const long o = g.fine_index(F[0], F[1], F[2], d);
    zf[o] = (V)((double)zf[o] + val);
  }
}

__global__ __launch_bounds__(256) void k_restrict_state(Grid3 g, const double* __restrict__ xf,
                                                        double* __restrict__ xc) {
  const long cell = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long ncell = (long)g.N[0] * g.N[1] * g.N[2];
  if (cell >= ncell) return;
  int C[3];
  C[0] = (int)(cell % g.N[0]);
  C[1] = (int)((cell / g.N[0]) % g.N[1]);
  C[2] = (int)(cell / ((long)g.N[0] * g.N[1]));
  const int f[3] = {fine_count(g, 0), fine_count(g, 1), fine_count(g, 2)};
  for (int d = 0; d < g.dof; ++d) {
    double s = 0.0;
    int cnt = 0;
    if (d >= g.dim) {
      for (int c = 0; c < f[2]; ++c)
        for (int b = 0; b < f[1]; ++b)
          for (int a = 0; a < f[0]; ++a, ++cnt)
            s += xf[g.fine_index(f[0] * C[0] + a, f[1] * C[1] + b, f[2] * C[2] + c, d)];
    } else {
      const int ax = d, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
      for (int u = 0; u < f[t2]; ++u)
        for (int w = 0; w < f[t1]; ++w, ++cnt) {
          int F[3];
          F[ax] = 2 * C[ax] + 1;
          F[t1] = f[t1] * C[t1] + w;
          F[t2] = f[t2] * C[t2] + u;
          s += xf[g.fine_index(F[0], F[1], F[2], d)];
        }
    }
    xc[g.coarse_index(C[0], C[1], C[2], d)] = s / cnt;
  }
}

template <typename V>
__global__ __launch_bounds__(256) void k_sub(long n, const V* __restrict__ a, V* __restrict__ b) {
  // b = a - b
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i] - b[i];
}

template <typename V>
__global__ __launch_bounds__(256) void k_add(long n, const V* __restrict__ a, V* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] += a[i];
}

int grid_for(long n) {
  long g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

bool make_grid(int nx, int ny, int nz, int dim, int dof, Grid3* g) {
  const int n[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    g->n[a] = n[a];
    const bool coarsen = a < dim;
    if (coarsen && (n[a] < 2 || n[a] % 2)) return false;
    g->N[a] = coarsen ? n[a] / 2 : n[a];
  }
  if (dim == 2 && nz != 1) return false;
  g->dim = dim;
  g->dof = dof;
  return true;
}

}  // namespace

struct tf_mg {
  Grid3 g;
  int n, nc, single, gauge_f, gauge_c;
  size_t es;
  tf_lineilu* smoother;
  tf_bcr* coarse;
  const int* rp;
  const int* ci;
  const void* val;
  void *t, *w, *rc, *ec;
};

namespace {

template <typename V>
int spmv(tf_mg* h, const V* x, V* y, hipStream_t st);
template <>
int spmv<double>(tf_mg* h, const double* x, double* y, hipStream_t st) {
  return tf_spmv_f64(h->n, h->rp, h->ci, (const double*)h->val, x, y, h->gauge_f, st);
}
template <>
int spmv<float>(tf_mg* h, const float* x, float* y, hipStream_t st) {
  return tf_spmv_f32(h->n, h->rp, h->ci, (const float*)h->val, x, y, h->gauge_f, st);
}

int smooth(tf_mg* h, const double* r, double* z, hipStream_t st) { return tf_lineilu_solve_f64(h->smoother, r, z, st); }
int smooth(tf_mg* h, const float* r, float* z, hipStream_t st) { return tf_lineilu_solve_f32(h->smoother, r, z, st); }
int coarse_solve(tf_mg* h, const double* r, double* z, hipStream_t st) { return tf_bcr_solve_f64(h->coarse, r, z, st); }
int coarse_solve(tf_mg* h, const float* r, float* z, hipStream_t st) { return tf_bcr_solve_f32(h->coarse, r, z, st); }

template <typename V>
int vcycle(tf_mg* h, const V* r, V* z, hipStream_t st) {
  V* t = (V*)h->t;
  V* w = (V*)h->w;
  V* rc = (V*)h->rc;
  V* ec = (V*)h->ec;
  const long n = h->n;
  if (int e = smooth(h, r, z, st)) return e;                  // z = S r
  if (int e = spmv<V>(h, z, t, st)) return e;                 // t = r - A z
  hipLaunchKernelGGL(k_sub<V>, dim3(grid_for(n)), dim3(256), 0, st, n, r, t);
  hipLaunchKernelGGL(k_restrict<V>, dim3(grid_for((long)h->g.N[0] * h->g.N[1] * h->g.N[2])), dim3(256), 0, st, h->g,
                     h->gauge_c, (const V*)t, rc);
  TF_LAUNCH_CHECK("two-grid restriction");
  if (int e = coarse_solve(h, rc, ec, st)) return e;          // e_c = B R t
  hipLaunchKernelGGL(k_prolong_add<V>, dim3(grid_for((long)h->g.n[0] * h->g.n[1] * h->g.n[2])), dim3(256), 0, st,
                     h->g, (const V*)ec, z);
  TF_LAUNCH_CHECK("two-grid prolongation");
  if (int e = spmv<V>(h, z, t, st)) return e;                 // t = r - A z
  hipLaunchKernelGGL(k_sub<V>, dim3(grid_for(n)), dim3(256), 0, st, n, r, t);
  if (int e = smooth(h, t, w, st)) return e;                  // z += S t
  hipLaunchKernelGGL(k_add<V>, dim3(grid_for(n)), dim3(256), 0, st, n, (const V*)w, z);
  TF_LAUNCH_CHECK("two-grid update");
  return 0;
}

}  // namespace

extern "C" {

int tf_mg_restrict_state(int nx, int ny, int nz, int dim, int dof, const double* fine, double* coarse, void* stream) {
  Grid3 g;
  if (!make_grid(nx, ny, nz, dim, dof, &g)) return tf::fail("tf_mg_restrict_state: the grid does not halve");
  const long nc = (long)g.N[0] * g.N[1] * g.N[2];
  hipLaunchKernelGGL(k_restrict_state, dim3(grid_for(nc)), dim3(256), 0, (hipStream_t)stream, g, fine, coarse);
  TF_LAUNCH_CHECK("k_restrict_state");
  return 0;
}

int tf_mg_create(tf_mg** out, int nx, int ny, int nz, int dim, int dof, int single) {
  if (!out) return tf::fail("null handle");
  *out = nullptr;
  Grid3 g;
  if (!make_grid(nx, ny, nz, dim, dof, &g)) return tf::fail("tf_mg_create: every coarsened direction needs an even size");
  tf_mg* h = new tf_mg();
  h->g = g;
  h->n = nx * ny * nz * dof;
  h->nc = g.N[0] * g.N[1] * g.N[2] * dof;
  h->single = single;
  h->es = single ? sizeof(float) : sizeof(double);
  auto alloc = [&]() -> int {
    TF_HIP(hipMalloc(&h->t, h->es * (size_t)h->n));
    TF_HIP(hipMalloc(&h->w, h->es * (size_t)h->n));
    TF_HIP(hipMalloc(&h->rc, h->es * (size_t)h->nc));
    TF_HIP(hipMalloc(&h->ec, h->es * (size_t)h->nc));
    return 0;
  };
  if (int e = alloc()) {
    tf_mg_destroy(h);
    return e;
  }
  *out = h;
  return 0;
}

int tf_mg_destroy(tf_mg* h) {
  if (!h) return 0;
  for (void* p : {h->t, h->w, h->rc, h->ec}) (void)hipFree(p);
  delete h;
  return 0;
}

int tf_mg_set(tf_mg* h, tf_lineilu* smoother, const int* row_ptr, const int* col, const void* val, int gauge_fine,
              tf_bcr* coarse, int gauge_coarse) {
  if (!h || !smoother || !coarse || !row_ptr || !col || !val) return tf::fail("tf_mg_set: null argument");
  if (gauge_fine >= h->n || gauge_coarse >= h->nc) return tf::fail("tf_mg_set: gauge row outside the system");
  h->smoother = smoother;
  h->rp = row_ptr;
  h->ci = col;
  h->val = val;
  h->gauge_f = gauge_fine;
  h->coarse = coarse;
  h->gauge_c = gauge_coarse;
  return 0;
}

int tf_mg_solve_f64(tf_mg* h, const double* r, double* z, void* stream) {
  if (!h || !h->smoother) return tf::fail("tf_mg_solve_f64: preconditioner not set up");
  if (h->single) return tf::fail("tf_mg_solve_f64 on a single-precision preconditioner");
  if (r == z) return tf::fail("tf_mg_solve_f64: r and z must differ");
  return vcycle<double>(h, r, z, (hipStream_t)stream);
}

int tf_mg_solve_f32(tf_mg* h, const float* r, float* z, void* stream) {
  if (!h || !h->smoother) return tf::fail("tf_mg_solve_f32: preconditioner not set up");
  if (!h->single) return tf::fail("tf_mg_solve_f32 on a double-precision preconditioner");
  if (r == z) return tf::fail("tf_mg_solve_f32: r and z must differ");
  return vcycle<float>(h, r, z, (hipStream_t)stream);
}

}  // extern "C"
