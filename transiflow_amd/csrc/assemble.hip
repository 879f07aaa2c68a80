// FP64 stencil assembly on gfx950: F(x), the Jacobian in CSR and the mass matrix.
//
// Launch layout: one workgroup = 64 consecutive cells x DOF wavefronts; wave w
// builds equation d1 = w of its 64 cells, so d1 is uniform per wave and the
// row atom is specialised at compile time (see stencil.h).  The workgroup's
// rows form the contiguous range [64*b*DOF, 64*(b+1)*DOF), which lets the CSR
// count/scan/fill passes use a workgroup-local prefix sum plus one scan of
// the per-workgroup totals:
//
//   k_assemble<COUNT>  per-row entry counts + per-workgroup totals
//   k_scan_blocks      exclusive scan of the totals, nnz -> row_ptr[n]
//   k_assemble<FILL>   row_ptr, col_idx, values (atoms recomputed, no atom in HBM)
//
// Reference: Discretization.rhs/jacobian/mass_matrix and assemble_* (Discretization.py:367-548),
// CrsMatrix.compress (CrsMatrix.py:52-71).
#include "stencil.h"
#include "tf_common.h"
#include "transiflow_hip.h"

using namespace tf;

enum : int { MODE_RHS = 0, MODE_COUNT = 1, MODE_FILL = 2 };

struct tf_problem {
  int n[3];
  int periodic[3];
  int dim, dof, buoy;
  int ncell, nrow, nblocks;
  double* coord_d[3];
  BcOp* ops_d;
  Physics ph;
  int* counts_d;      // nrow
  int* block_sums_d;  // nblocks
  int* block_off_d;   // nblocks
  int* nnz_pinned;    // host-visible nnz
};

namespace {

template <int DOF>
__device__ __forceinline__ Cell<DOF> make_cell(const Grid& g, const double* state, int cell) {
  Cell<DOF> c;
  c.n[0] = g.n[0];
  c.n[1] = g.n[1];
  c.n[2] = g.n[2];
  c.periodic[0] = g.periodic[0];
  c.periodic[1] = g.periodic[1];
  c.periodic[2] = g.periodic[2];
  c.coord[0] = g.coord[0];
  c.coord[1] = g.coord[1];
  c.coord[2] = g.coord[2];
  c.state = state;
  c.idx[0] = cell % g.n[0];
  const int rest = cell / g.n[0];
  c.idx[1] = rest % g.n[1];
  c.idx[2] = rest / g.n[1];
  return c;
}

template <int DIM, int DOF, int BUOY, int D1, int MODE>
__device__ __forceinline__ int row_work(const Cell<DOF>& c, const Physics& ph, int row, double* __restrict__ F,
                                        int* __restrict__ col, double* __restrict__ val, int offset) {
  double A[DOF * 27];
  double frc;
  if constexpr (MODE == MODE_RHS) {
    build_row<DIM, DOF, BUOY, D1, ATOM_F>(c, ph, A, frc);
    F[row] = apply_row<DOF>(c, A, frc);
    return 0;
  } else {
    build_row<DIM, DOF, BUOY, D1, ATOM_J>(c, ph, A, frc);
    if constexpr (MODE == MODE_COUNT)
      return emit_row<DOF, false>(c, A, nullptr, nullptr);
    else
      return emit_row<DOF, true>(c, A, col + offset, val + offset);
  }
}

template <int DIM, int DOF, int BUOY, int MODE>
__device__ __forceinline__ int dispatch_row(int d1, const Cell<DOF>& c, const Physics& ph, int row, double* F,
                                            int* col, double* val, int offset) {
  switch (d1) {
    case 0: return row_work<DIM, DOF, BUOY, 0, MODE>(c, ph, row, F, col, val, offset);
    case 1: return row_work<DIM, DOF, BUOY, 1, MODE>(c, ph, row, F, col, val, offset);
    case 2: return row_work<DIM, DOF, BUOY, 2, MODE>(c, ph, row, F, col, val, offset);
    case 3: if constexpr (DOF > 3) return row_work<DIM, DOF, BUOY, 3, MODE>(c, ph, row, F, col, val, offset); break;
    case 4: if constexpr (DOF > 4) return row_work<DIM, DOF, BUOY, 4, MODE>(c, ph, row, F, col, val, offset); break;
    default: break;
  }
  return 0;
}

// Inclusive scan of s[0..N) in LDS by all N threads of the block.
template <int N>
__device__ __forceinline__ void block_inclusive_scan(int* s, int t) {
#pragma unroll
  for (int off = 1; off < N; off <<= 1) {
    int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
}

template <int DIM, int DOF, int BUOY, int MODE>
__global__ __launch_bounds__(64 * DOF) void k_assemble(Grid g, Physics ph, const double* __restrict__ state,
                                                       double* __restrict__ F, int* __restrict__ counts,
                                                       int* __restrict__ block_sums, const int* __restrict__ block_off,
                                                       int* __restrict__ row_ptr, int* __restrict__ col,
                                                       double* __restrict__ val, int ncell) {
  constexpr int NT = 64 * DOF;
  __shared__ int s[NT];
  const int t = threadIdx.x;
  const int d1 = __builtin_amdgcn_readfirstlane(t >> 6);
  const int lane = t & 63;
  const int cell = blockIdx.x * 64 + lane;
  const bool active = cell < ncell;
  const int row = cell * DOF + d1;
  const int lr = lane * DOF + d1;  // row index local to the workgroup
  const Cell<DOF> c = make_cell<DOF>(g, state, active ? cell : 0);

  if constexpr (MODE == MODE_RHS) {
    if (active) dispatch_row<DIM, DOF, BUOY, MODE>(d1, c, ph, row, F, nullptr, nullptr, 0);
  } else if constexpr (MODE == MODE_COUNT) {
    int cnt = active ? dispatch_row<DIM, DOF, BUOY, MODE>(d1, c, ph, row, nullptr, nullptr, nullptr, 0) : 0;
    if (active) counts[row] = cnt;
    s[lr] = cnt;
    __syncthreads();
    block_inclusive_scan<NT>(s, t);
    if (t == NT - 1) block_sums[blockIdx.x] = s[NT - 1];
  } else {
    // The workgroup's rows are contiguous in the CSR arrays: when its entries fit,
    // the rows are emitted into LDS and the segment is written out coalesced.  The
    // capacity covers the mean row length with margin (2D about 8.7 entries per row,
    // 3D about 14): with 12 per row the 3D segments overflowed and the rows were
    // written in place, uncoalesced (1.7x the bytes written at 64^3).
    constexpr int CAP = NT * (DIM == 3 ? 16 : 12);
    __shared__ int s_col[CAP];
    __shared__ double s_val[CAP];
    s[lr] = active ? counts[row] : 0;
    __syncthreads();
    block_inclusive_scan<NT>(s, t);
    const int base = block_off[blockIdx.x];
    const int total = s[NT - 1];
    const int loc = lr > 0 ? s[lr - 1] : 0;
    if (active) row_ptr[row] = base + loc;
    if (total <= CAP) {
      if (active) dispatch_row<DIM, DOF, BUOY, MODE>(d1, c, ph, row, nullptr, s_col, s_val, loc);
      __syncthreads();
      for (int e = t; e < total; e += NT) {
        col[base + e] = s_col[e];
        val[base + e] = s_val[e];
      }
    } else if (active) {
      dispatch_row<DIM, DOF, BUOY, MODE>(d1, c, ph, row, nullptr, col, val, base + loc);
    }
  }
}

// Exclusive scan of the per-workgroup totals by one workgroup; writes the
// grand total to row_ptr[nrow] and to host-visible memory.
__global__ __launch_bounds__(1024) void k_scan_blocks(const int* __restrict__ sums, int* __restrict__ off, int nb,
                                                      int* __restrict__ row_ptr_end, int* __restrict__ total_host) {
  __shared__ int s[1024];
  const int t = threadIdx.x;
  const int chunk = (nb + 1023) / 1024;
  const int b0 = min(t * chunk, nb), b1 = min(b0 + chunk, nb);
  int acc = 0;
  for (int b = b0; b < b1; ++b) acc += sums[b];
  s[t] = acc;
  __syncthreads();
  block_inclusive_scan<1024>(s, t);
  int run = t > 0 ? s[t - 1] : 0;
  for (int b = b0; b < b1; ++b) {
    off[b] = run;
    run += sums[b];
  }
  if (t == 1023) {
    *row_ptr_end = s[1023];
    if (total_host) *total_host = s[1023];
  }
}

// Mass matrix (Discretization.py:417-437,1120-1176): diagonal cell volumes of
// u, v, w (staggered) and T, S (centred); pressure rows are empty.
template <int DOF>
__device__ __forceinline__ double mass_value(const Cell<DOF>& c, int d, int dim) {
  if (d == 0) return c.hdiff(0, -1, 1) * c.diff(1, -1, 0) * c.diff(2, -1, 0);
  if (d == 1) return c.hdiff(1, -1, 1) * c.diff(0, -1, 0) * c.diff(2, -1, 0);
  if (d == 2 && dim == 3) return c.hdiff(2, -1, 1) * c.diff(1, -1, 0) * c.diff(0, -1, 0);
  if (d == dim + 1 || d == dim + 2) return c.diff(0, -1, 0) * c.diff(1, -1, 0) * c.diff(2, -1, 0);
  return 0.0;
}

template <int DOF, int MODE>
__global__ __launch_bounds__(64 * DOF) void k_mass(Grid g, int dim, int* __restrict__ counts,
                                                   int* __restrict__ block_sums, const int* __restrict__ block_off,
                                                   int* __restrict__ row_ptr, int* __restrict__ col,
                                                   double* __restrict__ val, int ncell) {
  constexpr int NT = 64 * DOF;
  __shared__ int s[NT];
  const int t = threadIdx.x;
  const int d1 = t >> 6;
  const int lane = t & 63;
  const int cell = blockIdx.x * 64 + lane;
  const bool active = cell < ncell;
  const int row = cell * DOF + d1;
  const int lr = lane * DOF + d1;
  const Cell<DOF> c = make_cell<DOF>(g, nullptr, active ? cell : 0);
  const double v = mass_value<DOF>(c, d1, dim);
  const int cnt = (active && fabs(v) > DROP) ? 1 : 0;
  if constexpr (MODE == MODE_COUNT) {
    if (active) counts[row] = cnt;
    s[lr] = cnt;
    __syncthreads();
    block_inclusive_scan<NT>(s, t);
    if (t == NT - 1) block_sums[blockIdx.x] = s[NT - 1];
  } else {
    s[lr] = cnt;
    __syncthreads();
    block_inclusive_scan<NT>(s, t);
    const int off = block_off[blockIdx.x] + (lr > 0 ? s[lr - 1] : 0);
    if (active) {
      row_ptr[row] = off;
      if (cnt) {
        col[off] = row;
        val[off] = v;
      }
    }
  }
}

Grid grid_of(const tf_problem* p) {
  Grid g;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = p->n[a];
    g.periodic[a] = p->periodic[a];
    g.coord[a] = p->coord_d[a];
  }
  return g;
}

template <int MODE>
int launch_assemble(tf_problem* p, const double* state, double* F, int* row_ptr, int* col, double* val,
                    hipStream_t st) {
  const Grid g = grid_of(p);
  const dim3 grid(p->nblocks), block(64 * p->dof);
#define TF_ASM(DIM_, DOF_, BUOY_)                                                                              \
  hipLaunchKernelGGL((k_assemble<DIM_, DOF_, BUOY_, MODE>), grid, block, 0, st, g, p->ph, state, F, p->counts_d, \
                     p->block_sums_d, p->block_off_d, row_ptr, col, val, p->ncell)
  const int key = p->dim * 100 + p->dof * 10 + p->buoy;
  const int pid = MODE == MODE_RHS ? PROF_RHS : MODE == MODE_COUNT ? PROF_JAC_COUNT : PROF_JAC_FILL;
  TF_PROF_BEGIN(pid, st);
  switch (key) {
    case 231: TF_ASM(2, 3, 1); break;
    case 241: TF_ASM(2, 4, 1); break;
    case 251: TF_ASM(2, 5, 1); break;
    case 341: TF_ASM(3, 4, 1); break;
    case 342: TF_ASM(3, 4, 2); break;
    case 351: TF_ASM(3, 5, 1); break;
    case 352: TF_ASM(3, 5, 2); break;
    default: return fail("unsupported (dim, dof) combination for the HIP assembly");
  }
#undef TF_ASM
  TF_PROF_END(pid, st);
  TF_LAUNCH_CHECK("k_assemble");
  return 0;
}

int scan_blocks(tf_problem* p, int* row_ptr, hipStream_t st) {
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, p->block_sums_d, p->block_off_d, p->nblocks,
                     row_ptr + p->nrow, p->nnz_pinned);
  TF_LAUNCH_CHECK("k_scan_blocks");
  return 0;
}

}  // namespace

// Device buffers of a problem; on failure the caller frees the partial set
// (every pointer starts null, and hipFree/hipHostFree accept null).
static int problem_alloc(tf_problem* p, const double* const host[3]) {
  for (int a = 0; a < 3; ++a) {
    const size_t bytes = sizeof(double) * (p->n[a] + 3);
    TF_HIP(hipMalloc(&p->coord_d[a], bytes));
    TF_HIP(tf::upload(p->coord_d[a], host[a], bytes));
  }
  TF_HIP(hipMalloc(&p->ops_d, sizeof(BcOp) * MAX_BC_OPS));
  TF_HIP(hipMalloc(&p->counts_d, sizeof(int) * p->nrow));
  TF_HIP(hipMalloc(&p->block_sums_d, sizeof(int) * p->nblocks));
  TF_HIP(hipMalloc(&p->block_off_d, sizeof(int) * p->nblocks));
  TF_HIP(hipHostMalloc(&p->nnz_pinned, sizeof(int), hipHostMallocDefault));
  return 0;
}

extern "C" {

const char* tf_last_error(void) { return tf::last_error().c_str(); }

// Number of visible devices (0 when the runtime sees none; -1 on error).
int tf_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return -1;
  return n;
}

int tf_problem_create(tf_problem** out, int nx, int ny, int nz, int dim, int dof, const double* x, const double* y,
                      const double* z, int x_periodic, int y_periodic, int z_periodic) {
  if (!out) return fail("null output handle");
  *out = nullptr;
  if (nx < 1 || ny < 1 || nz < 1) return fail("grid sizes must be positive");
  if (dim != 2 && dim != 3) return fail("dim must be 2 or 3");
  if (dof < dim + 1 || dof > 5) return fail("dof must be in [dim+1, 5]");
  if ((long long)nx * ny * nz * dof >= (1LL << 31)) return fail("problem too large for int32 indices");
  tf_problem* p = new tf_problem();
  p->n[0] = nx;
  p->n[1] = ny;
  p->n[2] = nz;
  p->periodic[0] = x_periodic;
  p->periodic[1] = y_periodic;
  p->periodic[2] = z_periodic;
  p->dim = dim;
  p->dof = dof;
  p->buoy = (dim == 3 && nz > 1) ? 2 : 1;
  p->ncell = nx * ny * nz;
  p->nrow = p->ncell * dof;
  p->nblocks = (p->ncell + 63) / 64;
  const double* host[3] = {x, y, z};
  if (int e = problem_alloc(p, host)) {
    tf_problem_destroy(p);  // frees what was allocated before the failure
    return e;
  }
  p->ph = Physics{};
  p->ph.ops = p->ops_d;
  *out = p;
  return 0;
}

int tf_problem_destroy(tf_problem* p) {
  if (!p) return 0;
  for (int a = 0; a < 3; ++a) (void)hipFree(p->coord_d[a]);
  (void)hipFree(p->ops_d);
  (void)hipFree(p->counts_d);
  (void)hipFree(p->block_sums_d);
  (void)hipFree(p->block_off_d);
  (void)hipHostFree(p->nnz_pinned);
  delete p;
  return 0;
}

int tf_problem_set_physics(tf_problem* p, double lin_scale, double T_scale, double S_scale, double rbp_coef, int rbp,
                           int nonlinear) {
  if (!p) return fail("null problem");
  p->ph.lin_scale = lin_scale;
  p->ph.T_scale = T_scale;
  p->ph.S_scale = S_scale;
  p->ph.rbp_coef = rbp_coef;
  p->ph.rbp = rbp;
  p->ph.nonlinear = nonlinear;
  return 0;
}

int tf_problem_set_bcs(tf_problem* p, int nops, const int* kind, const int* axis, const int* side, const int* var,
                       const double* sign, const double* fc, const double* ac) {
  if (!p) return fail("null problem");
  if (nops < 0 || nops > MAX_BC_OPS) return fail("too many boundary-condition operations");
  BcOp ops[MAX_BC_OPS];
  for (int o = 0; o < nops; ++o) {
    if (kind[o] != BC_WALL && kind[o] != BC_FORCE) return fail("unknown boundary-condition kind");
    if (axis[o] < 0 || axis[o] > 2 || (side[o] != 1 && side[o] != -1)) return fail("bad boundary face");
    if (kind[o] == BC_FORCE && (var[o] < 0 || var[o] >= p->dof)) return fail("bad forcing variable");
    ops[o] = BcOp{kind[o], axis[o], side[o], var[o], sign[o], fc[o], ac[o]};
  }
  if (nops) TF_HIP(tf::upload(p->ops_d, ops, sizeof(BcOp) * nops));
  p->ph.nops = nops;
  return 0;
}

int tf_problem_rows(const tf_problem* p) { return p ? p->nrow : -1; }

int tf_rhs(tf_problem* p, const double* state, double* F, void* stream) {
  if (!p || !state || !F) return fail("null argument");
  return launch_assemble<MODE_RHS>(p, state, F, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int tf_jacobian_count(tf_problem* p, const double* state, int* row_ptr, int* nnz_host, void* stream) {
  if (!p || !state || !row_ptr) return fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = launch_assemble<MODE_COUNT>(p, state, nullptr, row_ptr, nullptr, nullptr, st);
  if (rc) return rc;
  rc = scan_blocks(p, row_ptr, st);
  if (rc) return rc;
  if (nnz_host) {
    TF_HIP(hipStreamSynchronize(st));
    *nnz_host = *p->nnz_pinned;
  }
  return 0;
}

int tf_jacobian_fill(tf_problem* p, const double* state, int* row_ptr, int* col, double* val, void* stream) {
  if (!p || !state || !row_ptr || !col || !val) return fail("null argument");
  return launch_assemble<MODE_FILL>(p, state, nullptr, row_ptr, col, val, (hipStream_t)stream);
}

int tf_mass_count(tf_problem* p, int* row_ptr, int* nnz_host, void* stream) {
  if (!p || !row_ptr) return fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  const Grid g = grid_of(p);
  const dim3 grid(p->nblocks), block(64 * p->dof);
  switch (p->dof) {
    case 3: hipLaunchKernelGGL((k_mass<3, MODE_COUNT>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, nullptr, nullptr, p->ncell); break;
    case 4: hipLaunchKernelGGL((k_mass<4, MODE_COUNT>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, nullptr, nullptr, p->ncell); break;
    case 5: hipLaunchKernelGGL((k_mass<5, MODE_COUNT>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, nullptr, nullptr, p->ncell); break;
    default: return fail("unsupported dof");
  }
  TF_LAUNCH_CHECK("k_mass count");
  int rc = scan_blocks(p, row_ptr, st);
  if (rc) return rc;
  if (nnz_host) {
    TF_HIP(hipStreamSynchronize(st));
    *nnz_host = *p->nnz_pinned;
  }
  return 0;
}

int tf_mass_fill(tf_problem* p, int* row_ptr, int* col, double* val, void* stream) {
  if (!p || !row_ptr || !col || !val) return fail("null argument");
  hipStream_t st = (hipStream_t)stream;
  const Grid g = grid_of(p);
  const dim3 grid(p->nblocks), block(64 * p->dof);
  switch (p->dof) {
    case 3: hipLaunchKernelGGL((k_mass<3, MODE_FILL>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, col, val, p->ncell); break;
    case 4: hipLaunchKernelGGL((k_mass<4, MODE_FILL>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, col, val, p->ncell); break;
    case 5: hipLaunchKernelGGL((k_mass<5, MODE_FILL>), grid, block, 0, st, g, p->dim, p->counts_d, p->block_sums_d, p->block_off_d, row_ptr, col, val, p->ncell); break;
    default: return fail("unsupported dof");
  }
  TF_LAUNCH_CHECK("k_mass fill");
  return 0;
}

}  // extern "C"
