/*
 * C ABI of libtransiflow_hip.so -- the MI355X (gfx950) backend of TransiFlow's
 * continuation hot path.
 *
 * The reference has no FFI: its backend boundary is the Python class
 * transiflow/interface/SciPy.py (contract: transiflow/interface/BaseInterface.py).
 * transiflow_amd/interface.py presents that Python contract and calls the
 * functions below through ctypes.  Each entry point names the reference
 * function it replaces.
 *
 * Conventions: every function returns 0 on success and a negative value on
 * failure (the message is then available from tf_last_error()).  All array
 * arguments are device pointers (HBM) unless named *_host; ``stream`` is a
 * hipStream_t passed as void*.  Row/column indices are int32, values FP64
 * (FP32 for the *_f32 solver variants).  No function allocates or frees
 * caller memory.
 */
#ifndef TRANSIFLOW_HIP_H
#define TRANSIFLOW_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tf_problem tf_problem;

/* Message of the last failed call on this thread. */
const char* tf_last_error(void);

/* Devices visible to the library's HIP runtime (-1 on error).  No reference
 * counterpart; lets the host layer fail loudly without a GPU. */
int tf_device_count(void);

/* ---- problem (replaces Discretization.__init__, Discretization.py:106-143) ---- */

/* Grid nx*ny*nz with dim (2|3) and dof (dim+1..5); x, y, z are host
 * coordinate vectors of length n+3 (utils.create_uniform_coordinate_vector,
 * utils.py:175-197); *_periodic as Discretization.py:123-128. */
int tf_problem_create(tf_problem** out, int nx, int ny, int nz, int dim, int dof,
                      const double* x_host, const double* y_host, const double* z_host,
                      int x_periodic, int y_periodic, int z_periodic);
int tf_problem_destroy(tf_problem* p);
int tf_problem_rows(const tf_problem* p);

/* Linear-part scalars (Discretization._linear_part_2D/3D, Discretization.py:229-317):
 * lin_scale = 1/(Re sqrt(Gr)), T_scale = 1/(Pr sqrt(Gr)), S_scale = 1/(Le Pr sqrt(Gr)),
 * rbp_coef = Bi/(Bi+1) when rbp; nonlinear = 0 when Re == 0 and dof <= dim+1
 * (Discretization.py:333-335). */
int tf_problem_set_physics(tf_problem* p, double lin_scale, double T_scale, double S_scale,
                           double rbp_coef, int rbp, int nonlinear);

/* Boundary program: the BoundaryConditions primitives in call order
 * (Discretization.boundaries, Discretization.py:704-714).  kind 0 = wall
 * (no_slip sign -1 / free_slip sign +1, BoundaryConditions.py:55-233),
 * kind 1 = _constant_forcing on variable var with constants fc, ac
 * (BoundaryConditions.py:471-544).  side +1 = east/north/top.  Host arrays. */
int tf_problem_set_bcs(tf_problem* p, int nops, const int* kind_host, const int* axis_host,
                       const int* side_host, const int* var_host, const double* sign_host,
                       const double* fc_host, const double* ac_host);

/* ---- assembly ---- */

/* F(state) (Discretization.rhs, Discretization.py:367-390; SciPy.Interface.rhs, SciPy.py:40-41). */
int tf_rhs(tf_problem* p, const double* state, double* F, void* stream);

/* Jacobian in CSR (Discretization.jacobian + assemble_jacobian + CrsMatrix.compress,
 * Discretization.py:392-415,474-524, CrsMatrix.py:52-71; SciPy.Interface.jacobian, SciPy.py:43-45).
 * count: row_ptr[0..n] is filled by the fill pass except row_ptr[n] = nnz,
 * which count writes; nnz_host (optional) receives nnz after a stream sync.
 * fill: writes row_ptr[0..n-1], col[0..nnz), val[0..nnz). */
int tf_jacobian_count(tf_problem* p, const double* state, int* row_ptr, int* nnz_host, void* stream);
int tf_jacobian_fill(tf_problem* p, const double* state, int* row_ptr, int* col, double* val,
                     void* stream);

/* Diagonal mass matrix (Discretization.mass_matrix, Discretization.py:417-437;
 * SciPy.Interface.mass_matrix, SciPy.py:47-49), same count/fill protocol. */
int tf_mass_count(tf_problem* p, int* row_ptr, int* nnz_host, void* stream);
int tf_mass_fill(tf_problem* p, int* row_ptr, int* col, double* val, void* stream);

/* ---- vectors (numpy arithmetic and utils.dot/norm, utils.py:5-21) ---- */

int tf_axpby_f64(long n, double alpha, const double* x, double beta, const double* y, double* out,
                 void* stream);
int tf_scale_f64(long n, double alpha, const double* x, double* out, void* stream);
int tf_div_f64(long n, const double* x, double a, double* out, void* stream);
int tf_f64_to_f32(long n, const double* x, float* out, void* stream);
int tf_f32_to_f64(long n, const float* x, double* out, void* stream);
int tf_reduce_work_size(void);
int tf_dot_f64(long n, const double* x, const double* y, double* work, double* out_dev,
               double* out_host, int norm, void* stream);

/* y = A x; gauge >= 0 applies the pressure gauge of SciPy.Interface.compute_bordered_matrix
 * (SciPy.py:95-103): column gauge removed, row gauge -> -x[gauge]. */
int tf_spmv_f64(int n, const int* row_ptr, const int* col, const double* val, const double* x,
                double* y, int gauge, void* stream);
int tf_spmv_f32(int n, const int* row_ptr, const int* col, const float* val, const float* x,
                float* y, int gauge, void* stream);
/* x[i] = v for 0 <= i < n (an error otherwise). */
int tf_set_entry_f64(double* x, long n, long i, double v, void* stream);

/* One round of the FP64 iterative refinement around an FP32 inner solve (the
 * mixed-precision variant of SciPy.Interface.solve, SciPy.py:343-370), before the
 * inner solve: r = b - A x (FP64, gauged as tf_spmv_f64), |r| into norm_dev (as
 * tf_dot_f64 with norm = 1, work of tf_reduce_work_size() doubles), r32 = (float) r
 * and d32 = 0.  The kernels of those calls one by one, in one call; nothing is
 * synchronised. */
int tf_refine_residual(int n, const int* row_ptr, const int* col, const double* val, int gauge, const double* x,
                       const double* b, double* r, float* r32, float* d32, double* work, double* norm_dev,
                       void* stream);
/* After the inner solve: x += (double) d32, through d = (double) d32 (tf_f32_to_f64
 * then tf_axpby_f64 in one call). */
int tf_refine_update(long n, const float* d32, double* d, double* x, void* stream);
/* count doubles from device memory to host memory; synchronises the stream. */
int tf_read_f64(const double* dev, int count, double* host, void* stream);

/* C = alpha A + beta B for CSR matrices (scipy sparse addition, as used by
 * TimeIntegration._newton, TimeIntegration.py:56: jacobian(x) - mass / (theta dt)).
 * A's entries keep their order; B's columns missing in A's row are appended. */
int tf_csr_add_count(int n, const int* ra, const int* ca, const int* rb, const int* cb, int* cnt, int* rc,
                     int* nnz_host, void* stream);
/* tf_csr_add_count without the cnt argument: the per-row counts go to a scratch array
 * the library keeps per device (16 MiB, n <= 4,194,304; larger n is an error) and the call
 * allocates nothing.  No counterpart in the reference: for callers that cannot allocate
 * (graph capture, allocation-free serving loops).  Calls that overlap on one device race. */
int tf_csr_add_count_noalloc(int n, const int* ra, const int* ca, const int* rb, const int* cb, int* rc,
                             int* nnz_host, void* stream);
int tf_csr_add_fill(int n, const int* ra, const int* ca, const double* va, double alpha, const int* rb,
                    const int* cb, const double* vb, double beta, const int* rc, int* cc, double* vc, void* stream);
/* Theta-method step residual, divided by theta dt (TimeIntegration._newton,
 * TimeIntegration.py:47-49: M (u - v) + dt theta F(v) + dt (1 - theta) F(u)):
 * out = Fv + cb Fu + cm mdiag .* (u - v) with cb = (1 - theta) / theta, cm = 1 / (theta dt)
 * and mdiag the diagonal of the mass matrix. */
int tf_theta_residual(long n, const double* Fv, const double* Fu, const double* mdiag, const double* u,
                      const double* v, double cb, double cm, double* out, void* stream);
/* Real interleaved form of the complex matrix A + iB (CSR, one row count n): unknown c becomes
 * (2c, 2c+1) = (re, im); row 2r holds (a, -b), row 2r+1 (b, a) over the union of the row's
 * columns; row and column ``gauge`` become -e in both halves (the pressure gauge of
 * SciPy.py:95-103 applied to a complex shifted matrix beta J - alpha M, as the reference's
 * complex splu solve of JaDa's shifted systems sees it, JaDa.py:146-151).  count: cnt (2n ints
 * of scratch), rc[0..2n] and nnz_host; fill: cc, vc. */
int tf_csr_interleave_count(int n, const int* ra, const int* ca, const int* rb, const int* cb, int gauge, int* cnt,
                            int* rc, int* nnz_host, void* stream);
int tf_csr_interleave_fill(int n, const int* ra, const int* ca, const double* va, const int* rb, const int* cb,
                           const double* vb, int gauge, const int* rc, int* cc, double* vc, void* stream);
/* diag[r] = A[r][r] of a CSR matrix, 0 for rows without a diagonal entry (the mass
 * matrix of Discretization.mass_matrix, Discretization.py:417-437, is diagonal). */
int tf_csr_diagonal(int n, const int* row_ptr, const int* col, const double* val, double* diag, void* stream);
int tf_fill_f64(long n, double v, double* x, void* stream);
/* FP32 vectors of the mixed-precision inner solve (no reference counterpart). */
int tf_fill_f32(long n, double v, float* x, void* stream);
int tf_axpby_f32(long n, double alpha, const float* x, double beta, const float* y, float* out, void* stream);
int tf_scale_f32(long n, double alpha, const float* x, float* out, void* stream);

/* ---- Krylov building blocks (scipy.sparse.linalg.gmres in SciPy.py:343-370) ----
 * V is a row-major basis with leading dimension ldv; h/coef are device FP64. */
int tf_krylov_work_size(void);
int tf_krylov_max_basis(void);
int tf_multidot_f64(int n, int k, const double* V, long ldv, const double* w, double* work, double* h,
                    int accumulate, void* stream);
int tf_multidot_f32(int n, int k, const float* V, long ldv, const float* w, double* work, double* h,
                    int accumulate, void* stream);
int tf_multiaxpy_f64(int n, int k, const double* V, long ldv, const double* coef, double* w, double sign,
                     void* stream);
int tf_multiaxpy_f32(int n, int k, const float* V, long ldv, const double* coef, float* w, double sign,
                     void* stream);
int tf_normalize_f64(int n, const double* w, double* work, double* nrm_dev, double* out, void* stream);
int tf_normalize_f32(int n, const float* w, double* work, double* nrm_dev, float* out, void* stream);

/* Native restarted FGMRES (the whole solve in one call; Givens rotations on the host).
 * single != 0: FP32 basis and operands.  Preconditioner kind 0 none, 1 tf_bcr, 2 tf_linelu,
 * 3 cell-block inverse (prec = the FP64 inverse blocks, ncell, dof), 4 tf_mg, 5 tf_lineilu.  The matrix values
 * have the handle's precision.  Replaces scipy.sparse.linalg.gmres (SciPy.py:343-370). */
typedef struct tf_krylov tf_krylov;
int tf_krylov_create(tf_krylov** out, int n, int restart, int single);
int tf_krylov_destroy(tf_krylov* k);
/* on != 0 (the default): each Arnoldi step (SpMV through a device matrix descriptor, two
 * CGS passes, normalisation, the read-back of the Hessenberg column) is captured once per
 * basis index into a hipGraph and replayed; 0: the same launches are issued one by one. */
int tf_krylov_set_graph(tf_krylov* k, int on);
int tf_fgmres(tf_krylov* k, const int* row_ptr, const int* col, const void* val, int gauge, int prec_kind,
              void* prec, int ncell, int dof, const void* b, void* x, double tol, int maxit, int x_is_zero,
              int* its_out, double* rel_out, void* stream);
/* Two systems with one matrix and preconditioner (A x0 = b0, A x1 = b1; the two solves
 * with one Jacobian of a Newton-corrector iteration, Continuation.py:158-159), each
 * with its own Krylov handle (same n, restart, precision) and tolerance, in lockstep:
 * while both are in a restart cycle one tf_bcr_solve2_* serves both.  its_out,
 * rel_out: two entries each. */
int tf_fgmres2(tf_krylov* k0, tf_krylov* k1, const int* row_ptr, const int* col, const void* val, int gauge,
               int prec_kind, void* prec, int ncell, int dof, const void* b0, const void* b1, void* x0, void* x1,
               double tol0, double tol1, int maxit, int x_is_zero, int* its_out, double* rel_out, void* stream);

/* Cell-block Jacobi preconditioner (dof x dof diagonal blocks of the gauged matrix). */
int tf_cellblock_setup(int ncell, int dof, const int* row_ptr, const int* col, const double* val,
                       int gauge, double* inv, void* stream);
int tf_cellblock_apply_f64(int ncell, int dof, const double* inv, const double* r, double* z, void* stream);
int tf_cellblock_apply_f32(int ncell, int dof, const double* inv, const float* r, float* z, void* stream);

/* ---- exact block-tridiagonal factorisation over grid lines ----
 * Replaces the sparse LU of SciPy._compute_factorization / _lu_solve (SciPy.py:131-162,194-202)
 * as the preconditioner of GMRES.  n = lines * m; single != 0 stores the factor in FP32. */
typedef struct tf_linelu tf_linelu;
int tf_linelu_create(tf_linelu** out, int n, int m, int single);
int tf_linelu_destroy(tf_linelu* h);
int tf_linelu_factor(tf_linelu* h, const int* row_ptr, const int* col, const double* val, int gauge,
                     void* stream);
int tf_linelu_status(tf_linelu* h, int* flags_host, void* stream);
int tf_linelu_solve_f64(tf_linelu* h, const double* r, double* x, void* stream);
int tf_linelu_solve_f32(tf_linelu* h, const float* r, float* x, void* stream);

/* ---- block cyclic reduction over grid lines (same role as tf_linelu, log-depth) ----
 * Odd-even reduction of the block-tridiagonal gauged Jacobian; every level is a batch of
 * dense inversions and GEMMs.  single != 0 factors and stores in FP32.  ylines != 0 (2D)
 * takes the grid columns as lines (m = ny*dof; nx and dof describe the grid) so that wide
 * domains factor short lines; vectors stay in the natural order. */
typedef struct tf_bcr tf_bcr;
int tf_bcr_create(tf_bcr** out, int n, int m, int single, int nx, int dof, int ylines);
int tf_bcr_destroy(tf_bcr* h);
int tf_bcr_factor(tf_bcr* h, const int* row_ptr, const int* col, const double* val, int gauge, void* stream);
int tf_bcr_status(tf_bcr* h, int* flags_host, void* stream);
int tf_bcr_solve_f64(tf_bcr* h, const double* r, double* x, void* stream);
int tf_bcr_solve_f32(tf_bcr* h, const float* r, float* x, void* stream);
/* Two right-hand sides in one pass over the factor blocks: x0 = M r0, x1 = M r1, each
 * bit-identical to tf_bcr_solve_*; for the two solves with one Jacobian of every
 * Newton-corrector iteration (Continuation.py:158-159: F_x z1 = -F, F_x z2 = F_mu). */
int tf_bcr_solve2_f64(tf_bcr* h, const double* r0, const double* r1, double* x0, double* x1, void* stream);
int tf_bcr_solve2_f32(tf_bcr* h, const float* r0, const float* r1, float* x0, float* x1, void* stream);
/* on != 0: the level sweeps of the apply and the scatter of its result are captured once
 * (per vector type and column count) into a hipGraph and replayed with one launch after the
 * gather kernel, which writes the call's output pointers into a device table the scatter
 * reads (no reference counterpart; launch overhead on small grids and batches). */
int tf_bcr_set_graph(tf_bcr* h, int on);
/* on != 0 would run the levels with fewer than 12288 rows per launch as one ordered-task
 * launch (no reference counterpart; the launch floor of the upper levels); it returns an
 * error: on the GPU that launch did not finish within the time limit and the cause is not
 * established.  0 (default) keeps one launch per level.
 * tf_bcr_top_errors: waits of that launch that gave up since the last call (then reset);
 * synchronises the stream. */
int tf_bcr_set_top(tf_bcr* h, int on);
/* on != 0 (default): levels with at most 8 lines get their line indices and block row
 * bases as kernel arguments instead of reading them from the schedule's device lists;
 * results equal bit for bit (no reference counterpart; latency of the upper levels). */
int tf_bcr_set_small(tf_bcr* h, int on);
int tf_bcr_top_errors(tf_bcr* h, int* count_host, void* stream);
/* on != 0 (single != 0 only): the FP32 apply (tf_bcr_solve_f32) reads FP16 copies of the
 * factor blocks with one power-of-two FP32 scale per row, remade after every factorisation:
 * half the bytes of the HBM-bound apply for the inner solves of the mixed-precision solver
 * (no reference counterpart; the reference's splu factor is FP64).  The first call
 * allocates the copies; an existing factor is converted on ``stream``. */
int tf_bcr_set_half(tf_bcr* h, int on, void* stream);
/* split 0 (default): the level GEMMs of a single-precision factor run on FP32 MFMA; 3: on BF16
 * MFMA over three-term BF16 splits of the FP32 operands (hi*hi + hi*lo + lo*hi, FP32
 * accumulation, about 2^-16 relative per product); 1: one term (no reference counterpart). */
int tf_bcr_set_gemm_split(tf_bcr* h, int split);
/* on != 0 (before the first factorisation): the factorisation also forms Y L and Y U of the
 * lines eliminated at levels >= 1 in place of their L and U, and each backward level of the apply
 * is one kernel (x_e = Y r_e - (Y L) x_a - (Y U) x_b) instead of two (no reference counterpart:
 * worth it for factors applied many times, such as the lagged FP32 factor of the mixed solver). */
int tf_bcr_set_fused(tf_bcr* h, int on);
/* lines in 2..4 (fused factors): once the reduction is down to at most that many lines, each
 * factorisation also forms the dense inverse Z of their reduced system (about 20 m x m x m
 * products at four lines; FP16 copies for single != 0), and each apply maps those lines'
 * right-hand sides to their solutions in one launch instead of the forward and backward levels
 * above them (five launches at four lines).  0 (default): off.  Allocates Z; takes effect at the
 * next factorisation (no reference counterpart: launch latency of the few-line levels). */
int tf_bcr_set_dense_top(tf_bcr* h, int lines);

/* ---- block ILU over grid x-lines, four-colour line order (two colours in 2D) ----
 * M = (P + L) P^-1 (P + U) with dense m x m pivots P (m = nx dof) and the Jacobian's own
 * sparse couplings between lines; all lines of one colour are factored and applied as one
 * batch.  single != 0: FP32 factor and FP32 vectors.  Replaces, as a preconditioner, the
 * incomplete LU of SciPy._compute_preconditioner (scipy spilu, SciPy.py:164-192). */
typedef struct tf_lineilu tf_lineilu;
int tf_lineilu_create(tf_lineilu** out, int nx, int ny, int nz, int dof, int single);
int tf_lineilu_destroy(tf_lineilu* h);
int tf_lineilu_factor(tf_lineilu* h, const int* row_ptr, const int* col, const double* val, int gauge, void* stream);
/* flags[0]: couplings beyond the lists (structure error); flags[1]: perturbed pivots.  Synchronises. */
int tf_lineilu_status(tf_lineilu* h, int* flags_host, void* stream);
int tf_lineilu_solve_f64(tf_lineilu* h, const double* r, double* x, void* stream);
int tf_lineilu_solve_f32(tf_lineilu* h, const float* r, float* x, void* stream);

/* ---- two-grid preconditioner (3D cavity grids whose exact factor does not fit) ----
 * V(1,1) cycle: line block ILU smoothing on the grid, exact solve (tf_bcr) of the Jacobian
 * re-assembled on the grid coarsened by two in every direction at the restricted state.
 * Replaces, as the preconditioner of GMRES, scipy's spilu (SciPy.py:164-192,260-309).
 * tf_mg_restrict_state: the coarse state (scalars: mean of the fine cells; velocities: mean
 * of the coincident fine faces).  tf_mg_set: the smoother, the fine gauged Jacobian (values
 * in the handle's precision) and the coarse factor the cycle uses (pointers are kept). */
typedef struct tf_mg tf_mg;
int tf_mg_restrict_state(int nx, int ny, int nz, int dim, int dof, const double* fine, double* coarse, void* stream);
int tf_mg_create(tf_mg** out, int nx, int ny, int nz, int dim, int dof, int single);
int tf_mg_destroy(tf_mg* h);
int tf_mg_set(tf_mg* h, tf_lineilu* smoother, const int* row_ptr, const int* col, const void* val, int gauge_fine,
              tf_bcr* coarse, int gauge_coarse);
int tf_mg_solve_f64(tf_mg* h, const double* r, double* z, void* stream);
int tf_mg_solve_f32(tf_mg* h, const float* r, float* z, void* stream);

/* ---- learned preconditioner: convolutions on the cell grid (no reference counterpart) ----
 * Activations cell-major [ncell][C] in BF16 (uint16 storage); layer weights in the
 * padded MFMA layout [kpad][npad] (K = taps*cin rows, row tap*cin + ci, taps = 3^dim,
 * tap = (dx+1) + 3(dy+1) [+ 9(dz+1) in 3D]); zero padding outside the grid; FP32 accumulation. */
int tf_conv_fwd(int nx, int ny, int nz, int dim, int cin, int cout, int kpad, int npad, const void* in,
                const void* W, const float* bias, int relu, float* out32, void* out16, void* stream);
int tf_conv_wgrad_work_size(long ncell, int kpad, int npad);
int tf_conv_wgrad(int nx, int ny, int nz, int dim, int cin, int cout, int kpad, int npad, const void* in,
                  const void* gout16, const float* g32, float* work, float* dW, float* db, void* stream);
int tf_mse_grad(long n, const float* y, const float* t, float* g, float* loss_partial, void* stream);
int tf_relu_bwd(long n, const float* gin, const void* act16, int relu, float* g32, void* g16, void* stream);
int tf_conv_pack(int dim, int cin, int cout, int kpad, int npad, int kpad_t, int npad_t, const float* W, void* Wf,
                 void* Wt, void* stream);
int tf_adam(long n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps, int step,
            void* stream);
int tf_to_bf16(long n, const float* x, float scale, void* y, void* stream);

/* ---- tile autotuner of the batched MFMA GEMM (no reference counterpart) ----
 * mode 1 (default): the first GEMM of each (type, M, N, K, batch) times the tile
 * candidates and keeps the fastest; mode 0: fixed rule.  tf_gemm_table copies rows
 * (single, M, N, K, batch, tile) and returns the row count. */
int tf_gemm_tuning(int mode);
int tf_gemm_table(int* out_host, int max_rows);

/* ---- opt-in kernel timing with HIP events (no reference counterpart; used by bench.py) ---- */
int tf_prof_enable(int on);
/* Graph captures (line-factor applies, Arnoldi steps) that failed -- work of another
 * thread can invalidate a capture -- and whose call ran eagerly with the same kernels
 * instead, since the process started (no reference counterpart). */
int tf_capture_fallbacks(void);
int tf_prof_read(int id, long* count, double* ms_total);
int tf_prof_reset(void);
int tf_prof_work(int id, double* work_total);

#ifdef __cplusplus
}
#endif

#endif /* TRANSIFLOW_HIP_H */
