"""Linear solves of the HIP backend: preconditioned restarted (F)GMRES.

``Interface.solve(jac, rhs)`` solves the pressure-gauged system of the
reference's direct solve (``transiflow/interface/SciPy.py:204-258``): the
right-hand side's ``pressure_row`` entry is zeroed, row and column
``pressure_row`` of J are replaced by -e (SciPy.py:95-103), and the result
satisfies ``||b - A x|| <= tol ||b||`` with tol = 1e-10 by default (the
reference's splu solve is matched to that relative residual).

Options (``parameters['HIP Solver']``, all optional):

``'Preconditioner'``   ``'auto'`` (default: ``'line'`` when its dense blocks fit
                       the memory budget, else ``'multigrid'`` for 3D grids,
                       else a SolverError naming the sizes), ``'line'`` (exact
                       block-tridiagonal factorisation over grid lines -- 2D --
                       or planes -- 3D -- by block cyclic reduction),
                       ``'thomas'`` (the same factorisation by the sequential
                       block Thomas recursion), ``'multigrid'`` (two grids:
                       line block ILU smoothing, exact factor of the Jacobian
                       re-assembled on the grid coarsened by two),
                       ``'line-ilu'`` (the four-colour line block ILU alone),
                       ``'block-jacobi'`` (cell blocks; does not converge on
                       these saddle-point systems beyond small cases);
``'Precision'``        ``'fp64'`` (default) or ``'mixed'``: FP64 iterative
                       refinement (residual b - A x, update and convergence test in
                       FP64) around an FP32 inner FGMRES (FP32 basis, SpMV with FP32
                       values, FP32 preconditioner factor and apply); the FP32
                       factor is reused across Jacobians while the inner solve
                       converges quickly;
``'Inner Tolerance'``  (mixed) relative residual of each inner solve (5e-4: FP32
                       solves stall near 1e-5 on these Jacobians, so the FP64
                       refinement does the rest); a step whose remaining
                       reduction to ``'Tolerance'`` is at least 1e-5 asks for
                       half of it instead, and so finishes the solve;
``'Inner Iterations'`` (mixed) FGMRES iterations per inner solve (12);
``'Apply Graph'``      replay the line-factor apply as one captured hipGraph
                       (True);
``'Step Graphs'``      replay each native FGMRES Arnoldi step (SpMV, two CGS
                       passes, normalisation) as one captured hipGraph (True)
                       or launch its kernels one by one;
``'Pair Solves'``      (True) ``solve_pair`` -- the two solves of a Newton-
                       corrector iteration -- runs both in lockstep with one
                       pass over the line factor per application for both;
                       False: two solves;
``'Apply Storage'``    (mixed, line factor) ``'fp16'`` (default) or ``'fp32'``:
                       with ``'fp16'`` the inner FP32 apply reads FP16 copies of
                       the factor blocks (one FP32 scale per row), half the bytes
                       of the HBM-bound apply; the FP64 refinement keeps the
                       result to the same tolerance;
``'Learned Network'``  (mixed) a ``learned.ConvNet`` (or the path of its saved
                       weights): the inner preconditioner becomes the line
                       factor followed by the judged network correction
                       (``learned.LearnedCorrection``);
``'Maximum Refinements'`` (mixed) outer refinement steps (20);
``'Tolerance'``        relative residual target (1e-10);
``'Restart'``          Krylov basis size (30);
``'Maximum Iterations'`` (300);
``'Reuse Limit'``      (mixed) GMRES iterations above which a lagged
                       preconditioner is refactored (10, for 'Inner Tolerance';
                       scaled by the reduction a step asks for; at least 1.5 times
                       the iterations the preconditioner needed on the Jacobian it
                       was made from); a refinement step that
                       stagnates on a lagged factor also refactors;
``'Factor GEMM'``      (mixed, line factor) ``'fp32'`` (default), ``'bf16x3'`` or
                       ``'bf16'``: the level GEMMs of the FP32 factor on FP32
                       MFMA, or on BF16 MFMA over three-term (one-term) BF16
                       splits of the FP32 operands;
``'Reuse Factor'``     (fp64) False (default) or True: reuse the factor of an
                       earlier Jacobian as the mixed precision does, refactoring
                       when a solve takes more than 'Reuse Limit' iterations
                       scaled from 'Inner Tolerance' to 'Tolerance' (30 by
                       default) or does not converge;
``'Fused Apply'``      (line factor) True or False (default: True for a factor
                       that is reused -- mixed, or 'Reuse Factor' -- else
                       False): the factorisation also forms Y L and Y U so
                       that each backward level of the apply is one kernel;
``'Dense Top'``        (line factor with 'Fused Apply') 0 (default), 2, 3 or 4:
                       the factorisation also forms the dense inverse of the
                       reduced system once the reduction is down to that many
                       lines, and the apply maps those lines in one launch;
``'Background Factor'`` (mixed, line factor) False (default) or True: when the
                       lagged factor exceeds the reuse limit, a second handle
                       factors the Jacobian on a side stream while the inner
                       solves go on with the lagged factor, and replaces it
                       between refinement steps once finished (the compute-bound
                       factorisation overlaps the HBM-bound applies; iteration
                       counts then depend on timing, results still meet
                       ``'Tolerance'``).

With the reference's ``'Use Iterative Solver': True`` the defaults become
those of its iterative path (SciPy.py:260-309): ``'Iterative Solver'``'s
``'Convergence Tolerance'`` (1e-6), ``'Restart'`` (100) and ``'Maximum
Iterations'`` (1000), and a solve that misses the tolerance returns its last
iterate instead of raising (``'Fail On Tolerance'`` False); the reference's
``'Preconditioner'`` settings belong to scipy's spilu and have no counterpart
here (the preconditioner is chosen as above).

Every solve appends (iterations, relative residual, preconditioner name,
factorisations it made -- a bool, or a count for the first solve of a pair) to
``Solver.log`` for measurement.
"""

import ctypes
import math
import time
import weakref

import numpy

from . import _lib
from .vector import ComplexCsrMatrix, DeviceCsrMatrix, DeviceVector, ptr, stream, torch

LINE_BUDGET_BYTES = 64 * 2 ** 30
INNER_FLOOR = 1e-5    # FP32 inner solves stall near this relative residual on these Jacobians


class SolverError(RuntimeError):
    pass


def _opts(parameters):
    o = dict(parameters.get('HIP Solver', {}))
    # the reference's iterative-solver settings (SciPy.py:260-309: restarted GMRES
    # with restart 100, at most 1000 iterations, relative tolerance 1e-6 from
    # 'Iterative Solver', failures ignored) when 'Use Iterative Solver' is set;
    # 'HIP Solver' entries still take precedence
    it = parameters.get('Use Iterative Solver', False)
    ip = dict(parameters.get('Iterative Solver', {})) if it else {}
    return {
        'Preconditioner': o.get('Preconditioner', 'auto'),
        'Precision': o.get('Precision', 'fp64'),
        'Tolerance': float(o.get('Tolerance', ip.get('Convergence Tolerance', 1e-6) if it else 1e-10)),
        'Restart': int(o.get('Restart', ip.get('Restart', 100) if it else 30)),
        'Maximum Iterations': int(o.get('Maximum Iterations', ip.get('Maximum Iterations', 1000) if it else 300)),
        'Fail On Tolerance': bool(o.get('Fail On Tolerance', not it)),
        'Reuse Limit': int(o.get('Reuse Limit', 10)),
        'Inner Tolerance': float(o.get('Inner Tolerance', 5e-4)),
        'Inner Iterations': int(o.get('Inner Iterations', 12)),
        'Maximum Refinements': int(o.get('Maximum Refinements', 20)),
        'Learned Network': o.get('Learned Network'),
        'Apply Graph': bool(o.get('Apply Graph', True)),
        'Step Graphs': bool(o.get('Step Graphs', True)),
        'Pair Solves': bool(o.get('Pair Solves', True)),
        'Apply Storage': o.get('Apply Storage', 'fp16'),
        'Background Factor': bool(o.get('Background Factor', False)),
        'Factor GEMM': o.get('Factor GEMM', 'fp32'),
        'Reuse Factor': bool(o.get('Reuse Factor', False)),
        'Fused Apply': o.get('Fused Apply'),
        'Dense Top': o.get('Dense Top'),
    }


_SPLITS = {'fp32': 0, 'bf16x3': 3, 'bf16': 1}


def _fused_apply(opts, single):
    """One-kernel backward levels (two more GEMMs per line at factor time) pay for a
    factor applied many times: the lagged factor of the mixed or factor-reusing
    solves; a factor made for every Jacobian is applied only a few times."""
    o = opts['Fused Apply']
    return (single or opts['Reuse Factor']) if o is None else bool(o)


def _dense_top(opts, fused):
    """Lines of the dense top of a block-cyclic-reduction factor (0: off, the
    default).  It costs about 20 m^3 products per factorisation and saves four
    launches per apply at four lines; it needs the fused backward."""
    o = opts['Dense Top']
    return 0 if o is None or not fused else int(o)


def _configure_line(f, opts, single):
    """Backward form and dense top of a line factor (before its first factorisation)."""
    fused = _fused_apply(opts, single)
    f.set_fused(fused)
    f.set_dense_top(_dense_top(opts, fused))


def _configure_single(f, opts):
    """Apply-storage and GEMM options of an FP32 line factor."""
    storage = opts['Apply Storage']
    if storage not in ('fp32', 'fp16'):
        raise ValueError('unknown apply storage %r' % storage)
    if opts['Factor GEMM'] not in _SPLITS:
        raise ValueError('unknown factor GEMM %r' % opts['Factor GEMM'])
    f.set_half(storage == 'fp16')
    f.set_gemm_split(_SPLITS[opts['Factor GEMM']])


class CellBlockJacobi:
    """Inverse of each cell's dof x dof diagonal block of the gauged matrix."""
    name = 'block-jacobi'

    def __init__(self, itf, jac, gauge):
        T = torch()
        self.ncell = itf.n // itf.dof
        self.dof = itf.dof
        self.inv = T.empty(self.ncell * self.dof * self.dof, dtype=T.float64, device='cuda')
        _lib.call('tf_cellblock_setup', self.ncell, self.dof, ptr(jac.row_ptr), ptr(jac.col), ptr(jac.val),
                  int(gauge), ptr(self.inv), stream())

    def apply(self, r, out):
        _lib.call('tf_cellblock_apply_f64', self.ncell, self.dof, ptr(self.inv), ptr(r), ptr(out), stream())
        return out

    def apply32(self, r, out):
        _lib.call('tf_cellblock_apply_f32', self.ncell, self.dof, ptr(self.inv), ptr(r), ptr(out), stream())
        return out

    def native(self):
        """(kind, handle, ncell, dof) for tf_fgmres."""
        return 3, ptr(self.inv), self.ncell, self.dof


class LineFactor:
    """Exact factorisation of the gauged Jacobian, block tridiagonal over grid
    lines: block cyclic reduction (``'line'``, csrc/bcr.hip) or the sequential
    Thomas recursion (``'thomas'``, csrc/linelu.hip)."""
    _ABI = {'line': 'tf_bcr', 'thomas': 'tf_linelu'}

    def __init__(self, itf, jac, gauge, single=False, kind='line', factor=True, layout=None):
        """``layout`` (n, m, nx, dof, ylines) overrides the problem's own (the
        interleaved real form of a complex matrix has twice the rows per line)."""
        self.name = kind
        self._p = self._ABI[kind]
        h = ctypes.c_void_p()
        if kind == 'line':
            m, ylines = line_layout(itf)
            n, nx, dof = itf.n, itf.nx, itf.dof
            if layout is not None:
                n, m, nx, dof, ylines = layout
            _lib.call('tf_bcr_create', ctypes.byref(h), n, m, 1 if single else 0, nx, dof, ylines)
        else:
            m = line_block_size(itf)
            _lib.call(self._p + '_create', ctypes.byref(h), itf.n, m, 1 if single else 0)
        self._h = h
        self.single = single
        self.m = m
        self.owner = None
        if factor:
            self.refactor(jac, gauge)

    def refactor(self, jac, gauge):
        """Factor ``jac`` into this handle's buffers (the previous owner's factor is gone)."""
        self.refactor_async(jac, gauge, stream())
        self.check(stream())

    def refactor_async(self, jac, gauge, st):
        """Queue the factorisation of ``jac`` on stream ``st`` (a ctypes stream pointer)."""
        old = self.owner() if self.owner is not None else None
        if old is not None and old is not jac and old.prec is self:
            old.prec = None
            old.prec_exact = False
        self.owner = weakref.ref(jac)
        _lib.call(self._p + '_factor', self._h, ptr(jac.row_ptr), ptr(jac.col), ptr(jac.val), int(gauge), st)

    def check(self, st):
        """Structure flags of the last factorisation (synchronises ``st``)."""
        flags = (ctypes.c_int * 2)()
        _lib.call(self._p + '_status', self._h, flags, st)
        if flags[0]:
            raise SolverError('line factorisation: matrix is not block tridiagonal over grid lines (flags %d)'
                              % flags[0])
        self.perturbed = flags[1]

    def check_applies(self, st):
        """Raise if a wait of the one-launch upper levels (tf_bcr_set_top, which is
        disabled) gave up since the last check (synchronises ``st``)."""
        if self._p == 'tf_bcr':
            n = ctypes.c_int()
            _lib.call('tf_bcr_top_errors', self._h, ctypes.byref(n), st)
            if n.value:
                raise SolverError('line-factor apply: %d upper-level waits gave up' % n.value)

    def set_small(self, on):
        """Levels with at most 8 lines get their operands as kernel arguments (block
        cyclic reduction); the results do not change."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_set_small', self._h, 1 if on else 0)

    def set_top(self, on):
        """The upper levels of the apply in one launch (block cyclic reduction with
        the fused backward); the results do not change."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_set_top', self._h, 1 if on else 0)

    def apply(self, r, out):
        _lib.call(self._p + '_solve_f64', self._h, ptr(r), ptr(out), stream())
        return out

    def apply32(self, r, out):
        _lib.call(self._p + '_solve_f32', self._h, ptr(r), ptr(out), stream())
        return out

    def apply_pair(self, r0, r1, out0, out1, single=False):
        """Two right-hand sides: one pass over the blocks of the block cyclic
        reduction (tf_bcr_solve2_*), else two applies."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_solve2_f32' if single else 'tf_bcr_solve2_f64', self._h, ptr(r0), ptr(r1), ptr(out0),
                      ptr(out1), stream())
        else:
            (self.apply32 if single else self.apply)(r0, out0)
            (self.apply32 if single else self.apply)(r1, out1)
        return out0, out1

    def set_half(self, on):
        """FP32 apply from FP16 copies of the blocks (block cyclic reduction, FP32 factor only)."""
        if self._p == 'tf_bcr' and self.single and bool(on) != getattr(self, 'half', False):
            _lib.call('tf_bcr_set_half', self._h, 1 if on else 0, stream())
            self.half = bool(on)

    def set_fused(self, on):
        """Form Y L and Y U at factor time so that each backward level of the apply
        is one kernel (block cyclic reduction; before the first factorisation)."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_set_fused', self._h, 1 if on else 0)
            self.fused = bool(on)

    def set_dense_top(self, lines):
        """Form the dense inverse of the last ``lines`` (2..4) lines of the reduction at
        factor time and apply it in one launch (block cyclic reduction, fused factors;
        0: off; from the next factorisation on)."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_set_dense_top', self._h, int(lines))
            self.dense_top = int(lines)

    def set_gemm_split(self, split):
        """Level GEMMs of the FP32 factor on BF16 terms (0 off, 1 or 3 terms); applies
        from the next factorisation on."""
        if self._p == 'tf_bcr' and self.single:
            _lib.call('tf_bcr_set_gemm_split', self._h, int(split))

    def native(self):
        """(kind, handle, ncell, dof) for tf_fgmres."""
        return (1 if self._p == 'tf_bcr' else 2), self._h, 0, 0

    def set_graph(self, on):
        """Replay the apply as one captured hipGraph (block cyclic reduction only)."""
        if self._p == 'tf_bcr':
            _lib.call('tf_bcr_set_graph', self._h, 1 if on else 0)
            self.graph = bool(on)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                getattr(_lib.load(), self._p + '_destroy')(h)
            except Exception:
                pass
            self._h = None


class LineILU:
    """Block ILU over grid x-lines with a four-colour line order (csrc/lineilu.hip):
    dense inverted line pivots, the Jacobian's own sparse couplings between lines."""
    name = 'line-ilu'

    def __init__(self, itf, single):
        h = ctypes.c_void_p()
        _lib.call('tf_lineilu_create', ctypes.byref(h), itf.nx, itf.ny, itf.nz, itf.dof, 1 if single else 0)
        self._h = h
        self.single = single
        self.owner = None
        self.perturbed = 0

    def refactor(self, jac, gauge):
        _take_ownership(self, jac)
        _lib.call('tf_lineilu_factor', self._h, ptr(jac.row_ptr), ptr(jac.col), ptr(jac.val), int(gauge), stream())
        flags = (ctypes.c_int * 2)()
        _lib.call('tf_lineilu_status', self._h, flags, stream())
        if flags[0]:
            raise SolverError('line ILU: more couplings between lines than the factor keeps (flags %d)' % flags[0])
        self.perturbed = flags[1]

    def apply(self, r, out):
        _lib.call('tf_lineilu_solve_f64', self._h, ptr(r), ptr(out), stream())
        return out

    def apply32(self, r, out):
        _lib.call('tf_lineilu_solve_f32', self._h, ptr(r), ptr(out), stream())
        return out

    def native(self):
        return 5, self._h, 0, 0

    def __del__(self):
        _destroy(self, 'tf_lineilu_destroy')


def coarse_coordinates(x, n):
    """Coordinate vector (length n/2 + 3, the layout of utils.py:175-232) of the
    grid coarsened by two: coarse face I is fine face 2I+1; the ghost entries
    mirror the first and last coarse cells."""
    N = n // 2
    X = numpy.empty(N + 3)
    X[:N] = x[1:n:2]
    X[-1] = x[-1]
    X[-2] = X[-1] - (X[0] - X[-1])
    X[N] = X[N - 1] + (X[N - 1] - X[N - 2]) if N > 1 else X[N - 1] + (X[N - 1] - X[-1])
    return X


def multigrid_fits(itf, single=False):
    """The two-grid method needs a 3D grid with even sizes and non-periodic
    walls, and the exact factor of the coarse grid within the memory budget."""
    d = itf.discretization
    if itf.nz == 1 or itf.dim != 3 or any(n % 2 for n in (itf.nx, itf.ny, itf.nz)):
        return False
    if d.x_periodic or d.y_periodic or d.z_periodic:
        return False
    m = (itf.nx // 2) * (itf.ny // 2) * itf.dof
    return 5 * (itf.nz // 2) * m * m * (4 if single else 8) <= LINE_BUDGET_BYTES


class Multigrid:
    """Two-grid preconditioner (csrc/mg.hip): line block ILU smoothing on the grid
    and the exact plane factorisation of the Jacobian re-assembled on the grid
    coarsened by two at the restricted state, one V(1,1) cycle per application.
    For 3D grids whose own plane factor does not fit (the 64^3 cavity: 16384-row
    planes); the line ILU alone converges ever more slowly as the grid grows."""
    name = 'multigrid'

    def __init__(self, itf, single):
        from .interface import Interface
        T = torch()
        self.itf = itf
        self.single = single
        self.owner = None
        nc = (itf.nx // 2, itf.ny // 2, itf.nz // 2)
        d = itf.discretization
        self.coarse = Interface(dict(itf.parameters), *nc, itf.dim, itf.dof,
                                x=coarse_coordinates(itf.x, itf.nx), y=coarse_coordinates(itf.y, itf.ny),
                                z=coarse_coordinates(itf.z, itf.nz), boundary_conditions=d.boundary_conditions)
        self.xc = T.empty(self.coarse.n, dtype=T.float64, device='cuda')
        self.smoother = LineILU(itf, single)
        self.cf = None
        h = ctypes.c_void_p()
        _lib.call('tf_mg_create', ctypes.byref(h), itf.nx, itf.ny, itf.nz, itf.dim, itf.dof, 1 if single else 0)
        self._h = h
        self._keep = None

    def refactor(self, jac, gauge):
        itf, c = self.itf, self.coarse
        _take_ownership(self, jac)
        _lib.call('tf_mg_restrict_state', itf.nx, itf.ny, itf.nz, itf.dim, itf.dof, ptr(jac.state), ptr(self.xc),
                  stream())
        c.parameters.clear()
        c.parameters.update(jac.parameters)
        Jc = c.jacobian(DeviceVector(self.xc))
        gc = c.pressure_row if (c.dof > c.dim and c.pressure_row is not None) else -1
        if self.cf is None:
            self.cf = LineFactor(c, Jc, gc, single=self.single, kind='line', factor=False)
            opts = _opts(itf.parameters)
            _configure_line(self.cf, opts, self.single)
            self.cf.set_graph(opts['Apply Graph'])
            if self.single:
                _configure_single(self.cf, opts)
        self.cf.refactor(Jc, gc)
        self.smoother.refactor(jac, gauge)
        val = jac.values32() if self.single else jac.val
        # the cycle reads these arrays: keep them alive while this factor is in use
        self._keep = (jac.row_ptr, jac.col, val, Jc)
        _lib.call('tf_mg_set', self._h, self.smoother._h, ptr(jac.row_ptr), ptr(jac.col), ptr(val), int(gauge),
                  self.cf._h, int(gc))
        self.perturbed = self.smoother.perturbed + getattr(self.cf, 'perturbed', 0)

    def apply(self, r, out):
        _lib.call('tf_mg_solve_f64', self._h, ptr(r), ptr(out), stream())
        return out

    def apply32(self, r, out):
        _lib.call('tf_mg_solve_f32', self._h, ptr(r), ptr(out), stream())
        return out

    def native(self):
        return 4, self._h, 0, 0

    def __del__(self):
        _destroy(self, 'tf_mg_destroy')


def _take_ownership(f, jac):
    """A refactored handle belongs to ``jac``; the Jacobian that held it drops it."""
    old = f.owner() if f.owner is not None else None
    if old is not None and old is not jac and old.prec is f:
        old.prec = None
        old.prec_exact = False
    f.owner = weakref.ref(jac)


def _destroy(obj, fn):
    h = getattr(obj, '_h', None)
    if h is not None and h.value:
        try:
            getattr(_lib.load(), fn)(h)
        except Exception:
            pass
        obj._h = None


def line_block_size(itf):
    """Rows of one grid line (2D x-line) or grid plane (3D)."""
    if itf.nz == 1:
        return itf.nx * itf.dof
    return itf.nx * itf.ny * itf.dof


def line_layout(itf):
    """(rows per line, y-lines?) of the block cyclic reduction: in 2D the
    shorter of the x- and y-lines (fewer rows per dense block)."""
    if itf.nz == 1 and itf.ny < itf.nx:
        return itf.ny * itf.dof, 1
    return line_block_size(itf), 0


def line_fits(itf, single=False):
    """Block cyclic reduction keeps five dense m x m blocks per line (D, L, U and
    the two elimination multipliers)."""
    m = line_layout(itf)[0]
    nlines = itf.n // m
    return 5 * nlines * m * m * (4 if single else 8) <= LINE_BUDGET_BYTES


def auto_preconditioner(itf, single):
    """'line' when the exact line (plane) factor fits the memory budget, else
    'multigrid' for 3D grids it covers; otherwise no preconditioner that
    converges is available and the solve fails with a clear error."""
    if line_fits(itf, single):
        return 'line'
    if multigrid_fits(itf, single):
        return 'multigrid'
    m, _ = line_layout(itf)
    raise SolverError('no preconditioner for this grid: the exact line factor needs %.1f GB (budget %.1f GB) and '
                      'the two-grid method needs a 3D grid with even sizes and walls on every side'
                      % (5 * itf.n * m * (4 if single else 8) / 1e9, LINE_BUDGET_BYTES / 1e9))


class Krylov:
    """Restarted flexible GMRES (right preconditioning, CGS2 orthogonalisation)
    on device-resident vectors of one precision: FP64, or FP32 for the inner
    solve of the mixed-precision refinement (basis, SpMV, preconditioner apply
    and block-level reductions in FP32).  The (j+2)-entry Hessenberg column is
    reduced across blocks in FP64 and the Givens rotations run on the host."""

    _K = {
        'float64': ('tf_axpby_f64', 'tf_scale_f64', 'tf_fill_f64', 'tf_multidot_f64', 'tf_multiaxpy_f64',
                    'tf_normalize_f64'),
        'float32': ('tf_axpby_f32', 'tf_scale_f32', 'tf_fill_f32', 'tf_multidot_f32', 'tf_multiaxpy_f32',
                    'tf_normalize_f32'),
    }

    def __init__(self, n, restart, dtype='float64'):
        T = torch()
        self.n = n
        self.m = restart
        self.dtype = dtype
        dt = getattr(T, dtype)
        self.V = T.empty((restart + 1, n), dtype=dt, device='cuda')
        self.Z = T.empty((restart, n), dtype=dt, device='cuda')
        self.w = T.empty(n, dtype=dt, device='cuda')
        self.h = T.zeros(restart + 2, dtype=T.float64, device='cuda')
        self.y = T.empty(restart + 1, dtype=T.float64, device='cuda')
        nw = _lib.load().tf_krylov_work_size()
        self.work = T.empty(nw, dtype=T.float64, device='cuda')
        self.h_host = T.empty(restart + 2, dtype=T.float64).pin_memory()
        (self._axpby, self._scale, self._fill, self._multidot, self._multiaxpy,
         self._normalize) = self._K[dtype]

    def norm(self, x):
        if self.dtype == 'float64':
            return _norm(x, self.work)
        _lib.call('tf_normalize_f32', x.numel(), ptr(x), ptr(self.work), ptr(self.h), ptr(self.w), stream())
        self.h_host[:1].copy_(self.h[:1], non_blocking=True)
        torch().cuda.current_stream().synchronize()
        return float(self.h_host[0])

    def solve(self, matvec, precond, b, x, tol, maxit, x_is_zero=False):
        """Solve A x = b (x holds the initial guess, updated in place).
        Returns (iterations, relative residual)."""
        n, m = self.n, self.m
        st = stream()
        V, Z, w, h = self.V, self.Z, self.w, self.h
        lib = _lib
        bnrm = self.norm(b)
        if bnrm == 0.0:
            lib.call(self._fill, n, 0.0, ptr(x), st)
            return 0, 0.0
        its = 0
        rel = math.inf
        first = True
        while True:
            # r = b - A x into V[0]
            if first and x_is_zero:
                V[0].copy_(b)
                beta = bnrm
            else:
                matvec(x, w)
                lib.call(self._axpby, n, 1.0, ptr(b), -1.0, ptr(w), ptr(V[0]), st)
                beta = self.norm(V[0])
            rel = beta / bnrm
            if rel <= tol or its >= maxit:
                return its, rel
            if not first and beta == 0.0:
                return its, rel
            first = False
            lib.call(self._scale, n, 1.0 / beta, ptr(V[0]), ptr(V[0]), st)
            H = numpy.zeros((m + 1, m))
            cs = numpy.zeros(m)
            sn = numpy.zeros(m)
            g = numpy.zeros(m + 1)
            g[0] = beta
            k = 0
            for j in range(m):
                precond(V[j], Z[j])
                matvec(Z[j], w)
                lib.call(self._multidot, n, j + 1, ptr(V), n, ptr(w), ptr(self.work), ptr(h), 0, st)
                lib.call(self._multiaxpy, n, j + 1, ptr(V), n, ptr(h), ptr(w), -1.0, st)
                lib.call(self._multidot, n, j + 1, ptr(V), n, ptr(w), ptr(self.work), ptr(self.y), 0, st)
                lib.call(self._multiaxpy, n, j + 1, ptr(V), n, ptr(self.y), ptr(w), -1.0, st)
                lib.call('tf_axpby_f64', j + 1, 1.0, ptr(h), 1.0, ptr(self.y), ptr(h), st)
                lib.call(self._normalize, n, ptr(w), ptr(self.work), ctypes.c_void_p(h.data_ptr() + 8 * (j + 1)),
                         ptr(V[j + 1]), st)
                self.h_host[:j + 2].copy_(h[:j + 2], non_blocking=True)
                torch().cuda.current_stream().synchronize()
                col = self.h_host[:j + 2].numpy()
                H[:j + 2, j] = col
                for i in range(j):
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                d = math.hypot(H[j, j], H[j + 1, j])
                if d == 0.0:
                    k = j
                    break
                cs[j] = H[j, j] / d
                sn[j] = H[j + 1, j] / d
                H[j, j] = d
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                its += 1
                k = j + 1
                if abs(g[j + 1]) / bnrm <= tol or its >= maxit or not numpy.isfinite(g[j + 1]):
                    break
            if k == 0:
                return its, rel
            yk = _back_substitute(H[:k, :k], g[:k])
            self.y[:k].copy_(torch().from_numpy(yk))
            lib.call(self._multiaxpy, n, k, ptr(Z), n, ptr(self.y), ptr(x), 1.0, st)
            if its >= maxit and self.dtype == 'float32':
                # inner solve of the refinement: the caller forms the true FP64 residual
                return its, abs(g[k]) / bnrm


class NativeKrylov:
    """The same FGMRES as :class:`Krylov`, run inside the library (tf_fgmres):
    one call per solve, no interpreter in the Arnoldi loop."""

    def __init__(self, n, restart, dtype='float64', graphs=True):
        self.n, self.m, self.dtype = n, restart, dtype
        h = ctypes.c_void_p()
        _lib.call('tf_krylov_create', ctypes.byref(h), n, restart, 1 if dtype == 'float32' else 0)
        self._h = h
        self.set_graphs(graphs)

    def set_graphs(self, on):
        """Replay each Arnoldi step as a captured graph (on) or launch it eagerly."""
        self.graphs = bool(on)
        _lib.call('tf_krylov_set_graph', self._h, 1 if on else 0)

    def solve(self, jac, val, gauge, prec, b, x, tol, maxit, x_is_zero=False):
        kind, handle, ncell, dof = prec.native()
        its = ctypes.c_int()
        rel = ctypes.c_double()
        _lib.call('tf_fgmres', self._h, ptr(jac.row_ptr), ptr(jac.col), ptr(val), int(gauge), kind, handle, ncell,
                  dof, ptr(b), ptr(x), float(tol), int(maxit), 1 if x_is_zero else 0, ctypes.byref(its),
                  ctypes.byref(rel), stream())
        return its.value, rel.value

    def solve2(self, other, jac, val, gauge, prec, b0, b1, x0, x1, tol0, tol1, maxit, x_is_zero=False):
        """Two systems with one matrix and preconditioner in lockstep (tf_fgmres2);
        ``other`` is a second NativeKrylov of the same size, restart and precision."""
        kind, handle, ncell, dof = prec.native()
        its = (ctypes.c_int * 2)()
        rel = (ctypes.c_double * 2)()
        _lib.call('tf_fgmres2', self._h, other._h, ptr(jac.row_ptr), ptr(jac.col), ptr(val), int(gauge), kind, handle,
                  ncell, dof, ptr(b0), ptr(b1), ptr(x0), ptr(x1), float(tol0), float(tol1), int(maxit),
                  1 if x_is_zero else 0, its, rel, stream())
        return (its[0], rel[0]), (its[1], rel[1])

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                _lib.load().tf_krylov_destroy(h)
            except Exception:
                pass
            self._h = None


def _back_substitute(R, g):
    k = len(g)
    y = numpy.zeros(k)
    for i in range(k - 1, -1, -1):
        y[i] = (g[i] - R[i, i + 1:] @ y[i + 1:]) / R[i, i]
    return y


def _norm(x, work):
    out = ctypes.c_double()
    _lib.call('tf_dot_f64', x.numel(), ptr(x), ptr(x), ptr(work), ctypes.c_void_p(work.data_ptr() + 8 * (work.numel() - 1)),
              ctypes.byref(out), 1, stream())
    return out.value


class BackgroundFactor:
    """Two FP32 line-factor handles for the mixed-precision solve: the inner
    solves use ``cur`` (the factor of an earlier Jacobian, or of this one)
    while the other handle factors the newest Jacobian on a side stream; a
    finished factor becomes ``cur`` between refinement steps.  Before it starts,
    the side stream waits for everything the solve stream has queued (the
    Jacobian's assembly and the last applies of the handle it overwrites), and
    the Jacobian is kept alive until the factorisation has finished."""

    def __init__(self, solver, kind):
        self.solver = solver
        self.kind = kind
        self.cur = None
        self.spare = None
        self.pending = None     # (factor, jac, done event)
        self.users = []         # weak references of Jacobians whose prec is cur
        self.stream = torch().cuda.Stream()
        self.started = 0        # background factorisations queued
        self.swaps = 0          # ... and taken into use

    def _configure(self, f):
        opts = _opts(self.solver.itf.parameters)
        _configure_line(f, opts, True)
        f.set_graph(opts['Apply Graph'])
        _configure_single(f, opts)
        return f

    def _adopt(self, f, jac, exact):
        """Make ``f`` the current factor; Jacobians that held the old one drop it."""
        old = self.cur
        self.cur, self.spare = f, old
        for ref in self.users:
            j = ref()
            if j is not None and j.prec is old:
                j.prec = None
                j.prec_exact = False
        self.users = []
        if jac is not None:
            self.assign(jac, exact)

    def assign(self, jac, exact=None):
        jac.prec = self.cur
        jac.prec_exact = (self.cur.owner is not None and self.cur.owner() is jac) if exact is None else exact
        self.users.append(weakref.ref(jac))

    def poll(self, jac=None, wait=False):
        """Take a finished background factor into use; True if one was."""
        p = self.pending
        if p is None:
            return False
        f, fjac, done = p
        if not wait and not done.query():
            return False
        done.synchronize()
        f.check(ctypes.c_void_p(self.stream.cuda_stream))
        self.pending = None
        self._adopt(f, None, False)
        for j in (fjac, jac):
            if j is not None and j.prec is None:
                self.assign(j)
        self.swaps += 1
        return True

    def start(self, jac):
        """Factor ``jac`` on the side stream unless a factorisation is running or
        the current factor is already this Jacobian's."""
        if self.pending is not None or self.cur is None:
            return
        if self.cur.owner is not None and self.cur.owner() is jac:
            return
        s = self.solver
        if self.spare is None:
            self.spare = self._configure(LineFactor(s.itf, jac, s.gauge(), single=True, kind=self.kind,
                                                    factor=False))
        T = torch()
        queued = T.cuda.Event()
        queued.record()
        self.stream.wait_event(queued)
        self.spare.refactor_async(jac, s.gauge(), ctypes.c_void_p(self.stream.cuda_stream))
        done = T.cuda.Event()
        done.record(self.stream)
        self.pending = (self.spare, jac, done)
        self.spare = None
        self.started += 1

    def now(self, jac):
        """The factor of ``jac`` itself, made on the solve stream if no background
        factorisation provides it."""
        self.poll(jac, wait=True)
        if self.cur is not None and self.cur.owner is not None and self.cur.owner() is jac:
            self.assign(jac, True)
            return self.cur
        s = self.solver
        f = self.spare
        if f is None:
            f = self._configure(LineFactor(s.itf, jac, s.gauge(), single=True, kind=self.kind, factor=False))
        f.refactor(jac, s.gauge())
        self._adopt(f, jac, True)
        return f


class Solver:
    def __init__(self, itf):
        self.itf = itf
        self.log = []
        self.refinements = []  # (mixed) outer refinement steps per solve
        self.fallbacks = 0     # (mixed) solves finished in FP64
        self.stale_refactors = 0   # (mixed) lagged factors replaced because refinement stagnated
        self.limit_refactors = 0   # (mixed) ... because an inner solve exceeded the reuse limit
        self.trace = []        # (mixed) per refinement step: (inner its, inner estimate, outer FP64 rel)
        self.observer = None   # optional callable(jac, precond32, v) on every inner preconditioner input
        self._learned = None
        self._krylov = {}
        self._lagged = None   # (precond) reused across Jacobians in mixed precision
        self._buf = None
        self._factors = {}    # (kind, single) -> LineFactor, refactored in place
        self._bg = None       # BackgroundFactor (mixed, 'Background Factor')

    def _krylov_for(self, restart, dtype='float64', native=False, slot=0):
        key = (dtype, native, slot)
        k = self._krylov.get(key)
        if k is None or k.m != restart:
            k = self._krylov[key] = (NativeKrylov if native else Krylov)(self.itf.n, restart, dtype)
        if native:
            graphs = _opts(self.itf.parameters)['Step Graphs']
            if k.graphs != graphs:
                k.set_graphs(graphs)
        return k

    def gauge(self):
        itf = self.itf
        return itf.pressure_row if (itf.dof > itf.dim and itf.pressure_row is not None) else -1

    def _build(self, jac, kind, single):
        f = self._build_prec(jac, kind, single)
        opts = _opts(self.itf.parameters)
        if hasattr(f, 'set_graph'):
            f.set_graph(opts['Apply Graph'])
        if single and hasattr(f, 'set_half'):
            _configure_single(f, opts)
        return f

    def _build_prec(self, jac, kind, single):
        g = self.gauge()
        if kind == 'auto':
            kind = auto_preconditioner(self.itf, single)
        if kind in ('multigrid', 'line-ilu'):
            f = self._factors.get((kind, single))
            if f is None:
                f = self._factors[(kind, single)] = (Multigrid if kind == 'multigrid' else LineILU)(self.itf, single)
            f.refactor(jac, g)
            return f
        if kind in ('line', 'thomas'):
            # one factor handle per (kind, precision), refactored in place: the
            # dense blocks are GBs, so a new Jacobian reuses the buffers instead
            # of a fresh allocation (the factor of the previous Jacobian goes)
            f = self._factors.get((kind, single))
            if f is None:
                opts = _opts(self.itf.parameters)
                f = self._factors[(kind, single)] = LineFactor(self.itf, jac, g, single=single, kind=kind,
                                                               factor=False)
                _configure_line(f, opts, single)
                if single:
                    _configure_single(f, opts)
            f.refactor(jac, g)
            return f
        if kind == 'block-jacobi':
            return CellBlockJacobi(self.itf, jac, g)
        raise ValueError('unknown preconditioner %r' % kind)

    def _background(self, opts):
        """The BackgroundFactor when enabled and the preconditioner is a line factor."""
        if opts['Precision'] != 'mixed' or not opts['Background Factor']:
            return None
        if self._bg is None:
            kind = opts['Preconditioner']
            if kind == 'auto':
                kind = auto_preconditioner(self.itf, True)
            if kind not in ('line', 'thomas'):
                return None
            self._bg = BackgroundFactor(self, kind)
        return self._bg

    def preconditioner(self, jac, opts):
        mixed = opts['Precision'] == 'mixed'
        lag = mixed or opts['Reuse Factor']
        bg = self._background(opts)
        if bg is not None:
            bg.poll(jac)
            if bg.cur is None:
                return bg.now(jac), True
            if jac.prec is None:
                bg.assign(jac)
            return jac.prec, False
        if jac.prec is not None:
            return jac.prec, False
        if lag and self._lagged is not None:
            jac.prec = self._lagged
            jac.prec_exact = False   # factor of an earlier Jacobian
            return jac.prec, False
        t0 = time.perf_counter()
        jac.prec = self._build(jac, opts['Preconditioner'], single=mixed)
        jac.prec_exact = True
        jac.prec_time = time.perf_counter() - t0
        if lag:
            self._lagged = jac.prec
        return jac.prec, True

    def solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        """y with (gauged J) y = rhs.  ``rhs`` of length n gives a DeviceVector;
        an (n, k) array gives an (n, k) numpy array whose columns are solved
        one after another, each with its ``pressure_row`` entry zeroed
        (SciPy.py:209-216, where splu solves all columns of the gauged rhs)."""
        if isinstance(jac, ComplexCsrMatrix):
            return self.solve_complex(jac, rhs)
        if rhs2 is not None or V is not None:
            return self.solve_bordered(jac, rhs, rhs2, V, W, C)
        n = self.itf.n
        if isinstance(rhs, DeviceVector):
            if rhs.size != n:
                raise ValueError('rhs has %d entries, the system has %d rows' % (rhs.size, n))
            return DeviceVector(self._solve_device(jac, rhs.data, zero_gauge=True))
        if numpy.iscomplexobj(rhs):
            # a real matrix and a complex rhs: the real and imaginary parts apart
            c = numpy.asarray(rhs)
            parts = [self.solve(jac, numpy.ascontiguousarray(p)) for p in (c.real, c.imag)]
            parts = [numpy.asarray(p.numpy() if isinstance(p, DeviceVector) else p) for p in parts]
            return parts[0] + 1j * parts[1]
        a = numpy.asarray(rhs, dtype=numpy.float64)
        if a.ndim == 1 and a.shape[0] == n:
            return DeviceVector(self._solve_device(jac, DeviceVector.from_array(a).data, zero_gauge=True))
        if a.ndim == 2 and a.shape[0] == n:
            out = numpy.empty_like(a)
            for j in range(a.shape[1]):
                col = DeviceVector.from_array(a[:, j]).data
                out[:, j] = self._solve_device(jac, col, zero_gauge=True).cpu().numpy()
            return out
        raise ValueError('rhs must have shape (%d,) or (%d, k), got %s' % (n, n, a.shape))

    def _solve_device(self, jac, rhs, zero_gauge):
        """x with (gauged J) x = rhs; ``zero_gauge`` zeroes rhs[pressure_row] first
        (SciPy.py:211-216; the bordered solve keeps it, SciPy.py:226-227)."""
        opts = _opts(self.itf.parameters)
        T = torch()
        g = self.gauge()
        if rhs.dim() != 1 or rhs.numel() != self.itf.n or rhs.dtype != T.float64:
            raise ValueError('right-hand side of shape %s (%s) for a system of %d FP64 rows'
                             % (tuple(rhs.shape), rhs.dtype, self.itf.n))
        b = rhs.clone()
        if g >= 0 and zero_gauge:
            _lib.call('tf_set_entry_f64', ptr(b), b.numel(), g, 0.0, stream())
        x = T.empty_like(b)
        _lib.call('tf_fill_f64', x.numel(), 0.0, ptr(x), stream())
        if opts['Precision'] == 'mixed':
            its, rel, prec, fresh = self._solve_mixed(jac, b, x, opts)
        else:
            its, rel, prec, fresh = self._solve_fp64(jac, b, x, opts)
        self.log.append((its, rel, getattr(prec, 'name', '?'), fresh))
        if not rel <= opts['Tolerance'] and opts['Fail On Tolerance']:
            raise SolverError('GMRES did not reach the tolerance: relative residual %.3e after %d iterations'
                              % (rel, its))
        return x

    def _solve_fp64(self, jac, b, x, opts):
        g = self.gauge()
        prec, fresh = self.preconditioner(jac, opts)

        def run(prec, first):
            if hasattr(prec, 'native'):
                kry = self._krylov_for(opts['Restart'], native=True)
                return kry.solve(jac, jac.val, g, prec, b, x, opts['Tolerance'], opts['Maximum Iterations'],
                                 x_is_zero=first)
            kry = self._krylov_for(opts['Restart'])

            def matvec(v, out):
                jac.matvec(v, out=out, gauge=g)

            return kry.solve(matvec, prec.apply, b, x, opts['Tolerance'], opts['Maximum Iterations'])

        its, rel = run(prec, True)
        if opts['Reuse Factor'] and not jac.prec_exact:
            # a lagged factor: refactor once it costs more iterations than the reuse
            # limit (scaled from 'Inner Tolerance' to 'Tolerance'), and finish the
            # solve with the new factor when the lagged one did not reach the tolerance
            limit = opts['Reuse Limit'] * math.log(opts['Tolerance']) / math.log(opts['Inner Tolerance'])
            if its > limit or not rel <= opts['Tolerance']:
                self.limit_refactors += 1
                prec, fresh = self._refactor_now(jac, opts)
                if not rel <= opts['Tolerance']:
                    its2, rel = run(prec, False)
                    its += its2
        return its, rel, prec, fresh

    def _learned_for(self, net):
        if self._learned is None:
            from .learned import ConvNet, LearnedCorrection
            if isinstance(net, str):
                net = ConvNet.load(net)
            itf = self.itf
            self._learned = LearnedCorrection(net, (itf.nx, itf.ny, itf.nz), itf.n)
        return self._learned

    def _fp64_fallback(self, jac, b, x, opts):
        prec = getattr(jac, 'prec64', None)
        if prec is None or prec.owner is None or prec.owner() is not jac:
            kind = opts['Preconditioner']
            prec = jac.prec64 = self._build(jac, kind, single=False)
        kry = self._krylov_for(opts['Restart'])
        g = self.gauge()

        def matvec(v, out):
            jac.matvec(v, out=out, gauge=g)

        if hasattr(prec, 'native'):
            nk = self._krylov_for(opts['Restart'], native=True)
            its, rel = nk.solve(jac, jac.val, g, prec, b, x, opts['Tolerance'], opts['Maximum Iterations'])
        else:
            its, rel = kry.solve(matvec, prec.apply, b, x, opts['Tolerance'], opts['Maximum Iterations'])
        return its, rel, prec, True

    def _refactor_now(self, jac, opts):
        """Replace a lagged factor by this Jacobian's own (mixed precision)."""
        bg = self._background(opts)
        if bg is not None:
            return bg.now(jac), True
        jac.prec = None
        self._lagged = None
        return self.preconditioner(jac, opts)

    def _buffers(self, n):
        T = torch()
        if self._buf is None or self._buf[0].numel() != n:
            self._buf = (T.empty(n, dtype=T.float64, device='cuda'), T.empty(n, dtype=T.float32, device='cuda'),
                         T.empty(n, dtype=T.float32, device='cuda'), T.empty(n, dtype=T.float64, device='cuda'))
        return self._buf

    def _refine_residuals(self, jac, g, items, work):
        """|b - A x| (FP64) for up to two (x, b, buffers) items, with r32 = (float) r
        and d32 = 0 queued for the inner solve (tf_refine_residual per item); the
        norms come back with one synchronisation."""
        T = torch()
        n, st = self.itf.n, stream()
        if getattr(self, '_normbuf', None) is None:
            self._normbuf = T.empty(2, dtype=T.float64, device='cuda')
        nd = self._normbuf.data_ptr()
        for q, (x, b, (r, r32, d32, _)) in enumerate(items):
            _lib.call('tf_refine_residual', n, ptr(jac.row_ptr), ptr(jac.col), ptr(jac.val), int(g), ptr(x), ptr(b),
                      ptr(r), ptr(r32), ptr(d32), ptr(work), ctypes.c_void_p(nd + 8 * q), st)
        out = (ctypes.c_double * len(items))()
        _lib.call('tf_read_f64', ctypes.c_void_p(nd), len(items), out, st)
        return list(out)

    def _solve_mixed(self, jac, b, x, opts):
        """FP64 iterative refinement: r = b - A x (FP64), A d = r by FP32 FGMRES
        with the FP32 factor, x += d (FP64), until ||r|| <= tol ||b|| in FP64."""
        n = self.itf.n
        g = self.gauge()
        st = stream()
        r, r32, d32, d = self._buffers(n)
        kry = self._krylov_for(opts['Restart'], 'float32')
        work = self._krylov_for(opts['Restart']).work
        val32 = jac.values32()
        bg = self._background(opts)
        prec, fresh = self.preconditioner(jac, opts)

        def matvec32(v, out):
            _lib.call('tf_spmv_f32', n, ptr(jac.row_ptr), ptr(jac.col), ptr(val32), ptr(v), ptr(out), int(g), st)

        bnrm = _norm(b, work)
        if bnrm == 0.0:
            return 0, 0.0, prec, fresh

        def inner_prec():
            p = prec
            net = opts['Learned Network']
            if net is not None:
                lc = self._learned_for(net)
                if not lc.off:   # once switched off, the base runs alone (natively)
                    p = lc.bind(p, matvec32)
            if self.observer is not None:
                obs, base = self.observer, p

                class _Observed:
                    name = getattr(base, 'name', '?')

                    @staticmethod
                    def apply32(v, out):
                        obs(jac, base, v)
                        return base.apply32(v, out)
                p = _Observed
            return p
        total, outer, rel, prev = 0, 0, 1.0, math.inf
        inner_max = min(opts['Restart'], opts['Inner Iterations'])
        while outer < opts['Maximum Refinements'] and total < opts['Maximum Iterations']:
            if outer > 0:
                # r = b - A x, r32 = (float) r, d32 = 0
                rel = self._refine_residuals(jac, g, [(x, b, (r, r32, d32, d))], work)[0] / bnrm
                if rel <= opts['Tolerance']:
                    break
                if bg is not None and bg.poll(jac):
                    prec = jac.prec
                if rel > 0.5 * prev and not jac.prec_exact:
                    # the lagged factor no longer contracts: factor this Jacobian
                    self.stale_refactors += 1
                    prec, fresh = self._refactor_now(jac, opts)
                    prev = math.inf
                elif rel > 0.5 * prev and jac.prec_exact:
                    # refinement stagnates with this Jacobian's own FP32 factor (FP32 too
                    # coarse for its conditioning): finish in FP64 from the current x
                    its64, rel, prec, _ = self._fp64_fallback(jac, b, x, opts)
                    total += its64
                    self.fallbacks += 1
                    self.refinements.append(outer)
                    return total, rel, prec, True
                prev = rel
            else:
                _lib.call('tf_f64_to_f32', n, ptr(b), ptr(r32), st)
                _lib.call('tf_fill_f32', n, 0.0, ptr(d32), st)
            # the last refinement step: when the remaining reduction is within what an
            # FP32 inner solve reaches, ask for it (with margin) instead of 'Inner Tolerance'
            itol = opts['Inner Tolerance']
            if outer > 0:
                need = 0.5 * opts['Tolerance'] / rel
                if INNER_FLOOR <= need < itol:
                    itol = need
            p_inner = inner_prec()
            if p_inner is prec and hasattr(prec, 'native'):
                nk = self._krylov_for(opts['Restart'], 'float32', native=True)
                its, est = nk.solve(jac, val32, g, prec, r32, d32, itol, inner_max, x_is_zero=True)
            else:
                its, est = kry.solve(matvec32, p_inner.apply32, r32, d32, itol, inner_max,
                                     x_is_zero=True)
            self.trace.append((its, est, rel))
            total += its
            _lib.call('tf_refine_update', n, ptr(d32), ptr(d), ptr(x), st)   # x += (double) d32
            outer += 1
            # the limit is for 'Inner Tolerance': a tighter last step may take
            # proportionally more iterations without the factor being staler.  A
            # preconditioner that is not exact even when fresh (the two-grid method)
            # is judged against 1.5 times what it needed on its own Jacobian.
            if jac.prec_exact and outer == 1 and itol == opts['Inner Tolerance']:
                prec.fresh_inner_its = its
            base = max(opts['Reuse Limit'], 1.5 * getattr(prec, 'fresh_inner_its', 0))
            limit = base * math.log(itol) / math.log(opts['Inner Tolerance'])
            # (a lagged factor whose inner solve ran out of iterations short of the
            # inner tolerance is stale too, also when the limit is above the cap)
            capped = its >= inner_max and est > itol and itol == opts['Inner Tolerance']
            if not jac.prec_exact and (its > limit or capped):
                # lagged factor too stale: refactor this Jacobian and keep refining
                # (in the background: with the lagged factor until the new one is ready)
                self.limit_refactors += 1
                if bg is not None:
                    bg.start(jac)
                else:
                    prec, fresh = self._refactor_now(jac, opts)
                    prev = math.inf
        if outer >= opts['Maximum Refinements'] or total >= opts['Maximum Iterations']:
            jac.matvec(x, out=r, gauge=g)
            _lib.call('tf_axpby_f64', n, 1.0, ptr(b), -1.0, ptr(r), ptr(r), st)
            rel = _norm(r, work) / bnrm
        self.refinements.append(outer)
        return total, rel, prec, fresh

    def solve_pair(self, jac, rhs0, rhs1):
        """(y0, y1) with (gauged J) y = rhs for two right-hand sides and one Jacobian:
        the two solves of a Newton-corrector iteration (Continuation.py:158-159).
        Solved together where the preconditioner is the native line factor: the two
        FGMRES runs advance in lockstep and every preconditioner application
        serves both (tf_fgmres2 / tf_bcr_solve2_*: one pass over the factor
        blocks for two vectors); in mixed precision the FP64 refinement of the two
        runs in lockstep too.  Each result meets 'Tolerance' as a single solve
        does, and each is logged as a solve.  Otherwise: two solves."""
        n = self.itf.n
        T = torch()
        for r in (rhs0, rhs1):
            if not isinstance(r, DeviceVector) or r.size != n:
                raise ValueError('solve_pair needs two DeviceVectors of %d entries' % n)
        opts = _opts(self.itf.parameters)
        mixed = opts['Precision'] == 'mixed'
        # the learned correction runs in the single solves until it has switched itself off
        learned = opts['Learned Network'] is not None and not (self._learned is not None and self._learned.off)
        if (not opts['Pair Solves'] or learned or self.observer is not None
                or self._background(opts) is not None or (not mixed and opts['Reuse Factor'])):
            return self.solve(jac, rhs0), self.solve(jac, rhs1)
        prec, fresh = self.preconditioner(jac, opts)
        if not hasattr(prec, 'native') or prec.native()[0] != 1:
            y0 = self.solve(jac, rhs0)
            if fresh:   # the factor made above served this solve
                self.log[-1] = self.log[-1][:3] + (True,)
            return y0, self.solve(jac, rhs1)
        g = self.gauge()
        bs = [rhs0.data.clone(), rhs1.data.clone()]
        xs = [T.empty_like(bs[0]), T.empty_like(bs[0])]
        for b, x in zip(bs, xs):
            if g >= 0:
                _lib.call('tf_set_entry_f64', ptr(b), b.numel(), g, 0.0, stream())
            _lib.call('tf_fill_f64', x.numel(), 0.0, ptr(x), stream())
        res = (self._solve_mixed_pair if mixed else self._solve_fp64_pair)(jac, bs, xs, opts, prec, fresh)
        for its, rel, p, fr in res:
            self.log.append((its, rel, getattr(p, 'name', '?'), fr))
        for its, rel, _, _ in res:
            if not rel <= opts['Tolerance'] and opts['Fail On Tolerance']:
                raise SolverError('GMRES did not reach the tolerance: relative residual %.3e after %d iterations'
                                  % (rel, its))
        return DeviceVector(xs[0]), DeviceVector(xs[1])

    def _solve_fp64_pair(self, jac, bs, xs, opts, prec, fresh):
        g = self.gauge()
        k0 = self._krylov_for(opts['Restart'], native=True)
        k1 = self._krylov_for(opts['Restart'], native=True, slot=1)
        (i0, r0), (i1, r1) = k0.solve2(k1, jac, jac.val, g, prec, bs[0], bs[1], xs[0], xs[1], opts['Tolerance'],
                                       opts['Tolerance'], opts['Maximum Iterations'], x_is_zero=True)
        return [(i0, r0, prec, fresh), (i1, r1, prec, False)]

    def _pair_buffers(self, n):
        T = torch()
        if getattr(self, '_buf2', None) is None or self._buf2[0][0].numel() != n:
            self._buf2 = tuple((T.empty(n, dtype=T.float64, device='cuda'), T.empty(n, dtype=T.float32, device='cuda'),
                                T.empty(n, dtype=T.float32, device='cuda'), T.empty(n, dtype=T.float64, device='cuda'))
                               for _ in range(2))
        return self._buf2

    def _solve_mixed_pair(self, jac, bs, xs, opts, prec, fresh):
        """The FP64 iterative refinement of _solve_mixed for two right-hand sides in
        lockstep: each refinement round computes both FP64 residuals, solves both
        corrections with one tf_fgmres2 (FP32, lagged FP32 factor), and applies
        the factor-reuse rules once for the pair (a refactorisation serves both)."""
        n = self.itf.n
        g = self.gauge()
        st = stream()
        bufs = self._pair_buffers(n)
        k0 = self._krylov_for(opts['Restart'], 'float32', native=True)
        k1 = self._krylov_for(opts['Restart'], 'float32', native=True, slot=1)
        work = self._krylov_for(opts['Restart']).work
        val32 = jac.values32()
        factored = 1 if fresh else 0   # factorisations made for this pair (logged with the first column)
        tol, itol0 = opts['Tolerance'], opts['Inner Tolerance']
        inner_max = min(opts['Restart'], opts['Inner Iterations'])
        cols = []
        for c in range(2):
            bnrm = _norm(bs[c], work)
            cols.append({'bnrm': bnrm, 'total': 0, 'outer': 0, 'rel': 1.0 if bnrm > 0 else 0.0, 'prev': math.inf,
                         'done': bnrm == 0.0, 'prec': prec})
        while True:
            live = [c for c in range(2) if not cols[c]['done']]
            for c in live:
                C = cols[c]
                if C['outer'] >= opts['Maximum Refinements'] or C['total'] >= opts['Maximum Iterations']:
                    C['done'] = True
            live = [c for c in range(2) if not cols[c]['done']]
            if not live:
                break
            refactor = False
            # r = b - A x, r32 = (float) r, d32 = 0 for the columns past their first round
            need = [c for c in live if cols[c]['outer'] > 0]
            norms = self._refine_residuals(jac, g, [(xs[c], bs[c], bufs[c]) for c in need], work) if need else []
            for c, nrm in zip(need, norms):
                C = cols[c]
                C['rel'] = nrm / C['bnrm']
                if C['rel'] <= tol:
                    C['done'] = True
                    continue
                if C['rel'] > 0.5 * C['prev'] and not jac.prec_exact:
                    refactor = True     # the lagged factor no longer contracts
                elif C['rel'] > 0.5 * C['prev'] and jac.prec_exact:
                    # stagnation with this Jacobian's own FP32 factor: finish in FP64
                    its64, C['rel'], p64, _ = self._fp64_fallback(jac, bs[c], xs[c], opts)
                    C['total'] += its64
                    C['prec'] = p64
                    self.fallbacks += 1
                    C['done'] = True
                    continue
                C['prev'] = C['rel']
            live = [c for c in range(2) if not cols[c]['done']]
            if not live:
                break
            if refactor:
                self.stale_refactors += 1
                prec, fr = self._refactor_now(jac, opts)
                factored += 1 if fr else 0
                for c in live:
                    cols[c]['prev'] = math.inf
                    cols[c]['prec'] = prec
            itols = {}
            for c in live:
                C = cols[c]
                r, r32, d32, d = bufs[c]
                if C['outer'] == 0:
                    _lib.call('tf_f64_to_f32', n, ptr(bs[c]), ptr(r32), st)
                    _lib.call('tf_fill_f32', n, 0.0, ptr(d32), st)
                itol = itol0
                if C['outer'] > 0:
                    need = 0.5 * tol / C['rel']
                    if INNER_FLOOR <= need < itol:
                        itol = need
                itols[c] = itol
            if len(live) == 2:
                (i0, e0), (i1, e1) = k0.solve2(k1, jac, val32, g, prec, bufs[0][1], bufs[1][1], bufs[0][2],
                                               bufs[1][2], itols[0], itols[1], inner_max, x_is_zero=True)
                inner = {0: (i0, e0), 1: (i1, e1)}
            else:
                c = live[0]
                inner = {c: k0.solve(jac, val32, g, prec, bufs[c][1], bufs[c][2], itols[c], inner_max,
                                     x_is_zero=True)}
            fresh_its = []
            for c in live:
                C = cols[c]
                r, r32, d32, d = bufs[c]
                its, est = inner[c]
                self.trace.append((its, est, C['rel']))
                C['total'] += its
                _lib.call('tf_refine_update', n, ptr(d32), ptr(d), ptr(xs[c]), st)   # x += (double) d32
                C['outer'] += 1
                if jac.prec_exact and C['outer'] == 1 and itols[c] == itol0:
                    fresh_its.append(its)
            if fresh_its:
                prec.fresh_inner_its = max(fresh_its)
            limit_hit = False
            for c in live:
                base = max(opts['Reuse Limit'], 1.5 * getattr(prec, 'fresh_inner_its', 0))
                limit = base * math.log(itols[c]) / math.log(itol0)
                capped = inner[c][0] >= inner_max and inner[c][1] > itols[c] and itols[c] == itol0
                if not jac.prec_exact and (inner[c][0] > limit or capped):
                    limit_hit = True
            if limit_hit:
                self.limit_refactors += 1
                prec, fr = self._refactor_now(jac, opts)
                factored += 1 if fr else 0
                for c in live:
                    cols[c]['prev'] = math.inf
                    cols[c]['prec'] = prec
        out = []
        for c in range(2):
            C = cols[c]
            r, _, _, _ = bufs[c]
            if C['bnrm'] > 0 and not C['rel'] <= tol:
                jac.matvec(xs[c], out=r, gauge=g)
                _lib.call('tf_axpby_f64', n, 1.0, ptr(bs[c]), -1.0, ptr(r), ptr(r), st)
                C['rel'] = _norm(r, work) / C['bnrm']
            self.refinements.append(C['outer'])
            out.append((C['total'], C['rel'], C['prec'], factored if c == 0 else 0))
        return out

    def solve_complex(self, A, rhs):
        """y with (gauged A) y = rhs for a complex matrix A = Ar + i Ai (the shifted
        beta J - alpha M of the eigenvalue glue, JaDa.py:146-151, which the
        reference solves with a complex splu, SciPy.py:204-258): row and column
        ``pressure_row`` of A become -e and the rhs's ``pressure_row`` entry is
        zeroed.

        Solved in real arithmetic on the interleaved form K (unknown c becomes
        the pair (re, im), ``tf_csr_interleave_*``): K keeps the grid-line block
        structure with twice the rows per line, so the exact line factorisation
        (block cyclic reduction) factors it in FP64 and FGMRES needs one or two
        iterations.  When that factor does not fit the memory budget the solve
        runs on [[Ar, -Ai], [Ai, Ar]] preconditioned block-diagonally by the
        factor of Ar (slower as |Ai| grows).  rhs of shape (n,) or (n, k);
        returns a complex numpy array of the same shape."""
        T = torch()
        n = self.itf.n
        b = numpy.asarray(rhs.numpy() if isinstance(rhs, DeviceVector) else rhs)
        if b.ndim == 2 and b.shape[0] == n:
            return numpy.stack([self.solve_complex(A, b[:, j]) for j in range(b.shape[1])], axis=1)
        if b.ndim != 1 or b.shape[0] != n:
            raise ValueError('rhs must have shape (%d,) or (%d, k), got %s' % (n, n, b.shape))
        opts = _opts(self.itf.parameters)
        g = self.gauge()
        b = numpy.ascontiguousarray(b, dtype=numpy.complex128)
        if g >= 0:
            b[g] = 0.0
        setup = getattr(A, '_real_form', None)    # made once per matrix (the glue reuses its shifted matrices)
        if setup is None:
            setup = A._real_form = self._real_form(A)
        K, factor = setup
        if factor is None:
            return self._solve_complex_blockdiag(A, b, opts)
        b2 = T.from_numpy(b.view(numpy.float64).copy()).cuda()   # (re, im) pairs: the interleaved layout
        x2 = T.zeros(2 * n, dtype=T.float64, device='cuda')
        kry = getattr(self, '_complex_krylov', None)
        if kry is None or kry.m != opts['Restart'] or kry.n != 2 * n:
            kry = self._complex_krylov = NativeKrylov(2 * n, opts['Restart'])
        its, rel = kry.solve(K, K.val, -1, factor, b2, x2, opts['Tolerance'], opts['Maximum Iterations'],
                             x_is_zero=True)
        self.log.append((its, rel, 'complex-line', False))
        if not rel <= opts['Tolerance'] and opts['Fail On Tolerance']:
            raise SolverError('complex GMRES did not reach the tolerance: relative residual %.3e after %d iterations'
                              % (rel, its))
        return x2.cpu().numpy().view(numpy.complex128).copy()

    def _real_form(self, A):
        """(K, factor): the interleaved real form of A with the gauge applied, and
        its exact FP64 line factorisation (None when it does not fit)."""
        T = torch()
        itf = self.itf
        n = itf.n
        g = self.gauge()
        st = stream()
        cnt = T.empty(2 * n, dtype=T.int32, device='cuda')
        rc = T.empty(2 * n + 1, dtype=T.int32, device='cuda')
        nnz = ctypes.c_int()
        Ar, Ai = A.real, A.imag
        _lib.call('tf_csr_interleave_count', n, ptr(Ar.row_ptr), ptr(Ar.col), ptr(Ai.row_ptr), ptr(Ai.col), int(g),
                  ptr(cnt), ptr(rc), ctypes.byref(nnz), st)
        cc = T.empty(max(nnz.value, 1), dtype=T.int32, device='cuda')[:nnz.value]
        vc = T.empty(max(nnz.value, 1), dtype=T.float64, device='cuda')[:nnz.value]
        _lib.call('tf_csr_interleave_fill', n, ptr(Ar.row_ptr), ptr(Ar.col), ptr(Ar.val), ptr(Ai.row_ptr),
                  ptr(Ai.col), ptr(Ai.val), int(g), ptr(rc), ptr(cc), ptr(vc), st)
        K = DeviceCsrMatrix(rc, cc, vc, 2 * n)
        m, ylines = line_layout(itf)
        if 5 * (n // m) * (2 * m) ** 2 * 8 > LINE_BUDGET_BYTES:
            return K, None
        f = LineFactor(itf, K, -1, single=False, kind='line', factor=False,
                       layout=(2 * n, 2 * m, itf.nx, 2 * itf.dof, ylines))
        f.set_graph(_opts(itf.parameters)['Apply Graph'])
        f.refactor(K, -1)
        return K, f

    def _solve_complex_blockdiag(self, A, b, opts):
        """[[Ar, -Ai], [Ai, Ar]] by FGMRES with the factor of Ar on both halves."""
        T = torch()
        n = self.itf.n
        g = self.gauge()
        st = stream()
        Ar, B = A.real, A.imag
        prec = getattr(Ar, 'prec64', None)
        if prec is None or prec.owner is None or prec.owner() is not Ar:
            prec = Ar.prec64 = self._build(Ar, opts['Preconditioner'], single=False)
        tmp = T.empty(n, dtype=T.float64, device='cuda')

        def bmul(v, out):
            B.matvec(v, out=out, gauge=g)
            if g >= 0:
                _lib.call('tf_set_entry_f64', ptr(out), n, g, 0.0, st)   # the gauge row of Ai is zero

        def matvec(v, out):
            vr, vi = v[:n], v[n:]
            Ar.matvec(vr, out=out[:n], gauge=g)
            bmul(vi, tmp)
            _lib.call('tf_axpby_f64', n, 1.0, ptr(out[:n]), -1.0, ptr(tmp), ptr(out[:n]), st)
            Ar.matvec(vi, out=out[n:], gauge=g)
            bmul(vr, tmp)
            _lib.call('tf_axpby_f64', n, 1.0, ptr(out[n:]), 1.0, ptr(tmp), ptr(out[n:]), st)

        def precond(v, out):
            prec.apply(v[:n], out[:n])
            prec.apply(v[n:], out[n:])

        b2 = T.from_numpy(numpy.concatenate([b.real, b.imag])).cuda()
        x2 = T.zeros(2 * n, dtype=T.float64, device='cuda')
        kry = getattr(self, '_complex_krylov_bd', None)
        if kry is None or kry.m != opts['Restart'] or kry.n != 2 * n:
            kry = self._complex_krylov_bd = Krylov(2 * n, opts['Restart'])
        its, rel = kry.solve(matvec, precond, b2, x2, opts['Tolerance'], opts['Maximum Iterations'], x_is_zero=True)
        self.log.append((its, rel, 'complex+' + getattr(prec, 'name', '?'), True))
        if not rel <= opts['Tolerance'] and opts['Fail On Tolerance']:
            raise SolverError('complex GMRES did not reach the tolerance: relative residual %.3e after %d iterations'
                              % (rel, its))
        x = x2.cpu().numpy()
        return x[:n] + 1j * x[n:]

    def line_stats(self):
        """Shape of the line factorisation for this problem (rows per line, lines)."""
        m, ylines = line_layout(self.itf)
        return {'m': m, 'lines': self.itf.n // m, 'ylines': bool(ylines)}

    def solve_bordered(self, jac, rhs, rhs2, V, W, C):
        """Bordered solve of SciPy.direct_solve (SciPy.py:51-129,204-258):

            [J~      s V  ] [y1]   [rhs     ]
            [s W~^T  s^2 C] [z ] = [s rhs2  ],   returns (y1, s z)

        with s = border_scaling, J~ the gauged Jacobian, W~ = W with the gauge
        entry dropped (the border row skips the gauge column) and the gauge
        entry of rhs kept (the bordered right-hand side is not gauged).  Solved
        by block elimination with the same factor: J~ a = rhs, J~ B = s V,
        z = (s^2 C - s W~^T B)^-1 (s rhs2 - s W~^T a), y1 = a - B z."""
        itf = self.itf
        n = itf.n
        s = itf.border_scaling
        g = self.gauge()
        T = torch()

        def columns(M):
            if M is None:
                return None
            if isinstance(M, DeviceVector):
                return [M.data]
            a = numpy.asarray(M, dtype=float)
            if a.ndim == 1:
                a = a[:, None]
            return [T.from_numpy(numpy.ascontiguousarray(a[:, j])).cuda() for j in range(a.shape[1])]

        Vc = columns(V)
        Wc = columns(W) if W is not None else Vc
        k = len(Vc)
        scalar = numpy.ndim(rhs2) == 0
        r2 = numpy.atleast_1d(numpy.asarray(rhs2, dtype=float))
        Cm = numpy.zeros((k, k)) if C is None else numpy.atleast_2d(numpy.asarray(C, dtype=float))
        b = rhs.data if isinstance(rhs, DeviceVector) else T.from_numpy(numpy.ascontiguousarray(rhs, dtype=float)).cuda()
        a = self._solve_device(jac, b, zero_gauge=False)
        Bc = []
        for v in Vc:
            sv = T.empty_like(v)
            _lib.call('tf_scale_f64', n, s, ptr(v), ptr(sv), stream())
            Bc.append(self._solve_device(jac, sv, zero_gauge=False))
        Wt = []
        for w in Wc:
            wt = w.clone()
            if g >= 0:
                _lib.call('tf_set_entry_f64', ptr(wt), wt.numel(), g, 0.0, stream())
            Wt.append(DeviceVector(wt))
        S = s * s * Cm - s * numpy.array([[Wt[i].dot(DeviceVector(Bc[j])) for j in range(k)] for i in range(k)])
        rr = s * r2 - s * numpy.array([Wt[i].dot(DeviceVector(a)) for i in range(k)])
        z = numpy.linalg.solve(S, rr)
        for j in range(k):
            _lib.call('tf_axpby_f64', n, 1.0, ptr(a), -float(z[j]), ptr(Bc[j]), ptr(a), stream())
        y2 = z * s
        return DeviceVector(a), (float(y2[0]) if scalar else y2)
