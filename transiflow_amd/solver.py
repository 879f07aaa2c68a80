"""Linear solves of the HIP backend: preconditioned restarted (F)GMRES.

``Interface.solve(jac, rhs)`` solves the pressure-gauged system of the
reference's direct solve (``transiflow/interface/SciPy.py:204-258``): the
right-hand side's ``pressure_row`` entry is zeroed, row and column
``pressure_row`` of J are replaced by -e (SciPy.py:95-103), and the result
satisfies ``||b - A x|| <= tol ||b||`` with tol = 1e-10 by default (the
reference's splu solve is matched to that relative residual).

Options (``parameters['HIP Solver']``, all optional):

``'Preconditioner'``   ``'line'`` (exact block-tridiagonal factorisation over
                       grid lines by block cyclic reduction, default when the
                       dense line blocks fit the memory budget), ``'thomas'``
                       (the same factorisation by the sequential block Thomas
                       recursion), ``'block-jacobi'`` (cell blocks);
``'Precision'``        ``'fp64'`` (default) or ``'mixed'``: the preconditioner is
                       stored in FP32 and reused across Newton iterations while
                       GMRES converges quickly; GMRES and the residual stay FP64;
``'Tolerance'``        relative residual target (1e-10);
``'Restart'``          Krylov basis size (30);
``'Maximum Iterations'`` (300);
``'Reuse Limit'``      (mixed) GMRES iterations above which a lagged
                       preconditioner is refactored (8).

Every solve appends (iterations, relative residual, preconditioner name,
factorised?) to ``Solver.log`` for measurement.
"""

import ctypes
import math
import time

import numpy

from . import _lib
from .vector import DeviceVector, ptr, stream, torch

LINE_BUDGET_BYTES = 64 * 2 ** 30


class SolverError(RuntimeError):
    pass


def _opts(parameters):
    o = dict(parameters.get('HIP Solver', {}))
    return {
        'Preconditioner': o.get('Preconditioner', 'auto'),
        'Precision': o.get('Precision', 'fp64'),
        'Tolerance': float(o.get('Tolerance', 1e-10)),
        'Restart': int(o.get('Restart', 30)),
        'Maximum Iterations': int(o.get('Maximum Iterations', 300)),
        'Reuse Limit': int(o.get('Reuse Limit', 8)),
    }


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


class LineFactor:
    """Exact factorisation of the gauged Jacobian, block tridiagonal over grid
    lines: block cyclic reduction (``'line'``, csrc/bcr.hip) or the sequential
    Thomas recursion (``'thomas'``, csrc/linelu.hip)."""
    _ABI = {'line': 'tf_bcr', 'thomas': 'tf_linelu'}

    def __init__(self, itf, jac, gauge, single=False, kind='line'):
        m = line_block_size(itf)
        self.name = kind
        self._p = self._ABI[kind]
        h = ctypes.c_void_p()
        _lib.call(self._p + '_create', ctypes.byref(h), itf.n, m, 1 if single else 0)
        self._h = h
        self.single = single
        self.m = m
        self.refactor(jac, gauge)

    def refactor(self, jac, gauge):
        _lib.call(self._p + '_factor', self._h, ptr(jac.row_ptr), ptr(jac.col), ptr(jac.val), int(gauge), stream())
        flags = (ctypes.c_int * 2)()
        _lib.call(self._p + '_status', self._h, flags, stream())
        if flags[0]:
            raise SolverError('line factorisation: matrix is not block tridiagonal over grid lines (flags %d)'
                              % flags[0])
        self.perturbed = flags[1]

    def apply(self, r, out):
        _lib.call(self._p + '_solve_f64', self._h, ptr(r), ptr(out), stream())
        return out

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value:
            try:
                getattr(_lib.load(), self._p + '_destroy')(h)
            except Exception:
                pass
            self._h = None


def line_block_size(itf):
    """Rows of one grid line (2D) or grid plane (3D)."""
    if itf.nz == 1:
        return itf.nx * itf.dof
    return itf.nx * itf.ny * itf.dof


def line_fits(itf, single=False):
    """Block cyclic reduction keeps five dense m x m blocks per line (D, L, U and
    the two elimination multipliers)."""
    m = line_block_size(itf)
    nlines = itf.n // m
    return 5 * nlines * m * m * (4 if single else 8) <= LINE_BUDGET_BYTES


class Krylov:
    """FP64 restarted flexible GMRES with CGS2 on device-resident vectors."""

    def __init__(self, n, restart):
        T = torch()
        self.n = n
        self.m = restart
        self.V = T.empty((restart + 1, n), dtype=T.float64, device='cuda')
        self.Z = T.empty((restart, n), dtype=T.float64, device='cuda')
        self.w = T.empty(n, dtype=T.float64, device='cuda')
        self.h = T.zeros(restart + 2, dtype=T.float64, device='cuda')
        self.y = T.empty(restart + 1, dtype=T.float64, device='cuda')
        nw = _lib.load().tf_krylov_work_size()
        self.work = T.empty(nw, dtype=T.float64, device='cuda')
        self.h_host = T.empty(restart + 2, dtype=T.float64).pin_memory()

    def solve(self, matvec, precond, b, x, tol, maxit):
        """Solve A x = b (x holds the initial guess, updated in place).
        Returns (iterations, relative residual)."""
        n, m = self.n, self.m
        st = stream()
        V, Z, w, h = self.V, self.Z, self.w, self.h
        lib = _lib
        bnrm = _norm(b, self.work)
        if bnrm == 0.0:
            lib.call('tf_fill_f64', n, 0.0, ptr(x), st)
            return 0, 0.0
        its = 0
        rel = math.inf
        first = True
        while True:
            # r = b - A x into V[0]
            matvec(x, w)
            lib.call('tf_axpby_f64', n, 1.0, ptr(b), -1.0, ptr(w), ptr(V[0]), st)
            beta = _norm(V[0], self.work)
            rel = beta / bnrm
            if rel <= tol or its >= maxit:
                return its, rel
            if not first and beta == 0.0:
                return its, rel
            first = False
            lib.call('tf_scale_f64', n, 1.0 / beta, ptr(V[0]), ptr(V[0]), st)
            H = numpy.zeros((m + 1, m))
            cs = numpy.zeros(m)
            sn = numpy.zeros(m)
            g = numpy.zeros(m + 1)
            g[0] = beta
            k = 0
            for j in range(m):
                precond(V[j], Z[j])
                matvec(Z[j], w)
                lib.call('tf_multidot_f64', n, j + 1, ptr(V), n, ptr(w), ptr(self.work), ptr(h), 0, st)
                lib.call('tf_multiaxpy_f64', n, j + 1, ptr(V), n, ptr(h), ptr(w), -1.0, st)
                lib.call('tf_multidot_f64', n, j + 1, ptr(V), n, ptr(w), ptr(self.work), ptr(self.y), 0, st)
                lib.call('tf_multiaxpy_f64', n, j + 1, ptr(V), n, ptr(self.y), ptr(w), -1.0, st)
                lib.call('tf_axpby_f64', j + 1, 1.0, ptr(h), 1.0, ptr(self.y), ptr(h), st)
                lib.call('tf_normalize_f64', n, ptr(w), ptr(self.work), ctypes.c_void_p(h.data_ptr() + 8 * (j + 1)),
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
            lib.call('tf_multiaxpy_f64', n, k, ptr(Z), n, ptr(self.y), ptr(x), 1.0, st)


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


class Solver:
    def __init__(self, itf):
        self.itf = itf
        self.log = []
        self._krylov = None
        self._lagged = None   # (precond) reused across Jacobians in mixed precision

    def _krylov_for(self, restart):
        if self._krylov is None or self._krylov.m != restart:
            self._krylov = Krylov(self.itf.n, restart)
        return self._krylov

    def gauge(self):
        itf = self.itf
        return itf.pressure_row if (itf.dof > itf.dim and itf.pressure_row is not None) else -1

    def _build(self, jac, kind, single):
        g = self.gauge()
        if kind == 'auto':
            kind = 'line' if line_fits(self.itf, single) else 'block-jacobi'
        if kind in ('line', 'thomas'):
            return LineFactor(self.itf, jac, g, single=single, kind=kind)
        if kind == 'block-jacobi':
            return CellBlockJacobi(self.itf, jac, g)
        raise ValueError('unknown preconditioner %r' % kind)

    def preconditioner(self, jac, opts):
        mixed = opts['Precision'] == 'mixed'
        if jac.prec is not None:
            return jac.prec, False
        if mixed and self._lagged is not None:
            jac.prec = self._lagged
            return jac.prec, False
        t0 = time.perf_counter()
        jac.prec = self._build(jac, opts['Preconditioner'], single=mixed)
        jac.prec_time = time.perf_counter() - t0
        if mixed:
            self._lagged = jac.prec
        return jac.prec, True

    def solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        if rhs2 is not None or V is not None:
            return self.solve_bordered(jac, rhs, rhs2, V, W, C)
        itf = self.itf
        opts = _opts(itf.parameters)
        T = torch()
        b_vec = rhs if isinstance(rhs, DeviceVector) else DeviceVector.from_array(rhs)
        g = self.gauge()
        b = b_vec.data.clone()
        if g >= 0:
            _lib.call('tf_set_entry_f64', ptr(b), g, 0.0, stream())   # SciPy.py:211-216
        x = T.empty_like(b)
        _lib.call('tf_fill_f64', x.numel(), 0.0, ptr(x), stream())
        prec, fresh = self.preconditioner(jac, opts)
        kry = self._krylov_for(opts['Restart'])

        def matvec(v, out):
            jac.matvec(v, out=out, gauge=g)

        its, rel = kry.solve(matvec, prec.apply, b, x, opts['Tolerance'], opts['Maximum Iterations'])
        if opts['Precision'] == 'mixed' and not fresh and (its > opts['Reuse Limit'] or rel > opts['Tolerance']):
            # lagged factor too stale: refactor this Jacobian and finish from the current iterate
            jac.prec = None
            self._lagged = None
            prec, fresh = self.preconditioner(jac, opts)
            its2, rel = kry.solve(matvec, prec.apply, b, x, opts['Tolerance'], opts['Maximum Iterations'])
            its += its2
        self.log.append((its, rel, getattr(prec, 'name', '?'), fresh))
        if not rel <= opts['Tolerance']:
            raise SolverError('GMRES did not reach the tolerance: relative residual %.3e after %d iterations'
                              % (rel, its))
        return DeviceVector(x)

    def line_stats(self):
        """Shape of the line factorisation for this problem (rows per line, lines)."""
        m = line_block_size(self.itf)
        return {'m': m, 'lines': self.itf.n // m}

    def solve_bordered(self, jac, rhs, rhs2, V, W, C):
        raise NotImplementedError('bordered solves are not available yet in the HIP backend')
