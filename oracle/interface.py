"""CPU oracle: the SciPy backend Interface.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Restates
``transiflow/interface/SciPy.py`` and the parts of
``transiflow/interface/BaseInterface.py`` on the hot path: vector, rhs,
jacobian (CSC), mass_matrix, the pressure gauge and border of
``compute_bordered_matrix`` (SciPy.py:51-129), the cached sparse LU
(SciPy.py:131-162) and ``solve`` (SciPy.py:204-315).

It is also the CPU baseline that ``bench.py`` times (scipy's SuperLU, which
is single threaded).
"""

import numpy
from scipy import sparse
from scipy.sparse import linalg

from .discretization import Discretization


def _value(V, i, j):
    # SciPy.py:53-60
    if not hasattr(V, 'shape') or len(V.shape) < 1:
        return V
    if len(V.shape) < 2:
        return V[i]
    return V[i, j]


class Interface:
    def __init__(self, parameters, nx, ny, nz=1, dim=None, dof=None,
                 x=None, y=None, z=None, boundary_conditions=None):
        self.discretization = Discretization(parameters, nx, ny, nz, dim, dof, x, y, z,
                                             boundary_conditions)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dim = dim or self.discretization.dim
        self.dof = dof or self.discretization.dof
        self.x = self.discretization.x
        self.y = self.discretization.y
        self.z = self.discretization.z
        self.parameters = parameters
        self.pressure_row = self.dim
        self.border_scaling = 1e-3
        self._lu = None
        self._prec = None

    def vector(self):
        return numpy.zeros(self.nx * self.ny * self.nz * self.dof)

    def vector_from_array(self, array):
        return array

    def array_from_vector(self, vector):
        return vector

    def set_parameter(self, name, value):
        self.discretization.set_parameter(name, value)

    def get_parameter(self, name):
        return self.discretization.get_parameter(name)

    def rhs(self, state):
        return self.discretization.rhs(state)

    def jacobian(self, state):
        jac = self.discretization.jacobian(state)
        n = jac.m
        # The reference lets scipy infer the shape from the largest column
        # (SciPy.py:45); the explicit (n, n) is the same matrix.
        return sparse.csr_matrix((jac.coA, jac.jcoA, jac.begA), shape=(n, n)).tocsc()

    def mass_matrix(self):
        mass = self.discretization.mass_matrix()
        return sparse.csr_matrix((mass.coA, mass.jcoA, mass.begA), (mass.n, mass.n)).tocsc()

    def compute_bordered_matrix(self, jac, V=None, W=None, C=None, fix_pressure_row=False):
        """[J, s V; s W^T, s^2 C] from the CSC Jacobian with, optionally, the
        gauge row/column ``pressure_row`` replaced by -e (SciPy.py:51-129)."""
        jac = sparse.csc_matrix(jac)
        n = jac.shape[0]
        s = self.border_scaling
        border = 0
        dtype = jac.dtype
        if V is not None:
            border = V.shape[1] if len(V.shape) > 1 else 1
            dtype = V.dtype
        if W is None:
            W = V
        if C is None and border:
            C = numpy.zeros((border, border), dtype=jac.data.dtype)

        cols = numpy.repeat(numpy.arange(n), numpy.diff(jac.indptr))
        rows = jac.indices.astype(numpy.int64)
        vals = jac.data.astype(dtype)
        p = self.pressure_row
        if fix_pressure_row:
            keep = (cols != p) & (rows != p)
            rows, cols, vals = rows[keep], cols[keep], vals[keep]
            rows = numpy.append(rows, p)
            cols = numpy.append(cols, p)
            vals = numpy.append(vals, numpy.array([-1.0], dtype=dtype))
        r_l, c_l, v_l = [rows], [cols], [vals]
        for j in range(border):
            cc = numpy.array([i for i in range(n) if not (fix_pressure_row and i == p)])
            r_l.append(numpy.full(len(cc), n + j))
            c_l.append(cc)
            v_l.append(numpy.array([s * _value(W, i, j) for i in cc], dtype=dtype))
        for i in range(border):
            r_l.append(numpy.arange(n))
            c_l.append(numpy.full(n, n + i))
            v_l.append(numpy.array([s * _value(V, j, i) for j in range(n)], dtype=dtype))
            r_l.append(n + numpy.arange(border))
            c_l.append(numpy.full(border, n + i))
            v_l.append(numpy.array([s * s * _value(C, j, i) for j in range(border)], dtype=dtype))
        N = n + border
        return sparse.csc_matrix((numpy.concatenate(v_l),
                                  (numpy.concatenate(r_l), numpy.concatenate(c_l))), shape=(N, N))

    def _compute_factorization(self, jac, V, W, C):
        # SciPy.py:131-162; the LU is cached on the matrix object
        if V is not None and getattr(jac, 'lu', None) is not None and jac.bordered_lu:
            return
        if V is None and getattr(jac, 'lu', None) is not None and not jac.bordered_lu:
            return
        jac.lu = None
        fix = self.dof > self.dim and self.pressure_row is not None
        A = self.compute_bordered_matrix(jac, V, W, C, fix)
        jac.lu = linalg.splu(A)
        jac.bordered_lu = V is not None
        self._lu = jac.lu
        self._prec = linalg.LinearOperator(jac.shape, matvec=self._lu.solve, dtype=jac.dtype)

    def direct_solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        x = rhs.copy()
        if self.dof > self.dim:
            if len(x.shape) < 2:
                x[self.pressure_row] = 0
            else:
                x[self.pressure_row, :] = 0
        if rhs2 is not None:
            # the bordered right-hand side uses the un-gauged rhs (SciPy.py:226-227)
            x = numpy.append(rhs, rhs2 * self.border_scaling)
        self._compute_factorization(jac, V, W, C)
        y = jac.lu.solve(x)
        if jac.bordered_lu:
            border = 1
            if hasattr(rhs2, 'shape') and len(rhs2.shape) > 0:
                border = rhs2.shape[0]
            y1 = y[:-border]
            y2 = y[-border:] * self.border_scaling
            if numpy.isscalar(rhs2):
                y2 = y2.item()
            return y1, y2
        return y

    def iterative_solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        # SciPy.py:260-309 (spilu preconditioned restarted GMRES)
        x = rhs.copy()
        if self.dof > self.dim:
            x[self.pressure_row] = 0
        if rhs2 is not None:
            x = numpy.append(rhs, rhs2 * self.border_scaling)
        fix = self.dof > self.dim and self.pressure_row is not None
        A = self.compute_bordered_matrix(jac, V, W, C, fix)
        p = self.parameters.get('Preconditioner', {})
        lu = linalg.spilu(A, drop_tol=p.get('Drop Tolerance', None),
                          fill_factor=p.get('Fill Factor', None),
                          drop_rule=p.get('Drop Rule', None))
        M = linalg.LinearOperator(A.shape, matvec=lu.solve, dtype=A.dtype)
        ip = self.parameters.get('Iterative Solver', {})
        restart = ip.get('Restart', 100)
        maxit = ip.get('Maximum Iterations', 1000)
        tol = ip.get('Convergence Tolerance', 1e-6)
        y, _info = linalg.gmres(A, x, restart=restart, maxiter=(maxit - 1) // restart + 1,
                                rtol=tol, atol=0, M=M)
        if V is not None:
            border = rhs2.shape[0] if hasattr(rhs2, 'shape') and len(rhs2.shape) > 0 else 1
            return y[:-border], y[-border:] * self.border_scaling
        return y

    def solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        if self.parameters.get('Use Iterative Solver', False):
            return self.iterative_solve(jac, rhs, rhs2, V, W, C)
        return self.direct_solve(jac, rhs, rhs2, V, W, C)
