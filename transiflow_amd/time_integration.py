"""Implicit theta-method time stepping on the HIP backend.

Contract: ``transiflow/TimeIntegration.py:6-113`` (same constructor
arguments, ``integration(x0, dt, tmax, callback)`` returning ``(x, t)``,
``callback(interface, x, t)`` before the first step and after every step).

One step from u to v solves, with the mass matrix M (diagonal, zero on the
pressure rows),

    G(v) = F(v) + (1 - theta)/theta F(u) + M (u - v) / (theta dt) = 0,

which is the reference's step residual divided by theta dt, by Newton's
method with dG/dv = J(v) - M / (theta dt).  G is formed by one fused kernel
(``tf_theta_residual``) from F(v), F(u), diag(M), u and v; F(u) and diag(M)
are computed once per step and once per integration.  The Newton matrix is
the backend's Jacobian minus the scaled mass matrix in the device CSR
arithmetic, so any preconditioner of ``Interface.solve`` applies.  Only the
norms used by the convergence test reach the host.
"""

from . import _lib
from .vector import DeviceVector, ptr, stream, torch


class TimeIntegration:
    def __init__(self, interface, theta=1.0, newton_tolerance=1e-10, maximum_newton_iterations=10,
                 residual_check='F', verbose=False, printer=print):
        if not 0.0 < theta <= 1.0:
            raise ValueError('theta must be in (0, 1], got %r' % theta)
        self.interface = interface
        self.theta = theta
        self.newton_tolerance = newton_tolerance
        self.maximum_newton_iterations = maximum_newton_iterations
        self.residual_check = residual_check
        self.verbose = verbose
        self._print = printer or (lambda *a, **k: None)
        self.newton_iterations = []   # Jacobian solves of every step
        self._mass = None

    def _mass_matrix(self):
        """(M, diag(M)) of the interface, made once."""
        if self._mass is None:
            M = self.interface.mass_matrix()
            d = torch().empty(M.n, dtype=torch().float64, device='cuda')
            _lib.call('tf_csr_diagonal', M.n, ptr(M.row_ptr), ptr(M.col), ptr(M.val), ptr(d), stream())
            self._mass = (M, d)
        return self._mass

    def step(self, u, dt):
        """One theta step of size dt from u; returns the new state (u is not changed)."""
        itf = self.interface
        M, mdiag = self._mass_matrix()
        cm = 1.0 / (self.theta * dt)
        cb = (1.0 - self.theta) / self.theta
        u = u if isinstance(u, DeviceVector) else itf.vector_from_array(u)
        n = u.size
        Fu = itf.rhs(u)
        v = u.copy()
        G = DeviceVector(torch().empty_like(u.data))
        solves = 0
        check_g = self.residual_check == 'F'
        while True:
            Fv = itf.rhs(v)
            _lib.call('tf_theta_residual', n, ptr(Fv.data), ptr(Fu.data), ptr(mdiag), ptr(u.data), ptr(v.data),
                      cb, cm, ptr(G.data), stream())
            gnorm = G.norm() if (check_g or self.verbose) else None
            if check_g and gnorm < self.newton_tolerance:
                self._print('theta step: |G| = %.3e after %d Newton solve(s)' % (gnorm, solves), flush=True)
                break
            if solves == self.maximum_newton_iterations:
                self._print('theta step: Newton stopped after %d solves without reaching %.1e'
                            % (solves, self.newton_tolerance), flush=True)
                break
            dv = itf.solve(itf.jacobian(v) - M * cm, -G)
            v += dv
            solves += 1
            if not check_g or self.verbose:
                dnorm = dv.norm()
                if not check_g and dnorm < self.newton_tolerance:
                    self._print('theta step: |dv| = %.3e after %d Newton solve(s)' % (dnorm, solves), flush=True)
                    break
                if self.verbose:
                    self._print('theta step, solve %d: |G| = %.3e, |dv| = %.3e' % (solves, gnorm, dnorm),
                                flush=True)
        self.newton_iterations.append(solves)
        return v

    def integration(self, x0, dt, tmax, callback=None):
        """Step from t = 0 while t < tmax (t grows by dt each step); returns (x, t)."""
        x, t = x0, 0
        if callback:
            callback(self.interface, x, t)
        while t < tmax:
            x = self.step(x, dt)
            t += dt
            self._print('time %g reached' % t, flush=True)
            if callback:
                callback(self.interface, x, t)
        return x, t
