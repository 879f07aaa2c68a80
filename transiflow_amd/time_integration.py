"""Theta-method time integration on the HIP backend.

Restates ``transiflow/TimeIntegration.py:6-113``: every time step solves

    M (u_n - u_{n+1}) + dt theta F(u_{n+1}) + dt (1 - theta) F(u_n) = 0

by Newton's method with the Jacobian J(u) - M / (theta dt), using the
interface's ``rhs``, ``jacobian``, ``mass_matrix`` and ``solve``.  The
vectors and matrices stay in HBM (``DeviceVector``, ``DeviceCsrMatrix``
with device sparse addition and scaling); only the norms reach the host.
"""

from .continuation import norm


class TimeIntegration:
    def __init__(self, interface, theta=1.0, newton_tolerance=1e-10, maximum_newton_iterations=10,
                 residual_check='F', verbose=False, printer=print):
        self.interface = interface
        self.verbose = verbose
        self.theta = theta
        self.maximum_newton_iterations = maximum_newton_iterations
        self.newton_tolerance = newton_tolerance
        self.residual_check = residual_check
        self._print = printer or (lambda *a, **k: None)
        self.newton_iterations = []

    # TimeIntegration.py:38-76
    def _newton(self, x0, dt):
        x = x0
        b0 = self.interface.rhs(x0)
        mass = self.interface.mass_matrix()
        k = 0
        for k in range(self.maximum_newton_iterations):
            fval = mass @ (x0 - x) + dt * self.theta * self.interface.rhs(x) + dt * (1 - self.theta) * b0
            fval /= self.theta * dt
            if self.residual_check == 'F' or self.verbose:
                fnorm = norm(fval)
            if self.residual_check == 'F' and fnorm < self.newton_tolerance:
                self._print('Newton converged i %d iterations with ||F||=%e' % (k, fnorm), flush=True)
                break
            jac = self.interface.jacobian(x) - mass / (self.theta * dt)
            dx = self.interface.solve(jac, -fval)
            x = x + dx
            if self.residual_check != 'F' or self.verbose:
                dxnorm = norm(dx)
            if self.residual_check != 'F' and dxnorm < self.newton_tolerance:
                self._print('Newton converged in %d iterations with ||dx||=%e' % (k, dxnorm), flush=True)
                break
            if self.verbose:
                self._print('Newton status at iteration %d: ||F||=%e, ||dx||=%e' % (k, fnorm, dxnorm), flush=True)
        self.newton_iterations.append(k)
        return x

    # TimeIntegration.py:78-113
    def integration(self, x0, dt, tmax, callback=None):
        x = x0
        t = 0
        if callback:
            callback(self.interface, x, t)
        while t < tmax:
            x = self._newton(x, dt)
            t += dt
            self._print('t = %f' % t, flush=True)
            if callback:
                callback(self.interface, x, t)
        return x, t
