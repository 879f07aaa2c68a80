"""Newton and pseudo-arclength continuation driver.

Restates ``transiflow/Continuation.py`` (Continuation.py:6-482): the same
Newton iteration, bordered-by-elimination corrector (eq. 2.2.8-2.2.13),
step-size control and secant convergence onto the target, operating on any
interface with ``rhs/jacobian/solve/set_parameter``.  Vectors only need
``size``, ``dot``, arithmetic with scalars and in-place ``+=``/``/=``; the
HIP backend's :class:`~transiflow_amd.vector.DeviceVector` keeps them in HBM
so that only the scalars of the control logic cross to the host.

Bifurcation detection needs an eigenvalue solver; it raises
``NotImplementedError`` here (the reference delegates it to jadapy).
"""

import math

import numpy


def norm(x):
    """2-norm of a state vector (utils.py:13-21)."""
    if hasattr(x, 'norm'):
        return x.norm()
    if len(x.shape) > 1:
        return numpy.array([math.sqrt(abs(x[:, i].T.conj() @ x[:, i])) for i in range(x.shape[1])])
    return math.sqrt(abs(x.T.conj() @ x))


class Continuation:
    """Pseudo-arclength continuation (Continuation.py:47-65 for the arguments).

    ``printer`` receives the progress lines the reference prints to stdout;
    pass ``None`` to silence them.
    """

    def __init__(self, interface, delta=1.0, destination_tolerance=1e-4,
                 newton_tolerance=1e-4, maximum_newton_iterations=10,
                 optimal_newton_iterations=3, residual_check='F',
                 bordered_solver=False, verbose=False, printer=print):
        self.interface = interface
        self.verbose = verbose
        self.destination_tolerance = destination_tolerance
        self.delta = delta
        self.zeta = None
        self.maximum_newton_iterations = maximum_newton_iterations
        self.optimal_newton_iterations = optimal_newton_iterations
        self.newton_tolerance = newton_tolerance
        self.newton_iterations = 0
        self.residual_check = residual_check
        self.bordered_solver = bordered_solver
        self._print = printer or (lambda *a, **k: None)
        # bookkeeping for measurements: every corrector outcome and Newton count
        self.history = []

    def _say(self, msg):
        self._print(msg, flush=True)

    # Continuation.py:67-114
    def newton(self, x0, tol=1e-10):
        x = x0
        fnorm = dxnorm = None
        k = 0
        for k in range(self.maximum_newton_iterations):
            fval = self.interface.rhs(x)
            if self.residual_check == 'F' or self.verbose:
                fnorm = norm(fval)
            if self.residual_check == 'F' and fnorm < tol:
                self._say('Newton converged in %d iterations with ||F||=%e' % (k, fnorm))
                break
            jac = self.interface.jacobian(x)
            dx = self.interface.solve(jac, -fval)
            x = x + dx
            if self.residual_check != 'F' or self.verbose:
                dxnorm = norm(dx)
            if self.residual_check != 'F' and dxnorm < tol:
                self._say('Newton converged in %d iterations with ||dx||=%e' % (k, dxnorm))
                break
            if self.verbose:
                self._say('Newton status at iteration %d: ||F||=%e, ||dx||=%e' % (k, fnorm, dxnorm))
        self.newton_iterations = k
        return x

    # Continuation.py:116-195
    def _newton_corrector(self, parameter_name, ds, x, x0, mu, mu0):
        self.newton_iterations = 0
        fnorm = None
        dxnorm = None
        prev_norm = None
        for k in range(self.maximum_newton_iterations):
            self.interface.set_parameter(parameter_name, mu)
            fval = self.interface.rhs(x)
            if self.residual_check == 'F' or self.verbose:
                prev_norm = fnorm
                fnorm = norm(fval)
            if self.residual_check == 'F' and fnorm < self.newton_tolerance:
                self._say('Newton corrector converged in %d iterations with ||F||=%e' % (k, fnorm))
                break
            if self.residual_check == 'F' and prev_norm is not None and prev_norm < fnorm:
                self.newton_iterations = self.maximum_newton_iterations
                break

            diff = x - x0
            rnp1 = self.zeta * diff.dot(diff) + (1 - self.zeta) * (mu - mu0) ** 2 - ds ** 2

            self.interface.set_parameter(parameter_name, mu + self.delta)
            dflval = (self.interface.rhs(x) - fval) / self.delta
            self.interface.set_parameter(parameter_name, mu)

            jac = self.interface.jacobian(x)

            if self.bordered_solver:
                dx, dmu = self.interface.solve(jac, -fval, -rnp1, dflval, 2 * self.zeta * diff,
                                               2 * (1 - self.zeta) * (mu - mu0))
            else:
                # both right-hand sides are known before either solve: a backend with
                # solve_pair solves them together (one pass over its factor per
                # preconditioner application serves both)
                pair = getattr(self.interface, 'solve_pair', None)
                if pair is not None:
                    z1, z2 = pair(jac, -fval, dflval)
                else:
                    z1 = self.interface.solve(jac, -fval)
                    z2 = self.interface.solve(jac, dflval)
                dmu = (-rnp1 - 2 * self.zeta * diff.dot(z1)) / (2 * (1 - self.zeta) * (mu - mu0)
                                                                - 2 * self.zeta * diff.dot(z2))
                dx = z1 - dmu * z2

            x += dx
            mu += dmu
            self.newton_iterations += 1

            if self.residual_check != 'F' or self.verbose:
                prev_norm = dxnorm
                dxnorm = norm(dx)
            if self.residual_check != 'F' and dxnorm < self.newton_tolerance:
                self._say('Newton corrector converged in %d iterations with ||dx||=%e' % (k, dxnorm))
                break
            if self.verbose:
                self._say('Newton corrector status at iteration %d: ||F||=%e, ||dx||=%e'
                          % (k, fnorm, dxnorm))
            if self.residual_check != 'F' and prev_norm is not None and prev_norm < dxnorm:
                self.newton_iterations = self.maximum_newton_iterations
                break

        if self.newton_iterations == self.maximum_newton_iterations:
            self._say('Newton did not converge. Adjusting step size and trying again')
            self.history.append(('rejected', self.newton_iterations, mu0, ds))
            return x0, mu0

        self.interface.set_parameter(parameter_name, mu)
        self.history.append(('accepted', self.newton_iterations, mu, ds))
        return x, mu

    # Continuation.py:197-209
    def _adjust_step_size(self, ds, ds_min, ds_max):
        factor = self.optimal_newton_iterations / max(self.newton_iterations, 1)
        factor = min(max(factor, 0.5), 2.0)
        ds *= factor
        ds = math.copysign(min(max(abs(ds), ds_min), ds_max), ds)
        if self.verbose:
            self._say('New stepsize: ds=%e, factor=%e' % (ds, factor))
        return ds

    # Continuation.py:234-249
    def _converge(self, parameter_name, x, mu, dx, dmu, target, ds, ds_min, ds_max, maxit):
        for _ in range(maxit):
            if abs(target - mu) < self.destination_tolerance:
                self.interface.set_parameter(parameter_name, target)
                self._say('Convergence achieved onto target %s = %f' % (parameter_name, mu))
                break
            ds = 1 / dmu * (target - mu)
            x, mu, dx, dmu, ds = self._step(parameter_name, x, mu, dx, dmu, ds, ds_min, ds_max)
        return x, mu

    # Continuation.py:251-288
    def _step(self, parameter_name, x, mu, dx, dmu, ds, ds_min, ds_max):
        mu0 = mu
        x0 = x
        mu = mu0 + ds * dmu
        x = x0 + ds * dx
        x, mu = self._newton_corrector(parameter_name, ds, x, x0, mu, mu0)
        if mu == mu0:
            prev_ds = ds
            ds = self._adjust_step_size(ds, ds_min, ds_max)
            if prev_ds == ds:
                raise Exception('Newton cannot achieve convergence')
            return self._step(parameter_name, x0, mu0, dx, dmu, ds, ds_min, ds_max)

        self._say('%s: %f' % (parameter_name, mu))
        dmu = mu - mu0
        dx = x - x0
        if abs(dmu) < 1e-12:
            raise Exception('dmu too small')
        dx /= ds
        dmu /= ds
        return x, mu, dx, dmu, ds

    # Continuation.py:330-350
    def _initial_tangent(self, x, parameter_name, mu):
        self.interface.set_parameter(parameter_name, mu + self.delta)
        dflval = self.interface.rhs(x)
        self.interface.set_parameter(parameter_name, mu)
        fval = self.interface.rhs(x)
        dflval = (dflval - fval) / self.delta
        jac = self.interface.jacobian(x)
        dx = self.interface.solve(jac, -dflval)
        dmu = 1
        nrm = math.sqrt(self.zeta * dx.dot(dx) + dmu ** 2)
        dmu /= nrm
        dx /= nrm
        return dx, dmu

    # Continuation.py:352-359
    def _return(self, x, mu, dx, dmu, ds, callback, return_step):
        if callback is not None:
            callback(self.interface, x, mu)
        if return_step:
            return x, mu, dx * ds, dmu * ds
        return x, mu

    # Continuation.py:361-482
    def continuation(self, x0, parameter_name, start, target, ds, ds_min=0.01, ds_max=1000,
                     dx=None, dmu=None, maxit=1000, detect_bifurcations=False,
                     switch_branches=False, return_step=False, callback=None):
        if detect_bifurcations or switch_branches:
            raise NotImplementedError('bifurcation detection needs an eigenvalue solver')
        x = x0
        mu = start
        self.zeta = 1 / x.size
        if dx is None or dmu is None:
            dx, dmu = self._initial_tangent(x, parameter_name, mu)
        else:
            dx /= ds
            dmu /= ds
        for j in range(maxit):
            mu0 = mu
            if callback is not None:
                callback(self.interface, x, mu)
            x, mu, dx, dmu, ds = self._step(parameter_name, x, mu, dx, dmu, ds, ds_min, ds_max)
            if (mu >= target and mu0 < target) or (mu <= target and mu0 > target):
                x, mu = self._converge(parameter_name, x, mu, dx, dmu, target,
                                       ds, ds_min, ds_max, maxit - j)
                return self._return(x, mu, dx, dmu, ds, callback, return_step)
            ds = self._adjust_step_size(ds, ds_min, ds_max)
        return self._return(x, mu, dx, dmu, ds, callback, return_step)
