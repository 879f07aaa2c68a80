"""CPU oracle: theta-method time integration.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Restates
``transiflow/TimeIntegration.py:40-113`` on numpy vectors and the oracle
Interface: each step solves

    M (u - v) + dt theta F(v) + dt (1 - theta) F(u) = 0

for v by Newton's method with the matrix J(v) - M / (theta dt) and the
gauged sparse LU of ``Interface.solve``; ``integration`` advances t by dt
while t < tmax.  ``newton_iterations`` records the Newton solves of every
step (the reference keeps no such record).

Pinned by tests/test_oracle.py without the reference: a steady state is a
fixed point of a step, backward Euler (theta = 1) and Crank-Nicolson
(theta = 1/2) converge at first and second order in dt on a linear (Stokes)
problem, and a long integration ends at the steady Newton solution.
"""

import numpy


class TimeIntegration:
    def __init__(self, interface, theta=1.0, newton_tolerance=1e-10, maximum_newton_iterations=10):
        self.interface = interface
        self.theta = theta
        self.newton_tolerance = newton_tolerance
        self.maximum_newton_iterations = maximum_newton_iterations
        self.newton_iterations = []

    def step(self, u, dt):
        itf = self.interface
        th = self.theta
        mass = itf.mass_matrix()
        Fu = itf.rhs(u)
        v = numpy.array(u, dtype=float)
        solves = 0
        while True:
            g = (mass @ (u - v) + dt * th * itf.rhs(v) + dt * (1 - th) * Fu) / (th * dt)
            if numpy.linalg.norm(g) < self.newton_tolerance or solves == self.maximum_newton_iterations:
                break
            v = v + itf.solve(itf.jacobian(v) - mass / (th * dt), -g)
            solves += 1
        self.newton_iterations.append(solves)
        return v

    def integration(self, x0, dt, tmax):
        x, t = numpy.array(x0, dtype=float), 0
        while t < tmax:
            x = self.step(x, dt)
            t += dt
        return x, t
