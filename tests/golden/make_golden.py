"""Write golden vectors for the parts of the hot path that no reference fixture pins.

Run once, in the development container where the reference checkout exists:

    python tests/golden/make_golden.py /root/reference

It imports the reference package (pure Python, numpy/scipy only) and stores
inputs and outputs as small ``.npz`` files next to this script.  Nothing in
the test suite, the benchmark or the product imports the reference; they read
only these ``.npz`` files.

Cases (sizes kept small so the files stay a few hundred KB):

* ``ldc2d_16.npz``     -- LDC 2D 16x16, Re=100: F and J (CSR) at the state
  ``default_rng(0).uniform(-1, 1, 768)``, and Newton from zero (tol 1e-8).
* ``rb2d_16x8.npz``    -- Rayleigh-Benard 2D 16x8, dof 4, Pr=1, Ra=1000,
  Biot=1, X-max=4 (the heat flux with Biot path has no reference fixture):
  F and J at ``default_rng(1).uniform(-1, 1)``.
* ``ldc3d_semi.npz``   -- LDC with dim=3 on nz=1 (duplicate columns merged by
  the z wrap): F and J at ``default_rng(3).uniform(-1, 1)``.
* ``cont_ldc2d_8.npz`` -- pseudo-arclength continuation of LDC 2D 8x8 from
  Re 0 to 2000 with ds=100 and Newton tolerance 1e-8: the accepted mu
  sequence, Newton iterations per accepted step and the final state.
"""
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))


def csr_arrays(jac):
    nnz = jac.begA[-1]
    return dict(begA=numpy.asarray(jac.begA, dtype=numpy.int64),
                jcoA=numpy.asarray(jac.jcoA[:nnz], dtype=numpy.int64),
                coA=numpy.asarray(jac.coA[:nnz], dtype=float))


def main(reference_root):
    sys.path.insert(0, reference_root)
    from transiflow import Discretization, Continuation
    from transiflow.interface import create

    # LDC 2D 16x16 at a seeded state, and Newton from zero at Re=100
    params = {'Reynolds Number': 100}
    disc = Discretization(params, 16, 16, 1, 2, 3)
    state = numpy.random.default_rng(0).uniform(-1, 1, 16 * 16 * 3)
    F = disc.rhs(state)
    J = disc.jacobian(state)
    interface = create({'Reynolds Number': 100}, 16, 16, 1, 2, 3)
    cont = Continuation(interface, newton_tolerance=1e-8)
    x_newton = cont.newton(interface.vector(), tol=1e-8)
    numpy.savez_compressed(os.path.join(HERE, 'ldc2d_16.npz'), state=state, F=F,
                           x_newton=x_newton, **csr_arrays(J))

    # Rayleigh-Benard 2D with Biot number and aspect ratio
    params = {'Problem Type': 'Rayleigh-Benard', 'Prandtl Number': 1.0,
              'Rayleigh Number': 1000.0, 'Biot Number': 1.0, 'X-max': 4.0,
              'Reynolds Number': 1.0}
    disc = Discretization(params, 16, 8, 1, 2, 4)
    state = numpy.random.default_rng(1).uniform(-1, 1, 16 * 8 * 4)
    numpy.savez_compressed(os.path.join(HERE, 'rb2d_16x8.npz'), state=state,
                           F=disc.rhs(state), **csr_arrays(disc.jacobian(state)))

    # semi-2D: dim 3 on a single z layer
    disc = Discretization({'Reynolds Number': 50}, 6, 5, 1, 3, 4)
    state = numpy.random.default_rng(3).uniform(-1, 1, 6 * 5 * 4)
    numpy.savez_compressed(os.path.join(HERE, 'ldc3d_semi.npz'), state=state,
                           F=disc.rhs(state), **csr_arrays(disc.jacobian(state)))

    # continuation trajectory LDC 2D 8x8, Re 0 -> 2000
    interface = create({}, 8, 8, 1, 2, 3)
    cont = Continuation(interface, newton_tolerance=1e-8)
    x0 = cont.newton(interface.vector())
    mus, its = [], []

    def callback(_interface, _x, mu):
        mus.append(mu)
        its.append(cont.newton_iterations)

    x, mu = cont.continuation(x0, 'Reynolds Number', 0, 2000, 100, callback=callback)
    numpy.savez_compressed(os.path.join(HERE, 'cont_ldc2d_8.npz'), x0=x0, x=x, mu=mu,
                           mus=numpy.array(mus), newton_its=numpy.array(its))


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
