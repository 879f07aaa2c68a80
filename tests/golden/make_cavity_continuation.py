"""Write the oracle's lid-driven-cavity continuation at full size as a fixture.

    python tests/golden/make_cavity_continuation.py [n]

Runs the CPU oracle (numpy assembly, scipy splu with the reference's
pressure gauge) under the product's continuation driver: LDC 2D n x n
(default 256, n = 196,608 unknowns), Newton from zero at Re = 0 with
tolerance 1e-8, then pseudo-arclength continuation in the Reynolds number
from 0 to 2000 with ds = 100 and Newton tolerance 1e-8 -- the
``Continuation(interface, newton_tolerance=1e-8).continuation(x0,
'Reynolds Number', 0, 2000, 100)`` run of the reference.  About 45 sparse LU
factorisations; a few minutes on one core.

Stores ``cavity_continuation_<n>.npz`` next to this script: the accepted
Reynolds numbers ``mus`` (the callback's, starting with the initial state), the Newton
iterations of each accepted step ``newton_its``, the final state ``x`` and
its Reynolds number ``mu``.  The GPU tests compare the HIP backend's FP64 and
mixed-precision continuations with it.
"""
import os
import sys
import time

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main(n):
    from oracle.interface import Interface
    from transiflow_amd.continuation import Continuation

    itf = Interface({}, n, n, 1, 2, 3)
    cont = Continuation(itf, newton_tolerance=1e-8, printer=print)
    t0 = time.perf_counter()
    x0 = cont.newton(itf.vector(), tol=1e-8)
    mus, its = [], []

    def callback(_interface, _x, mu):
        mus.append(mu)
        its.append(cont.newton_iterations)
        print('accepted Re = %.6f after %d Newton iterations (%.0f s)' % (mu, cont.newton_iterations,
                                                                          time.perf_counter() - t0), flush=True)

    x, mu = cont.continuation(x0, 'Reynolds Number', 0, 2000, 100, callback=callback)
    path = os.path.join(HERE, 'cavity_continuation_%d.npz' % n)
    numpy.savez_compressed(path, x=x, mu=mu, mus=numpy.array(mus), newton_its=numpy.array(its))
    print('wrote %s (%.0f s)' % (path, time.perf_counter() - t0))


if __name__ == '__main__':
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 256)
