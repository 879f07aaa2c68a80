"""Many independent continuations on one GPU.

Each instance is a full reference-style run -- Newton from zero at
Re = 0, then pseudo-arclength continuation in the Reynolds number to the
instance's target (``Continuation.continuation``, Continuation.py:361-482) --
with its own Interface, Jacobians and line factor, so an instance's result
does not depend on the others.  Workers (threads) each own a HIP stream of
default priority and take instances from a queue, largest first; ctypes
releases the interpreter lock while a launch is issued, so the small grids'
kernels from different instances run concurrently on the GPU instead of one
instance leaving most of the 256 CUs idle.
"""
import queue
import threading
import time

import numpy

from .continuation import Continuation
from .interface import Interface
from .vector import torch


def c4_instances(count=64, seed=2, sizes=tuple(range(32, 257, 32)), re_range=(100.0, 1500.0)):
    """The seeded batch: square cavity grids drawn from ``sizes`` and target
    Reynolds numbers uniform in ``re_range``."""
    rng = numpy.random.default_rng(seed)
    ns = rng.choice(numpy.asarray(sizes), count)
    res = rng.uniform(re_range[0], re_range[1], count)
    return [{'n': int(n), 'Re': float(r)} for n, r in zip(ns, res)]


def run_instance(inst, precision='mixed', ds=100.0, tol=1e-8):
    """One instance on the calling thread's current stream.  Returns the final
    state (host copy), mu, accepted steps and corrector iterations."""
    T = torch()
    itf = Interface({'HIP Solver': {'Precision': precision}}, inst['n'], inst['n'])
    cont = Continuation(itf, newton_tolerance=tol, printer=None)
    t0 = time.perf_counter()
    x0 = cont.newton(itf.vector(), tol=tol)
    x, mu = cont.continuation(x0, 'Reynolds Number', 0.0, inst['Re'], ds)
    xh = itf.array_from_vector(x)
    T.cuda.current_stream().synchronize()
    return {'x': xh, 'mu': mu, 'seconds': time.perf_counter() - t0,
            'accepted': sum(1 for h in cont.history if h[0] == 'accepted'),
            'iterations': sum(h[1] for h in cont.history) + cont.newton_iterations,
            'factorizations': sum(int(entry[3]) for entry in itf.solver.log)}


def run_batch(instances, workers=8, precision='mixed', ds=100.0, tol=1e-8):
    """Run all instances with ``workers`` threads, one stream each.  Returns
    (results in instance order, wall seconds)."""
    T = torch()
    order = sorted(range(len(instances)), key=lambda i: -instances[i]['n'])
    q = queue.Queue()
    for i in order:
        q.put(i)
    results = [None] * len(instances)
    errors = []

    def work():
        stream = T.cuda.Stream()   # default priority
        with T.cuda.stream(stream):
            while True:
                try:
                    i = q.get_nowait()
                except queue.Empty:
                    return
                try:
                    results[i] = run_instance(instances[i], precision, ds, tol)
                except Exception as e:   # reported after the batch
                    errors.append((i, e))
        stream.synchronize()

    T.cuda.synchronize()
    t0 = time.perf_counter()
    threads = [threading.Thread(target=work) for _ in range(max(1, workers))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    T.cuda.synchronize()
    wall = time.perf_counter() - t0
    if errors:
        i, e = errors[0]
        raise RuntimeError('instance %d (%s) failed: %s' % (i, instances[i], e)) from e
    return results, wall
