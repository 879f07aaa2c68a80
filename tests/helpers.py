"""Shared helpers for the parity tests (HIP backend vs the CPU oracle)."""
import numpy

REL_TOL = 1e-12   # parity bound for FP64 values (relative to the largest entry)


def csr_of_hip(jac):
    return jac.begA, jac.jcoA, jac.coA


def assert_csr_equal(hip_csr, ref_begA, ref_jcoA, ref_coA, rel=REL_TOL):
    """row_ptr/col_idx bit-exact, values within ``rel`` of the largest entry.
    Returns the fraction of values that are bit-identical."""
    begA, jcoA, coA = hip_csr
    assert numpy.array_equal(begA, numpy.asarray(ref_begA, dtype=numpy.int64)), 'row_ptr differs'
    nnz = int(ref_begA[-1])
    assert numpy.array_equal(jcoA, numpy.asarray(ref_jcoA[:nnz], dtype=numpy.int64)), 'col_idx differs'
    ref = numpy.asarray(ref_coA[:nnz])
    scale = max(numpy.max(numpy.abs(ref)) if nnz else 0.0, 1e-300)
    err = numpy.max(numpy.abs(coA - ref)) / scale if nnz else 0.0
    assert err <= rel, 'CSR values differ: max rel err %g' % err
    return float(numpy.mean(coA == ref)) if nnz else 1.0


def assert_vec_close(a, b, rel=REL_TOL):
    a = numpy.asarray(a)
    b = numpy.asarray(b)
    scale = max(numpy.max(numpy.abs(b)), 1e-300)
    err = numpy.max(numpy.abs(a - b)) / scale
    assert err <= rel, 'max rel err %g' % err
    return float(numpy.mean(a == b))


def seeded_state(n, seed, lo=-1.0, hi=1.0):
    return numpy.random.default_rng(seed).uniform(lo, hi, n)
