"""Readers for the reference's golden fixtures (tests/data/*.txt of TransiFlow).

The files in ``reference_data/`` are data copied verbatim from the reference's
own test suite (Apache-2.0, BIMAU/transiflow v1.0.1 ``tests/data``): 1-based
``row col value`` CSR triples in row order, and one value per line for rhs
vectors.
"""
import os

import numpy

DATA = os.path.join(os.path.dirname(__file__), 'reference_data')


def read_matrix(name):
    """Return (begA, jcoA, coA) with 0-based indices, as the reference reader does
    (tests/test_fvm.py:380-407)."""
    rows, cols, vals = [], [], []
    with open(os.path.join(DATA, name)) as f:
        for line in f:
            r, c, v = line.split()
            rows.append(int(r) - 1)
            cols.append(int(c) - 1)
            vals.append(float(v))
    rows = numpy.array(rows, dtype=numpy.int64)
    assert numpy.all(numpy.diff(rows) >= 0)
    n = int(rows.max()) + 1 if len(rows) else 0
    begA = numpy.zeros(n + 1, dtype=numpy.int64)
    numpy.add.at(begA, rows + 1, 1)
    return numpy.cumsum(begA), numpy.array(cols, dtype=numpy.int64), numpy.array(vals)


def read_vector(name):
    with open(os.path.join(DATA, name)) as f:
        return numpy.array([float(line) for line in f if line.strip()])
