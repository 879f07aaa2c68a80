"""CPU oracle: CSR assembly of atoms and the reference's compression rule.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Restates
``Discretization.assemble_jacobian`` (Discretization.py:474-524),
``assemble_mass_matrix`` (:526-548) and ``CrsMatrix.compress``
(CrsMatrix.py:52-71), vectorised over rows:

* row ``((k*ny + j)*nx + i)*dof + d1``;
* candidates visited in (d2, sx, sy, sz) order, kept when |a| > 1e-14;
* column ``wrap(i+sx-1)*dof + wrap(j+sy-1)*nx*dof + wrap(k+sz-1)*nx*ny*dof + d2``
  (every axis wraps modulo its length);
* per row: columns sorted ascending, duplicates summed in order of
  appearance starting from 0.0, sums with |v| <= 1e-14 dropped.
"""

import numpy

DROP = 1e-14


class CrsMatrix:
    """Minimal CSR container with the reference attribute names (coA, jcoA, begA)."""

    def __init__(self, coA, jcoA, begA, n=None):
        self.coA = numpy.asarray(coA, dtype=float)
        self.jcoA = numpy.asarray(jcoA, dtype=numpy.int64)
        self.begA = numpy.asarray(begA, dtype=numpy.int64)
        self._n = n

    @property
    def m(self):
        return len(self.begA) - 1

    @property
    def n(self):
        return self._n if self._n else self.m

    @property
    def shape(self):
        return (self.m, self.n)

    @property
    def nnz(self):
        return int(self.begA[-1])

    def to_scipy(self):
        from scipy import sparse
        nnz = self.nnz
        return sparse.csr_matrix((self.coA[:nnz], self.jcoA[:nnz], self.begA), shape=self.shape)

    def __matmul__(self, x):
        return self.to_scipy() @ x


def merge_sorted_rows(rows, cols, vals, order, nrows):
    """Sort (row, col) with appearance ``order`` as tie-break, sum duplicates
    sequentially and drop |sum| <= 1e-14 (CrsMatrix.py:57-71)."""
    perm = numpy.lexsort((order, cols, rows))
    rows = rows[perm]
    cols = cols[perm]
    vals = vals[perm]
    if len(rows) == 0:
        return CrsMatrix(numpy.zeros(0), numpy.zeros(0, dtype=numpy.int64),
                         numpy.zeros(nrows + 1, dtype=numpy.int64))
    new = numpy.ones(len(rows), dtype=bool)
    new[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    starts = numpy.flatnonzero(new)
    lengths = numpy.diff(numpy.append(starts, len(rows)))
    sums = numpy.zeros(len(starts)) + vals[starts]
    for k in range(1, int(lengths.max())):
        sel = lengths > k
        sums[sel] = sums[sel] + vals[starts[sel] + k]
    keep = numpy.abs(sums) > DROP
    r = rows[starts][keep]
    c = cols[starts][keep]
    v = sums[keep]
    begA = numpy.zeros(nrows + 1, dtype=numpy.int64)
    numpy.add.at(begA, r + 1, 1)
    begA = numpy.cumsum(begA)
    return CrsMatrix(v, c, begA)


def assemble_from_atom(atom, wrap=True):
    """Assemble and compress; ``wrap=False`` gives the unwrapped column rule
    ``row + (sx-1)*dof + (sy-1)*nx*dof + (sz-1)*nx*ny*dof + d2 - d1`` of the
    reference's slow test assembler (tests/test_fvm.py:431-454), against
    which the ``*_lin`` fixtures were written."""
    nx, ny, nz, dof = atom.shape[:4]
    n = nx * ny * nz * dof
    I, J, K = numpy.meshgrid(numpy.arange(nx), numpy.arange(ny), numpy.arange(nz), indexing='ij')
    cell_row = ((K * ny + J) * nx + I) * dof
    rows_l, cols_l, vals_l, ord_l = [], [], [], []
    order = 0
    for d2, sx, sy, sz in numpy.ndindex(dof, 3, 3, 3):
        v = atom[:, :, :, :, d2, sx, sy, sz]
        mask = numpy.abs(v) > DROP
        if mask.any():
            idx = numpy.nonzero(mask)
            if wrap:
                col = ((I + sx - 1) % nx) * dof + ((J + sy - 1) % ny) * nx * dof \
                    + ((K + sz - 1) % nz) * nx * ny * dof + d2
                cols_l.append(col[idx[:3]])
            else:
                cols_l.append(cell_row[idx[:3]] + (sx - 1) * dof + (sy - 1) * nx * dof
                              + (sz - 1) * nx * ny * dof + d2)
            rows_l.append(cell_row[idx[:3]] + idx[3])
            vals_l.append(v[idx])
            ord_l.append(numpy.full(len(idx[0]), order))
        order += 1
    if not rows_l:
        return CrsMatrix(numpy.zeros(0), numpy.zeros(0, dtype=numpy.int64),
                         numpy.zeros(n + 1, dtype=numpy.int64))
    return merge_sorted_rows(numpy.concatenate(rows_l), numpy.concatenate(cols_l),
                             numpy.concatenate(vals_l), numpy.concatenate(ord_l), n)


def assemble_diagonal(m):
    """assemble_mass_matrix: diagonal entries with |v| > 1e-14, rows in state order."""
    nx, ny, nz, dof = m.shape
    vec = numpy.ascontiguousarray(m.transpose(2, 1, 0, 3)).reshape(-1)
    n = len(vec)
    keep = numpy.abs(vec) > DROP
    idx = numpy.flatnonzero(keep)
    begA = numpy.zeros(n + 1, dtype=numpy.int64)
    begA[1:] = numpy.cumsum(keep)
    return CrsMatrix(vec[idx], idx, begA, n=n)
