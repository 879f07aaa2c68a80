"""tf_csr_add_count_noalloc (scratch in a library-owned device array, no
allocation) against tf_csr_add_count (caller-provided scratch): the same
row pointers and nnz, and the same combined matrix after tf_csr_add_fill,
on two calls in a row, the second on a smaller matrix."""
import ctypes

import numpy
import pytest

from tests.helpers import seeded_state

pytestmark = pytest.mark.gpu


def test_csr_add_count_noalloc_matches_count(hip):
    from transiflow_amd import _lib
    from transiflow_amd.vector import ptr, stream, torch
    T = torch()
    for shape in ((64, 48), (20, 12)):
        itf = hip.Interface({'Reynolds Number': 300.0}, *shape)
        n = itf.n
        J = itf.jacobian(seeded_state(n, 61) * 0.5)
        M = itf.mass_matrix()
        print('case %s: n %d, nnz J %d, nnz M %d' % (shape, n, J.nnz, M.nnz), flush=True)
        cnt = T.empty(n, dtype=T.int32, device='cuda')
        rc = [T.full((n + 1,), -7, dtype=T.int32, device='cuda') for _ in range(2)]
        nnz = [ctypes.c_int(-1), ctypes.c_int(-1)]
        _lib.call('tf_csr_add_count', n, ptr(J.row_ptr), ptr(J.col), ptr(M.row_ptr), ptr(M.col), ptr(cnt),
                  ptr(rc[0]), ctypes.byref(nnz[0]), stream())
        _lib.call('tf_csr_add_count_noalloc', n, ptr(J.row_ptr), ptr(J.col), ptr(M.row_ptr), ptr(M.col),
                  ptr(rc[1]), ctypes.byref(nnz[1]), stream())
        assert nnz[0].value == nnz[1].value > 0
        assert T.equal(rc[0], rc[1])
        out = []
        for r in rc:
            cc = T.empty(nnz[0].value, dtype=T.int32, device='cuda')
            vc = T.empty(nnz[0].value, dtype=T.float64, device='cuda')
            _lib.call('tf_csr_add_fill', n, ptr(J.row_ptr), ptr(J.col), ptr(J.val), 1.0, ptr(M.row_ptr),
                      ptr(M.col), ptr(M.val), -2.5, ptr(r), ptr(cc), ptr(vc), stream())
            out.append((cc.cpu().numpy(), vc.cpu().numpy()))
        assert numpy.array_equal(out[0][0], out[1][0])
        assert numpy.array_equal(out[0][1], out[1][1])
    with pytest.raises(_lib.BackendError, match='scratch'):
        _lib.call('tf_csr_add_count_noalloc', (1 << 22) + 1, None, None, None, None, None, None, None)
