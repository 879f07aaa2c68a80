"""The Krylov building blocks against numpy: CSR SpMV with the pressure gauge,
the fused multi-dot and multi-axpy of classical Gram-Schmidt and the
normalisation, in FP64 and FP32, on sizes that take the 16-byte vector path
(n a multiple of 4) and the one-entry-per-lane path (n odd, or a leading
dimension that breaks the alignment), and on bases of 1 to 17 vectors (the
multi-dot works in passes of 8)."""

import numpy
import pytest
from scipy import sparse

pytestmark = pytest.mark.gpu

SIZES = [1, 7, 4096, 200003, 3 * 2 ** 18]


def _cuda(a, dt):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(a)).to(dt).cuda()


def _csr(n, seed, per_row=9.0):
    rng = numpy.random.default_rng(seed)
    A = sparse.random(n, n, density=min(1.0, per_row / n), random_state=rng, format='csr')
    A = (A + sparse.eye(n, format='csr')).tocsr()
    A.sort_indices()
    return A


@pytest.mark.parametrize('n,per_row', [(5, 9), (3000, 9), (196608 + 3, 9), (3000, 45)])
@pytest.mark.parametrize('single', [False, True])
def test_spmv_gauge_matches_scipy(hip, n, per_row, single):
    """y = A x with column g left out of every row and y[g] = -x[g]; 45 entries
    per row take the path for workgroups whose entries exceed the LDS stage."""
    import torch
    from transiflow_amd import _lib
    from transiflow_amd.vector import ptr, stream
    A = _csr(n, 11, per_row)
    g = min(2, n - 1)
    x = numpy.random.default_rng(12).uniform(-1, 1, n)
    B = A.tolil()
    B[:, g] = 0
    B[g, :] = 0
    B[g, g] = -1
    ref = B.tocsr() @ x
    dt = torch.float32 if single else torch.float64
    rp, ci = _cuda(A.indptr.astype(numpy.int32), torch.int32), _cuda(A.indices.astype(numpy.int32), torch.int32)
    v, xd = _cuda(A.data, dt), _cuda(x, dt)
    y = torch.empty(n, dtype=dt, device='cuda')
    _lib.call('tf_spmv_f32' if single else 'tf_spmv_f64', n, ptr(rp), ptr(ci), ptr(v), ptr(xd), ptr(y), g, stream())
    got = y.double().cpu().numpy()
    err = numpy.max(numpy.abs(got - ref)) / numpy.max(numpy.abs(ref))
    print('spmv n=%d %s: %.2e' % (n, 'fp32' if single else 'fp64', err), flush=True)
    assert err <= (2e-6 if single else 1e-14)


@pytest.mark.parametrize('n', SIZES)
@pytest.mark.parametrize('k', [1, 5, 8, 9, 17])
@pytest.mark.parametrize('single', [False, True])
@pytest.mark.parametrize('pad', [0, 1])
def test_multidot_multiaxpy_normalize(hip, n, k, single, pad):
    import torch
    from transiflow_amd import _lib
    from transiflow_amd.vector import ptr, stream
    dt = torch.float32 if single else torch.float64
    sfx = 'f32' if single else 'f64'
    ldv = n + pad
    rng = numpy.random.default_rng(1000 * k + n % 1000 + pad)
    Vh = rng.uniform(-1, 1, (k, ldv))
    wh = rng.uniform(-1, 1, n)
    if single:   # the reference sees the FP32-rounded operands
        Vh, wh = Vh.astype(numpy.float32).astype(numpy.float64), wh.astype(numpy.float32).astype(numpy.float64)
    V, w = _cuda(Vh, dt), _cuda(wh, dt)
    work = torch.empty(_lib.load().tf_krylov_work_size(), dtype=torch.float64, device='cuda')
    h = torch.full((k + 1,), 0.5, dtype=torch.float64, device='cuda')
    st = stream()
    # h = V^T w, then h += V^T w (accumulate)
    _lib.call('tf_multidot_' + sfx, n, k, ptr(V), ldv, ptr(w), ptr(work), ptr(h), 0, st)
    dots = Vh[:, :n] @ wh
    hh = h.cpu().numpy()
    tol = (2e-6 if single else 1e-13) * max(1.0, numpy.sqrt(n))
    assert numpy.allclose(hh[:k], dots, rtol=0, atol=tol), (hh[:k], dots)
    assert hh[k] == 0.5
    _lib.call('tf_multidot_' + sfx, n, k, ptr(V), ldv, ptr(w), ptr(work), ptr(h), 1, st)
    assert numpy.allclose(h.cpu().numpy()[:k], 2 * dots, rtol=0, atol=2 * tol)
    # w -= V^T c
    c = rng.uniform(-1, 1, k)
    cd = _cuda(c, torch.float64)
    _lib.call('tf_multiaxpy_' + sfx, n, k, ptr(V), ldv, ptr(cd), ptr(w), -1.0, st)
    ref = wh - c @ Vh[:, :n]
    got = w.double().cpu().numpy()
    assert numpy.max(numpy.abs(got - ref)) <= (1e-6 if single else 1e-14) * k
    # out = w / ||w||
    out = torch.empty(n, dtype=dt, device='cuda')
    _lib.call('tf_normalize_' + sfx, n, ptr(w), ptr(work), ptr(h), ptr(out), st)
    nrm = numpy.linalg.norm(got)
    hn = h.cpu().numpy()[0]
    assert abs(hn - nrm) <= (1e-5 if single else 1e-13) * nrm
    assert numpy.max(numpy.abs(out.double().cpu().numpy() - got / nrm)) <= (1e-5 if single else 1e-15)


def test_multidot_deterministic(hip):
    """The same inputs give bit-identical sums (fixed partition, fixed order)."""
    import torch
    from transiflow_amd import _lib
    from transiflow_amd.vector import ptr, stream
    n, k = 3 * 2 ** 18, 16
    V = torch.rand((k, n), dtype=torch.float32, device='cuda')
    w = torch.rand(n, dtype=torch.float32, device='cuda')
    work = torch.empty(_lib.load().tf_krylov_work_size(), dtype=torch.float64, device='cuda')
    hs = []
    for _ in range(3):
        h = torch.zeros(k, dtype=torch.float64, device='cuda')
        _lib.call('tf_multidot_f32', n, k, ptr(V), n, ptr(w), ptr(work), ptr(h), 0, stream())
        hs.append(h.cpu().numpy())
    assert all(numpy.array_equal(hs[0], x) for x in hs[1:])
