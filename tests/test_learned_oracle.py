"""CPU pinning of the convolution oracle (oracle/conv.py) that checks the
learned preconditioner's BF16 MFMA kernels."""
import numpy
import pytest
import scipy.signal

from oracle import conv as oc


@pytest.mark.parametrize('dim,shape', [(2, (7, 5, 1)), (3, (4, 5, 3))])
def test_single_channel_matches_scipy_correlate(dim, shape):
    """A 1 -> 1 channel layer is a textbook cross-correlation with zero padding."""
    rng = numpy.random.default_rng(1)
    nx, ny, nz = shape
    x = rng.standard_normal((nx * ny * nz, 1))
    W = rng.standard_normal((oc.taps(dim), 1))
    y = oc.forward(x, W, None, shape, dim, relu=False)
    grid = x.reshape(nz, ny, nx) if dim == 3 else x.reshape(ny, nx)
    kern = W.reshape(3, 3, 3) if dim == 3 else W.reshape(3, 3)   # [dz][dy][dx] from tap order
    ref = scipy.signal.correlate(grid, kern, mode='same', method='direct')
    numpy.testing.assert_allclose(y.reshape(ref.shape), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('dim,shape,cin,cout', [(2, (5, 4, 1), 3, 6), (3, (3, 4, 2), 4, 5)])
def test_forward_matches_loop(dim, shape, cin, cout):
    rng = numpy.random.default_rng(2)
    n = shape[0] * shape[1] * shape[2]
    x = rng.standard_normal((n, cin))
    W = rng.standard_normal((oc.taps(dim) * cin, cout))
    b = rng.standard_normal(cout)
    for relu in (False, True):
        numpy.testing.assert_allclose(oc.forward(x, W, b, shape, dim, relu), oc.forward_loop(x, W, b, shape, dim, relu),
                                      rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('dim,shape,cin,cout', [(2, (5, 4, 1), 3, 4), (3, (3, 3, 2), 2, 3)])
def test_backward_matches_finite_differences(dim, shape, cin, cout):
    rng = numpy.random.default_rng(3)
    n = shape[0] * shape[1] * shape[2]
    x = rng.standard_normal((n, cin))
    W = rng.standard_normal((oc.taps(dim) * cin, cout))
    b = rng.standard_normal(cout)
    gy = rng.standard_normal((n, cout))
    relu = False

    def L(x_, W_, b_):
        return float((gy * oc.forward(x_, W_, b_, shape, dim, relu)).sum())

    y = oc.forward(x, W, b, shape, dim, relu)
    dW, db, dx = oc.backward(x, W, y, gy, shape, dim, relu)
    eps = 1e-6
    for idx in [(0, 0), (W.shape[0] - 1, cout - 1), (W.shape[0] // 2, 1)]:
        Wp, Wm = W.copy(), W.copy()
        Wp[idx] += eps
        Wm[idx] -= eps
        assert (L(x, Wp, b) - L(x, Wm, b)) / (2 * eps) == pytest.approx(dW[idx], rel=1e-6, abs=1e-8)
    for idx in [(0, 0), (n - 1, cin - 1), (n // 2, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        assert (L(xp, W, b) - L(xm, W, b)) / (2 * eps) == pytest.approx(dx[idx], rel=1e-6, abs=1e-8)
    bp = b.copy()
    bp[0] += eps
    bm = b.copy()
    bm[0] -= eps
    assert (L(x, W, bp) - L(x, W, bm)) / (2 * eps) == pytest.approx(db[0], rel=1e-6)


def test_relu_mask_in_backward():
    rng = numpy.random.default_rng(4)
    shape, dim = (4, 4, 1), 2
    x = rng.standard_normal((16, 2))
    W = rng.standard_normal((18, 3))
    y = oc.forward(x, W, None, shape, dim, relu=True)
    gy = numpy.ones_like(y)
    _, db, _ = oc.backward(x, W, y, gy, shape, dim, relu=True)
    assert numpy.array_equal(db, (y > 0).sum(axis=0))


def test_bf16_rounding_known_values():
    v = numpy.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e38, numpy.inf], dtype=numpy.float32)
    r = oc.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie -> even mantissa
    assert r[2] == 1 + 2.0 ** -6             # tie -> even mantissa (up)
    assert r[3] == -2.5
    # 3e38 = 0x7f61b1e6: the low half 0xb1e6 is above the tie -> rounds up to 0x7f62
    assert r[4] == numpy.array([0x7f620000], dtype=numpy.uint32).view(numpy.float32)[0]
    assert numpy.isinf(r[5])
