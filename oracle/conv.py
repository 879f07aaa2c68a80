"""CPU oracle of the learned preconditioner's convolutions (test infrastructure
only; never called by the product path).

The network has no reference counterpart: it exists only in the GPU
backend.  These functions restate its arithmetic in
numpy so that the BF16 MFMA kernels (transiflow_amd/csrc/conv.hip) can be
checked: forward, the gradients with respect to weights, bias and input, and
BF16 rounding.  Pinning (tests/test_learned_oracle.py): single-channel layers against
scipy.signal.correlate (a textbook cross-correlation with zero padding), the
brute-force per-cell loop against the vectorised forward, finite differences
of ``forward`` against ``backward``, and BF16 rounding against known bit
patterns.

Layout: activations [ncell, C] with cell = i + nx (j + ny k); weights
[taps * Cin, Cout] with row tap * Cin + ci and tap = (dx+1) + 3 (dy+1) [+ 9 (dz+1)
in 3D; dz = 0 in 2D]; zero padding outside the grid.
"""
import numpy


def taps(dim):
    return 27 if dim == 3 else 9


def offsets(dim):
    out = []
    for t in range(taps(dim)):
        out.append((t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1 if dim == 3 else 0))
    return out


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = numpy.ascontiguousarray(x, dtype=numpy.float32)
    u = a.view(numpy.uint32).astype(numpy.uint64)
    finite = (u & 0x7f800000) != 0x7f800000
    r = numpy.where(finite, (u + 0x7fff + ((u >> 16) & 1)) >> 16, u >> 16)
    return (r.astype(numpy.uint32) << 16).view(numpy.float32).reshape(a.shape)


def _shift(x, shape, off):
    """y[cell] = x[cell + off] (zero outside the grid); x [ncell, C]."""
    nx, ny, nz = shape
    C = x.shape[1]
    g = x.reshape(nz, ny, nx, C)
    y = numpy.zeros_like(g)
    dx, dy, dz = off

    def sl(d, n):
        return (slice(max(0, -d), n - max(0, d)), slice(max(0, d), n - max(0, -d)))
    (xo, xi), (yo, yi), (zo, zi) = sl(dx, nx), sl(dy, ny), sl(dz, nz)
    y[zo, yo, xo] = g[zi, yi, xi]
    return y.reshape(-1, C)


def patches(x, shape, dim):
    """[ncell, taps * C]: the GEMM operand of the convolution."""
    return numpy.concatenate([_shift(x, shape, o) for o in offsets(dim)], axis=1)


def forward(x, W, b, shape, dim, relu):
    y = patches(x, shape, dim) @ W
    if b is not None:
        y = y + b
    return numpy.maximum(y, 0) if relu else y


def backward(x, W, y, gy, shape, dim, relu):
    """Gradients of sum(gy * y) for y = forward(x, W, b): (dW, db, dx)."""
    g = gy * (y > 0) if relu else gy
    P = patches(x, shape, dim)
    dW = P.T @ g
    db = g.sum(axis=0)
    dP = g @ W.T                       # [ncell, taps * C]
    C = x.shape[1]
    dx = numpy.zeros_like(x)
    for t, (ox, oy, oz) in enumerate(offsets(dim)):
        # patch column block t read x[cell + off]: scatter back with -off
        dx += _shift(dP[:, t * C:(t + 1) * C], shape, (-ox, -oy, -oz))
    return dW, db, dx


def forward_loop(x, W, b, shape, dim, relu):
    """Brute-force restatement of ``forward`` (small grids)."""
    nx, ny, nz = shape
    C = x.shape[1]
    out = numpy.zeros((nx * ny * nz, W.shape[1]))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                cell = i + nx * (j + ny * k)
                acc = numpy.array(b, dtype=float) if b is not None else numpy.zeros(W.shape[1])
                for t, (dx, dy, dz) in enumerate(offsets(dim)):
                    I, J, K = i + dx, j + dy, k + dz
                    if 0 <= I < nx and 0 <= J < ny and 0 <= K < nz:
                        acc = acc + x[I + nx * (J + ny * K)] @ W[t * C:(t + 1) * C]
                out[cell] = numpy.maximum(acc, 0) if relu else acc
    return out
