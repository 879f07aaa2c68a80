"""Learned preconditioner: a small convolutional residual-to-correction network.

The network maps the residual that a base preconditioner leaves behind to a
correction of the iterate:

    z0 = M r              (base: line factor -- possibly of an earlier Jacobian -- or cell-block Jacobi)
    r1 = r - A z0
    c  = s * Net(r1 / s)  (s = rms of r1)
    z  = z0 + alpha c,    alpha = <r1, A c> / <A c, A c>

``alpha`` is the least-squares step along the correction, so the corrected
residual r1 - alpha A c is never larger than r1: the base preconditioner is
both the fallback (alpha <= 0 or a negligible gain keeps z = z0) and the
measure of the network's error (the residual it leaves decides).  FGMRES accepts the
resulting non-linear preconditioner.

Network: dof -> 32 -> 32 -> dof channels, 3x3 (2D) or 3x3x3 (3D) stencils on
the cell grid, ReLU between layers, BF16 MFMA convolutions with FP32
accumulation (csrc/conv.hip), FP32 master weights and Adam.  It is trained on
(Jacobian, residual, correction) triples taken from the plain FP64 method's
own continuation runs: the residual is what the lagged FP32 line factor
leaves of a Newton right-hand side, the correction is the exact FP64 solve
of the same Jacobian for that residual.

There is no reference counterpart: the reference solves with SuperLU
(transiflow/interface/SciPy.py:131-162).
"""
import ctypes
import math

import numpy

from . import _lib
from .vector import ptr, stream, torch


def _pad16(v):
    return (v + 15) // 16 * 16


class ConvLayer:
    def __init__(self, dim, cin, cout, relu, rng):
        T = torch()
        self.dim, self.cin, self.cout, self.relu = dim, cin, cout, relu
        self.taps = 27 if dim == 3 else 9
        K = self.taps * cin
        self.kpad, self.npad = _pad16(K), _pad16(cout)
        self.kpad_t, self.npad_t = _pad16(self.taps * cout), _pad16(cin)
        std = math.sqrt(2.0 / K) if relu else math.sqrt(1.0 / K)
        w = numpy.zeros((self.kpad, self.npad), dtype=numpy.float32)   # padded MFMA layout
        w[:K, :cout] = rng.standard_normal((K, cout)).astype(numpy.float32) * std
        self.W = T.from_numpy(w).cuda()
        self.b = T.zeros(cout, dtype=T.float32, device='cuda')
        self.mW, self.vW = T.zeros_like(self.W), T.zeros_like(self.W)
        self.mb, self.vb = T.zeros_like(self.b), T.zeros_like(self.b)
        self.dW = T.zeros((self.kpad, self.npad), dtype=T.float32, device='cuda')
        self.db = T.zeros(cout, dtype=T.float32, device='cuda')
        self.Wf = T.zeros(self.kpad * self.npad, dtype=T.int16, device='cuda')
        self.Wt = T.zeros(self.kpad_t * self.npad_t, dtype=T.int16, device='cuda')
        self.pack()

    def pack(self):
        _lib.call('tf_conv_pack', self.dim, self.cin, self.cout, self.kpad, self.npad, self.kpad_t, self.npad_t,
                  ptr(self.W), ptr(self.Wf), ptr(self.Wt), stream())


class ConvNet:
    """dof -> hidden -> hidden -> dof convolutional network on an nx x ny x nz grid."""

    def __init__(self, dim, dof, hidden=32, seed=0):
        rng = numpy.random.default_rng(seed)
        self.dim, self.dof, self.hidden = dim, dof, hidden
        self.layers = [ConvLayer(dim, dof, hidden, True, rng), ConvLayer(dim, hidden, hidden, True, rng),
                       ConvLayer(dim, hidden, dof, False, rng)]
        self.step_count = 0
        self._shape = None

    # -- buffers -------------------------------------------------------------
    def _buffers(self, shape):
        if self._shape == shape:
            return
        T = torch()
        ncell = shape[0] * shape[1] * shape[2]
        self._shape = shape
        self.x16 = T.empty(ncell * self.dof, dtype=T.int16, device='cuda')
        self.act16 = [T.empty(ncell * L.cout, dtype=T.int16, device='cuda') for L in self.layers]
        self.y32 = T.empty(ncell * self.dof, dtype=T.float32, device='cuda')
        cmax = max(self.hidden, self.dof)
        self.g32 = T.empty(ncell * cmax, dtype=T.float32, device='cuda')
        self.gin32 = T.empty(ncell * cmax, dtype=T.float32, device='cuda')
        self.g16 = T.empty(ncell * cmax, dtype=T.int16, device='cuda')
        nw = max(_lib.load().tf_conv_wgrad_work_size(ncell, L.kpad, L.npad) for L in self.layers)
        self.work = T.empty(nw, dtype=T.float32, device='cuda')
        self.loss_partial = T.empty((ncell * self.dof + 255) // 256, dtype=T.float32, device='cuda')

    # -- forward / backward ----------------------------------------------------
    def forward(self, x32, shape, scale=1.0, out=None):
        """y = Net(scale * x); x32, y: FP32 [ncell * dof] device tensors."""
        self._buffers(shape)
        nx, ny, nz = shape
        st = stream()
        _lib.call('tf_to_bf16', x32.numel(), ptr(x32), float(scale), ptr(self.x16), st)
        src = self.x16
        out = self.y32 if out is None else out
        for i, L in enumerate(self.layers):
            last = i == len(self.layers) - 1
            _lib.call('tf_conv_fwd', nx, ny, nz, self.dim, L.cin, L.cout, L.kpad, L.npad, ptr(src), ptr(L.Wf),
                      ptr(L.b), 1 if L.relu else 0, ptr(out) if last else None, ptr(self.act16[i]), st)
            src = self.act16[i]
        return out

    def backward(self, gy32, shape):
        """Accumulate nothing: sets each layer's dW, db for the loss gradient gy32
        with respect to the last forward's output."""
        nx, ny, nz = shape
        ncell = nx * ny * nz
        st = stream()
        g_in = gy32
        for i in range(len(self.layers) - 1, -1, -1):
            L = self.layers[i]
            n = ncell * L.cout
            _lib.call('tf_relu_bwd', n, ptr(g_in), ptr(self.act16[i]), 1 if L.relu else 0, ptr(self.g32),
                      ptr(self.g16), st)
            src = self.x16 if i == 0 else self.act16[i - 1]
            _lib.call('tf_conv_wgrad', nx, ny, nz, self.dim, L.cin, L.cout, L.kpad, L.npad, ptr(src), ptr(self.g16),
                      ptr(self.g32), ptr(self.work), ptr(L.dW), ptr(L.db), st)
            if i > 0:
                # dx = conv of g with the flipped, transposed weights
                _lib.call('tf_conv_fwd', nx, ny, nz, self.dim, L.cout, L.cin, L.kpad_t, L.npad_t, ptr(self.g16),
                          ptr(L.Wt), None, 0, ptr(self.gin32), None, st)
                g_in = self.gin32

    def grads(self):
        """(dW [taps*cin, cout], db) per layer as host arrays."""
        out = []
        for L in self.layers:
            out.append((L.dW.cpu().numpy()[:L.taps * L.cin, :L.cout], L.db.cpu().numpy()))
        return out

    def weights(self):
        """(W [taps*cin, cout], b) per layer as host arrays."""
        return [(L.W.cpu().numpy()[:L.taps * L.cin, :L.cout], L.b.cpu().numpy()) for L in self.layers]

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        st = stream()
        for L in self.layers:
            # W and dW share the padded layout; padding has zero gradient and stays zero
            _lib.call('tf_adam', L.W.numel(), ptr(L.W), ptr(L.dW), ptr(L.mW), ptr(L.vW), lr, b1, b2, eps,
                      self.step_count, st)
            _lib.call('tf_adam', L.b.numel(), ptr(L.b), ptr(L.db), ptr(L.mb), ptr(L.vb), lr, b1, b2, eps,
                      self.step_count, st)
            L.pack()

    def mse(self, y32, t32):
        """Returns the loss (host float) and leaves dL/dy in self.g_loss."""
        T = torch()
        n = y32.numel()
        if getattr(self, 'g_loss', None) is None or self.g_loss.numel() != n:
            self.g_loss = T.empty(n, dtype=T.float32, device='cuda')
        _lib.call('tf_mse_grad', n, ptr(y32), ptr(t32), ptr(self.g_loss), ptr(self.loss_partial), stream())
        return float(self.loss_partial[:(n + 255) // 256].cpu().numpy().astype(numpy.float64).sum()) / n

    # -- persistence -----------------------------------------------------------
    def save(self, path):
        arrs = {'dim': self.dim, 'dof': self.dof, 'hidden': self.hidden}
        for i, (W, b) in enumerate(self.weights()):
            arrs['W%d' % i] = W
            arrs['b%d' % i] = b
        numpy.savez(path, **arrs)

    @classmethod
    def load(cls, path):
        f = numpy.load(path)
        net = cls(int(f['dim']), int(f['dof']), int(f['hidden']))
        for i in range(len(net.layers)):
            net.set_layer(i, f['W%d' % i], f['b%d' % i])
        return net

    def set_layer(self, i, W, b):
        """Load host weights W [taps*cin, cout] and b [cout] into layer i."""
        T = torch()
        L = self.layers[i]
        w = numpy.zeros((L.kpad, L.npad), dtype=numpy.float32)
        w[:L.taps * L.cin, :L.cout] = W
        L.W.copy_(T.from_numpy(w))
        L.b.copy_(T.from_numpy(numpy.asarray(b, dtype=numpy.float32)))
        L.pack()


def train(net, samples, shape, steps, lr=1e-3, seed=0, log=None):
    """Adam on the mean squared error of Net(r1 / s) against e1 / s over the
    (r1, e1) FP32 device pairs (already divided by s).  Returns the loss history."""
    rng = numpy.random.default_rng(seed)
    hist = []
    for it in range(steps):
        r1, e1 = samples[rng.integers(len(samples))]
        y = net.forward(r1, shape)
        loss = net.mse(y, e1)
        net.backward(net.g_loss, shape)
        net.adam_step(lr)
        hist.append(loss)
        if log is not None and (it % 100 == 0 or it == steps - 1):
            log('step %d loss %.4e' % (it, loss))
    return hist


class LearnedCorrection:
    """FP32 preconditioner z = M r + alpha s Net(r1 / s) judged by its residual
    (see the module docstring).  ``bind(base, matvec32)`` attaches the base
    preconditioner (with ``apply32``) and the gauged FP32 Jacobian product of
    the current solve."""

    name = 'learned'

    def __init__(self, net, shape, n, min_gain=0.0, window=16, off_below=0.02):
        T = torch()
        self.net, self.shape, self.n = net, shape, n
        self.min_gain = min_gain
        # the correction switches itself off when its mean gain over the last
        # ``window`` applications is below ``off_below`` (the base alone then runs)
        self.window, self.off_below = window, off_below
        self.recent = []
        self.off = False
        self.rw = T.empty((2, n), dtype=T.float32, device='cuda')   # rows: r1, A c
        self.c = T.empty(n, dtype=T.float32, device='cuda')
        self.h = T.zeros(4, dtype=T.float64, device='cuda')
        self.h_host = T.empty(4, dtype=T.float64).pin_memory()
        self.work = T.empty(_lib.load().tf_krylov_work_size(), dtype=T.float64, device='cuda')
        self.base = None
        self.matvec32 = None
        self.applied = 0
        self.accepted = 0
        self.gain_sum = 0.0

    def bind(self, base, matvec32):
        self.base, self.matvec32 = base, matvec32
        self.name = 'learned+' + getattr(base, 'name', '?')
        return self

    def _dots(self, k, w):
        _lib.call('tf_multidot_f32', self.n, k, ptr(self.rw), self.n, ptr(w), ptr(self.work), ptr(self.h), 0,
                  stream())
        self.h_host[:k].copy_(self.h[:k], non_blocking=True)
        torch().cuda.current_stream().synchronize()
        return [float(v) for v in self.h_host[:k]]

    def residual(self, r, z0):
        """r1 = r - A z0 into self.rw[0]; returns ||r1||^2."""
        r1, Ac = self.rw[0], self.rw[1]
        self.matvec32(z0, Ac)
        _lib.call('tf_axpby_f32', self.n, 1.0, ptr(r), -1.0, ptr(Ac), ptr(r1), stream())
        return self._dots(1, r1)[0]

    def _record(self, gain):
        self.recent.append(gain)
        if len(self.recent) > self.window:
            self.recent.pop(0)
            if sum(self.recent) / self.window < self.off_below:
                self.off = True

    def apply32(self, r, out):
        n = self.n
        self.base.apply32(r, out)                             # z0
        if self.off:
            return out
        rr = self.residual(r, out)                            # r1 = r - A z0
        self.applied += 1
        if not rr > 0.0:
            return out
        s = math.sqrt(rr / n)
        self.net.forward(self.rw[0], self.shape, scale=1.0 / s, out=self.c)   # c / s
        Ac = self.rw[1]
        self.matvec32(self.c, Ac)
        r1w, ww = self._dots(2, Ac)
        if not (ww > 0.0 and r1w > 0.0):
            self._record(0.0)
            return out
        alpha = r1w / ww
        gain = r1w * r1w / ww / rr    # fraction of ||r1||^2 removed
        self._record(gain)
        if gain <= self.min_gain:
            return out
        self.accepted += 1
        self.gain_sum += gain
        _lib.call('tf_axpby_f32', n, 1.0, ptr(out), alpha, ptr(self.c), ptr(out), stream())
        return out


# -- training data from the plain method's runs ------------------------------------

def collect(interface_factory, run, every=1, max_samples=256):
    """Run ``run(itf)`` (a continuation) on ``interface_factory()`` with the
    mixed solver and record, for every ``every``-th inner preconditioner input
    v, the residual r1 = v - A M v left by the base preconditioner M and the
    Jacobian.  Returns (itf, [(jac, r1 FP32 device tensor)])."""
    itf = interface_factory()
    recs = []
    count = [0]

    def observer(jac, base, v):
        count[0] += 1
        if len(recs) >= max_samples or count[0] % every:
            return
        T = torch()
        z0 = T.empty_like(v)
        base.apply32(v, z0)
        r1 = T.empty_like(v)
        val32 = jac.values32()
        g = itf.solver.gauge()
        _lib.call('tf_spmv_f32', itf.n, ptr(jac.row_ptr), ptr(jac.col), ptr(val32), ptr(z0), ptr(r1), int(g), stream())
        _lib.call('tf_axpby_f32', itf.n, 1.0, ptr(v), -1.0, ptr(r1), ptr(r1), stream())
        recs.append((jac, r1))

    itf.solver.observer = observer
    run(itf)
    itf.solver.observer = None
    return itf, recs


def make_pairs(itf, recs):
    """(r1 / s, e1 / t) FP32 pairs with e1 the FP64 solve of the recorded
    Jacobian for r1 (gauged system), s the rms of r1 and t the rms of e1.  The
    network learns the direction of the correction: its scale is fixed by the
    least-squares step at every application."""
    from .solver import _norm
    from .vector import DeviceVector
    T = torch()
    n = itf.n
    st = stream()
    pairs = []
    by_jac = {}
    for jac, r1 in recs:
        by_jac.setdefault(id(jac), (jac, []))[1].append(r1)
    work = T.empty(_lib.load().tf_reduce_work_size() + 1, dtype=T.float64, device='cuda')
    saved = itf.parameters.get('HIP Solver')
    itf.parameters['HIP Solver'] = {'Precision': 'fp64'}
    try:
        for jac, rs in by_jac.values():
            old = (jac.prec, jac.prec_exact)
            jac.prec = None
            for r1 in rs:
                b = T.empty(n, dtype=T.float64, device='cuda')
                _lib.call('tf_f32_to_f64', n, ptr(r1), ptr(b), st)
                s = _norm(b, work) / math.sqrt(n)
                if not s > 0.0:
                    continue
                e1 = itf.solver.solve(jac, DeviceVector(b)).data   # the solver zeroes the gauge entry itself
                t = _norm(e1, work) / math.sqrt(n)
                if not t > 0.0:
                    continue
                x_in = T.empty(n, dtype=T.float32, device='cuda')
                y_out = T.empty(n, dtype=T.float32, device='cuda')
                _lib.call('tf_scale_f32', n, 1.0 / s, ptr(r1), ptr(x_in), st)
                _lib.call('tf_f64_to_f32', n, ptr(e1), ptr(y_out), st)
                _lib.call('tf_scale_f32', n, 1.0 / t, ptr(y_out), ptr(y_out), st)
                pairs.append((x_in, y_out))
            jac.prec, jac.prec_exact = old
    finally:
        if saved is None:
            itf.parameters.pop('HIP Solver', None)
        else:
            itf.parameters['HIP Solver'] = saved
    T.cuda.synchronize()
    return pairs


def evaluate(net, itf, recs, max_samples=64):
    """Mean fraction of ||r1||^2 that one judged correction removes on recorded
    residuals (0: no gain; 1: exact)."""
    lc = LearnedCorrection(net, (itf.nx, itf.ny, itf.nz), itf.n, off_below=-1.0)
    g = itf.solver.gauge()
    gains = []
    for jac, r1 in recs[:max_samples]:
        val32 = jac.values32()

        def mv(v, out, jac=jac, val32=val32):
            _lib.call('tf_spmv_f32', itf.n, ptr(jac.row_ptr), ptr(jac.col), ptr(val32), ptr(v), ptr(out), int(g),
                      stream())

        class _Zero:
            @staticmethod
            def apply32(v, out):
                _lib.call('tf_fill_f32', v.numel(), 0.0, ptr(out), stream())
                return out
        lc.bind(_Zero, mv)
        T = torch()
        out = T.empty_like(r1)
        before = lc.gain_sum
        lc.apply32(r1, out)   # base zero: the correction alone acts on r1
        gains.append(lc.gain_sum - before)
    return float(numpy.mean(gains)) if gains else 0.0
