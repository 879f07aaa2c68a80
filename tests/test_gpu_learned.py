"""GPU tests of the learned preconditioner: the BF16 MFMA convolution kernels
against the numpy oracle (oracle/conv.py, with the same BF16 roundings and
FP64 accumulation), training by the project's own kernels, and the judged
correction inside the solver."""
import numpy
import pytest

from oracle import conv as oc

pytestmark = pytest.mark.gpu


def _bits(x):
    return (oc.bf16(x).view(numpy.uint32) >> 16).astype(numpy.uint16).view(numpy.int16)


def _pad16(v):
    return (v + 15) // 16 * 16


@pytest.mark.parametrize('shape,dim,cin,cout,relu', [((13, 9, 1), 2, 3, 32, 1), ((6, 5, 4), 3, 32, 4, 0),
                                                     ((17, 7, 1), 2, 32, 32, 1), ((5, 4, 3), 3, 4, 32, 1),
                                                     ((20, 17, 9), 3, 32, 32, 1), ((19, 6, 7), 3, 4, 32, 1)])
def test_conv_layer_parity(hip, shape, dim, cin, cout, relu):
    import torch
    from transiflow_amd import _lib
    from transiflow_amd.vector import ptr, stream
    rng = numpy.random.default_rng(0)
    nx, ny, nz = shape
    n = nx * ny * nz
    taps = oc.taps(dim)
    K = taps * cin
    kpad, npad, kpt, npt = _pad16(K), _pad16(cout), _pad16(taps * cout), _pad16(cin)
    x = oc.bf16(rng.standard_normal((n, cin)).astype(numpy.float32))
    W = numpy.zeros((kpad, npad), numpy.float32)
    W[:K, :cout] = rng.standard_normal((K, cout)) * 0.3
    b = rng.standard_normal(cout).astype(numpy.float32)
    g = oc.bf16(rng.standard_normal((n, cout)).astype(numpy.float32))
    d = {k: torch.from_numpy(v).cuda() for k, v in
         dict(x=_bits(x), W=W, b=b, g16=_bits(g), g=g).items()}
    Wf = torch.zeros(kpad * npad, dtype=torch.int16, device='cuda')
    Wt = torch.zeros(kpt * npt, dtype=torch.int16, device='cuda')
    out = torch.zeros((n, cout), dtype=torch.float32, device='cuda')
    _lib.call('tf_conv_pack', dim, cin, cout, kpad, npad, kpt, npt, ptr(d['W']), ptr(Wf), ptr(Wt), stream())
    _lib.call('tf_conv_fwd', nx, ny, nz, dim, cin, cout, kpad, npad, ptr(d['x']), ptr(Wf), ptr(d['b']), relu,
              ptr(out), None, stream())
    Wb = oc.bf16(W[:K, :cout]).astype(float)
    ref = oc.forward(x.astype(float), Wb, b.astype(float), shape, dim, relu)
    assert numpy.abs(out.cpu().numpy() - ref).max() <= 1e-5 * numpy.abs(ref).max()
    work = torch.zeros(_lib.load().tf_conv_wgrad_work_size(n, kpad, npad), dtype=torch.float32, device='cuda')
    dW = torch.zeros((kpad, npad), dtype=torch.float32, device='cuda')
    db = torch.zeros(cout, dtype=torch.float32, device='cuda')
    _lib.call('tf_conv_wgrad', nx, ny, nz, dim, cin, cout, kpad, npad, ptr(d['x']), ptr(d['g16']), ptr(d['g']),
              ptr(work), ptr(dW), ptr(db), stream())
    rdW, rdb, rdx = oc.backward(x.astype(float), Wb, None, g.astype(float), shape, dim, False)
    assert numpy.abs(dW.cpu().numpy()[:K, :cout] - rdW).max() <= 1e-5 * numpy.abs(rdW).max()
    dWh = dW.cpu().numpy()
    assert not dWh[K:, :].any() and not dWh[:, cout:].any()   # padding of the MFMA layout stays zero
    assert numpy.abs(db.cpu().numpy() - rdb).max() <= 1e-5 * numpy.abs(rdb).max()
    dx = torch.zeros((n, cin), dtype=torch.float32, device='cuda')
    _lib.call('tf_conv_fwd', nx, ny, nz, dim, cout, cin, kpt, npt, ptr(d['g16']), ptr(Wt), None, 0, ptr(dx), None,
              stream())
    assert numpy.abs(dx.cpu().numpy() - rdx).max() <= 1e-5 * numpy.abs(rdx).max()


def _oracle_net(net, x, shape, g_y):
    """The network's forward and backward in FP64 with the GPU's BF16 roundings."""
    Ws = [(oc.bf16(W).astype(float), b.astype(float)) for W, b in net.weights()]
    dim = net.dim
    a = [oc.bf16(x.astype(numpy.float32)).astype(float)]
    for i, (W, b) in enumerate(Ws):
        y = oc.forward(a[-1], W, b, shape, dim, relu=i < 2)
        a.append(oc.bf16(y.astype(numpy.float32)).astype(float) if i < 2 else y)
    grads = [None] * 3
    g = g_y.astype(float)
    for i in range(2, -1, -1):
        g16 = oc.bf16(g.astype(numpy.float32)).astype(float)
        dW, _, dx = oc.backward(a[i], Ws[i][0], None, g16, shape, dim, False)
        grads[i] = (dW, g.sum(axis=0))
        if i > 0:
            g = dx * (a[i] > 0)
    return a[-1], grads


@pytest.mark.parametrize('dim,dof,shape', [(2, 3, (16, 12, 1)), (3, 4, (6, 5, 4))])
def test_network_gradients_match_oracle(hip, dim, dof, shape):
    import torch
    from transiflow_amd.learned import ConvNet
    rng = numpy.random.default_rng(1)
    net = ConvNet(dim, dof, seed=3)
    n = shape[0] * shape[1] * shape[2]
    x = rng.standard_normal((n, dof)).astype(numpy.float32)
    t = rng.standard_normal((n, dof)).astype(numpy.float32)
    xd, td = torch.from_numpy(x.ravel()).cuda(), torch.from_numpy(t.ravel()).cuda()
    y = net.forward(xd, shape)
    net.mse(y, td)
    net.backward(net.g_loss, shape)
    g_y = net.g_loss.cpu().numpy().reshape(n, dof)
    y_ref, grads_ref = _oracle_net(net, x, shape, g_y)
    assert numpy.abs(y.cpu().numpy().reshape(n, dof) - y_ref).max() <= 1e-4 * numpy.abs(y_ref).max()
    for (dW, db), (rW, rb) in zip(net.grads(), grads_ref):
        assert numpy.abs(dW - rW).max() <= 2e-3 * numpy.abs(rW).max()
        assert numpy.abs(db - rb).max() <= 2e-3 * numpy.abs(rb).max()


def test_training_reduces_loss(hip):
    """Student network fits a teacher network's outputs with the GPU kernels."""
    import torch
    from transiflow_amd.learned import ConvNet, train
    rng = numpy.random.default_rng(2)
    shape = (32, 32, 1)
    teacher = ConvNet(2, 3, seed=11)
    student = ConvNet(2, 3, seed=12)
    samples = []
    for _ in range(8):
        x = torch.from_numpy(rng.standard_normal(32 * 32 * 3).astype(numpy.float32)).cuda()
        samples.append((x, teacher.forward(x, shape).clone()))
    hist = train(student, samples, shape, steps=600, lr=3e-3)
    first, last = numpy.mean(hist[:20]), numpy.mean(hist[-20:])
    print('teacher-student loss %.3e -> %.3e' % (first, last))
    assert last < 0.25 * first


def test_learned_correction_ab(hip):
    """A/B of the learned correction inside the mixed solver.  The network is
    trained on (Jacobian, residual, correction) triples recorded from the plain
    mixed continuation of a 48x48 cavity (Re 0 -> 1000) and evaluated on the
    held-out later part of the trajectory.  Reported: the fraction of the base
    preconditioner's residual that one correction removes, and the inner
    iterations of the continuation with and without the network.  Asserted:
    training lowers the loss, the judged correction never removes a negative
    amount, both runs reach the same converged states, and the correction does
    not cost convergence (it switches itself off when it does not help)."""
    from transiflow_amd import learned
    import transiflow_amd as tf
    N = 48

    def run(itf):
        cont = tf.Continuation(itf, newton_tolerance=1e-8, printer=None)
        x0 = cont.newton(itf.vector(), tol=1e-8)
        x, mu = cont.continuation(x0, 'Reynolds Number', 0, 1000, 100)
        run.x = itf.array_from_vector(x)
        run.mu = mu

    itf, recs = learned.collect(lambda: tf.Interface({'HIP Solver': {'Precision': 'mixed'}}, N, N), run,
                                every=2, max_samples=120)
    pairs = learned.make_pairs(itf, recs)
    ntr = int(0.7 * len(pairs))
    net = learned.ConvNet(2, 3, seed=0)
    hist = learned.train(net, pairs[:ntr], (N, N, 1), steps=400, lr=1e-3)
    gain_train = learned.evaluate(net, itf, recs[:ntr])
    gain_held = learned.evaluate(net, itf, recs[ntr:])
    out = {}
    for label, p in (('plain', {'Precision': 'mixed'}), ('learned', {'Precision': 'mixed', 'Learned Network': net})):
        it2 = tf.Interface({'HIP Solver': p}, N, N)
        run(it2)
        out[label] = (run.x, run.mu, sum(entry[0] for entry in it2.solver.log), it2.solver._learned)
    lc = out['learned'][3]
    print('triples %d, loss %.3f -> %.3f, gain (fraction of |r1|^2 removed) train %.4f held-out %.4f; '
          'inner its plain %d learned %d; correction applied %d accepted %d switched off %s'
          % (len(pairs), numpy.mean(hist[:20]), numpy.mean(hist[-20:]), gain_train, gain_held, out['plain'][2],
             out['learned'][2], lc.applied, lc.accepted, lc.off))
    assert numpy.mean(hist[-20:]) < numpy.mean(hist[:20])
    assert 0.0 <= gain_held <= 1.0 and 0.0 <= gain_train <= 1.0
    xp, xl = out['plain'][0], out['learned'][0]
    assert numpy.linalg.norm(xl - xp) / numpy.linalg.norm(xp) <= 1e-8
    assert out['learned'][1] == pytest.approx(out['plain'][1], abs=1e-8)
    assert out['learned'][2] <= out['plain'][2] + 5
