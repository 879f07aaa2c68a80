"""Parity at the full sizes of the benchmark configurations.

* Lid-driven cavity 2D 256 x 256: the whole continuation Re 0 -> 2000
  (ds = 100, Newton tolerance 1e-8) against the oracle's run of the same
  continuation (scipy splu with the reference's gauge), stored by
  tests/golden/make_cavity_continuation.py: accepted Reynolds numbers, Newton
  iterations per step and the final state, for the FP64 and the
  mixed-precision solver.
* Rayleigh-Benard 2D 512 x 128 (dof 4, Pr 1, Biot 1, X-max 4): bit-exact
  CSR pattern and values within 1e-12 against the oracle at a seeded state,
  one solve against the oracle's splu, and the configured continuation
  Ra 0 -> 5000 from the seeded 1e-3 temperature perturbation against the
  oracle's.
* The 64-instance batch (square cavities n in {32..256}, Re in U[100, 1500],
  default_rng(2)): every instance run concurrently reaches its Reynolds
  number with the same result as its run alone on one stream.
* Multi-column right-hand sides: an (n, k) rhs against splu's solve of the
  gauged (n, k) rhs.
"""
import os

import numpy
import pytest

from tests.helpers import assert_csr_equal, assert_vec_close, seeded_state

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
pytestmark = pytest.mark.gpu

RB = {'Problem Type': 'Rayleigh-Benard', 'Prandtl Number': 1.0, 'Biot Number': 1.0, 'X-max': 4.0}


def _continuation(hip, itf, x0_host, param, target, ds):
    cont = hip.Continuation(itf, newton_tolerance=1e-8, printer=None)
    mus, its = [], []

    def cb(_i, _x, mu):
        mus.append(mu)
        its.append(cont.newton_iterations)

    x0 = cont.newton(itf.vector_from_array(x0_host), tol=1e-8)
    x, mu = cont.continuation(x0, param, 0, target, ds, callback=cb)
    return itf.array_from_vector(x), mu, numpy.array(mus), numpy.array(its)


@pytest.mark.timeout(300)
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_cavity_continuation_256_matches_oracle(hip, precision):
    g = numpy.load(os.path.join(GOLDEN, 'cavity_continuation_256.npz'))
    itf = hip.Interface({'HIP Solver': {'Precision': precision}}, 256, 256)
    x, mu, mus, its = _continuation(hip, itf, numpy.zeros(itf.n), 'Reynolds Number', 2000, 100)
    err = numpy.linalg.norm(x - g['x']) / numpy.linalg.norm(g['x'])
    print('%s: %d accepted steps, Newton its %s, final Re %.10f, |x - x_oracle|/|x_oracle| %.2e, '
          'max |mu - mu_oracle| %.2e' % (precision, len(mus) - 1, its.tolist(), mu, err,
                                         numpy.max(numpy.abs(mus - g['mus'])) if len(mus) == len(g['mus'])
                                         else numpy.inf), flush=True)
    assert numpy.array_equal(its, g['newton_its'])
    numpy.testing.assert_allclose(mus, g['mus'], rtol=1e-9, atol=1e-9)
    assert mu == pytest.approx(float(g['mu']), rel=1e-9)
    assert err <= 1e-8


@pytest.mark.timeout(300)
def test_rb_512x128_assembly_and_solve(hip):
    from oracle.interface import Interface as OracleInterface
    params = dict(RB, **{'Rayleigh Number': 5000.0})
    itf = hip.Interface(dict(params), 512, 128)
    ref = OracleInterface(dict(params), 512, 128)
    state = seeded_state(itf.n, 31) * 0.1
    F = itf.array_from_vector(itf.rhs(state))
    fe = assert_vec_close(F, ref.rhs(state))
    J = itf.jacobian(state)
    Jo = ref.discretization.jacobian(state)
    je = assert_csr_equal((J.begA, J.jcoA, J.coA), Jo.begA, Jo.jcoA, Jo.coA)
    rhs = seeded_state(itf.n, 32)
    y = itf.array_from_vector(itf.solve(J, rhs))
    Jc = ref.jacobian(state)
    yo = ref.solve(Jc, rhs)
    A = ref.compute_bordered_matrix(Jc, fix_pressure_row=True)
    b = rhs.copy()
    b[ref.pressure_row] = 0
    res = numpy.linalg.norm(A @ y - b) / numpy.linalg.norm(b)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    print('RB 512x128: F bit-exact %.4f, J bit-exact %.4f, nnz %d, solve rel res %.2e, vs splu %.2e'
          % (fe, je, Jo.nnz, res, err), flush=True)
    assert res <= 1e-9 and err <= 1e-8


@pytest.mark.timeout(400)
def test_rb_512x128_continuation_matches_oracle(hip):
    """The configured Rayleigh continuation: Newton at Ra = 0 from the seeded
    1e-3 perturbation of T (linear, so it lands on the conduction state), then
    Ra 0 -> 5000 with ds = 100, mixed precision; the oracle runs the same."""
    from oracle.interface import Interface as OracleInterface
    from transiflow_amd.continuation import Continuation
    params = dict(RB, **{'Rayleigh Number': 0.0})
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Precision': 'mixed'}}), 512, 128)
    h = numpy.zeros(itf.n)
    h[itf.dof - 1::itf.dof] = 1e-3 * numpy.random.default_rng(1).uniform(-1, 1, itf.n // itf.dof)
    x, mu, mus, its = _continuation(hip, itf, h, 'Rayleigh Number', 5000, 100)
    ref = OracleInterface(dict(params), 512, 128)
    cont = Continuation(ref, newton_tolerance=1e-8, printer=print)   # progress lines: the CPU run takes a minute
    rmus, rits = [], []

    def cb(_i, _x, m):
        rmus.append(m)
        rits.append(cont.newton_iterations)

    xo, muo = cont.continuation(cont.newton(h, tol=1e-8), 'Rayleigh Number', 0, 5000, 100, callback=cb)
    err = numpy.linalg.norm(x - xo) / numpy.linalg.norm(xo)
    print('RB continuation: %d steps, Newton its %s (oracle %s), final Ra %.6f, rel err %.2e'
          % (len(mus) - 1, its.tolist(), rits, mu, err), flush=True)
    assert numpy.array_equal(its, numpy.array(rits))
    numpy.testing.assert_allclose(mus, rmus, rtol=1e-9, atol=1e-9)
    assert err <= 1e-8


@pytest.mark.timeout(400)
def test_batch_64_instances(hip):
    from transiflow_amd.batch import c4_instances, run_batch
    inst = c4_instances()
    one, _ = run_batch(inst, workers=1)
    many, wall = run_batch(inst, workers=8)
    worst = 0.0
    for i, s, b in zip(inst, one, many):
        err = numpy.linalg.norm(b['x'] - s['x']) / numpy.linalg.norm(s['x'])
        worst = max(worst, err)
        assert b['accepted'] == s['accepted'] and b['mu'] == s['mu'], i
        assert err <= 1e-12, (i, err)
        assert abs(b['mu'] - i['Re']) <= 1e-4 * i['Re'], (i, b['mu'])   # reached its target
    print('64 instances, %d unknowns: %.2f s on 8 streams, worst rel diff to one stream %.1e'
          % (sum(3 * i['n'] ** 2 for i in inst), wall, worst), flush=True)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_solve_multi_column_rhs(hip, precision):
    """SciPy.py:209-216: every column of an (n, k) rhs has its pressure_row
    entry zeroed and is solved with the same factor."""
    from oracle.interface import Interface as OracleInterface
    params = {'Reynolds Number': 700}
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Precision': precision}}), 40, 32)
    ref = OracleInterface(dict(params), 40, 32)
    state = seeded_state(itf.n, 41) * 0.3
    B = numpy.stack([seeded_state(itf.n, 42 + j) for j in range(3)], axis=1)
    Y = itf.solve(itf.jacobian(state), B)
    Yo = ref.solve(ref.jacobian(state), B)
    assert isinstance(Y, numpy.ndarray) and Y.shape == B.shape
    err = numpy.linalg.norm(Y - Yo, axis=0) / numpy.linalg.norm(Yo, axis=0)
    print('(n, 3) rhs, %s: per-column rel err vs splu %s' % (precision, err), flush=True)
    assert numpy.all(err <= 1e-8)
    with pytest.raises(ValueError):
        itf.solve(itf.jacobian(state), B[:-1])


@pytest.mark.timeout(300)
def test_rb_convective_branch_matches_oracle(hip):
    """Rayleigh-Benard 64 x 16 (Pr 1, Biot 1, X-max 4) on the convective branch,
    where the continuation does real corrector work: from the conduction state at
    Ra = 5000 plus a seeded temperature perturbation, 150 backward-Euler steps
    (dt = 1) grow the rolls, Newton converges them, and the continuation runs
    Ra 5000 -> 6000 (ds 100, tol 1e-8).  The HIP backend (mixed precision) and
    the oracle land on the same steady rolls and take the same steps."""
    from oracle.interface import Interface as OracleInterface
    from oracle.time_integration import TimeIntegration as OracleTI
    from transiflow_amd.continuation import Continuation
    params = dict(RB, **{'Rayleigh Number': 5000.0})
    out = {}
    for name in ('hip', 'oracle'):
        if name == 'hip':
            itf = hip.Interface(dict(params, **{'HIP Solver': {'Precision': 'mixed'}}), 64, 16)
            ti = hip.TimeIntegration(itf, theta=1.0, newton_tolerance=1e-8, printer=None)
            conv = hip.Continuation
        else:
            itf = OracleInterface(dict(params), 64, 16)
            ti = OracleTI(itf, theta=1.0, newton_tolerance=1e-8)
            conv = Continuation
        n = itf.vector().size
        cont = conv(itf, newton_tolerance=1e-8, printer=None)
        x = cont.newton(itf.vector(), tol=1e-10)
        h = numpy.zeros(n)
        h[3::4] = 1e-2 * numpy.random.default_rng(1).uniform(-1, 1, n // 4)
        x = x + (itf.vector_from_array(h) if name == 'hip' else h)
        for _ in range(150):
            x = ti.step(x, 1.0)
        x = cont.newton(x, tol=1e-10)
        xs = numpy.asarray(itf.array_from_vector(x) if name == 'hip' else x).copy()
        mus, its = [], []
        x, mu = cont.continuation(x, 'Rayleigh Number', 5000, 6000, 100,
                                  callback=lambda _i, _x, m: (mus.append(m), its.append(cont.newton_iterations)))
        out[name] = (xs, numpy.asarray(itf.array_from_vector(x) if name == 'hip' else x), mu, mus, its)
    (sh, xh, mh, mush, itsh), (so, xo, mo, muso, itso) = out['hip'], out['oracle']
    vel = numpy.arange(len(so)) % 4 < 2
    e0 = numpy.linalg.norm(sh - so) / numpy.linalg.norm(so)
    e1 = numpy.linalg.norm(xh - xo) / numpy.linalg.norm(xo)
    print('convective rolls: max |u| %.3f, start rel diff %.2e; continuation Newton its %s (oracle %s), '
          'final rel diff %.2e' % (numpy.abs(so[vel]).max(), e0, itsh, itso, e1), flush=True)
    assert numpy.abs(so[vel]).max() > 0.05          # rolls, not the conduction state
    assert e0 <= 1e-8 and e1 <= 1e-8
    assert itsh == itso and sum(itso[1:]) > 0        # real corrector work
    numpy.testing.assert_allclose(mush, muso, rtol=1e-9)
