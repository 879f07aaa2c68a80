"""GPU linear solves and Newton/continuation parity.

solve(): restarted GMRES on the pressure-gauged Jacobian, preconditioned by
the exact block-tridiagonal line factorisation (block cyclic reduction, and
the Thomas recursion for comparison), must reach a relative
residual of 1e-10 and agree with the oracle's sparse-LU (scipy splu) solution
of the same gauged system (SciPy.py:204-258 semantics) to 1e-8 relative.
Newton and the pseudo-arclength continuation are compared with vectors
written from the reference (tests/golden/*.npz).
"""
import os

import numpy
import pytest

from tests.helpers import seeded_state

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
pytestmark = pytest.mark.gpu


def _solve_case(hip, params, shape, dim, dof, seed, prec='line', precision='fp64', amplitude=1.0, storage='fp32',
                factor_gemm='fp32'):
    from oracle.interface import Interface as OracleInterface
    p = dict(params)
    p['HIP Solver'] = {'Preconditioner': prec, 'Precision': precision, 'Apply Storage': storage,
                       'Factor GEMM': factor_gemm}
    itf = hip.Interface(p, *shape, dim, dof)
    ref = OracleInterface(dict(params), *shape, dim, dof)
    state = seeded_state(itf.n, seed) * amplitude
    rhs = seeded_state(itf.n, seed + 100)
    J = itf.jacobian(state)
    y = itf.array_from_vector(itf.solve(J, rhs))
    yo = ref.solve(ref.jacobian(state), rhs)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    # residual of the gauged system, computed on the host
    A = ref.compute_bordered_matrix(ref.jacobian(state), fix_pressure_row=True)
    b = rhs.copy()
    b[ref.pressure_row] = 0
    res = numpy.linalg.norm(A @ y - b) / numpy.linalg.norm(b)
    its, rel, name, fresh = itf.solver.log[-1]
    print('%s %s: GMRES its %d, rel res %.2e (host %.2e), |y - y_splu|/|y_splu| %.2e'
          % (shape, name, its, rel, res, err), flush=True)
    assert res <= 1e-9
    assert err <= 1e-8
    return itf


@pytest.mark.parametrize('label,params,shape,dim,dof,seed', [
    ('ldc2d_16', {'Reynolds Number': 100}, (16, 16, 1), 2, 3, 1),
    ('ldc2d_64x40', {'Reynolds Number': 2000}, (64, 40, 1), 2, 3, 2),
    ('ldc2d_stretched', {'Reynolds Number': 500, 'Grid Stretching': True}, (48, 48, 1), 2, 3, 3),
    ('rb2d_64x16', {'Problem Type': 'Rayleigh-Benard', 'Prandtl Number': 1.0, 'Rayleigh Number': 3000.0,
                    'Biot Number': 1.0, 'X-max': 4.0}, (64, 16, 1), 2, 4, 4),
    ('ldc3d_6', {'Reynolds Number': 100}, (6, 6, 6), 3, 4, 5),
])
def test_solve_line(hip, label, params, shape, dim, dof, seed):
    _solve_case(hip, params, shape, dim, dof, seed)


@pytest.mark.parametrize('shape', [(24, 17, 1), (20, 5, 1), (12, 2, 1)])
def test_solve_line_odd_counts(hip, shape):
    """Reduction schedules with odd active counts at several levels (17 = 8+8+1 ...),
    and a two-line grid."""
    _solve_case(hip, {'Reynolds Number': 400}, shape, 2, 3, 11)


@pytest.mark.parametrize('label,params,shape,dim,dof,seed', [
    ('ldc2d_64x40', {'Reynolds Number': 2000}, (64, 40, 1), 2, 3, 2),
    ('ldc3d_6', {'Reynolds Number': 100}, (6, 6, 6), 3, 4, 5),
])
def test_solve_thomas(hip, label, params, shape, dim, dof, seed):
    _solve_case(hip, params, shape, dim, dof, seed, prec='thomas')


def test_solve_full_size_ldc256(hip):
    """C2 size: one Jacobian at a seeded state, Re = 2000."""
    _solve_case(hip, {'Reynolds Number': 2000}, (256, 256, 1), 2, 3, 6)


@pytest.mark.parametrize('shape,re', [((64, 64, 1), 1000), ((128, 96, 1), 2000), ((256, 256, 1), 2000)])
def test_solve_mixed_refinement(hip, shape, re):
    """FP64 refinement around FP32 FGMRES with the FP32 line factor, at a state of
    flow-like magnitude (|u| <= 0.1): FP64 residual to 1e-10 and the splu solution
    to 1e-8, reached by the inner FP32 solves alone (>= 2 refinement steps, no
    FP64 fallback)."""
    itf = _solve_case(hip, {'Reynolds Number': re}, shape, 2, 3, 12, precision='mixed', amplitude=0.1)
    assert itf.solver.refinements[-1] >= 2
    assert itf.solver.fallbacks == 0


@pytest.mark.parametrize('shape,re', [((64, 64, 1), 1000), ((256, 256, 1), 2000)])
def test_solve_mixed_fp16_copies(hip, shape, re):
    """The same refinement with the inner apply reading FP16 copies of the FP32
    factor blocks: the FP64 bar is unchanged (1e-10 residual, splu to 1e-8)."""
    itf = _solve_case(hip, {'Reynolds Number': re}, shape, 2, 3, 12, precision='mixed', amplitude=0.1,
                      storage='fp16')
    assert itf.solver.fallbacks == 0


@pytest.mark.parametrize('factor_gemm', ['bf16x3', 'bf16'])
@pytest.mark.parametrize('shape,re', [((64, 40, 1), 1000), ((256, 256, 1), 2000)])
def test_solve_mixed_bf16_factor_gemm(hip, shape, re, factor_gemm):
    """The FP32 factor with its level GEMMs on BF16 MFMA (three-term splits or plain
    BF16 rounding of the operands) inside the same refinement: the FP64 bar is
    unchanged (1e-10 residual, splu to 1e-8)."""
    itf = _solve_case(hip, {'Reynolds Number': re}, shape, 2, 3, 12, precision='mixed', amplitude=0.1,
                      storage='fp16', factor_gemm=factor_gemm)
    assert itf.solver.fallbacks == 0


@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1), (24, 17, 1), (100, 100, 1), (120, 104, 1)])
def test_half_apply(hip, shape):
    """FP32 apply from FP16 copies of the factor blocks (one power-of-two FP32 scale
    per row): still an approximate inverse of the gauged Jacobian (residual of
    A z = r at FP16 level), close to the apply from the FP32 blocks (100 x 100 and
    120 x 104: the widest levels take two rows per lane group, with a ragged last
    block; m = 300 reads element-wise, m = 312 / 360 with 16-byte loads), bit-identical
    between eager and graph replays and across switching the copies off and on, and
    remade by a refactorisation (the residual bar holds for a second Jacobian
    factored into the same handle)."""
    import torch
    itf = hip.Interface({'Reynolds Number': 500, 'HIP Solver': {'Precision': 'mixed', 'Apply Storage': 'fp16'}},
                        *shape)
    n = itf.n
    g = itf.solver.gauge()
    rng = numpy.random.default_rng(41)
    r = torch.from_numpy(rng.uniform(-1, 1, n)).cuda()
    if g >= 0:
        r[g] = 0.0
    r32 = r.float()

    def apply32(P, graph):
        P.set_graph(graph)
        z = torch.empty_like(r32)
        P.apply32(r32, z)
        torch.cuda.synchronize()
        return z.cpu().numpy()

    def residual(J, z):
        out = J.matvec(torch.from_numpy(z.astype(numpy.float64)).cuda(), gauge=g)
        return float(torch.linalg.norm(out - r) / torch.linalg.norm(r))

    prev = None
    for k in range(2):
        state = itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1)
        J = itf.jacobian(state)
        P = itf.solver._build(J, 'line', True)
        assert P.half
        zh = [apply32(P, False), apply32(P, True), apply32(P, True)]
        P.set_half(False)
        z32 = [apply32(P, True), apply32(P, False)]
        P.set_half(True)
        zh.append(apply32(P, True))
        for z in zh[1:]:
            assert numpy.array_equal(z, zh[0])
        assert numpy.array_equal(z32[0], z32[1])
        res_h, res_32 = residual(J, zh[0]), residual(J, z32[0])
        diff = numpy.linalg.norm(zh[0] - z32[0]) / numpy.linalg.norm(z32[0])
        print('%s state %d: residual FP16 copies %.2e, FP32 blocks %.2e, |z16 - z32|/|z32| %.2e'
              % (shape, k, res_h, res_32, diff), flush=True)
        assert numpy.isfinite(zh[0]).all()
        # the FP16 residual grows with the grid's conditioning: 4e-3 .. 8e-3 up to
        # 48 x 48, 1.03e-2 at 100 x 100
        assert res_h <= (1e-2 if n <= 48 * 48 * 3 else 2.5e-2)
        assert diff <= 1e-2
        if prev is not None:
            assert not numpy.array_equal(zh[0], prev)
        prev = zh[0]
    P.set_graph(False)


def test_background_factor_continuation(hip):
    """Mixed precision with the next factor made on a side stream while the inner
    solves run with the lagged one: the continuation reaches the FP64 method's
    final state to 1e-8, every solve meets the 1e-10 bar, and background factors
    were taken into use."""
    import numpy as np
    out = {}
    for label, p in (('fp64', {'Precision': 'fp64'}),
                     ('bg', {'Precision': 'mixed', 'Apply Storage': 'fp16', 'Background Factor': True})):
        itf = hip.Interface({'HIP Solver': p}, 64, 64)
        c = hip.Continuation(itf, newton_tolerance=1e-8, printer=None)
        x0 = c.newton(itf.vector(), tol=1e-8)
        x, mu = c.continuation(x0, 'Reynolds Number', 0, 1000, 100)
        out[label] = (itf.array_from_vector(x), mu, itf)
    xb, mub, itf = out['bg']
    xf, muf, _ = out['fp64']
    err = np.linalg.norm(xb - xf) / np.linalg.norm(xf)
    bg = itf.solver._bg
    print('background factor: %d started, %d taken into use, %d solves, final state rel err %.2e'
          % (bg.started, bg.swaps, len(itf.solver.log), err), flush=True)
    assert abs(mub - muf) <= 1e-8 * max(1.0, abs(muf))
    assert err <= 1e-8
    assert bg.swaps >= 1
    assert max(entry[1] for entry in itf.solver.log) <= 1e-10


def test_solve_mixed_fallback(hip):
    """At a state of unit magnitude the line blocks reach condition numbers ~1e7 and
    the FP32 factor cannot refine: the solve finishes in FP64 and still meets the
    1e-10 bar."""
    itf = _solve_case(hip, {'Reynolds Number': 2000}, (128, 96, 1), 2, 3, 12, precision='mixed')
    assert itf.solver.fallbacks == 1


def test_solve_mixed_stale_factor_refactors(hip):
    """A lagged FP32 factor from a far-away Jacobian (Re = 0 at rest, used for
    Re = 2000 at a moving state) with the reuse limit out of reach: the solver
    still refactors that Jacobian once (its inner solves stop short of the
    inner tolerance, or the refinement stagnates) and the solve meets 1e-10 and
    agrees with splu."""
    from oracle.interface import Interface as OracleInterface
    params = {'Reynolds Number': 0.0}
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Precision': 'mixed', 'Reuse Limit': 1000}}), 48, 40)
    ref = OracleInterface(dict(params), 48, 40)
    rhs = seeded_state(itf.n, 51)
    itf.solve(itf.jacobian(itf.vector()), rhs)
    assert itf.solver.stale_refactors == 0
    itf.set_parameter('Reynolds Number', 2000.0)
    ref.set_parameter('Reynolds Number', 2000.0)
    state = seeded_state(itf.n, 52) * 0.3
    y = itf.array_from_vector(itf.solve(itf.jacobian(state), rhs))
    yo = ref.solve(ref.jacobian(state), rhs)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    print('stale factor: refactors %d (stagnation) %d (limit), refinement trace %s, vs splu %.2e'
          % (itf.solver.stale_refactors, itf.solver.limit_refactors, itf.solver.trace, err), flush=True)
    # the reuse limit is out of reach, so the factor is replaced either because an
    # inner solve ran out of iterations short of the inner tolerance ('limit') or
    # because the refinement stagnates ('stale'), exactly once
    assert itf.solver.stale_refactors + itf.solver.limit_refactors == 1
    assert itf.solver.log[-1][1] <= 1e-10 and itf.solver.log[-1][3]
    assert err <= 1e-8


def test_newton_ldc2d_16(hip):
    """C1: Newton from zero at Re = 100 against the reference's result."""
    g = numpy.load(os.path.join(GOLDEN, 'ldc2d_16.npz'))
    itf = hip.Interface({'Reynolds Number': 100}, 16, 16, 1, 2, 3)
    cont = hip.Continuation(itf, newton_tolerance=1e-8, printer=None)
    x = cont.newton(itf.vector(), tol=1e-8)
    xh = itf.array_from_vector(x)
    err = numpy.linalg.norm(xh - g['x_newton']) / numpy.linalg.norm(g['x_newton'])
    print('Newton 16x16: rel err vs reference %.2e' % err)
    assert err <= 1e-8


def test_continuation_ldc2d_8(hip):
    """LDC 2D 8x8, Re 0 -> 2000, ds = 100: accepted mu sequence, Newton counts and
    final state against the reference's trajectory."""
    g = numpy.load(os.path.join(GOLDEN, 'cont_ldc2d_8.npz'))
    itf = hip.Interface({}, 8, 8, 1, 2, 3)
    cont = hip.Continuation(itf, newton_tolerance=1e-8, printer=None)
    x0 = cont.newton(itf.vector())
    mus, its = [], []

    def cb(_itf, _x, mu):
        mus.append(mu)
        its.append(cont.newton_iterations)

    x, mu = cont.continuation(x0, 'Reynolds Number', 0, 2000, 100, callback=cb)
    assert len(mus) == len(g['mus'])
    numpy.testing.assert_allclose(mus, g['mus'], rtol=1e-7, atol=1e-7)
    assert list(its) == list(g['newton_its'])
    xh = itf.array_from_vector(x)
    err = numpy.linalg.norm(xh - g['x']) / numpy.linalg.norm(g['x'])
    print('continuation 8x8: final mu %.6f, rel err of final state %.2e' % (mu, err))
    assert abs(mu - float(g['mu'])) <= 1e-6
    assert err <= 1e-8


@pytest.mark.parametrize('k', [1, 2])
def test_bordered_solve_matches_oracle(hip, k):
    """[J~, sV; sW~^T, s^2 C] solve (SciPy.py:51-129,204-258) against the oracle's
    sparse LU of the assembled bordered matrix, with k border columns."""
    from oracle.interface import Interface as OracleInterface
    params = {'Reynolds Number': 300}
    itf = hip.Interface(dict(params), 24, 20)
    ref = OracleInterface(dict(params), 24, 20)
    n = itf.n
    rng = numpy.random.default_rng(21)
    state = rng.uniform(-1, 1, n) * 0.1
    f = rng.uniform(-1, 1, n)
    if k == 1:
        V, W, C, r2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), 0.7, 0.3
    else:
        V, W = rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (n, k))
        C, r2 = rng.uniform(-1, 1, (k, k)) + 2 * numpy.eye(k), rng.uniform(-1, 1, k)
    y1, y2 = itf.solve(itf.jacobian(state), f, r2, V, W, C)
    z1, z2 = ref.solve(ref.jacobian(state), f, r2, V, W, C)
    y1 = itf.array_from_vector(y1)
    err1 = numpy.linalg.norm(y1 - z1) / numpy.linalg.norm(z1)
    err2 = numpy.linalg.norm(numpy.atleast_1d(y2) - numpy.atleast_1d(z2)) / numpy.linalg.norm(numpy.atleast_1d(z2))
    print('bordered k=%d: |y1 - ref| %.2e, |y2 - ref| %.2e' % (k, err1, err2))
    assert err1 <= 1e-8 and err2 <= 1e-8
    assert numpy.ndim(y2) == (0 if k == 1 else 1)


def test_bordered_continuation_matches_oracle(hip):
    """Continuation(bordered_solver=True) (Continuation.py:152-155): HIP backend vs
    the same driver on the oracle backend, LDC 8x8, Re 0 -> 2000."""
    from oracle.interface import Interface as OracleInterface
    out = {}
    for name, itf in (('hip', hip.Interface({}, 8, 8)), ('oracle', OracleInterface({}, 8, 8, 1, 2, 3))):
        cont = hip.Continuation(itf, newton_tolerance=1e-8, bordered_solver=True, printer=None)
        mus, its = [], []

        def cb(_i, _x, mu, mus=mus, its=its, cont=cont):
            mus.append(mu)
            its.append(cont.newton_iterations)
        x0 = cont.newton(itf.vector())
        x, mu = cont.continuation(x0, 'Reynolds Number', 0, 2000, 100, callback=cb)
        out[name] = (numpy.asarray(itf.array_from_vector(x) if name == 'hip' else x), mus, its, mu)
    (xh, mh, ih, muh), (xo, mo, io, muo) = out['hip'], out['oracle']
    assert ih == io
    numpy.testing.assert_allclose(mh, mo, rtol=1e-8, atol=1e-8)
    assert numpy.linalg.norm(xh - xo) / numpy.linalg.norm(xo) <= 1e-8
    assert abs(muh - muo) <= 1e-8


@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1)])
def test_graph_apply_matches_eager(hip, shape):
    """The hipGraph replay of the line-factor apply gives the eager apply's bits:
    several replays interleaved with eager applies, then again after the handle
    is refactored with another Jacobian (the graph keeps its pointers)."""
    import torch
    itf = hip.Interface({'Reynolds Number': 500}, *shape)
    n = itf.n
    rng = numpy.random.default_rng(31)
    states = [itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1) for _ in range(2)]
    r = torch.from_numpy(rng.uniform(-1, 1, n)).cuda()
    r32 = r.float()
    for single in (False, True):
        P = None
        for state in states:
            J = itf.jacobian(state)
            P = itf.solver._build(J, 'line', single)
            outs = {}
            for graph in (False, True, False, True, True):
                P.set_graph(graph)
                z = torch.empty_like(r)
                z32 = torch.empty_like(r32)
                P.apply(r, z)
                P.apply32(r32, z32)
                torch.cuda.synchronize()
                outs.setdefault(graph, []).append((z.cpu().numpy(), z32.cpu().numpy()))
            ref64, ref32 = outs[False][0]
            assert numpy.isfinite(ref64).all()
            for z64, z32 in outs[False][1:] + outs[True]:
                assert numpy.array_equal(z64, ref64)
                assert numpy.array_equal(z32, ref32)
        P.set_graph(False)


@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1), (8, 8, 6), (100, 100, 1), (17, 9, 1)])
@pytest.mark.parametrize('storage', ['fp64', 'fp32', 'fp16'])
def test_small_levels_match_list_kernels(hip, shape, storage):
    """Levels with at most 8 lines through k_small (line indices and block row
    bases as kernel arguments) give, bit for bit, the list-reading kernels: one
    and two columns, eager and replayed, x- and y-line orders and 3D planes,
    FP64 / FP32 / FP16-copy blocks, fused backward (also for the FP64 factor)."""
    import torch
    itf = hip.Interface({'Reynolds Number': 400, 'HIP Solver': {'Fused Apply': True}}, *shape)
    n = itf.n
    rng = numpy.random.default_rng(37)
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1))
    single = storage != 'fp64'
    P = itf.solver._build(J, 'line', single)
    if single:
        P.set_half(storage == 'fp16')
    dt = torch.float32 if single else torch.float64
    r0 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    r1 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    apply = P.apply32 if single else P.apply
    outs = {}
    for small in (False, True, False, True):
        P.set_small(small)
        for graph in (False, True, True):
            P.set_graph(graph)
            z0 = torch.full_like(r0, numpy.nan)
            p0, p1 = torch.full_like(r0, numpy.nan), torch.full_like(r0, numpy.nan)
            apply(r0, z0)
            P.apply_pair(r0, r1, p0, p1, single=single)
            torch.cuda.synchronize()
            outs.setdefault(small, []).append((z0.cpu().numpy(), p0.cpu().numpy(), p1.cpu().numpy()))
    ref = outs[False][0]
    assert all(numpy.isfinite(a).all() for a in ref)
    for small in (False, True):
        for got in outs[small]:
            for a, b in zip(got, ref):
                assert numpy.array_equal(a, b), (small, storage)
    P.set_small(True)
    P.set_graph(True)


@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1), (8, 8, 6), (100, 100, 1), (17, 9, 1)])
@pytest.mark.parametrize('storage', ['fp64', 'fp32', 'fp16'])
def test_dense_top_matches_levels(hip, shape, storage):
    """The dense top (the inverse of the reduced system of the last 2..4 lines, formed
    at factor time in the factor's precision, applied in one launch; FP32 rows under
    the FP16-copy apply) is an inverse as good as the per-level
    sweeps it replaces: residuals of A z = r within twice theirs (or the precision's
    floor), for one and two columns, eager and replayed, at four (3 or 4 lines
    active), two and no lines, switched by refactorisations; the results of one
    setting are bit-identical across eager and graph replays and the pair's columns
    equal single applies.  Shapes give 3- and 4-line tops, x- and y-line orders and
    3D planes."""
    import torch
    itf = hip.Interface({'Reynolds Number': 400, 'HIP Solver': {'Fused Apply': True}}, *shape)
    n = itf.n
    g = itf.solver.gauge()
    rng = numpy.random.default_rng(43)
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1))
    single = storage != 'fp64'
    P = itf.solver._build(J, 'line', single)
    assert getattr(P, 'dense_top', 0) == 0
    if single:
        P.set_half(storage == 'fp16')
    with pytest.raises(Exception):
        P.set_dense_top(5)
    dt = torch.float32 if single else torch.float64
    r64 = [torch.from_numpy(rng.uniform(-1, 1, n)).cuda() for _ in range(2)]
    if g >= 0:
        for r in r64:
            r[g] = 0.0
    r0, r1 = (r.to(dt) for r in r64)
    apply = P.apply32 if single else P.apply

    def run(graph):
        P.set_graph(graph)
        z0 = torch.full_like(r0, numpy.nan)
        p0, p1 = torch.full_like(r0, numpy.nan), torch.full_like(r0, numpy.nan)
        apply(r0, z0)
        P.apply_pair(r0, r1, p0, p1, single=single)
        torch.cuda.synchronize()
        return tuple(v.cpu().numpy() for v in (z0, p0, p1))

    def residual(z, r):
        out = J.matvec(torch.from_numpy(z.astype(numpy.float64)).cuda(), gauge=g)
        return float(torch.linalg.norm(out - r) / torch.linalg.norm(r))

    outs = {}
    for lines in (0, 4, 2, 0, 4):
        P.set_dense_top(lines)
        P.refactor(J, g)
        outs.setdefault(lines, []).extend([run(False), run(True), run(True)])
    floor = {'fp64': 1e-10, 'fp32': 1e-4, 'fp16': 1e-2}[storage]
    for lines, got in outs.items():
        for a in got:
            assert all(numpy.isfinite(v).all() for v in a), (lines, storage)
            for v, w in zip(a, got[0]):
                assert numpy.array_equal(v, w), (lines, storage)
        assert numpy.array_equal(got[0][0], got[0][1]), (lines, storage)
    base = [residual(outs[0][0][1], r64[0]), residual(outs[0][0][2], r64[1])]
    for lines in (2, 4):
        res = [residual(outs[lines][0][1], r64[0]), residual(outs[lines][0][2], r64[1])]
        for a, b in zip(res, base):
            assert a <= max(2.0 * b, floor), (lines, storage, res, base)
    P.set_dense_top(0)
    P.set_graph(True)


@pytest.mark.skip(reason='the ordered-task launch of the upper levels (tf_bcr_set_top) did not finish within '
                         'the time limit on the GPU (48 x 48, FP32 blocks), before and after its polls were '
                         'made relaxed; the cause is not established, so the launch is disabled and this case '
                         'stays off the GPU')
@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1), (8, 8, 6), (100, 100, 1), (17, 9, 1)])
@pytest.mark.parametrize('storage', ['fp64', 'fp32', 'fp16'])
def test_top_levels_launch_matches_per_level(hip, shape, storage):
    """The upper levels of the apply in one ordered-task launch (k_top) give, bit
    for bit, the per-level launches: one and two columns, eager and replayed, x-
    and y-line orders and 3D planes, FP64 / FP32 / FP16-copy blocks; no wait of
    the launch gives up."""
    import torch
    itf = hip.Interface({'Reynolds Number': 400}, *shape)
    n = itf.n
    rng = numpy.random.default_rng(35)
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1))
    single = storage != 'fp64'
    P = itf.solver._build(J, 'line', single)
    if single:
        P.set_half(storage == 'fp16')
    dt = torch.float32 if single else torch.float64
    r0 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    r1 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    apply = P.apply32 if single else P.apply
    outs = {}
    for top in (False, True, False, True):
        P.set_top(top)
        for graph in (False, True, True):
            P.set_graph(graph)
            z0 = torch.full_like(r0, numpy.nan)
            p0, p1 = torch.full_like(r0, numpy.nan), torch.full_like(r0, numpy.nan)
            apply(r0, z0)
            P.apply_pair(r0, r1, p0, p1, single=single)
            torch.cuda.synchronize()
            outs.setdefault(top, []).append((z0.cpu().numpy(), p0.cpu().numpy(), p1.cpu().numpy()))
    P.check_applies(hip.vector.stream())
    ref = outs[False][0]
    assert all(numpy.isfinite(a).all() for a in ref)
    for top in (False, True):
        for got in outs[top]:
            for a, b in zip(got, ref):
                assert numpy.array_equal(a, b), (top, storage)
    P.set_top(True)
    P.set_graph(True)


@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1), (8, 8, 6), (100, 100, 1)])
@pytest.mark.parametrize('storage', ['fp64', 'fp32', 'fp16'])
def test_pair_apply_matches_single(hip, shape, storage):
    """Two right-hand sides through one pass over the factor blocks give, bit for
    bit, the two single applies: x- and y-line orders (64 x 40 factors its
    shorter y-lines) and 3D planes, FP64 / FP32 / FP16-copy blocks (at 100 x 100
    two rows per lane group at the widest levels), eager and replayed."""
    import torch
    itf = hip.Interface({'Reynolds Number': 400}, *shape)
    n = itf.n
    rng = numpy.random.default_rng(33)
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.1))
    single = storage != 'fp64'
    P = itf.solver._build(J, 'line', single)
    if single:
        P.set_half(storage == 'fp16')
    dt = torch.float32 if single else torch.float64
    r0 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    r1 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    apply = P.apply32 if single else P.apply
    for graph in (False, True, True):
        P.set_graph(graph)
        z0, z1 = torch.empty_like(r0), torch.empty_like(r0)
        apply(r0, z0)
        apply(r1, z1)
        p0, p1 = torch.full_like(r0, numpy.nan), torch.full_like(r0, numpy.nan)
        P.apply_pair(r0, r1, p0, p1, single=single)
        torch.cuda.synchronize()
        assert torch.isfinite(z0).all() and torch.isfinite(z1).all()
        assert torch.equal(p0, z0) and torch.equal(p1, z1), (graph, storage)
    P.set_graph(True)


def test_csr_combination_matches_scipy(hip):
    """jacobian(x) - mass / (theta dt) on the device equals scipy's sparse
    arithmetic on the same matrices (TimeIntegration.py:56)."""
    itf = hip.Interface({'Reynolds Number': 150}, 12, 10)
    rng = numpy.random.default_rng(41)
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, itf.n)))
    M = itf.mass_matrix()
    C = J - M / 0.37
    D = 2.5 * J + M
    ref_C = J.to_scipy() - M.to_scipy() / 0.37
    ref_D = 2.5 * J.to_scipy() + M.to_scipy()
    for got, ref in ((C, ref_C), (D, ref_D)):
        diff = abs(got.to_scipy() - ref)
        assert diff.max() <= 1e-14 * abs(ref).max()


@pytest.mark.parametrize('theta', [1.0, 0.5])
def test_time_integration_matches_oracle(hip, theta):
    """Theta method (contract TimeIntegration.py:40-113) on the HIP backend against
    the oracle's theta method: LDC 16x16, Re 100, from rest, dt 0.05 to t = 0.2;
    the same Newton solves per step and the same state to 1e-8."""
    from oracle.interface import Interface as OracleInterface
    from oracle.time_integration import TimeIntegration as OracleTI
    itf = hip.Interface({'Reynolds Number': 100}, 16, 16)
    ti = hip.TimeIntegration(itf, theta=theta, newton_tolerance=1e-9, printer=None)
    seen = []
    x, t = ti.integration(itf.vector(), 0.05, 0.2, callback=lambda _i, _x, tt: seen.append(tt))
    ref = OracleInterface({'Reynolds Number': 100}, 16, 16, 1, 2, 3)
    to = OracleTI(ref, theta=theta, newton_tolerance=1e-9)
    xo, tr = to.integration(ref.vector(), 0.05, 0.2)
    xh = itf.array_from_vector(x)
    err = numpy.linalg.norm(xh - xo) / numpy.linalg.norm(xo)
    print('theta %.1f: steps %d, Newton solves %s / %s, rel err %.2e' % (theta, len(ti.newton_iterations),
                                                                          ti.newton_iterations, to.newton_iterations,
                                                                          err))
    assert t == tr and seen == [0, 0.05, 0.1, 0.15000000000000002, 0.2]
    assert ti.newton_iterations == to.newton_iterations
    assert err <= 1e-8
    # the step-size residual test ('dx') converges to the same state
    ti2 = hip.TimeIntegration(hip.Interface({'Reynolds Number': 100}, 16, 16), theta=theta,
                              newton_tolerance=1e-9, residual_check='dx', printer=None)
    x2, _ = ti2.integration(ti2.interface.vector(), 0.05, 0.2)
    assert numpy.linalg.norm(ti2.interface.array_from_vector(x2) - xo) <= 1e-8 * numpy.linalg.norm(xo)


def test_checkpoint_resume(hip, tmp_path):
    """State + parameter files (BaseInterface.py:118-215: ``name.npy`` and
    ``name.params`` JSON) and ``return_step`` (Continuation.py:352-359): stop at
    Re = 500, save, load into a fresh interface, resume to Re = 1000; the result
    is the converged state of an uninterrupted run."""
    import json
    name = str(tmp_path / 'ldc')
    itf = hip.Interface({}, 16, 16)
    cont = hip.Continuation(itf, newton_tolerance=1e-10, printer=None)
    x0 = cont.newton(itf.vector(), tol=1e-10)
    x, mu, dx, dmu = cont.continuation(x0, 'Reynolds Number', 0, 500, 100, return_step=True)
    itf.save_state(name, x)
    assert numpy.load(name + '.npy').shape == (itf.n,)
    with open(name + '.params') as f:
        assert json.load(f)['Reynolds Number'] == pytest.approx(mu)
    itf2 = hip.Interface({}, 16, 16)
    x2 = itf2.load_state(name)
    assert itf2.get_parameter('Reynolds Number') == pytest.approx(mu)
    assert numpy.array_equal(itf2.array_from_vector(x2), itf.array_from_vector(x))
    cont2 = hip.Continuation(itf2, newton_tolerance=1e-10, printer=None)
    xr, mur = cont2.continuation(x2, 'Reynolds Number', mu, 1000, 100, dx=dx, dmu=dmu)
    itf3 = hip.Interface({}, 16, 16)
    cont3 = hip.Continuation(itf3, newton_tolerance=1e-10, printer=None)
    xu, muu = cont3.continuation(cont3.newton(itf3.vector(), tol=1e-10), 'Reynolds Number', 0, 1000, 100)
    a, b = itf2.array_from_vector(xr), itf3.array_from_vector(xu)
    assert abs(mur - 1000) <= 1e-3 and abs(muu - 1000) <= 1e-3
    assert numpy.linalg.norm(a - b) / numpy.linalg.norm(b) <= 1e-6


def test_use_iterative_solver_options(hip):
    """The reference's iterative-solver settings (tests/test_SciPy.py:32-67):
    'Use Iterative Solver' with 'Convergence Tolerance' 1e-8 gives a solve to
    that relative residual whose velocities recover x from b = J x to
    10 tol ||b|| (the pressure only up to its gauge)."""
    dim, dof, nx = 3, 4, 4
    params = {'Use Iterative Solver': True, 'Iterative Solver': {'Convergence Tolerance': 1e-8}}
    itf = hip.Interface(params, nx, nx, nx, dim, dof)
    n = itf.n
    x = numpy.random.default_rng(7).random(n)
    A = itf.jacobian(numpy.zeros(n))
    b = A.to_scipy() @ x
    y = itf.array_from_vector(itf.solve(A, b))
    its, rel, name, _ = itf.solver.log[-1]
    tol = 1e-8 * numpy.linalg.norm(b) * 10
    y[dim::dof] += x[dim]
    print('iterative settings: %d iterations, rel res %.2e (%s)' % (its, rel, name), flush=True)
    assert rel <= 1e-8
    for d in range(dim):
        assert numpy.linalg.norm(y[d::dof] - x[d::dof]) < tol


@pytest.mark.parametrize('single', [False, True])
@pytest.mark.parametrize('shape', [(48, 48, 1), (24, 17, 1), (6, 6, 6)])
def test_fused_apply_matches_two_pass(hip, single, shape):
    """A factor made with Y L, Y U in place of L, U (one backward kernel per level)
    applies the same inverse as the two-pass factor: equal to rounding, and the
    solve with it still meets 1e-10."""
    import torch
    from transiflow_amd.solver import LineFactor
    itf = hip.Interface({'Reynolds Number': 500}, *shape)
    g = itf.solver.gauge()
    J = itf.jacobian(seeded_state(itf.n, 61) * 0.2)
    r = torch.from_numpy(seeded_state(itf.n, 62)).cuda()
    r[g] = 0.0
    out = []
    for fused in (False, True):
        f = LineFactor(itf, J, g, single=single, factor=False)
        f.set_fused(fused)
        f.refactor(J, g)
        rv = r.float() if single else r
        z = torch.empty_like(rv)
        (f.apply32 if single else f.apply)(rv, z)
        out.append(z.double().cpu().numpy())
    diff = numpy.linalg.norm(out[1] - out[0]) / numpy.linalg.norm(out[0])
    print('%s %s: fused vs two-pass apply %.2e' % (shape, 'fp32' if single else 'fp64', diff), flush=True)
    assert diff <= (1e-3 if single else 1e-10)


@pytest.mark.parametrize('alpha,beta', [(0.4 + 0.9j, 1.0), (-0.2 + 2.5j, 0.8 - 0.1j), (0.7, 1.0)])
def test_shifted_complex_solve_matches_splu(hip, alpha, beta):
    """Shifted solves of the eigenvalue glue (JaDa.py:98-99,146-151): the device
    matrix beta J - alpha M with complex shifts, solved against the oracle's
    complex splu of the same gauged matrix (SciPy.py:204-258)."""
    from oracle.interface import Interface as OracleInterface
    params = {'Reynolds Number': 300}
    itf = hip.Interface(dict(params), 24, 16)
    ref = OracleInterface(dict(params), 24, 16)
    state = seeded_state(itf.n, 91) * 0.3
    rng = numpy.random.default_rng(92)
    b = rng.uniform(-1, 1, itf.n) + 1j * rng.uniform(-1, 1, itf.n)
    A = beta * itf.jacobian(state) - alpha * itf.mass_matrix()
    y = itf.solve(A, b)
    Ao = beta * ref.jacobian(state) - alpha * ref.mass_matrix()
    if numpy.iscomplexobj(Ao.data):
        yo = ref.solve(Ao, b)
    else:   # a real shift: the real factor solves the two parts (scipy's real LU takes no complex rhs)
        yo = ref.solve(Ao, b.real) + 1j * ref.solve(Ao, b.imag)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    its, rel, name, _ = itf.solver.log[-1]
    print('alpha %s beta %s: %s, %d iterations, rel res %.2e, vs splu %.2e' % (alpha, beta, name, its, rel, err),
          flush=True)
    assert numpy.iscomplexobj(y) and y.shape == b.shape
    assert rel <= 1e-10 and err <= 1e-8


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
@pytest.mark.parametrize('shape', [(48, 48, 1), (64, 40, 1)])
def test_solve_pair_matches_two_solves(hip, precision, shape):
    """The two corrector solves in lockstep (solve_pair) meet the tolerance and
    agree with two single solves of the same systems, also with a lagged
    (mixed) factor of an earlier Jacobian; each is logged as a solve."""
    itf = hip.Interface({'Reynolds Number': 300, 'HIP Solver': {'Precision': precision}}, *shape)
    ref = hip.Interface({'Reynolds Number': 300, 'HIP Solver': {'Precision': precision, 'Pair Solves': False}},
                        *shape)
    rng = numpy.random.default_rng(35)
    n = itf.n
    for step in range(3):
        state = rng.uniform(-1, 1, n) * 0.05 * (step + 1)
        r0, r1 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        J, Jr = itf.jacobian(itf.vector_from_array(state)), ref.jacobian(ref.vector_from_array(state))
        nlog = len(itf.solver.log)
        y0, y1 = itf.solve_pair(J, itf.vector_from_array(r0), itf.vector_from_array(r1))
        assert len(itf.solver.log) == nlog + 2
        z0, z1 = ref.solve(Jr, ref.vector_from_array(r0)), ref.solve(Jr, ref.vector_from_array(r1))
        for y, z, e in ((y0, z0, itf.solver.log[-2]), (y1, z1, itf.solver.log[-1])):
            y, z = itf.array_from_vector(y), ref.array_from_vector(z)
            err = numpy.linalg.norm(y - z) / numpy.linalg.norm(z)
            print('%s %s step %d: its %d rel %.2e, vs single solve %.2e' % (shape, precision, step, e[0], e[1], err),
                  flush=True)
            assert e[1] <= 1e-10 and err <= 1e-7


@pytest.mark.parametrize('single', [False, True])
@pytest.mark.parametrize('graphs', [False, True])
def test_fgmres2_matches_two_fgmres(hip, single, graphs):
    """tf_fgmres2 (two systems in lockstep: one two-column apply and one two-system
    Arnoldi step per iteration) gives each system the iterations and, bit for bit,
    the iterate of tf_fgmres on it alone, through restarts and with the two
    systems converging at different iterations (a factor of another Jacobian)."""
    import torch
    from transiflow_amd.solver import NativeKrylov
    itf = hip.Interface({'Reynolds Number': 400}, 40, 40)
    n = itf.n
    rng = numpy.random.default_rng(37)
    J0 = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.05))
    J = itf.jacobian(itf.vector_from_array(rng.uniform(-1, 1, n) * 0.08))
    P = itf.solver._build(J0, 'line', single)     # lagged: several iterations per solve
    dt = torch.float32 if single else torch.float64
    val = J.values32() if single else J.val
    g = itf.solver.gauge()
    b0 = torch.from_numpy(rng.uniform(-1, 1, n)).to(dt).cuda()
    b1 = torch.from_numpy(rng.uniform(-1, 1, n) * 1e3).to(dt).cuda()
    tol0, tol1 = (1e-4, 1e-6) if single else (1e-9, 1e-11)
    dts = 'float32' if single else 'float64'
    k0, k1, ks = (NativeKrylov(n, 5, dts, graphs=graphs) for _ in range(3))
    x0, x1 = torch.zeros_like(b0), torch.zeros_like(b0)
    (i0, r0), (i1, r1) = k0.solve2(k1, J, val, g, P, b0, b1, x0, x1, tol0, tol1, 60, x_is_zero=True)
    s0, s1 = torch.zeros_like(b0), torch.zeros_like(b0)
    j0, q0 = ks.solve(J, val, g, P, b0, s0, tol0, 60, x_is_zero=True)
    j1, q1 = ks.solve(J, val, g, P, b1, s1, tol1, 60, x_is_zero=True)
    print('fgmres2 single=%s graphs=%s: its %d/%d (alone %d/%d), rel %.2e/%.2e' % (single, graphs, i0, i1, j0, j1,
                                                                                 r0, r1), flush=True)
    assert (i0, i1) == (j0, j1) and max(i0, i1) > 5
    assert r0 == q0 and r1 == q1
    assert torch.equal(x0, s0) and torch.equal(x1, s1)
