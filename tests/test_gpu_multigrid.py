"""Line block ILU and the two-grid preconditioner for 3D grids.

* The four-colour line block ILU (csrc/lineilu.hip) applied on the GPU equals
  the dense numpy statement of the same factorisation (tests/linegrid.py), in
  FP64 and FP32.
* The two-grid transfers and one V(1,1) cycle (csrc/mg.hip) equal their numpy
  statements, with the coarse Jacobian assembled by the oracle at the
  restricted state and solved by scipy's splu.
* Solves preconditioned by the two-grid method reach the 1e-10 residual and
  agree with splu at 16^3, and with the exact plane factorisation at 32^3;
  a 3D continuation with it matches the one with the exact plane factor.
"""

import numpy
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from tests import linegrid
from tests.helpers import seeded_state

pytestmark = pytest.mark.gpu


def _gauged(ref, J):
    return sparse.csr_matrix(ref.compute_bordered_matrix(J, fix_pressure_row=True))


@pytest.mark.parametrize('shape', [(6, 6, 4), (5, 4, 3), (6, 5, 4)])
@pytest.mark.parametrize('single', [False, True])
def test_line_ilu_apply_matches_numpy(hip, shape, single):
    """3D grids whose pivots are all regular (smallest singular value above
    1e-5 of the largest), so the GPU Gauss-Jordan and numpy's inverse agree to
    the pivots' conditioning."""
    import torch
    from oracle.interface import Interface as OracleInterface
    from transiflow_amd.solver import LineILU
    params = {'Reynolds Number': 150}
    itf = hip.Interface(dict(params), *shape)
    ref = OracleInterface(dict(params), *shape)
    state = seeded_state(itf.n, 71) * 0.5
    J = itf.jacobian(state)
    P = LineILU(itf, single)
    P.refactor(J, itf.solver.gauge())
    r = seeded_state(itf.n, 72)
    z = linegrid.dilu_apply(_gauged(ref, ref.jacobian(state)), r, *shape, itf.dof)
    dt = torch.float32 if single else torch.float64
    rt = torch.from_numpy(r).to(dt).cuda()
    out = torch.empty_like(rt)
    (P.apply32 if single else P.apply)(rt, out)
    got = out.double().cpu().numpy()
    err = numpy.max(numpy.abs(got - z)) / numpy.max(numpy.abs(z))
    print('%s %s: line ILU apply vs numpy %.2e, perturbed pivots %d' % (shape, 'fp32' if single else 'fp64', err,
                                                                          P.perturbed), flush=True)
    assert P.perturbed == 0
    assert err <= (2e-2 if single else 1e-9)


def test_line_ilu_solve_2d(hip):
    """In 2D the top line has a single earlier neighbour, and its pivot keeps the
    line's pressure null space: the factorisation regularises that pivot, and
    GMRES with the line ILU still reaches 1e-10 and the splu solution."""
    from oracle.interface import Interface as OracleInterface
    params = {'Reynolds Number': 100}
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Preconditioner': 'line-ilu', 'Maximum Iterations': 600}}),
                        16, 12)
    ref = OracleInterface(dict(params), 16, 12)
    state = seeded_state(itf.n, 73) * 0.3
    rhs = seeded_state(itf.n, 74)
    J = itf.jacobian(state)
    y = itf.array_from_vector(itf.solve(J, rhs))
    yo = ref.solve(ref.jacobian(state), rhs)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    its, rel, name, _ = itf.solver.log[-1]
    print('2D line ILU: %d iterations, rel res %.2e, vs splu %.2e, regularised pivots %d'
          % (its, rel, err, J.prec.perturbed), flush=True)
    assert name == 'line-ilu' and rel <= 1e-10 and err <= 1e-8 and J.prec.perturbed >= 1


@pytest.mark.parametrize('shape,dim', [((8, 8, 8), 3), ((8, 6, 4), 3)])
def test_two_grid_cycle_matches_numpy(hip, shape, dim):
    import torch
    from oracle.interface import Interface as OracleInterface
    from transiflow_amd import _lib
    from transiflow_amd.solver import Multigrid
    from transiflow_amd.vector import ptr, stream
    params = {'Reynolds Number': 200}
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Preconditioner': 'multigrid'}}), *shape)
    dof = itf.dof
    state = seeded_state(itf.n, 81) * 0.3
    J = itf.jacobian(state)
    # state restriction
    nc = tuple(s // 2 for s in shape)
    xc = torch.empty(nc[0] * nc[1] * nc[2] * dof, dtype=torch.float64, device='cuda')
    _lib.call('tf_mg_restrict_state', *shape, dim, dof, ptr(J.state), ptr(xc), stream())
    xc_np = linegrid.restrict_state(state, shape, dim, dof)
    assert numpy.allclose(xc.cpu().numpy(), xc_np, rtol=1e-14, atol=1e-15)
    # one cycle against numpy: z = S r; z += P Bc R (r - A z); z += S (r - A z)
    M = Multigrid(itf, False)
    M.refactor(J, itf.solver.gauge())
    ref = OracleInterface(dict(params), *shape)
    A = _gauged(ref, ref.jacobian(state))
    refc = OracleInterface(dict(params), *nc)
    Ac = _gauged(refc, refc.jacobian(xc_np))
    r = seeded_state(itf.n, 82)
    S = lambda v: linegrid.dilu_apply(A, v, *shape, dof)
    z = S(r)
    rc = linegrid.restrict(r - A @ z, shape, dim, dof, itf.pressure_row)
    z = z + linegrid.prolong(spla.spsolve(sparse.csc_matrix(Ac), rc), shape, dim, dof)
    z = z + S(r - A @ z)
    out = torch.empty(itf.n, dtype=torch.float64, device='cuda')
    M.apply(torch.from_numpy(r).cuda(), out)
    err = numpy.max(numpy.abs(out.cpu().numpy() - z)) / numpy.max(numpy.abs(z))
    print('%s: two-grid cycle vs numpy %.2e' % (shape, err), flush=True)
    assert err <= 1e-10


def _solve_mg(hip, n, re, seed, precision):
    params = {'Reynolds Number': re}
    itf = hip.Interface(dict(params, **{'HIP Solver': {'Preconditioner': 'multigrid', 'Precision': precision,
                                                        'Maximum Iterations': 400}}), n, n, n)
    state = seeded_state(itf.n, seed) * 0.1
    rhs = seeded_state(itf.n, seed + 1)
    J = itf.jacobian(state)
    y = itf.array_from_vector(itf.solve(J, rhs))
    return itf, state, rhs, y


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_solve_multigrid_16_matches_splu(hip, precision):
    from oracle.interface import Interface as OracleInterface
    itf, state, rhs, y = _solve_mg(hip, 16, 300.0, 91, precision)
    ref = OracleInterface({'Reynolds Number': 300.0}, 16, 16, 16)
    Jo = ref.jacobian(state)
    yo = ref.solve(Jo, rhs)
    A = _gauged(ref, Jo)
    b = rhs.copy()
    b[ref.pressure_row] = 0
    res = numpy.linalg.norm(A @ y - b) / numpy.linalg.norm(b)
    err = numpy.linalg.norm(y - yo) / numpy.linalg.norm(yo)
    its, rel, name, _ = itf.solver.log[-1]
    print('16^3 %s: %s, %d iterations, rel res %.2e (host %.2e), vs splu %.2e' % (precision, name, its, rel, res,
                                                                                   err), flush=True)
    assert name == 'multigrid' and res <= 1e-9 and err <= 1e-8


@pytest.mark.timeout(300)
def test_solve_multigrid_32_matches_plane_factor(hip):
    itf, state, rhs, y = _solve_mg(hip, 32, 500.0, 93, 'fp64')
    its, rel, name, _ = itf.solver.log[-1]
    ref = hip.Interface({'Reynolds Number': 500.0, 'HIP Solver': {'Preconditioner': 'line'}}, 32, 32, 32)
    yl = ref.array_from_vector(ref.solve(ref.jacobian(state), rhs))
    J = ref.jacobian(state).to_scipy()
    g = ref.pressure_row
    A = J.tolil()
    A[g, :] = 0
    A[:, g] = 0
    A[g, g] = -1
    b = rhs.copy()
    b[g] = 0
    res = numpy.linalg.norm(A.tocsr() @ y - b) / numpy.linalg.norm(b)
    err = numpy.linalg.norm(y - yl) / numpy.linalg.norm(yl)
    print('32^3: %s, %d iterations, rel res %.2e (host %.2e), vs plane factor %.2e' % (name, its, rel, res, err),
          flush=True)
    assert res <= 1e-9 and err <= 1e-8


@pytest.mark.timeout(300)
def test_continuation_3d_multigrid_matches_plane_factor(hip):
    """3D cavity 16^3, Re 0 -> 300 (ds 100, tol 1e-8): the two-grid (mixed
    precision) and the exact plane factor (FP64) give the same steps and state."""
    out = {}
    for name, opts in (('mg', {'Preconditioner': 'multigrid', 'Precision': 'mixed'}),
                       ('line', {'Preconditioner': 'line'})):
        itf = hip.Interface({'HIP Solver': opts}, 16, 16, 16)
        cont = hip.Continuation(itf, newton_tolerance=1e-8, printer=None)
        mus = []
        x0 = cont.newton(itf.vector(), tol=1e-8)
        x, mu = cont.continuation(x0, 'Reynolds Number', 0, 300, 100, callback=lambda _i, _x, m: mus.append(m))
        out[name] = (itf.array_from_vector(x), mu, mus, sum(e[0] for e in itf.solver.log))
    (xm, mm, musm, itm), (xl, ml, musl, itl) = out['mg'], out['line']
    err = numpy.linalg.norm(xm - xl) / numpy.linalg.norm(xl)
    print('3D 16^3 continuation: steps %d/%d, GMRES its %d/%d, rel diff %.2e' % (len(musm), len(musl), itm, itl, err),
          flush=True)
    numpy.testing.assert_allclose(musm, musl, rtol=1e-9)
    assert abs(mm - 300) <= 1e-3 and err <= 1e-8
