"""CPU tests of the host logic: continuation driver, solve semantics of the
oracle backend, the boundary-condition recorder and the C ABI exports."""
import os
import re

import numpy
import pytest

from oracle.interface import Interface as OracleInterface
from transiflow_amd.continuation import Continuation

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# -- solve semantics (tests/test_SciPy.py, tests/test_interface.py) -----------

def test_direct_solve_recovers_modulo_pressure():
    """tests/test_SciPy.py:6-31: J x = b is recovered up to the pressure gauge."""
    itf = OracleInterface({}, 4, 4, 4, 3, 4)
    x = numpy.random.default_rng(1).random(itf.vector().size)
    A = itf.jacobian(numpy.zeros_like(x))
    y = itf.solve(A, A @ x)
    y[3::4] += x[3]
    assert numpy.linalg.norm(y - x) < 1e-11


def test_bordered_matrix_identity():
    """tests/test_SciPy.py:68-101."""
    itf = OracleInterface({}, 4, 4, 4, 3, 4)
    itf.border_scaling = 1
    rng = numpy.random.default_rng(2)
    n = itf.vector().size
    x, x2 = rng.random(n), rng.random(3)
    V, W, C = rng.random((n, 3)), rng.random((n, 3)), rng.random((3, 3))
    A = itf.jacobian(numpy.zeros(n))
    B = itf.compute_bordered_matrix(A, V, W, C)
    y = B @ numpy.append(x, x2)
    assert numpy.linalg.norm(y - numpy.append(A @ x + V @ x2, W.T @ x + C @ x2)) < 1e-11


def test_bordered_solve_matches_elimination():
    """The bordered direct solve equals block elimination with the gauged matrix."""
    itf = OracleInterface({'Reynolds Number': 50}, 6, 6)
    rng = numpy.random.default_rng(3)
    n = itf.vector().size
    state = rng.uniform(-1, 1, n)
    J = itf.jacobian(state)
    f = rng.uniform(-1, 1, n)
    v = rng.uniform(-1, 1, n)
    w = rng.uniform(-1, 1, n)
    c, r2 = 0.7, 0.3
    y1, y2 = itf.solve(J, f, r2, v, w, c)
    A = itf.compute_bordered_matrix(J, fix_pressure_row=True).toarray()
    p = itf.pressure_row
    # the reference's bordered system: row p keeps the border entry and the un-gauged rhs
    B = numpy.zeros((n + 1, n + 1))
    B[:n, :n] = A
    B[:n, n] = v * 1e-3
    B[n, :n] = w * 1e-3
    B[n, p] = 0.0
    B[n, n] = c * 1e-6
    sol = numpy.linalg.solve(B, numpy.append(f, r2 * 1e-3))
    assert numpy.allclose(y1, sol[:n], rtol=1e-8, atol=1e-10)
    assert y2 == pytest.approx(sol[n] * 1e-3, rel=1e-8)


# -- continuation driver (Continuation.py) ------------------------------------

def test_continuation_golden_trajectory():
    """Product Continuation + oracle backend reproduce the reference run bit for bit."""
    g = numpy.load(os.path.join(GOLDEN, 'cont_ldc2d_8.npz'))
    itf = OracleInterface({}, 8, 8, 1, 2, 3)
    cont = Continuation(itf, newton_tolerance=1e-8, printer=None)
    x0 = cont.newton(itf.vector())
    assert numpy.array_equal(x0, g['x0'])
    mus, its = [], []

    def cb(_i, _x, mu):
        mus.append(mu)
        its.append(cont.newton_iterations)

    x, mu = cont.continuation(x0, 'Reynolds Number', 0, 2000, 100, callback=cb)
    assert numpy.array_equal(numpy.array(mus), g['mus'])
    assert numpy.array_equal(numpy.array(its), g['newton_its'])
    assert numpy.array_equal(x, g['x'])
    assert mu == float(g['mu'])


def test_newton_golden_16():
    g = numpy.load(os.path.join(GOLDEN, 'ldc2d_16.npz'))
    itf = OracleInterface({'Reynolds Number': 100}, 16, 16, 1, 2, 3)
    x = Continuation(itf, newton_tolerance=1e-8, printer=None).newton(itf.vector(), tol=1e-8)
    assert numpy.array_equal(x, g['x_newton'])


@pytest.mark.parametrize('its,ds,expected', [(1, 10.0, 20.0), (3, 10.0, 10.0), (6, 10.0, 5.0),
                                             (10, 10.0, 5.0), (0, -4.0, -8.0), (2, 900.0, 1000.0),
                                             (9, 0.015, 0.01)])
def test_adjust_step_size(its, ds, expected):
    """Continuation.py:197-209: factor 3/its clamped to [1/2, 2], |ds| in [ds_min, ds_max]."""
    c = Continuation(None, printer=None)
    c.newton_iterations = its
    assert c._adjust_step_size(ds, 0.01, 1000) == pytest.approx(expected)


class _Diverging:
    """Interface wrapper whose residual grows: every corrector is rejected."""

    def __init__(self, itf):
        self.itf = itf
        self.calls = 0

    def __getattr__(self, name):
        return getattr(self.itf, name)

    def rhs(self, x):
        self.calls += 1
        return self.itf.rhs(x) * (10.0 ** self.calls)


def test_step_rejection_and_terminal_exception():
    """Continuation.py:265-273: a rejected corrector shrinks ds and retries; when ds
    cannot change any more the step raises."""
    itf = OracleInterface({}, 4, 4, 1, 2, 3)
    cont = Continuation(_Diverging(itf), newton_tolerance=1e-8, printer=None)
    x0 = itf.vector()
    with pytest.raises(Exception, match='Newton cannot achieve convergence'):
        cont.continuation(x0, 'Reynolds Number', 0, 100, 50, ds_min=50, ds_max=50)
    assert any(h[0] == 'rejected' for h in cont.history)


def test_maximum_newton_iterations_rejects():
    itf = OracleInterface({}, 6, 6, 1, 2, 3)
    cont = Continuation(itf, newton_tolerance=1e-14, maximum_newton_iterations=1, printer=None)
    with pytest.raises(Exception):
        cont.continuation(itf.vector(), 'Reynolds Number', 0, 100, 50, ds_min=50, ds_max=50)


# -- boundary-condition recorder (host side of the HIP backend) --------------

def test_bc_program_constants_match_oracle():
    """The recorded FORCE constants equal the oracle's heat-flux/lid constants."""
    from transiflow_amd.problem import BC_FORCE, BC_WALL, Discretization as HostDisc
    p = {'Problem Type': 'Rayleigh-Benard', 'Biot Number': 1.0, 'X-max': 4.0, 'Asymmetry Parameter': 0.3}
    d = HostDisc(dict(p), 16, 8)
    ops = d.boundary_program()
    kinds = [o[0] for o in ops]
    assert kinds == [BC_FORCE, BC_FORCE, BC_WALL, BC_WALL, BC_FORCE, BC_FORCE, BC_WALL, BC_WALL]
    h = (d.y[8] - d.y[6]) / 2
    north = ops[4]
    assert north[1:4] == (1, 1, 3)
    assert north[5] == 0.0
    assert north[6] == (1 - h * 1.0 / 2) / (1 + h * 1.0 / 2)
    hx = (d.x[16] - d.x[14]) / 2
    assert ops[0][5] == hx * 0.3 / (1 + hx * 0.0 / 2)
    ldc = HostDisc({'Lid Velocity': 2.5}, 8, 8).boundary_program()
    assert ldc[3] == (BC_FORCE, 1, 1, 0, 0.0, 5.0, -1.0)
    assert ldc[4] == (BC_WALL, 1, 1, 0, -1.0, 0.0, 0.0)


def test_physics_scalars_match_reference_formulas():
    from transiflow_amd.problem import Discretization as HostDisc
    d = HostDisc({'Problem Type': 'Rayleigh-Benard', 'Rayleigh Number': 3000.0, 'Prandtl Number': 7.0}, 8, 8)
    lin, ts, ss, rbc, rbp, nl = d.physics()
    Gr = 3000.0 / 7.0
    assert lin == 1 / (1.0 * numpy.sqrt(Gr)) and ts == 1 / (7.0 * numpy.sqrt(Gr))
    assert nl == 1 and rbp == 0
    d0 = HostDisc({'Reynolds Number': 0}, 8, 8)
    assert d0.physics()[0] == 1.0 and d0.physics()[5] == 0


def test_unsupported_problem_types_fail_loudly():
    from transiflow_amd.problem import Discretization as HostDisc
    with pytest.raises(NotImplementedError):
        HostDisc({'Problem Type': 'AMOC'}, 4, 4)
    with pytest.raises(Exception, match='Invalid problem type'):
        HostDisc({'Problem Type': 'nonsense'}, 4, 4)


# -- C ABI -----------------------------------------------------------------------

def _declared_symbols():
    names = set()
    inc = os.path.join(ROOT, 'include')
    for fn in os.listdir(inc):
        if fn.endswith('.h'):
            with open(os.path.join(inc, fn)) as f:
                txt = f.read()
            txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
            names.update(re.findall(r'\b(tf_[a-z0-9_]+)\s*\(', txt))
    return names


def test_library_exports_every_declared_symbol():
    from transiflow_amd import _lib
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) > 20
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes table declares exactly the exported ABI
    assert declared == set(_lib.SIGNATURES), sorted(declared ^ set(_lib.SIGNATURES))


def test_backend_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    import transiflow_amd
    from transiflow_amd._lib import BackendError
    with pytest.raises(BackendError):
        transiflow_amd.Interface({}, 8, 8)


def test_solve_rejects_bad_rhs_shapes():
    """Solver.solve accepts (n,) and (n, k) right-hand sides (SciPy.py:209-216)
    and raises ValueError for anything else, before touching the device."""
    from transiflow_amd.solver import Solver

    class _Itf:
        n = 12
        parameters = {}

    s = Solver(_Itf())
    for bad in (numpy.zeros(11), numpy.zeros((11, 2)), numpy.zeros((12, 2, 1)), numpy.zeros(0)):
        with pytest.raises(ValueError):
            s.solve(None, bad)


# -- preconditioner choice under 'auto' ------------------------------------------

def _grid(nx, ny, nz, periodic=False):
    from types import SimpleNamespace
    dim = 3 if nz > 1 else 2
    dof = dim + 1
    d = SimpleNamespace(x_periodic=periodic, y_periodic=False, z_periodic=False)
    return SimpleNamespace(nx=nx, ny=ny, nz=nz, dim=dim, dof=dof, n=nx * ny * nz * dof, discretization=d)


def test_auto_preconditioner_choice():
    """'auto' takes the exact line (plane) factor when it fits the memory budget,
    the two-grid method for 3D grids beyond it, and otherwise raises a
    SolverError that names the factor size, instead of falling back to a
    preconditioner that does not converge."""
    from transiflow_amd.solver import SolverError, auto_preconditioner
    assert auto_preconditioner(_grid(256, 256, 1), False) == 'line'
    assert auto_preconditioner(_grid(32, 32, 32), False) == 'line'
    assert auto_preconditioner(_grid(64, 64, 64), False) == 'multigrid'
    assert auto_preconditioner(_grid(64, 64, 64), True) == 'multigrid'
    for g in (_grid(63, 63, 63), _grid(64, 64, 64, periodic=True), _grid(4096, 4096, 1)):
        with pytest.raises(SolverError, match='no preconditioner for this grid.*GB'):
            auto_preconditioner(g, False)
