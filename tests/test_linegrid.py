"""CPU checks of the dense statements of the line block ILU and the two-grid
transfers (tests/linegrid.py) that the GPU tests compare the kernels with."""
import numpy
from scipy import sparse

from oracle.interface import Interface as OracleInterface
from tests import linegrid


def _gauged(ref, J):
    return sparse.csr_matrix(ref.compute_bordered_matrix(J, fix_pressure_row=True))


def test_two_grid_is_galerkin_on_smooth_fields():
    """R A_h P e agrees with the coarse Jacobian A_2h e to O(h^2) for a smooth e
    vanishing on the walls (Stokes, Re = 0): the property the coarse grid rests on."""
    errs = []
    for n in (16, 32):
        N = n // 2
        itf = OracleInterface({'Reynolds Number': 0}, n, n, n)
        A = _gauged(itf, itf.jacobian(numpy.zeros(itf.vector().size)))
        c = OracleInterface({'Reynolds Number': 0}, N, N, N)
        Ac = _gauged(c, c.jacobian(numpy.zeros(c.vector().size)))
        ctr, face = (numpy.arange(N) + 0.5) / N, (numpy.arange(N) + 1.0) / N
        S = numpy.sin
        e = numpy.zeros((N, N, N, 4))
        e[..., 0] = S(numpy.pi * ctr)[:, None, None] * S(numpy.pi * ctr)[None, :, None] * S(numpy.pi * face)[None, None, :]
        e[..., 1] = S(numpy.pi * ctr)[:, None, None] * S(numpy.pi * face)[None, :, None] * S(numpy.pi * ctr)[None, None, :]
        e[..., 2] = S(numpy.pi * face)[:, None, None] * S(numpy.pi * ctr)[None, :, None] * S(numpy.pi * ctr)[None, None, :]
        e = e.reshape(-1)
        lhs = linegrid.restrict(A @ linegrid.prolong(e, (n, n, n), 3, 4), (n, n, n), 3, 4, 3)
        rhs = Ac @ e
        vel = numpy.arange(e.size) % 4 < 3
        errs.append(numpy.linalg.norm((lhs - rhs)[vel]) / numpy.linalg.norm(rhs[vel]))
    print('Galerkin defect at 16^3 / 32^3: %.3e / %.3e' % tuple(errs), flush=True)
    assert errs[0] < 0.06 and errs[1] < 0.015 and errs[0] / errs[1] > 3.5



def test_line_ilu_is_exact_on_a_single_line_column():
    """With one line per colour class reachable (2D, two lines) the D-ILU is the
    exact block LU: M^-1 A = I."""
    itf = OracleInterface({'Reynolds Number': 100}, 6, 2, 1, 2, 3)
    n = itf.vector().size
    A = _gauged(itf, itf.jacobian(numpy.random.default_rng(3).uniform(-1, 1, n) * 0.5))
    r = numpy.random.default_rng(4).uniform(-1, 1, n)
    z = linegrid.dilu_apply(A, A @ r, 6, 2, 1, 3)
    assert numpy.allclose(z, r, rtol=1e-10, atol=1e-10)


def test_transfers_preserve_constants_and_sums():
    """Restriction sums the continuity residuals of the 8 fine cells (a coarse
    cell's mass balance); prolongation copies cell values and interpolates faces
    linearly, so a field linear in x along u's normal is reproduced inside."""
    n, dof = (8, 8, 8), 4
    rng = numpy.random.default_rng(5)
    r = rng.uniform(-1, 1, 8 * 8 * 8 * dof)
    rc = linegrid.restrict(r, n, 3, dof, -1).reshape(4, 4, 4, dof)
    R = r.reshape(8, 8, 8, dof)
    assert numpy.isclose(rc[1, 2, 3, 3], R[2:4, 4:6, 6:8, 3].sum())
    e = numpy.zeros((4, 4, 4, dof))
    e[..., 3] = 2.5
    e[..., 0] = (numpy.arange(4) + 1.0)[None, None, :]          # u = coarse face position
    z = linegrid.prolong(e.reshape(-1), n, 3, dof).reshape(8, 8, 8, dof)
    assert numpy.all(z[..., 3] == 2.5)
    # fine face i sits at (i + 1) / 2 coarse units: linear along x away from the y/z walls
    assert numpy.allclose(z[2:6, 2:6, :, 0], ((numpy.arange(8) + 1.0) / 2)[None, None, :])
