"""GPU parity of the fused stencil kernels (rhs, Jacobian CSR, mass matrix).

Checked against: the reference's golden matrices/rhs (tests/golden/reference_data),
the golden vectors written from the reference (tests/golden/*.npz) and the CPU
oracle on seeded random states up to the full 256x256 benchmark grid.  The
bar is bit-exact row_ptr/col_idx and values within 1e-12 of the largest
entry; the kernels are written to reproduce the reference's operation order,
so the tests also print the fraction of bit-identical values.
"""
import os

import numpy
import pytest

from tests.golden.fixtures import read_matrix, read_vector
from tests.helpers import assert_csr_equal, assert_vec_close, seeded_state

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')

pytestmark = pytest.mark.gpu


def _pair(hip, params, nx, ny, nz, dim=None, dof=None, bc=None, bc_oracle=None):
    from oracle.discretization import Discretization
    itf = hip.Interface(dict(params), nx, ny, nz, dim, dof, boundary_conditions=bc)
    disc = Discretization(dict(params), nx, ny, nz, dim, dof, boundary_conditions=bc_oracle)
    return itf, disc


def _check(itf, disc, state, label):
    F = itf.array_from_vector(itf.rhs(state))
    Fo = disc.rhs(state)
    fe = assert_vec_close(F, Fo)
    J = itf.jacobian(state)
    Jo = disc.jacobian(state)
    je = assert_csr_equal((J.begA, J.jcoA, J.coA), Jo.begA, Jo.jcoA, Jo.coA)
    print('%s: F bit-exact %.4f, J bit-exact %.4f, nnz %d' % (label, fe, je, Jo.nnz), flush=True)


@pytest.mark.parametrize('case', ['ldc2d_16', 'rb2d_16x8', 'ldc3d_semi'])
def test_golden_vectors(hip, case):
    g = numpy.load(os.path.join(GOLDEN, case + '.npz'))
    if case == 'ldc2d_16':
        itf = hip.Interface({'Reynolds Number': 100}, 16, 16, 1, 2, 3)
    elif case == 'rb2d_16x8':
        itf = hip.Interface({'Problem Type': 'Rayleigh-Benard', 'Prandtl Number': 1.0, 'Rayleigh Number': 1000.0,
                             'Biot Number': 1.0, 'X-max': 4.0, 'Reynolds Number': 1.0}, 16, 8, 1, 2, 4)
    else:
        itf = hip.Interface({'Reynolds Number': 50}, 6, 5, 1, 3, 4)
    F = itf.array_from_vector(itf.rhs(g['state']))
    assert numpy.array_equal(F, g['F']), 'rhs not bit-exact vs golden'
    J = itf.jacobian(g['state'])
    assert numpy.array_equal(J.begA, g['begA'])
    assert numpy.array_equal(J.jcoA, g['jcoA'])
    assert numpy.array_equal(J.coA, g['coA']), 'values not bit-exact vs golden'


FIXTURES = [
    ('ldc', {'Reynolds Number': 100}, 4),
    ('ldc_stretched', {'Reynolds Number': 100, 'Grid Stretching': True}, 4),
    ('bous', {'Reynolds Number': 1, 'Rayleigh Number': 100, 'Prandtl Number': 100,
              'Problem Type': 'Rayleigh-Benard'}, 5),
    ('bous_stretched', {'Reynolds Number': 1, 'Rayleigh Number': 100, 'Prandtl Number': 100,
                        'Problem Type': 'Rayleigh-Benard', 'Grid Stretching': True}, 5),
    ('dhc', {'Reynolds Number': 1, 'Rayleigh Number': 100, 'Prandtl Number': 100,
             'Problem Type': 'Differentially Heated Cavity'}, 5),
]


@pytest.mark.parametrize('name,params,dof', FIXTURES)
def test_reference_fixtures_3d(hip, name, params, dof):
    """The reference's own 4x4x4 known-answer matrices (tests/test_fvm.py:843-1137)."""
    n = 4 * 4 * 4 * dof
    state = numpy.arange(1, n + 1, dtype=float)
    itf, disc = _pair(hip, params, 4, 4, 4, 3, dof)
    _check(itf, disc, state, name)
    begB, jB, cB = read_matrix(name + '_4x4x4.txt')
    J = itf.jacobian(state)
    assert numpy.array_equal(J.begA, begB) and numpy.array_equal(J.jcoA, jB)
    numpy.testing.assert_allclose(J.coA, cB, rtol=1e-6, atol=1e-12)
    rhs_B = read_vector(name + '_rhs_4x4x4.txt')
    numpy.testing.assert_allclose(itf.array_from_vector(itf.rhs(state)), rhs_B, rtol=1e-6, atol=1e-12)


CASES = [
    ('ldc2d_64x48_Re1000', {'Reynolds Number': 1000}, (64, 48, 1), None, None, 11),
    ('ldc2d_stretched_40', {'Reynolds Number': 300, 'Grid Stretching': True}, (40, 40, 1), None, None, 12),
    ('ldc2d_Re0', {'Reynolds Number': 0}, (32, 32, 1), None, None, 13),
    ('rb2d_64x16', {'Problem Type': 'Rayleigh-Benard', 'Prandtl Number': 1.0, 'Rayleigh Number': 3000.0,
                    'Biot Number': 1.0, 'X-max': 4.0}, (64, 16, 1), None, None, 14),
    ('rbp2d_24x12', {'Problem Type': 'Rayleigh-Benard Perturbation', 'Prandtl Number': 7.0,
                     'Rayleigh Number': 2000.0, 'Biot Number': 2.0}, (24, 12, 1), None, None, 15),
    ('ldc3d_12', {'Reynolds Number': 500}, (12, 12, 12), None, None, 16),
    ('ldc3d_stretched_9x8x7', {'Reynolds Number': 200, 'Grid Stretching': True}, (9, 8, 7), None, None, 17),
    ('rb3d_8', {'Problem Type': 'Rayleigh-Benard', 'Rayleigh Number': 1500.0, 'Prandtl Number': 2.0,
                'Biot Number': 0.5}, (8, 8, 8), None, None, 18),
    ('dhc3d_6x5x4', {'Problem Type': 'Differentially Heated Cavity', 'Rayleigh Number': 1e4,
                     'Prandtl Number': 0.7}, (6, 5, 4), None, None, 19),
    ('semi2d_10x9', {'Reynolds Number': 150}, (10, 9, 1), 3, 4, 20),
]


@pytest.mark.parametrize('label,params,shape,dim,dof,seed', CASES, ids=[c[0] for c in CASES])
def test_random_states(hip, label, params, shape, dim, dof, seed):
    itf, disc = _pair(hip, params, *shape, dim, dof)
    state = seeded_state(itf.n, seed)
    _check(itf, disc, state, label)


def test_full_size_ldc256(hip):
    """C2 size (LDC 2D 256x256, n = 196,608) at a seeded state, Re = 2000."""
    itf, disc = _pair(hip, {'Reynolds Number': 2000}, 256, 256, 1)
    state = seeded_state(itf.n, 21)
    _check(itf, disc, state, 'ldc2d_256')


def test_parameter_changes_are_seen(hip):
    """set_parameter between calls (mu, mu + delta, mu as in the corrector)."""
    itf, disc = _pair(hip, {'Reynolds Number': 10}, 20, 20, 1)
    state = seeded_state(itf.n, 22)
    for re in (10.0, 11.0, 10.0, 0.0, 2000.0):
        itf.set_parameter('Reynolds Number', re)
        disc.set_parameter('Reynolds Number', re)
        assert_vec_close(itf.array_from_vector(itf.rhs(state)), disc.rhs(state))


@pytest.mark.parametrize('shape,dim,dof,params', [
    ((16, 16, 1), 2, 3, {}),
    ((8, 6, 5), 3, 4, {'Grid Stretching': True}),
    ((10, 5, 1), 2, 4, {'Problem Type': 'Rayleigh-Benard'}),
])
def test_mass_matrix(hip, shape, dim, dof, params):
    itf, disc = _pair(hip, params, *shape, dim, dof)
    M = itf.mass_matrix()
    Mo = disc.mass_matrix()
    assert_csr_equal((M.begA, M.jcoA, M.coA), Mo.begA, Mo.jcoA, Mo.coA)


def test_custom_boundary_conditions(hip):
    """User callback as in the reference's test_custom_bc (tests/test_interface.py:118-148)."""
    def boundaries(bc, atom):
        bc.heat_flux_east(atom, 0)
        bc.heat_flux_west(atom, 0)
        bc.no_slip_east(atom)
        bc.no_slip_west(atom)
        bc.heat_flux_north(atom, 0)
        bc.heat_flux_south(atom, 0)
        bc.no_slip_north(atom)
        bc.no_slip_south(atom)
        bc.heat_flux_top(atom, 0, 0)
        bc.temperature_bottom(atom, 0)
        bc.free_slip_top(atom)
        bc.no_slip_bottom(atom)
        return bc.get_forcing()

    params = {'Rayleigh Number': 100, 'Prandtl Number': 100}
    itf, disc = _pair(hip, params, 5, 4, 6, 3, 5, bc=boundaries, bc_oracle=boundaries)
    _check(itf, disc, seeded_state(itf.n, 23), 'custom_bc')
