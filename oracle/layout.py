"""CPU oracle: state layout and coordinate vectors.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Restates
``transiflow/utils.py:23-267``: the (nx, ny, nz, dof) view of a state vector,
its padded form used by every stencil, and the coordinate vectors of length
n+3 whose last two entries are the points left of the domain.
"""

import numpy


def state_mtx(state, nx, ny, nz, dof):
    """utils.create_state_mtx: index d + i*dof + j*dof*nx + k*dof*nx*ny (utils.py:23-60)."""
    return numpy.asarray(state).reshape(nz, ny, nx, dof).transpose(2, 1, 0, 3).copy()


def state_vec(mtx):
    """utils.create_state_vec (utils.py:135-173)."""
    return numpy.ascontiguousarray(numpy.asarray(mtx).transpose(2, 1, 0, 3)).reshape(-1).copy()


def padded_state_mtx(state, nx, ny, nz, dof, x_periodic, y_periodic, z_periodic):
    """utils.create_padded_state_mtx (utils.py:62-133).

    Ghost layers are zero, or copies of the opposite side when periodic.  In a
    non-periodic direction the wall-normal velocity of the last interior cell
    (padded index n) is zeroed: it sits on the wall.
    """
    sp = numpy.zeros((nx + 2, ny + 2, nz + 2, dof))
    sp[1:nx + 1, 1:ny + 1, 1:nz + 1, :] = state_mtx(state, nx, ny, nz, dof)
    if x_periodic:
        sp[0] = sp[nx]
        sp[nx + 1] = sp[1]
    else:
        sp[nx, :, :, 0] = 0
    if y_periodic:
        sp[:, 0] = sp[:, ny]
        sp[:, ny + 1] = sp[:, 1]
    else:
        sp[:, ny, :, 1] = 0
    if z_periodic:
        sp[:, :, 0] = sp[:, :, nz]
        sp[:, :, nz + 1] = sp[:, :, 1]
    else:
        sp[:, :, nz, 2] = 0
    return sp


def uniform_coordinate_vector(start, end, n):
    """utils.create_uniform_coordinate_vector (utils.py:175-197).

    Entry i (0 <= i <= n) is the right face of cell i; entries -2 and -1 are
    start - dx and start.
    """
    dx = (end - start) / n
    x = start + numpy.arange(-1, n + 2) * dx
    return numpy.roll(x, -2)


def stretched_coordinate_vector(start, end, n, sigma):
    """utils.create_stretched_coordinate_vector, tanh stretching (utils.py:199-232)."""
    x = uniform_coordinate_vector(0, 1, n)
    x = 0.5 * (1 + numpy.tanh(2 * sigma * (x - 0.5)) / numpy.tanh(sigma))
    x = start + x * (end - start)
    dx = x[0] - x[-1]
    if start == 0:
        x[-2] = x[-1] - dx
    if end == 1:
        x[-3] = x[-4] + dx
    return x


def stretched_coordinate_vector_sin(start, end, n, sigma):
    """utils.create_stretched_coordinate_vector2, sine stretching (utils.py:234-267)."""
    x = uniform_coordinate_vector(0, 1, n)
    x = x - sigma * numpy.sin(2 * numpy.pi * x)
    x = start + x * (end - start)
    dx = x[0] - x[-1]
    if start == 0:
        x[-2] = x[-1] - dx
    if end == 1:
        x[-3] = x[-4] + dx
    return x


def dot(x, y):
    """utils.dot (utils.py:5-11)."""
    return x.T.conj() @ y


def norm(x):
    """utils.norm (utils.py:13-21)."""
    x = numpy.asarray(x)
    if x.ndim > 1:
        return numpy.array([numpy.sqrt(abs(dot(x[:, i], x[:, i]))) for i in range(x.shape[1])])
    return float(numpy.sqrt(abs(dot(x, x))))
