"""CPU oracle: finite-volume discretization on the staggered C-grid.

TEST INFRASTRUCTURE ONLY.  This module restates the reference algorithm of
``transiflow/Discretization.py``, ``transiflow/BoundaryConditions.py`` and the
layout helpers of ``transiflow/utils.py`` in vectorised numpy so that the HIP
product path can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package never does.

Every per-cell loop of the reference is replaced by whole-grid array
arithmetic, but the floating-point operations applied to each atom entry, and
their order, are the reference's (left-to-right evaluation of
``1 / dx * dy * dz`` etc.), so the fixtures that the reference compares with
``==`` are met bit for bit.

Atom layout: ``atom[i, j, k, d1, d2, sx, sy, sz]`` -- equation ``d1`` of cell
``(i, j, k)`` refers to variable ``d2`` of the neighbour at slot offset
``(sx-1, sy-1, sz-1)`` (Discretization.py:71-102).
"""

import numpy

from . import layout


def _problem_is(parameters, name):
    """Discretization.problem_type_equals (Discretization.py:550-557)."""
    return parameters.get('Problem Type', 'Lid-driven Cavity').lower() == name.lower()


class _Geometry:
    """Coordinate lookups over the whole cell grid.

    ``role`` permutations express the reference's habit of calling one 1D
    stencil helper with swapped (i, j, k) / (x, y, z) arguments, e.g.
    ``_u_xx(atom, j, i, k, y, x, z)`` is role ``(1, 0, 2)``.
    """

    def __init__(self, nx, ny, nz, x, y, z):
        grids = numpy.meshgrid(numpy.arange(nx), numpy.arange(ny), numpy.arange(nz),
                               indexing='ij')
        self.idx = grids
        self.coord = (numpy.asarray(x, dtype=float), numpy.asarray(y, dtype=float),
                      numpy.asarray(z, dtype=float))

    def c(self, axis, offset):
        # Python-style negative indices (x[-1] is the start of the domain).
        return self.coord[axis][self.idx[axis] + offset]

    def diff(self, axis, lo, hi):
        return self.c(axis, hi) - self.c(axis, lo)

    def half_diff(self, axis, lo, hi):
        return (self.c(axis, hi) - self.c(axis, lo)) / 2


def _slot(axis, s):
    """Index tuple into the (sx, sy, sz) part of an atom for slot s along axis."""
    out = [1, 1, 1]
    out[axis] = s
    return tuple(out)


class Discretization:
    """Oracle restatement of ``transiflow.Discretization`` for the hot path.

    Supports the problem types whose boundary conditions are built from
    ``BoundaryConditions`` primitives: Lid-driven Cavity, Rayleigh-Benard,
    Rayleigh-Benard Perturbation and Differentially Heated Cavity, plus a
    user-supplied ``boundary_conditions(bc, atom)`` callback
    (Discretization.py:577-643,716-733).
    """

    def __init__(self, parameters, nx, ny, nz=1, dim=None, dof=None,
                 x=None, y=None, z=None, boundary_conditions=None):
        self.parameters = parameters
        self.old_parameters = None
        self.nx, self.ny, self.nz = nx, ny, nz

        self.dim = dim if dim is not None else (3 if nz > 1 else 2)
        self.dof = dof if dof is not None else self._default_dof()

        # Discretization.py:123-128
        self.x_periodic = False
        self.y_periodic = False
        self.z_periodic = nz == 1

        self.x = self.get_coordinate_vector(parameters.get('X-min', 0.0),
                                            parameters.get('X-max', 1.0), nx) if x is None else x
        self.y = self.get_coordinate_vector(parameters.get('Y-min', 0.0),
                                            parameters.get('Y-max', 1.0), ny) if y is None else y
        self.z = self.get_coordinate_vector(parameters.get('Z-min', 0.0),
                                            parameters.get('Z-max', 1.0), nz) if z is None else z

        self.atom = None
        self.boundary_conditions = boundary_conditions or self._problem_boundaries()
        self.geom = _Geometry(nx, ny, nz, self.x, self.y, self.z)

    # -- parameters (Discretization.py:145-184) --------------------------
    def set_parameter(self, name, value):
        if name in self.parameters and self.get_parameter(name) == value:
            return
        self.parameters[name] = value

    def get_parameter(self, name, default=0):
        if name not in self.parameters:
            return default
        return self.parameters.get(name, default)

    def problem_type_equals(self, name):
        return _problem_is(self.parameters, name)

    def _default_dof(self):
        # Discretization.py:559-573
        if self.problem_type_equals('Rayleigh-Benard') \
           or self.problem_type_equals('Rayleigh-Benard Perturbation') \
           or self.problem_type_equals('Differentially Heated Cavity'):
            return self.dim + 2
        if self.problem_type_equals('AMOC'):
            return self.dim + 3
        return self.dim + 1

    def get_coordinate_vector(self, start, end, n):
        # Discretization.py:186-208
        p = self.parameters
        if p.get('Grid Stretching', False) or 'Grid Stretching Factor' in p.keys():
            if p.get('Grid Stretching Method', 'tanh') == 'sin':
                return layout.stretched_coordinate_vector_sin(
                    start, end, n, p.get('Grid Stretching Factor', 0.1))
            return layout.stretched_coordinate_vector(
                start, end, n, p.get('Grid Stretching Factor', 1.5))
        return layout.uniform_coordinate_vector(start, end, n)

    def _zeros(self):
        return numpy.zeros((self.nx, self.ny, self.nz, self.dof, self.dof, 3, 3, 3))

    # -- 1D stencil helpers, vectorised over cells ------------------------
    # Each returns the coefficient arrays that the reference helper writes
    # into one atom line; ``r`` is the role permutation (see _Geometry).

    def _second_diff_face(self, r):
        # Discretization.py:741-754 (_u_xx)
        g = self.geom
        dx = g.diff(r[0], -1, 0)
        dxp1 = g.diff(r[0], 0, 1)
        dy = g.diff(r[1], -1, 0)
        dz = g.diff(r[2], -1, 0)
        a0 = 1 / dx * dy * dz
        a2 = 1 / dxp1 * dy * dz
        return a0, -a0 - a2, a2

    def _second_diff_center_y(self, r):
        # Discretization.py:778-791 (_u_yy)
        g = self.geom
        dy = g.half_diff(r[1], -2, 0)
        dyp1 = g.half_diff(r[1], -1, 1)
        dx = g.half_diff(r[0], -1, 1)
        dz = g.diff(r[2], -1, 0)
        a0 = 1 / dy * dx * dz
        a2 = 1 / dyp1 * dx * dz
        return a0, -a0 - a2, a2

    def _second_diff_center_z(self, r):
        # Discretization.py:815-828 (_u_zz)
        g = self.geom
        dz = g.half_diff(r[2], -2, 0)
        dzp1 = g.half_diff(r[2], -1, 1)
        dx = g.half_diff(r[0], -1, 1)
        dy = g.diff(r[1], -1, 0)
        a0 = 1 / dz * dx * dy
        a2 = 1 / dzp1 * dx * dy
        return a0, -a0 - a2, a2

    def _second_diff_scalar(self, r):
        # Discretization.py:852-865 (_C_xx)
        g = self.geom
        dx = g.half_diff(r[0], -2, 0)
        dxp1 = g.half_diff(r[0], -1, 1)
        dy = g.diff(r[1], -1, 0)
        dz = g.diff(r[2], -1, 0)
        a0 = 1 / dx * dy * dz
        a2 = 1 / dxp1 * dy * dz
        return a0, -a0 - a2, a2

    def _face_area(self, r):
        # dy * dz of _forward_u_x / _backward_u_x (Discretization.py:913-921,957-965)
        g = self.geom
        return g.diff(r[1], -1, 0) * g.diff(r[2], -1, 0)

    def _half_volume(self, r):
        # dx * dy * dz / 2 of _forward_average_x / _backward_average_x (:1017-1070)
        g = self.geom
        return g.half_diff(r[0], -1, 1) * g.diff(r[1], -1, 0) * g.diff(r[2], -1, 0) / 2

    # -- linear terms (Discretization.py:756-1084) -------------------------
    def _line(self, d1, d2, axis, coeffs):
        atom = self._zeros()
        for s, c in coeffs:
            atom[(slice(None),) * 3 + (d1, d2) + _slot(axis, s)] = c
        return atom

    def _lap_line(self, var, axis, kind, r):
        a0, a1, a2 = kind(r)
        return self._line(var, var, axis, ((0, a0), (2, a2), (1, a1)))

    def u_xx(self):
        return self._lap_line(0, 0, self._second_diff_face, (0, 1, 2))

    def v_yy(self):
        return self._lap_line(1, 1, self._second_diff_face, (1, 0, 2))

    def w_zz(self):
        return self._lap_line(2, 2, self._second_diff_face, (2, 1, 0))

    def u_yy(self):
        return self._lap_line(0, 1, self._second_diff_center_y, (0, 1, 2))

    def v_xx(self):
        return self._lap_line(1, 0, self._second_diff_center_y, (1, 0, 2))

    def w_yy(self):
        return self._lap_line(2, 1, self._second_diff_center_y, (2, 1, 0))

    def u_zz(self):
        return self._lap_line(0, 2, self._second_diff_center_z, (0, 1, 2))

    def v_zz(self):
        return self._lap_line(1, 2, self._second_diff_center_z, (1, 0, 2))

    def w_xx(self):
        return self._lap_line(2, 0, self._second_diff_center_z, (2, 1, 0))

    def C_xx(self, var):
        return self._lap_line(var, 0, self._second_diff_scalar, (0, 1, 2))

    def C_yy(self, var):
        return self._lap_line(var, 1, self._second_diff_scalar, (1, 0, 2))

    def C_zz(self, var):
        return self._lap_line(var, 2, self._second_diff_scalar, (2, 1, 0))

    _ROLES = ((0, 1, 2), (1, 0, 2), (2, 1, 0))

    def _gradient(self, axis, var):
        # C_x / C_y / C_z: forward difference in equation ``axis``
        a2 = self._face_area(self._ROLES[axis])
        return self._line(axis, var, axis, ((2, a2), (1, -a2)))

    def _divergence_part(self, axis):
        # u_x / v_y / w_z: backward difference of velocity ``axis``
        a1 = self._face_area(self._ROLES[axis])
        return self._line(self.dim, axis, axis, ((1, a1), (0, -a1)))

    def div(self):
        out = self._divergence_part(0) + self._divergence_part(1)
        if self.dim != 2:
            out = out + self._divergence_part(2)
        return out

    def _forward_average(self, axis, var):
        # forward_average_C_y / _C_z (Discretization.py:1029-1041)
        a1 = self._half_volume(self._ROLES[axis])
        return self._line(axis, var, axis, ((1, a1), (2, a1)))

    def _backward_average(self, axis):
        # backward_average_v_y / _w_z (Discretization.py:1072-1084)
        a0 = self._half_volume(self._ROLES[axis])
        return self._line(self.dim + 1, axis, axis, ((0, a0), (1, a0)))

    # -- linear part (Discretization.py:210-317) ---------------------------
    def linear_part(self):
        parameters = dict(self.parameters)
        if parameters != self.old_parameters:
            self.atom = self._linear_part_2D() if self.dim == 2 else self._linear_part_3D()
            self.old_parameters = parameters
        return self.atom

    def _scalings(self):
        Re = self.get_parameter('Reynolds Number', 1.0)
        Ra = self.get_parameter('Rayleigh Number', 1.0)
        Pr = self.get_parameter('Prandtl Number', 1.0)
        Gr = self.get_parameter('Grashof Number', Ra / Pr)
        Le = self.get_parameter('Lewis Number', 1.0)
        if Re == 0:
            Re = 1
        if Gr == 0:
            Gr = 1 / Pr
        return Re, Pr, Gr, Le

    def _linear_part_2D(self):
        Re, Pr, Gr, Le = self._scalings()
        if self.get_parameter('Rossby Parameter'):
            raise NotImplementedError('Coriolis term is outside the oracle scope')
        atom = 1 / (Re * numpy.sqrt(Gr)) * (self.u_xx() + self.u_yy() + self.v_xx() + self.v_yy()) \
            - (self._gradient(0, self.dim) + self._gradient(1, self.dim)) \
            + self.div()
        if self.dof > 3:
            atom += 1 / (Pr * numpy.sqrt(Gr)) * (self.C_xx(self.dim + 1) + self.C_yy(self.dim + 1))
            atom += self._forward_average(1, self.dim + 1)
        if self.dof > 4:
            atom += 1 / (Le * Pr * numpy.sqrt(Gr)) * (self.C_xx(self.dim + 2) + self.C_yy(self.dim + 2))
            atom -= self._forward_average(1, self.dim + 2)
        if self.problem_type_equals('Rayleigh-Benard Perturbation'):
            Bi = self.get_parameter('Biot Number')
            atom += Bi / (Bi + 1) * self._backward_average(1)
        return atom

    def _linear_part_3D(self):
        Re, Pr, Gr, Le = self._scalings()
        atom = 1 / (Re * numpy.sqrt(Gr)) * (self.u_xx() + self.u_yy() + self.u_zz()
                                            + self.v_xx() + self.v_yy() + self.v_zz()
                                            + self.w_xx() + self.w_yy() + self.w_zz()) \
            - (self._gradient(0, self.dim) + self._gradient(1, self.dim)
               + self._gradient(2, self.dim)) \
            + self.div()
        buoy_axis = 2 if self.nz > 1 else 1
        if self.dof > 4:
            T = self.dim + 1
            atom += 1 / (Pr * numpy.sqrt(Gr)) * (self.C_xx(T) + self.C_yy(T) + self.C_zz(T))
            atom += self._forward_average(buoy_axis, T)
        if self.dof > 5:
            S = self.dim + 2
            atom += 1 / (Le * Pr * numpy.sqrt(Gr)) * (self.C_xx(S) + self.C_yy(S) + self.C_zz(S))
            atom -= self._forward_average(buoy_axis, S)
        if self.problem_type_equals('Rayleigh-Benard Perturbation'):
            Bi = self.get_parameter('Biot Number')
            atom += Bi / (Bi + 1) * self._backward_average(buoy_axis)
        return atom

    # -- nonlinear part (Discretization.py:319-365,1178-1459) -------------
    def _padded(self, state):
        return layout.padded_state_mtx(state, self.nx, self.ny, self.nz, self.dof,
                                       self.x_periodic, self.y_periodic, self.z_periodic)

    @staticmethod
    def _avg(sp, axis):
        """average_x/y/z on the padded (nx+2, ny+2, nz+2) field.

        Averages along ``axis`` over all nx+2 padded points (n+1 results),
        cropped to the interior along the other two axes.
        """
        n = [s - 2 for s in sp.shape]
        crop = [slice(1, n[a] + 1) for a in range(3)]
        lo = list(crop)
        hi = list(crop)
        lo[axis] = slice(0, n[axis] + 1)
        hi[axis] = slice(1, n[axis] + 2)
        out = numpy.zeros(tuple(n[a] + 1 if a == axis else n[a] for a in range(3)))
        out += 1 / 2 * sp[tuple(lo)]
        out += 1 / 2 * sp[tuple(hi)]
        return out

    def _weights(self, axis):
        # _weighted_average (Discretization.py:1212-1221) for every index along axis
        c = numpy.asarray((self.x, self.y, self.z)[axis], dtype=float)
        i = numpy.arange((self.nx, self.ny, self.nz)[axis])
        dx = (c[i + 1] - c[i - 1]) / 2
        w0 = 1 / 2 * (c[i] - c[i - 1]) / dx
        w1 = 1 / 2 * (c[i + 1] - c[i]) / dx
        return w0, w1

    def _wavg(self, sp, axis):
        """weighted_average_x/y/z: (n) results along axis, n+1 along the others."""
        n = [s - 2 for s in sp.shape]
        w0, w1 = self._weights(axis)
        shape = [1, 1, 1]
        shape[axis] = -1
        w0 = w0.reshape(shape)
        w1 = w1.reshape(shape)
        crop = [slice(0, n[a] + 1) for a in range(3)]
        lo = list(crop)
        hi = list(crop)
        lo[axis] = slice(1, n[axis] + 1)
        hi[axis] = slice(2, n[axis] + 2)
        out = numpy.zeros(tuple(n[a] if a == axis else n[a] + 1 for a in range(3)))
        out += w0 * sp[tuple(lo)]
        out += w1 * sp[tuple(hi)]
        return out

    def _cells(self, arr, off):
        """arr[i + off0, j + off1, k + off2] over all cells."""
        return arr[off[0]:off[0] + self.nx, off[1]:off[1] + self.ny, off[2]:off[2] + self.nz]

    @staticmethod
    def _sub_line(atom, d1, d2, axis, slots, value):
        """atom[..., d1, d2, <slots along axis>] -= value (same value on each slot)."""
        for s in slots:
            atom[(slice(None),) * 3 + (d1, d2) + _slot(axis, s)] -= value

    def _self_advection(self, atomJ, atomF, sp, a):
        """u_u_x / v_v_y / w_w_z (Discretization.py:1268-1279,1351-1362,1434-1445)."""
        avg = self._avg(sp[..., a], a)
        area = self._face_area(self._ROLES[a])
        a1 = -area
        a2 = area
        off0 = [0, 0, 0]
        off1 = [0, 0, 0]
        off1[a] = 1
        v1 = a1 * self._cells(avg, off0) * 1 / 2
        v2 = a2 * self._cells(avg, off1) * 1 / 2
        for atom in (atomF, atomJ):
            self._sub_line(atom, a, a, a, (0, 1), v1)
            self._sub_line(atom, a, a, a, (1, 2), v2)

    def _cross_advection(self, atomJ, atomF, sp, a, b, F_off, J_off):
        """Transport of velocity ``b`` by velocity ``a`` (a != b).

        Covers u_v_x, v_u_y, u_w_x, v_w_y, w_u_z, w_v_z
        (Discretization.py:1281-1311,1335-1349,1364-1378,1402-1432).
        The face area is half_diff along ``b`` times the cell width along the
        third axis (the _backward_u_y/_u_z call of each term), ``F_off`` the offsets into the weighted average of velocity ``a``
        (transporting velocity, averaged to the ``b`` location) and ``J_off``
        the offsets into the plain average of ``b``.
        """
        wav = self._wavg(sp[..., a], b)
        pav = self._avg(sp[..., b], a)
        third = 3 - a - b
        g = self.geom
        area = g.half_diff(b, -1, 1) * g.diff(third, -1, 0)
        a0 = -area
        a1 = area
        f0 = a0 * self._cells(wav, F_off[0]) * 1 / 2
        f1 = a1 * self._cells(wav, F_off[1]) * 1 / 2
        self._sub_line(atomF, b, b, a, (0, 1), f0)
        self._sub_line(atomF, b, b, a, (1, 2), f1)

        w0, w1 = self._weights(b)
        shape = [1, 1, 1]
        shape[b] = -1
        w = (w0.reshape(shape), w1.reshape(shape))
        t0 = a0 * self._cells(pav, J_off[0])
        t1 = a1 * self._cells(pav, J_off[1])
        for slot_a, t in ((0, t0), (1, t1)):
            for s_b, ww in ((1, w[0]), (2, w[1])):
                idx = [1, 1, 1]
                idx[a] = slot_a
                idx[b] = s_b
                atomJ[(slice(None),) * 3 + (b, a) + tuple(idx)] -= t * ww

    def _scalar_advection(self, atomJ, atomF, sp, a, var):
        """u_C_x / v_C_y / w_C_z (Discretization.py:1313-1325,1380-1392,1447-1459)."""
        n = (self.nx, self.ny, self.nz)
        sl = [slice(1, n[0] + 1), slice(1, n[1] + 1), slice(1, n[2] + 1)]
        sl[a] = slice(0, n[a] + 1)
        vel = sp[tuple(sl) + (a,)]
        avgC = self._avg(sp[..., var], a)
        area = self._face_area(self._ROLES[a])
        a1 = area
        a0 = -a1
        off1 = [0, 0, 0]
        off1[a] = 1
        v0 = a0 * self._cells(vel, (0, 0, 0)) * 1 / 2
        v1 = a1 * self._cells(vel, off1) * 1 / 2
        self._sub_line(atomF, var, var, a, (0, 1), v0)
        self._sub_line(atomF, var, var, a, (1, 2), v1)
        atomJ[(slice(None),) * 3 + (var, a) + _slot(a, 0)] -= a0 * self._cells(avgC, (0, 0, 0))
        atomJ[(slice(None),) * 3 + (var, a) + _slot(a, 1)] -= a1 * self._cells(avgC, off1)

    def nonlinear_part(self, state):
        sp = self._padded(state)
        atomJ = self._zeros()
        atomF = self._zeros()

        Re = self.get_parameter('Reynolds Number')
        if Re == 0 and not self.dof > self.dim + 1:
            return atomJ, atomF

        self._self_advection(atomJ, atomF, sp, 0)                       # u_u_x
        self._cross_advection(atomJ, atomF, sp, 0, 1,  # u_v_x
                              ((0, 0, 1), (1, 0, 1)), ((0, 0, 0), (1, 0, 0)))
        self._cross_advection(atomJ, atomF, sp, 1, 0,  # v_u_y
                              ((0, 0, 1), (0, 1, 1)), ((0, 0, 0), (0, 1, 0)))
        self._self_advection(atomJ, atomF, sp, 1)                       # v_v_y

        if self.dim > 2:
            self._cross_advection(atomJ, atomF, sp, 0, 2,  # u_w_x
                              ((0, 1, 0), (1, 1, 0)), ((0, 0, 0), (1, 0, 0)))
            self._cross_advection(atomJ, atomF, sp, 1, 2,  # v_w_y
                              ((1, 0, 0), (1, 1, 0)), ((0, 0, 0), (0, 1, 0)))
            self._cross_advection(atomJ, atomF, sp, 2, 0,  # w_u_z
                              ((0, 1, 0), (0, 1, 1)), ((0, 0, 0), (0, 0, 1)))
            self._cross_advection(atomJ, atomF, sp, 2, 1,  # w_v_z
                              ((1, 0, 0), (1, 0, 1)), ((0, 0, 0), (0, 0, 1)))
            self._self_advection(atomJ, atomF, sp, 2)                   # w_w_z

        if self.dof > self.dim + 1:
            T = self.dim + 1
            self._scalar_advection(atomJ, atomF, sp, 0, T)
            self._scalar_advection(atomJ, atomF, sp, 1, T)
            if self.dim > 2:
                self._scalar_advection(atomJ, atomF, sp, 2, T)

        if self.dof > self.dim + 2:
            S = self.dim + 2
            self._scalar_advection(atomJ, atomF, sp, 0, S)
            self._scalar_advection(atomJ, atomF, sp, 1, S)
            if self.dim > 2:
                self._scalar_advection(atomJ, atomF, sp, 2, S)

        atomJ += atomF
        return atomJ, atomF

    # -- rhs / jacobian / mass (Discretization.py:367-548) -----------------
    def rhs(self, state):
        atomJ, atomF = self.nonlinear_part(state)
        atomF += self.linear_part()
        frc = self.boundaries(atomF)
        return self.assemble_rhs(state, atomF) + frc

    def jacobian(self, state):
        atomJ, atomF = self.nonlinear_part(state)
        atomJ += self.linear_part()
        self.boundaries(atomJ)
        return self.assemble_jacobian(atomJ)

    def assemble_rhs(self, state, atom):
        sp = self._padded(state)
        out = numpy.zeros((self.nx, self.ny, self.nz, self.dof))
        nx, ny, nz = self.nx, self.ny, self.nz
        for sz, sy, sx, d1, d2 in numpy.ndindex(3, 3, 3, self.dof, self.dof):
            out[:, :, :, d1] += atom[:, :, :, d1, d2, sx, sy, sz] \
                * sp[sx:sx + nx, sy:sy + ny, sz:sz + nz, d2]
        return layout.state_vec(out)

    def assemble_jacobian(self, atom):
        from .crs import assemble_from_atom
        return assemble_from_atom(atom)

    def mass_matrix(self):
        g = self.geom
        m = numpy.zeros((self.nx, self.ny, self.nz, self.dof))
        for a in range(self.dim):
            r = self._ROLES[a]
            m[..., a] += g.half_diff(r[0], -1, 1) * g.diff(r[1], -1, 0) * g.diff(r[2], -1, 0)
        vol = g.diff(0, -1, 0) * g.diff(1, -1, 0) * g.diff(2, -1, 0)
        if self.dof > self.dim + 1:
            m[..., self.dim + 1] += vol
        if self.dof > self.dim + 2:
            m[..., self.dim + 2] += vol
        from .crs import assemble_diagonal
        return assemble_diagonal(m)

    # -- boundary conditions ----------------------------------------------
    def boundaries(self, atom):
        bc = BoundaryConditions(self.nx, self.ny, self.nz, self.dim, self.dof,
                                self.x, self.y, self.z)
        return self.boundary_conditions(bc, atom)

    def _problem_boundaries(self):
        # Discretization.py:716-733
        if self.problem_type_equals('Lid-driven Cavity'):
            return self._lid_driven_cavity
        if self.problem_type_equals('Rayleigh-Benard') \
           or self.problem_type_equals('Rayleigh-Benard Perturbation'):
            return self._rayleigh_benard
        if self.problem_type_equals('Differentially Heated Cavity'):
            return self._differentially_heated_cavity
        if self.problem_type_equals('Double Gyre') or self.problem_type_equals('AMOC'):
            raise NotImplementedError('problem type outside the oracle scope')
        raise Exception('Invalid problem type %s' % self.get_parameter('Problem Type'))

    def _lid_driven_cavity(self, bc, atom):
        # Discretization.py:577-593
        v = self.get_parameter('Lid Velocity', 1)
        bc.no_slip_east(atom)
        bc.no_slip_west(atom)
        bc.no_slip_south(atom)
        if self.dim == 2 or self.nz <= 1:
            bc.moving_lid_north(atom, v)
            return bc.get_forcing()
        bc.no_slip_north(atom)
        bc.no_slip_bottom(atom)
        bc.moving_lid_top(atom, v)
        return bc.get_forcing()

    def _rayleigh_benard(self, bc, atom):
        # Discretization.py:595-623
        asym = self.get_parameter('Asymmetry Parameter')
        bc.heat_flux_east(atom, asym)
        bc.heat_flux_west(atom, 0)
        bc.no_slip_east(atom)
        bc.no_slip_west(atom)
        Bi = self.get_parameter('Biot Number')
        bottom = 1 if self.problem_type_equals('Rayleigh-Benard') else 0
        if self.dim == 2 or self.nz <= 1:
            bc.heat_flux_north(atom, 0, Bi)
            bc.temperature_south(atom, bottom)
            bc.free_slip_north(atom)
            bc.no_slip_south(atom)
            return bc.get_forcing()
        bc.heat_flux_north(atom, 0)
        bc.heat_flux_south(atom, 0)
        bc.no_slip_north(atom)
        bc.no_slip_south(atom)
        bc.heat_flux_top(atom, 0, Bi)
        bc.temperature_bottom(atom, bottom)
        bc.free_slip_top(atom)
        bc.no_slip_bottom(atom)
        return bc.get_forcing()

    def _differentially_heated_cavity(self, bc, atom):
        # Discretization.py:625-643
        bc.temperature_east(atom, -1 / 2)
        bc.temperature_west(atom, 1 / 2)
        bc.no_slip_east(atom)
        bc.no_slip_west(atom)
        bc.heat_flux_north(atom, 0)
        bc.heat_flux_south(atom, 0)
        bc.no_slip_north(atom)
        bc.no_slip_south(atom)
        if self.dim > 2 and self.nz > 1:
            bc.heat_flux_top(atom, 0)
            bc.heat_flux_bottom(atom, 0)
            bc.no_slip_top(atom)
            bc.no_slip_bottom(atom)
        return bc.get_forcing()


class BoundaryConditions:
    """Oracle restatement of ``transiflow.BoundaryConditions``.

    Faces are addressed by (axis, side): side +1 is east/north/top, -1 is
    west/south/bottom.  Every operation edits the atom in place and updates
    the accumulated forcing ``frc`` (BoundaryConditions.py:37-544).
    """

    def __init__(self, nx, ny, nz, dim, dof, x, y, z):
        self.n = (nx, ny, nz)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dim = dim
        self.dof = dof
        self.coords = (x, y, z)
        self.x, self.y, self.z = x, y, z
        self.frc = numpy.zeros((nx, ny, nz, dof))

    def get_forcing(self):
        return layout.state_vec(self.frc)

    # index helpers
    def _cell(self, axis, pos):
        idx = [slice(None)] * 3
        idx[axis] = pos
        return tuple(idx)

    def _at(self, axis, pos, d1=slice(None), d2=slice(None), slot=None):
        s = [slice(None)] * 3
        if slot is not None:
            s[axis] = slot
        return self._cell(axis, pos) + (d1, d2) + tuple(s)

    def _wall(self, atom, axis, side, sign):
        """no_slip (sign -1) / free_slip (sign +1) at face (axis, side).

        BoundaryConditions.py:55-233.  The high faces also pin the
        wall-normal velocity row to -1 and clear the neighbour's reference to
        it; ``frc`` of the normal component is cleared there (for the top
        face the reference clears component 1, reproduced here).
        """
        n = self.n[axis]
        if side > 0:
            c = n - 1
            atom[self._at(axis, c, slot=1)] += sign * atom[self._at(axis, c, slot=2)]
            atom[self._at(axis, c, d2=axis, slot=1)] = 0
            atom[self._at(axis, c, d1=axis)] = 0
            atom[self._at(axis, c, slot=2)] = 0
            centre = self._cell(axis, c) + (axis, axis, 1, 1, 1)
            atom[centre] = -1
            atom[self._at(axis, n - 2, d1=axis, d2=axis, slot=2)] = 0
            self.frc[self._cell(axis, c) + (axis if axis < 2 else 1,)] = 0
        else:
            atom[self._at(axis, 0, d2=axis, slot=0)] = 0
            atom[self._at(axis, 0, slot=1)] += sign * atom[self._at(axis, 0, slot=0)]
            atom[self._at(axis, 0, slot=0)] = 0

    def _constant_forcing(self, atom, axis, side, var, forcing_constant, atom_constant):
        """_constant_forcing_<face> (BoundaryConditions.py:471-544).

        The forcing of equation ``var`` sums its coefficients on variable
        ``var`` at the outside slot, over the two tangential slot axes (the
        higher-numbered one outermost), times the constant.
        """
        n = self.n[axis]
        c = n - 1 if side > 0 else 0
        out_slot = 2 if side > 0 else 0
        tang = [a for a in range(3) if a != axis]
        frc = numpy.zeros((self.nx, self.ny, self.nz, self.dof))
        acc = numpy.zeros(atom[self._cell(axis, c) + (0, 0, 0, 0, 0)].shape)
        for s_b in range(3):
            for s_a in range(3):
                sl = [0, 0, 0]
                sl[axis] = out_slot
                sl[tang[0]] = s_a
                sl[tang[1]] = s_b
                acc += atom[self._cell(axis, c) + (var, var) + tuple(sl)] * forcing_constant
        frc[self._cell(axis, c) + (var,)] = acc

        atom[self._at(axis, c, d2=var, slot=1)] += atom_constant * atom[self._at(axis, c, d2=var, slot=out_slot)]
        atom[self._at(axis, c, d2=var, slot=out_slot)] = 0
        return frc

    # public primitives with the reference's names
    def no_slip_east(self, atom):
        self._wall(atom, 0, 1, -1)

    def no_slip_west(self, atom):
        self._wall(atom, 0, -1, -1)

    def no_slip_north(self, atom):
        self._wall(atom, 1, 1, -1)

    def no_slip_south(self, atom):
        self._wall(atom, 1, -1, -1)

    def no_slip_top(self, atom):
        self._wall(atom, 2, 1, -1)

    def no_slip_bottom(self, atom):
        self._wall(atom, 2, -1, -1)

    def free_slip_east(self, atom):
        self._wall(atom, 0, 1, 1)

    def free_slip_west(self, atom):
        self._wall(atom, 0, -1, 1)

    def free_slip_north(self, atom):
        self._wall(atom, 1, 1, 1)

    def free_slip_south(self, atom):
        self._wall(atom, 1, -1, 1)

    def free_slip_top(self, atom):
        self._wall(atom, 2, 1, 1)

    def free_slip_bottom(self, atom):
        self._wall(atom, 2, -1, 1)

    # moving lids (BoundaryConditions.py:235-293): tangential component
    _LID_VAR = {0: 1, 1: 0, 2: 0}

    def _lid(self, atom, axis, side, velocity):
        self.frc += self._constant_forcing(atom, axis, side, self._LID_VAR[axis], 2 * velocity, -1)
        self._wall(atom, axis, side, -1)

    def moving_lid_east(self, atom, velocity):
        self._lid(atom, 0, 1, velocity)

    def moving_lid_west(self, atom, velocity):
        self._lid(atom, 0, -1, velocity)

    def moving_lid_north(self, atom, velocity):
        self._lid(atom, 1, 1, velocity)

    def moving_lid_south(self, atom, velocity):
        self._lid(atom, 1, -1, velocity)

    def moving_lid_top(self, atom, velocity):
        self._lid(atom, 2, 1, velocity)

    def moving_lid_bottom(self, atom, velocity):
        self._lid(atom, 2, -1, velocity)

    # temperature / heat flux / salinity flux (BoundaryConditions.py:297-469)
    def _temperature(self, atom, axis, side, temperature):
        self.frc += self._constant_forcing(atom, axis, side, self.dim + 1, 2 * temperature, -1)

    def temperature_east(self, atom, t):
        self._temperature(atom, 0, 1, t)

    def temperature_west(self, atom, t):
        self._temperature(atom, 0, -1, t)

    def temperature_north(self, atom, t):
        self._temperature(atom, 1, 1, t)

    def temperature_south(self, atom, t):
        self._temperature(atom, 1, -1, t)

    def temperature_top(self, atom, t):
        self._temperature(atom, 2, 1, t)

    def temperature_bottom(self, atom, t):
        self._temperature(atom, 2, -1, t)

    def _h(self, axis, side):
        c = self.coords[axis]
        n = self.n[axis]
        if side > 0:
            return (c[n] - c[n - 2]) / 2
        return (c[0] - c[-2]) / 2

    def _heat_flux(self, atom, axis, side, heat_flux, biot):
        h = self._h(axis, side)
        if side > 0:
            fc = h * heat_flux / (1 + h * biot / 2)
            ac = (1 - h * biot / 2) / (1 + h * biot / 2)
        else:
            fc = -h * heat_flux / (1 - h * biot / 2)
            ac = (1 + h * biot / 2) / (1 - h * biot / 2)
        self.frc += self._constant_forcing(atom, axis, side, self.dim + 1, fc, ac)

    def heat_flux_east(self, atom, q, biot=0.0):
        self._heat_flux(atom, 0, 1, q, biot)

    def heat_flux_west(self, atom, q, biot=0.0):
        self._heat_flux(atom, 0, -1, q, biot)

    def heat_flux_north(self, atom, q, biot=0.0):
        self._heat_flux(atom, 1, 1, q, biot)

    def heat_flux_south(self, atom, q, biot=0.0):
        self._heat_flux(atom, 1, -1, q, biot)

    def heat_flux_top(self, atom, q, biot=0.0):
        self._heat_flux(atom, 2, 1, q, biot)

    def heat_flux_bottom(self, atom, q, biot=0.0):
        self._heat_flux(atom, 2, -1, q, biot)

    def _salinity_flux(self, atom, axis, side, q):
        h = self._h(axis, side)
        fc = h * q if side > 0 else -h * q
        self.frc += self._constant_forcing(atom, axis, side, self.dim + 2, fc, 1)

    def salinity_flux_east(self, atom, q):
        self._salinity_flux(atom, 0, 1, q)

    def salinity_flux_west(self, atom, q):
        self._salinity_flux(atom, 0, -1, q)

    def salinity_flux_north(self, atom, q):
        self._salinity_flux(atom, 1, 1, q)

    def salinity_flux_south(self, atom, q):
        self._salinity_flux(atom, 1, -1, q)

    def salinity_flux_top(self, atom, q):
        self._salinity_flux(atom, 2, 1, q)

    def salinity_flux_bottom(self, atom, q):
        self._salinity_flux(atom, 2, -1, q)
