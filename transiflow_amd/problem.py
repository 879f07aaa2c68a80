"""Host-side description of a discretized problem for the HIP kernels.

Holds what ``transiflow.Discretization`` keeps on the host -- grid sizes,
coordinate vectors, periodicity, the parameter dict (held by reference) --
and turns it into the two inputs of the stencil kernel:

* the physics scalars of the linear part (Discretization.py:229-317);
* the boundary-condition program: the sequence of BoundaryConditions
  primitives that the problem's boundary function calls, recorded by
  :class:`BoundaryConditions` with the same constants the reference computes
  (BoundaryConditions.py:235-469).  User-supplied
  ``boundary_conditions(bc, atom)`` callbacks are recorded the same way.
"""

import numpy

BC_WALL = 0
BC_FORCE = 1


def uniform_coordinate_vector(start, end, n):
    """Coordinate vector of length n+3 (utils.py:175-197): entry i <= n is the
    right face of cell i, entries -2 and -1 are start - dx and start."""
    dx = (end - start) / n
    x = start + numpy.arange(-1, n + 2) * dx
    return numpy.roll(x, -2)


def stretched_coordinate_vector(start, end, n, sigma):
    """tanh-stretched coordinate vector (utils.py:199-232)."""
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
    """sine-stretched coordinate vector (utils.py:234-267)."""
    x = uniform_coordinate_vector(0, 1, n)
    x = x - sigma * numpy.sin(2 * numpy.pi * x)
    x = start + x * (end - start)
    dx = x[0] - x[-1]
    if start == 0:
        x[-2] = x[-1] - dx
    if end == 1:
        x[-3] = x[-4] + dx
    return x


class BoundaryConditions:
    """Records BoundaryConditions primitives as kernel ops.

    Same method names and arguments as ``transiflow.BoundaryConditions``; the
    ``atom`` argument is a placeholder (the atoms only exist inside the
    kernel).  ``get_forcing()`` returns this recorder, so a user callback
    written as ``...; return bc.get_forcing()`` works unchanged.
    """

    _LID_VAR = {0: 1, 1: 0, 2: 0}

    def __init__(self, nx, ny, nz, dim, dof, x, y, z):
        self.n = (nx, ny, nz)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dim = dim
        self.dof = dof
        self.coords = (x, y, z)
        self.x, self.y, self.z = x, y, z
        self.ops = []

    def get_forcing(self):
        return self

    def _wall(self, axis, side, sign):
        self.ops.append((BC_WALL, axis, side, 0, float(sign), 0.0, 0.0))

    def _force(self, axis, side, var, fc, ac):
        if not 0 <= var < self.dof:
            raise ValueError('boundary condition on a variable outside the state (%d)' % var)
        self.ops.append((BC_FORCE, axis, side, var, 0.0, float(fc), float(ac)))

    def _h(self, axis, side):
        c = self.coords[axis]
        n = self.n[axis]
        if side > 0:
            return (c[n] - c[n - 2]) / 2
        return (c[0] - c[-2]) / 2

    # walls (BoundaryConditions.py:55-233)
    def no_slip_east(self, atom=None): self._wall(0, 1, -1)
    def no_slip_west(self, atom=None): self._wall(0, -1, -1)
    def no_slip_north(self, atom=None): self._wall(1, 1, -1)
    def no_slip_south(self, atom=None): self._wall(1, -1, -1)
    def no_slip_top(self, atom=None): self._wall(2, 1, -1)
    def no_slip_bottom(self, atom=None): self._wall(2, -1, -1)
    def free_slip_east(self, atom=None): self._wall(0, 1, 1)
    def free_slip_west(self, atom=None): self._wall(0, -1, 1)
    def free_slip_north(self, atom=None): self._wall(1, 1, 1)
    def free_slip_south(self, atom=None): self._wall(1, -1, 1)
    def free_slip_top(self, atom=None): self._wall(2, 1, 1)
    def free_slip_bottom(self, atom=None): self._wall(2, -1, 1)

    # moving lids (BoundaryConditions.py:235-293)
    def _lid(self, axis, side, velocity):
        self._force(axis, side, self._LID_VAR[axis], 2 * velocity, -1)
        self._wall(axis, side, -1)

    def moving_lid_east(self, atom, velocity): self._lid(0, 1, velocity)
    def moving_lid_west(self, atom, velocity): self._lid(0, -1, velocity)
    def moving_lid_north(self, atom, velocity): self._lid(1, 1, velocity)
    def moving_lid_south(self, atom, velocity): self._lid(1, -1, velocity)
    def moving_lid_top(self, atom, velocity): self._lid(2, 1, velocity)
    def moving_lid_bottom(self, atom, velocity): self._lid(2, -1, velocity)

    # fixed temperature (BoundaryConditions.py:297-337)
    def _temperature(self, axis, side, t):
        self._force(axis, side, self.dim + 1, 2 * t, -1)

    def temperature_east(self, atom, t): self._temperature(0, 1, t)
    def temperature_west(self, atom, t): self._temperature(0, -1, t)
    def temperature_north(self, atom, t): self._temperature(1, 1, t)
    def temperature_south(self, atom, t): self._temperature(1, -1, t)
    def temperature_top(self, atom, t): self._temperature(2, 1, t)
    def temperature_bottom(self, atom, t): self._temperature(2, -1, t)

    # heat flux with optional Biot number (BoundaryConditions.py:339-421)
    def _heat_flux(self, axis, side, q, biot):
        h = self._h(axis, side)
        if side > 0:
            fc = h * q / (1 + h * biot / 2)
            ac = (1 - h * biot / 2) / (1 + h * biot / 2)
        else:
            fc = -h * q / (1 - h * biot / 2)
            ac = (1 + h * biot / 2) / (1 - h * biot / 2)
        self._force(axis, side, self.dim + 1, fc, ac)

    def heat_flux_east(self, atom, q, biot=0.0): self._heat_flux(0, 1, q, biot)
    def heat_flux_west(self, atom, q, biot=0.0): self._heat_flux(0, -1, q, biot)
    def heat_flux_north(self, atom, q, biot=0.0): self._heat_flux(1, 1, q, biot)
    def heat_flux_south(self, atom, q, biot=0.0): self._heat_flux(1, -1, q, biot)
    def heat_flux_top(self, atom, q, biot=0.0): self._heat_flux(2, 1, q, biot)
    def heat_flux_bottom(self, atom, q, biot=0.0): self._heat_flux(2, -1, q, biot)

    # salinity flux (BoundaryConditions.py:423-469)
    def _salinity_flux(self, axis, side, q):
        h = self._h(axis, side)
        self._force(axis, side, self.dim + 2, h * q if side > 0 else -h * q, 1)

    def salinity_flux_east(self, atom, q): self._salinity_flux(0, 1, q)
    def salinity_flux_west(self, atom, q): self._salinity_flux(0, -1, q)
    def salinity_flux_north(self, atom, q): self._salinity_flux(1, 1, q)
    def salinity_flux_south(self, atom, q): self._salinity_flux(1, -1, q)
    def salinity_flux_top(self, atom, q): self._salinity_flux(2, 1, q)
    def salinity_flux_bottom(self, atom, q): self._salinity_flux(2, -1, q)


class Discretization:
    """Host mirror of ``transiflow.Discretization``'s configuration.

    The arithmetic lives in the HIP kernels; this class decides the problem
    dimensions, coordinates, periodicity, physics scalars and boundary
    program (Discretization.py:106-143,186-208,229-317,550-733).
    """

    SUPPORTED = ('lid-driven cavity', 'rayleigh-benard', 'rayleigh-benard perturbation',
                 'differentially heated cavity')

    def __init__(self, parameters, nx, ny, nz=1, dim=None, dof=None,
                 x=None, y=None, z=None, boundary_conditions=None):
        self.parameters = parameters
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dim = dim if dim is not None else (3 if nz > 1 else 2)
        self.dof = dof if dof is not None else self._default_dof()
        self.x_periodic = False
        self.y_periodic = False
        self.z_periodic = nz == 1
        self.x = self.get_coordinate_vector(parameters.get('X-min', 0.0), parameters.get('X-max', 1.0), nx) \
            if x is None else numpy.asarray(x, dtype=float)
        self.y = self.get_coordinate_vector(parameters.get('Y-min', 0.0), parameters.get('Y-max', 1.0), ny) \
            if y is None else numpy.asarray(y, dtype=float)
        self.z = self.get_coordinate_vector(parameters.get('Z-min', 0.0), parameters.get('Z-max', 1.0), nz) \
            if z is None else numpy.asarray(z, dtype=float)
        self.boundary_conditions = boundary_conditions
        if boundary_conditions is None:
            ptype = self.get_parameter('Problem Type', 'Lid-driven Cavity').lower()
            if ptype not in self.SUPPORTED:
                if ptype in ('double gyre', 'amoc'):
                    raise NotImplementedError("problem type '%s' is not supported by the HIP backend" % ptype)
                raise Exception('Invalid problem type %s' % self.get_parameter('Problem Type'))

    def problem_type_equals(self, name):
        return self.get_parameter('Problem Type', 'Lid-driven Cavity').lower() == name.lower()

    def _default_dof(self):
        if self.problem_type_equals('Rayleigh-Benard') or self.problem_type_equals('Rayleigh-Benard Perturbation') \
           or self.problem_type_equals('Differentially Heated Cavity'):
            return self.dim + 2
        if self.problem_type_equals('AMOC'):
            return self.dim + 3
        return self.dim + 1

    def set_parameter(self, name, value):
        if name in self.parameters and self.get_parameter(name) == value:
            return
        self.parameters[name] = value

    def get_parameter(self, name, default=0):
        if name not in self.parameters:
            return default
        return self.parameters.get(name, default)

    def get_coordinate_vector(self, start, end, n):
        p = self.parameters
        if p.get('Grid Stretching', False) or 'Grid Stretching Factor' in p.keys():
            if p.get('Grid Stretching Method', 'tanh') == 'sin':
                return stretched_coordinate_vector_sin(start, end, n, p.get('Grid Stretching Factor', 0.1))
            return stretched_coordinate_vector(start, end, n, p.get('Grid Stretching Factor', 1.5))
        return uniform_coordinate_vector(start, end, n)

    # -- kernel inputs ----------------------------------------------------
    def physics(self):
        """(lin_scale, T_scale, S_scale, rbp_coef, rbp, nonlinear) with the
        reference's arithmetic (Discretization.py:236-267,333-335)."""
        Re = self.get_parameter('Reynolds Number', 1.0)
        Ra = self.get_parameter('Rayleigh Number', 1.0)
        Pr = self.get_parameter('Prandtl Number', 1.0)
        Gr = self.get_parameter('Grashof Number', Ra / Pr)
        Le = self.get_parameter('Lewis Number', 1.0)
        if Re == 0:
            Re = 1
        if Gr == 0:
            Gr = 1 / Pr
        if self.get_parameter('Rossby Parameter'):
            raise NotImplementedError('the Coriolis term is not supported by the HIP backend')
        lin = float(1 / (Re * numpy.sqrt(Gr)))
        t_scale = float(1 / (Pr * numpy.sqrt(Gr)))
        s_scale = float(1 / (Le * Pr * numpy.sqrt(Gr)))
        rbp = self.problem_type_equals('Rayleigh-Benard Perturbation')
        rbp_coef = 0.0
        if rbp:
            Bi = self.get_parameter('Biot Number')
            rbp_coef = float(Bi / (Bi + 1))
        nonlinear = not (self.get_parameter('Reynolds Number') == 0 and not self.dof > self.dim + 1)
        return lin, t_scale, s_scale, rbp_coef, int(rbp), int(nonlinear)

    def boundary_program(self):
        """Ops recorded from the problem's boundary function (or the user's)."""
        bc = BoundaryConditions(self.nx, self.ny, self.nz, self.dim, self.dof, self.x, self.y, self.z)
        fn = self.boundary_conditions or self._problem_boundaries()
        fn(bc, None)
        return bc.ops

    def _problem_boundaries(self):
        if self.problem_type_equals('Lid-driven Cavity'):
            return self._lid_driven_cavity
        if self.problem_type_equals('Rayleigh-Benard') or self.problem_type_equals('Rayleigh-Benard Perturbation'):
            return self._rayleigh_benard
        if self.problem_type_equals('Differentially Heated Cavity'):
            return self._differentially_heated_cavity
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
