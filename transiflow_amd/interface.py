"""The HIP backend Interface: a drop-in for ``transiflow.interface.SciPy.Interface``.

Same constructor and methods as the reference backend (contract in
``transiflow/interface/BaseInterface.py:32-386``, serial implementation
``transiflow/interface/SciPy.py:9-340``):

* ``vector()``            -> zero :class:`DeviceVector` of length n
* ``rhs(state)``          -> new F(state) (one fused stencil kernel)
* ``jacobian(state)``     -> :class:`DeviceCsrMatrix` with the reference's exact
                            CSR pattern (count / scan / fill kernels)
* ``mass_matrix()``       -> diagonal :class:`DeviceCsrMatrix`
* ``solve(jac, rhs, rhs2=None, V=None, W=None, C=None)`` -> preconditioned
  restarted GMRES on the pressure-gauged system (SciPy.py:204-258 semantics)
* ``set_parameter``/``get_parameter``, ``vector_from_array``/``array_from_vector``,
  ``save_state``/``load_state`` and the attributes ``nx, ny, nz, dim, dof, x,
  y, z, parameters, discretization, pressure_row, border_scaling``.

Solver options live in the same ``parameters`` dict as the reference's
(``'Iterative Solver'`` sub-dict) plus the keys documented in
:mod:`transiflow_amd.solver`.
"""

import ctypes
import json

import numpy

from . import _lib
from .problem import Discretization
from .vector import DeviceCsrMatrix, DeviceVector, ptr, stream, torch


class NumpyEncoder(json.JSONEncoder):
    """JSON encoding of numpy scalars/arrays and complex numbers (BaseInterface.py:10-23)."""

    def default(self, obj):
        if isinstance(obj, (numpy.integer,)):
            return int(obj)
        if isinstance(obj, (numpy.floating,)):
            return float(obj)
        if isinstance(obj, (complex, numpy.complexfloating)):
            return {'__complex__': True, 'real': obj.real, 'imag': obj.imag}
        if isinstance(obj, numpy.ndarray):
            return obj.tolist()
        return json.JSONEncoder.default(self, obj)


def numpy_decoder(obj):
    if '__complex__' in obj:
        return complex(obj['real'], obj['imag'])
    return obj


class Interface:
    def __init__(self, parameters, nx, ny, nz=1, dim=None, dof=None,
                 x=None, y=None, z=None, boundary_conditions=None):
        self.discretization = Discretization(parameters, nx, ny, nz, dim, dof, x, y, z,
                                             boundary_conditions)
        d = self.discretization
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dim = dim or d.dim
        self.dof = dof or d.dof
        self.x, self.y, self.z = d.x, d.y, d.z
        self.parameters = parameters
        self.pressure_row = self.dim
        self.border_scaling = 1e-3
        self.n = nx * ny * nz * self.dof

        torch()  # fail loudly without a GPU
        handle = ctypes.c_void_p()
        coords = [numpy.ascontiguousarray(c, dtype=numpy.float64) for c in (d.x, d.y, d.z)]
        _lib.call('tf_problem_create', ctypes.byref(handle), nx, ny, nz, self.dim, self.dof,
                  _lib.dptr(coords[0]), _lib.dptr(coords[1]), _lib.dptr(coords[2]),
                  int(d.x_periodic), int(d.y_periodic), int(d.z_periodic))
        self._handle = handle
        self._coords = coords
        self._pushed_physics = None
        self._pushed_bcs = None
        self._solver = None

    def __del__(self):
        h = getattr(self, '_handle', None)
        if h is not None and h.value:
            try:
                _lib.load().tf_problem_destroy(h)
            except Exception:
                pass
            self._handle = None

    # -- parameters -------------------------------------------------------
    def set_parameter(self, name, value):
        self.discretization.set_parameter(name, value)

    def get_parameter(self, name):
        return self.discretization.get_parameter(name)

    def _sync(self):
        """Push the physics scalars and boundary program if the parameters changed."""
        phys = self.discretization.physics()
        if phys != self._pushed_physics:
            _lib.call('tf_problem_set_physics', self._handle, *phys)
            self._pushed_physics = phys
        ops = tuple(self.discretization.boundary_program())
        if ops != self._pushed_bcs:
            if len(ops) > 32:
                raise ValueError('at most 32 boundary-condition operations are supported')
            cols = list(zip(*ops)) if ops else [()] * 7
            ia = [numpy.ascontiguousarray(c, dtype=numpy.int32) for c in cols[:4]]
            da = [numpy.ascontiguousarray(c, dtype=numpy.float64) for c in cols[4:]]
            _lib.call('tf_problem_set_bcs', self._handle, len(ops), *[_lib.iptr(a) for a in ia],
                      *[_lib.dptr(a) for a in da])
            self._pushed_bcs = ops

    # -- vectors ----------------------------------------------------------
    def vector(self):
        return DeviceVector.zeros(self.n)

    def vector_from_array(self, array):
        return DeviceVector.from_array(array)

    def array_from_vector(self, vector):
        if isinstance(vector, DeviceVector):
            return vector.numpy()
        return numpy.asarray(vector)

    def _device(self, state):
        if isinstance(state, DeviceVector):
            return state
        return DeviceVector.from_array(state)

    # -- model ------------------------------------------------------------
    def rhs(self, state):
        self._sync()
        x = self._device(state)
        F = torch().empty(self.n, dtype=torch().float64, device='cuda')
        _lib.call('tf_rhs', self._handle, ptr(x.data), ptr(F), stream())
        return DeviceVector(F)

    def jacobian(self, state):
        self._sync()
        T = torch()
        x = self._device(state)
        row_ptr = T.empty(self.n + 1, dtype=T.int32, device='cuda')
        nnz = ctypes.c_int()
        _lib.call('tf_jacobian_count', self._handle, ptr(x.data), ptr(row_ptr), ctypes.byref(nnz), stream())
        col = T.empty(max(nnz.value, 1), dtype=T.int32, device='cuda')[:nnz.value]
        val = T.empty(max(nnz.value, 1), dtype=T.float64, device='cuda')[:nnz.value]
        _lib.call('tf_jacobian_fill', self._handle, ptr(x.data), ptr(row_ptr), ptr(col), ptr(val), stream())
        J = DeviceCsrMatrix(row_ptr, col, val, self.n)
        # where the Jacobian was taken: the two-grid preconditioner re-assembles it
        # on the coarse grid at the restricted state with the same parameters
        J.state = x.data.clone()
        J.parameters = dict(self.parameters)
        return J

    def mass_matrix(self):
        T = torch()
        row_ptr = T.empty(self.n + 1, dtype=T.int32, device='cuda')
        nnz = ctypes.c_int()
        _lib.call('tf_mass_count', self._handle, ptr(row_ptr), ctypes.byref(nnz), stream())
        col = T.empty(max(nnz.value, 1), dtype=T.int32, device='cuda')[:nnz.value]
        val = T.empty(max(nnz.value, 1), dtype=T.float64, device='cuda')[:nnz.value]
        _lib.call('tf_mass_fill', self._handle, ptr(row_ptr), ptr(col), ptr(val), stream())
        return DeviceCsrMatrix(row_ptr, col, val, self.n)

    # -- linear solves ----------------------------------------------------
    @property
    def solver(self):
        if self._solver is None:
            from .solver import Solver
            self._solver = Solver(self)
        return self._solver

    def solve(self, jac, rhs, rhs2=None, V=None, W=None, C=None):
        return self.solver.solve(jac, rhs, rhs2, V, W, C)

    def solve_pair(self, jac, rhs0, rhs1):
        """(solve(jac, rhs0), solve(jac, rhs1)), the two solves of a Newton-corrector
        iteration (Continuation.py:158-159), solved together (Solver.solve_pair)."""
        return self.solver.solve_pair(jac, rhs0, rhs1)

    # -- persistence (BaseInterface.py:118-215) ---------------------------
    def save_json(self, name, obj):
        with open(name, 'w') as f:
            json.dump(obj, f, cls=NumpyEncoder)

    def save_parameters(self, name):
        self.save_json(name + '.params', self.parameters)

    def save_state(self, name, x):
        self.save_parameters(name)
        if not name.endswith('.npy'):
            name += '.npy'
        numpy.save(name, self.array_from_vector(x))

    def load_json(self, name):
        with open(name, 'r') as f:
            return json.load(f, object_hook=numpy_decoder)

    def load_parameters(self, name):
        self.parameters.update(self.load_json(name + '.params'))

    def load_state(self, name):
        self.load_parameters(name)
        if not name.endswith('.npy'):
            name += '.npy'
        return self.vector_from_array(numpy.load(name))

    def eigs(self, state, return_eigenvectors=False, enable_recycling=False):
        raise NotImplementedError('eigenvalue computation is not provided by the HIP backend')
