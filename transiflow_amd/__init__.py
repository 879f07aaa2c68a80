"""MI355X-native compute backend for TransiFlow's continuation hot path.

``Interface(parameters, nx, ny, nz=1, dim=None, dof=None,
boundary_conditions=None, backend='HIP')`` mirrors ``transiflow.Interface``
(``transiflow/interface/create.py:4-68``) and returns the HIP backend, whose
``rhs``, ``jacobian``, ``mass_matrix`` and ``solve`` run as hand-written HIP
kernels on gfx950 behind the C ABI of ``include/transiflow_hip.h``.
``Continuation`` and ``TimeIntegration`` restate ``transiflow.Continuation``
and ``transiflow.TimeIntegration``.
"""

from .continuation import Continuation, norm
from .time_integration import TimeIntegration

__all__ = ['Interface', 'create', 'Continuation', 'TimeIntegration', 'norm']


def create(parameters, nx, ny, nz=1, dim=None, dof=None, boundary_conditions=None, backend='HIP'):
    """Backend factory (create.py:4-68).  Only the HIP backend ships here;
    the reference's other backend names raise ImportError, which the
    reference's tests turn into a skip."""
    if backend.lower() in ('hip', 'mi355x', 'gpu'):
        from .interface import Interface as HipInterface
        return HipInterface(parameters, nx, ny, nz, dim, dof, boundary_conditions=boundary_conditions)
    raise ImportError("backend '%s' is not part of transiflow_amd" % backend)


Interface = create
