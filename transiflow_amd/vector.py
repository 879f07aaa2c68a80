"""Device-resident FP64 state vectors and CSR matrices.

:class:`DeviceVector` gives a state vector in HBM the numpy-like protocol the
continuation driver uses (Continuation.py:98-357): ``size``, ``shape``,
``dot``, ``+``, ``-``, unary ``-``, scalar ``*`` and ``/``, in-place ``+=``,
``-=``, ``*=``, ``/=``, and ``.T.conj() @ y`` for ``utils.norm``.  Every
operation is one of the backend's HIP kernels on the current stream;
``dot`` and ``norm`` return host floats (one synchronisation each).  PyTorch
only allocates the buffers.

:class:`DeviceCsrMatrix` is what ``Interface.jacobian`` returns: the CSR
arrays (int32 indices, FP64 values) in HBM, with the preconditioner cached on
the object like the reference caches its LU on the matrix (SciPy.py:134-152).
"""

import ctypes
import threading
import numbers

import numpy

from . import _lib

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        if not t.cuda.is_available():
            raise _lib.BackendError('the HIP backend needs a GPU (torch.cuda.is_available() is False)')
        if _lib.load().tf_device_count() < 1:
            raise _lib.BackendError('libtransiflow_hip.so sees no GPU: a second HIP runtime was loaded '
                                    'before PyTorch\'s')
        _torch = t
    return _torch


def stream():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class _Workspace:
    """Scratch for reductions, one per thread and device: threads that run
    instances on their own streams never share it."""
    _local = threading.local()

    @classmethod
    def get(cls):
        T = torch()
        dev = T.cuda.current_device()
        bufs = getattr(cls._local, 'bufs', None)
        if bufs is None:
            bufs = cls._local.bufs = {}
        buf = bufs.get(dev)
        if buf is None:
            nwork = _lib.load().tf_reduce_work_size()
            buf = T.empty(nwork + 8, dtype=T.float64, device='cuda')
            bufs[dev] = buf
        return buf


def device_dot(x, y, norm=False):
    """Deterministic FP64 dot product (or 2-norm) of two device tensors."""
    ws = _Workspace.get()
    out = ctypes.c_double()
    n = x.numel()
    _lib.call('tf_dot_f64', n, ptr(x), ptr(y), ptr(ws), ctypes.c_void_p(ws.data_ptr() + 8 * (ws.numel() - 1)),
              ctypes.byref(out), 1 if norm else 0, stream())
    return out.value


class DeviceVector:
    __array_priority__ = 1000

    def __init__(self, data):
        self.data = data

    @classmethod
    def from_array(cls, array):
        T = torch()
        a = numpy.ascontiguousarray(numpy.asarray(array, dtype=numpy.float64))
        return cls(T.from_numpy(a).to('cuda'))

    @classmethod
    def zeros(cls, n):
        T = torch()
        return cls(T.zeros(n, dtype=T.float64, device='cuda'))

    # shape protocol
    @property
    def size(self):
        return self.data.numel()

    @property
    def shape(self):
        return (self.data.numel(),)

    @property
    def ndim(self):
        return 1

    @property
    def dtype(self):
        return numpy.dtype(numpy.float64)

    def __len__(self):
        return self.data.numel()

    @property
    def T(self):
        return self

    def conj(self):
        return self

    def copy(self):
        return DeviceVector(self.data.clone())

    def numpy(self):
        return self.data.cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a if dtype is None else a.astype(dtype)

    # arithmetic
    def _axpby(self, alpha, other, beta, out=None):
        out = torch().empty_like(self.data) if out is None else out
        _lib.call('tf_axpby_f64', self.data.numel(), float(alpha), ptr(self.data), float(beta),
                  ptr(other.data), ptr(out), stream())
        return out

    def _scale(self, alpha, out=None):
        out = torch().empty_like(self.data) if out is None else out
        _lib.call('tf_scale_f64', self.data.numel(), float(alpha), ptr(self.data), ptr(out), stream())
        return out

    def __add__(self, other):
        if isinstance(other, DeviceVector):
            return DeviceVector(self._axpby(1.0, other, 1.0))
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, DeviceVector):
            return DeviceVector(self._axpby(1.0, other, -1.0))
        return NotImplemented

    def __neg__(self):
        return DeviceVector(self._scale(-1.0))

    def __mul__(self, alpha):
        if isinstance(alpha, numbers.Real):
            return DeviceVector(self._scale(alpha))
        return NotImplemented

    __rmul__ = __mul__

    def _div(self, alpha, out=None):
        out = torch().empty_like(self.data) if out is None else out
        _lib.call('tf_div_f64', self.data.numel(), ptr(self.data), float(alpha), ptr(out), stream())
        return out

    def __truediv__(self, alpha):
        if isinstance(alpha, numbers.Real):
            return DeviceVector(self._div(alpha))
        return NotImplemented

    def __iadd__(self, other):
        if isinstance(other, DeviceVector):
            self._axpby(1.0, other, 1.0, out=self.data)
            return self
        return NotImplemented

    def __isub__(self, other):
        if isinstance(other, DeviceVector):
            self._axpby(1.0, other, -1.0, out=self.data)
            return self
        return NotImplemented

    def __imul__(self, alpha):
        self._scale(alpha, out=self.data)
        return self

    def __itruediv__(self, alpha):
        self._div(alpha, out=self.data)
        return self

    def dot(self, other):
        return device_dot(self.data, other.data)

    def __matmul__(self, other):
        if isinstance(other, DeviceVector):
            return self.dot(other)
        return NotImplemented

    def norm(self):
        return device_dot(self.data, self.data, norm=True)


class DeviceCsrMatrix:
    """CSR matrix in HBM (row_ptr/col int32, values FP64)."""

    def __init__(self, row_ptr, col, val, n):
        self.row_ptr = row_ptr
        self.col = col
        self.val = val
        self.n = n
        self.lu = None
        self.bordered_lu = False
        self.prec = None
        self.prec_exact = False
        self._val32 = None

    def values32(self):
        """FP32 copy of the values (cached; used by the mixed-precision inner solve)."""
        if self._val32 is None:
            T = torch()
            self._val32 = T.empty(self.val.numel(), dtype=T.float32, device='cuda')
            _lib.call('tf_f64_to_f32', self.val.numel(), ptr(self.val), ptr(self._val32), stream())
        return self._val32

    @property
    def shape(self):
        return (self.n, self.n)

    @property
    def dtype(self):
        return numpy.dtype(numpy.float64)

    @property
    def nnz(self):
        return self.col.numel()

    # reference CrsMatrix attribute names, as host copies
    @property
    def begA(self):
        return self.row_ptr.cpu().numpy().astype(numpy.int64)

    @property
    def jcoA(self):
        return self.col.cpu().numpy().astype(numpy.int64)

    @property
    def coA(self):
        return self.val.cpu().numpy()

    def to_scipy(self):
        from scipy import sparse
        return sparse.csr_matrix((self.coA, self.jcoA, self.begA), shape=self.shape)

    def matvec(self, x, out=None, gauge=-1):
        T = torch()
        out = T.empty(self.n, dtype=T.float64, device='cuda') if out is None else out
        _lib.call('tf_spmv_f64', self.n, ptr(self.row_ptr), ptr(self.col), ptr(self.val), ptr(x), ptr(out),
                  int(gauge), stream())
        return out

    def __matmul__(self, x):
        if isinstance(x, DeviceVector):
            return DeviceVector(self.matvec(x.data))
        return self.to_scipy() @ numpy.asarray(x)

    # -- scipy-like arithmetic (TimeIntegration.py:56: jacobian(x) - mass / (theta dt)) --
    def _combine(self, alpha, other, beta):
        T = torch()
        st = stream()
        cnt = T.empty(self.n, dtype=T.int32, device='cuda')
        rc = T.empty(self.n + 1, dtype=T.int32, device='cuda')
        nnz = ctypes.c_int()
        _lib.call('tf_csr_add_count', self.n, ptr(self.row_ptr), ptr(self.col), ptr(other.row_ptr), ptr(other.col),
                  ptr(cnt), ptr(rc), ctypes.byref(nnz), st)
        cc = T.empty(max(nnz.value, 1), dtype=T.int32, device='cuda')[:nnz.value]
        vc = T.empty(max(nnz.value, 1), dtype=T.float64, device='cuda')[:nnz.value]
        _lib.call('tf_csr_add_fill', self.n, ptr(self.row_ptr), ptr(self.col), ptr(self.val), float(alpha),
                  ptr(other.row_ptr), ptr(other.col), ptr(other.val), float(beta), ptr(rc), ptr(cc), ptr(vc), st)
        return DeviceCsrMatrix(rc, cc, vc, self.n)

    def _scaled(self, alpha):
        T = torch()
        v = T.empty_like(self.val)
        _lib.call('tf_scale_f64', self.val.numel(), float(alpha), ptr(self.val), ptr(v), stream())
        return DeviceCsrMatrix(self.row_ptr, self.col, v, self.n)

    def __add__(self, other):
        if isinstance(other, DeviceCsrMatrix):
            return self._combine(1.0, other, 1.0)
        if isinstance(other, ComplexCsrMatrix):
            return ComplexCsrMatrix(self._combine(1.0, other.real, 1.0), other.imag)
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, DeviceCsrMatrix):
            return self._combine(1.0, other, -1.0)
        if isinstance(other, ComplexCsrMatrix):
            return ComplexCsrMatrix(self._combine(1.0, other.real, -1.0), -other.imag)
        return NotImplemented

    def __mul__(self, alpha):
        if isinstance(alpha, numbers.Real):
            return self._scaled(alpha)
        if isinstance(alpha, numbers.Complex):
            a = complex(alpha)
            return ComplexCsrMatrix(self._scaled(a.real), self._scaled(a.imag))
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, alpha):
        if isinstance(alpha, numbers.Number):
            return self * (1.0 / alpha)
        return NotImplemented

    def __neg__(self):
        return self._scaled(-1.0)


class ComplexCsrMatrix:
    """A + i B for real device CSR matrices A, B of one pattern size: what the
    scipy arithmetic of the reference's eigenvalue glue builds from complex
    shifts, ``beta * J - alpha * M`` (JaDa.py:98-99), and what
    ``Interface.solve`` then solves (in the interleaved real form)."""

    def __init__(self, real, imag):
        self.real, self.imag = real, imag
        self.n = real.n

    @property
    def shape(self):
        return (self.n, self.n)

    @property
    def dtype(self):
        return numpy.dtype(numpy.complex128)

    @staticmethod
    def _parts(other):
        if isinstance(other, ComplexCsrMatrix):
            return other.real, other.imag
        if isinstance(other, DeviceCsrMatrix):
            return other, None
        return None

    def _combine(self, other, sign):
        p = self._parts(other)
        if p is None:
            return NotImplemented
        re = self.real._combine(1.0, p[0], sign)
        im = self.imag if p[1] is None else self.imag._combine(1.0, p[1], sign)
        return ComplexCsrMatrix(re, im)

    def __add__(self, other):
        return self._combine(other, 1.0)

    def __sub__(self, other):
        return self._combine(other, -1.0)

    def __radd__(self, other):
        return self._combine(other, 1.0)

    def __rsub__(self, other):
        p = self._parts(other)
        if p is None:
            return NotImplemented
        return (-self)._combine(other, 1.0)

    def __neg__(self):
        return ComplexCsrMatrix(-self.real, -self.imag)

    def __mul__(self, alpha):
        if not isinstance(alpha, numbers.Number):
            return NotImplemented
        a = complex(alpha)
        # (A + iB)(a + ib) = (a A - b B) + i (b A + a B)
        return ComplexCsrMatrix(self.real._combine(a.real, self.imag, -a.imag),
                                self.real._combine(a.imag, self.imag, a.real))

    __rmul__ = __mul__

    def __truediv__(self, alpha):
        if isinstance(alpha, numbers.Number):
            return self * (1.0 / alpha)
        return NotImplemented

    def to_scipy(self):
        return self.real.to_scipy() + 1j * self.imag.to_scipy()
