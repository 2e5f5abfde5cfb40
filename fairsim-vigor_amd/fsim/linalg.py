"""Device vectors behind fairSIM's vector API (org.fairsim.linalg).

``AccelVectorFactory`` plays the role of the reference's accelerated
``VectorFactory`` slot (linalg/VectorFactory.java:28-171,
accel/AccelVectorFactory.java:31-149); ``Vec2dCplx`` / ``Vec2dReal`` mirror
``Vec2d.Cplx`` / ``Vec2d.Real`` (linalg/Vec2d.java:105-220).  Every method
is one call into libfsim; operations are in place with ``self`` as the
output, sizes are checked and a mismatch raises ``FsimError`` (the
reference's RuntimeException from Vec.failSize, linalg/Vec.java:100-149).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, check, lib


class _Vec:
    _cplx = False

    def __init__(self, w, h):
        _lib.init()
        self._h = C.c_void_p()
        call("fsim_vec_create", 1 if self._cplx else 0, int(w), int(h), C.byref(self._h))
        self.w, self.h = int(w), int(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            lib().fsim_vec_destroy(h)
            self._h = None

    # --- shape ---
    def vectorWidth(self):
        return self.w

    def vectorHeight(self):
        return self.h

    def vectorSize(self):
        return self.w * self.h

    @property
    def handle(self):
        return self._h

    def device_ptr(self):
        return lib().fsim_vec_device_ptr(self._h)

    # --- host mirror ---
    def _host_shape(self):
        return (self.h, self.w, 2) if self._cplx else (self.h, self.w)

    def vectorData(self):
        """Coherent host copy, interleaved complex (Vec.vectorData)."""
        out = np.empty(self._host_shape(), dtype=np.float32)
        call("fsim_vec_download", self._h, out.ctypes.data_as(C.c_void_p))
        return out

    def setData(self, arr):
        """Write host data and sync it to the device (vectorData + syncBuffer)."""
        a = np.ascontiguousarray(arr, dtype=np.float32)
        if a.size != int(np.prod(self._host_shape())):
            raise _lib.FsimError(_lib.FSIM_ERR_ARG, "setData: size mismatch")
        call("fsim_vec_upload", self._h, a.ctypes.data_as(C.c_void_p))

    # --- arithmetic (Vec.java:155-395) ---
    def zero(self):
        call("fsim_vec_zero", self._h)

    def copy(self, other):
        call("fsim_vec_copy", self._h, other._h)

    def add(self, *others):
        for o in others:
            call("fsim_vec_add", self._h, o._h)

    def axpy(self, a, x):
        if isinstance(a, complex):
            call("fsim_vec_axpy", self._h, float(a.real), float(a.imag), x._h)
        else:
            call("fsim_vec_axpy", self._h, float(a), 0.0, x._h)

    def times(self, x):
        call("fsim_vec_times", self._h, x._h, 0)

    def norm2(self):
        r = C.c_double()
        call("fsim_vec_norm2", self._h, C.byref(r))
        return r.value

    def setFrom16bitPixels(self, pixels):
        p = np.ascontiguousarray(pixels, dtype=np.uint16)
        if p.size != self.w * self.h:
            raise _lib.FsimError(_lib.FSIM_ERR_ARG,
                                 f"Short array to vector size mismatch {self.w}/{self.h}/{p.size}")
        call("fsim_vec_set_u16", self._h, p.ctypes.data_as(C.c_void_p))

    def setFrom16bitPixelsDevice(self, dev_ptr):
        call("fsim_vec_set_u16_device", self._h, C.c_void_p(dev_ptr))

    def duplicate(self):
        d = type(self)(self.w, self.h)
        d.copy(self)
        return d


class Vec2dReal(_Vec):
    """Vec2d.Real on the device."""
    _cplx = False

    def addConst(self, a):
        call("fsim_vec_add_const", self._h, float(a), 0.0)

    def scal(self, a):
        call("fsim_vec_scal", self._h, float(a), 0.0)

    def sumElements(self):
        re, im = C.c_double(), C.c_double()
        call("fsim_vec_sum", self._h, C.byref(re), C.byref(im))
        return re.value

    def get(self, x, y):
        re, im = C.c_float(), C.c_float()
        call("fsim_vec_get", self._h, int(x), int(y), C.byref(re), C.byref(im))
        return re.value

    def set(self, x, y, v):
        call("fsim_vec_set", self._h, int(x), int(y), float(v), 0.0)


class Vec2dCplx(_Vec):
    """Vec2d.Cplx on the device."""
    _cplx = True

    def addConst(self, a):
        a = complex(a)
        call("fsim_vec_add_const", self._h, a.real, a.imag)

    def scal(self, a):
        a = complex(a)
        call("fsim_vec_scal", self._h, a.real, a.imag)

    def timesConj(self, x):
        call("fsim_vec_times", self._h, x._h, 1)

    def sumElements(self):
        re, im = C.c_double(), C.c_double()
        call("fsim_vec_sum", self._h, C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def fft2d(self, inverse):
        call("fsim_vec_fft2d", self._h, 1 if inverse else 0)

    def fourierShift(self, kx, ky):
        call("fsim_vec_fourier_shift", self._h, float(kx), float(ky))

    def pasteFreq(self, inv, xOffset=0, yOffset=0):
        call("fsim_vec_paste_freq", self._h, inv._h, int(xOffset), int(yOffset))

    def get(self, x, y):
        re, im = C.c_float(), C.c_float()
        call("fsim_vec_get", self._h, int(x), int(y), C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def set(self, x, y, v):
        v = complex(v)
        call("fsim_vec_set", self._h, int(x), int(y), v.real, v.imag)

    def to_numpy(self):
        d = self.vectorData()
        return d[..., 0] + 1j * d[..., 1]

    def from_numpy(self, z):
        z = np.asarray(z)
        a = np.empty(z.shape + (2,), dtype=np.float32)
        a[..., 0] = z.real
        a[..., 1] = z.imag
        self.setData(a)


class AccelVectorFactory:
    """The accelerated VectorFactory (accel/AccelVectorFactory.java:31-149)."""

    _instance = None

    @classmethod
    def getFactory(cls):
        if cls._instance is None:
            cls._instance = cls()
        return cls._instance

    def __init__(self, device=0):
        _lib.init(device)

    def createReal2D(self, w, h):
        return Vec2dReal(w, h)

    def createCplx2D(self, w, h):
        return Vec2dCplx(w, h)

    def createArrayCplx2D(self, *dims):
        """createArrayCplx2D(n0, [n1, [n2,]] w, h) -> nested lists."""
        *outer, w, h = dims
        if not outer:
            return self.createCplx2D(w, h)
        return [self.createArrayCplx2D(*outer[1:], w, h) for _ in range(outer[0])]

    def createArrayReal2D(self, *dims):
        *outer, w, h = dims
        if not outer:
            return self.createReal2D(w, h)
        return [self.createArrayReal2D(*outer[1:], w, h) for _ in range(outer[0])]

    @staticmethod
    def syncConcurrent():
        call("fsim_sync")


class DeviceBuffer:
    """Raw device allocation (AccelVectorFactory.nativeAllocMemory)."""

    def __init__(self, nbytes):
        _lib.init()
        self.ptr = C.c_void_p()
        call("fsim_alloc_device", int(nbytes), C.byref(self.ptr))
        self.nbytes = int(nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise _lib.FsimError(_lib.FSIM_ERR_ARG, "upload larger than buffer")
        call("fsim_memcpy", self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 1)

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise _lib.FsimError(_lib.FSIM_ERR_ARG, "download larger than buffer")
        call("fsim_memcpy", out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, 2)
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr is not None and self.ptr.value:
            check(lib().fsim_free_device(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if _lib._lib is not None:
                self.free()
        except Exception:
            pass
