"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the CPU-baseline leg of bench.py import
this module.  It restates heFFTe's single-rank 3D FFT semantics in FP64; see
heffte_oracle.h for the reference file:line each routine follows.

Arrays are numpy: complex128 for complex data, float64 for real data, stored
flat in the order of their box (box3d order = (fast, mid, slow) axes).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HEFFTE_ORACLE_LIBRARY selects another build of the oracle (the sanitized one of heffte_amd/sanitize.py)
_LIB_PATH = os.environ.get("HEFFTE_ORACLE_LIBRARY") or os.path.join(_HERE, "liboracle.so")
_lib = None


class Box(ctypes.Structure):
    """Mirror of orc_box: inclusive low/high and the (fast, mid, slow) order."""
    _fields_ = [("low", ctypes.c_longlong * 3), ("high", ctypes.c_longlong * 3), ("order", ctypes.c_int * 3)]

    @classmethod
    def make(cls, low, high, order=(0, 1, 2)):
        b = cls()
        for d in range(3):
            b.low[d] = int(low[d])
            b.high[d] = int(high[d])
            b.order[d] = int(order[d])
        return b

    @property
    def size(self):
        return tuple(self.high[d] - self.low[d] + 1 for d in range(3))

    def count(self):
        s = self.size
        return 0 if min(s) <= 0 else s[0] * s[1] * s[2]

    def r2c(self, dim):
        """heffte_geometry.h:96-104: keep size/2 + 1 entries along dim."""
        high = list(self.high)
        high[dim] = self.low[dim] + self.size[dim] // 2
        return Box.make(self.low, high, self.order)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(Box)
        dp = ctypes.c_void_p
        L.orc_box_count.argtypes = [P]
        L.orc_box_count.restype = ctypes.c_longlong
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_transpose.argtypes = [P, P, ctypes.c_int, dp, dp]
        L.orc_dft1d_ld.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_fft1d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_fft3d_c2c.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_fft3d_real_c2c.argtypes = [P, P, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_fft3d_r2c.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_fft3d_c2r.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_r2r_cos.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_r2r_sin.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_r2r_cos1.argtypes = [P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]
        L.orc_r2r_1d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        for f in ("orc_transpose", "orc_fft3d_c2c", "orc_fft3d_real_c2c", "orc_fft3d_r2c",
                  "orc_fft3d_c2r", "orc_r2r_cos", "orc_r2r_sin", "orc_r2r_cos1"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(rc):
    if rc != 0:
        raise ValueError("oracle rejected the boxes")


def set_threads(n):
    return lib().orc_set_threads(int(n))


def transpose(inbox, outbox, data):
    data = np.ascontiguousarray(data)
    out = np.empty_like(data)
    _check(lib().orc_transpose(ctypes.byref(inbox), ctypes.byref(outbox), data.itemsize, _ptr(data), _ptr(out)))
    return out


def dft1d_ld(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().orc_dft1d_ld(x.size, sign, _ptr(x), _ptr(y))
    return y


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().orc_fft1d(x.size, sign, _ptr(x), _ptr(y))
    return y


def fft3d_c2c(inbox, outbox, x, sign=-1, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty(batch * outbox.count(), dtype=np.complex128)
    _check(lib().orc_fft3d_c2c(ctypes.byref(inbox), ctypes.byref(outbox), sign, scale, batch, _ptr(x), _ptr(y)))
    return y


def convolve3d(inbox, outbox, x, y, scale=1.0, batch=1):
    """backward(forward(x) * forward(y)) * scale, x and y in the inbox layout: the sequence of
    benchmarks/convolution.cpp:86-95 (forward, pointwise product, backward) with the scale
    on the backward transform; scale = 1/N gives the circular convolution of x and y."""
    xh = fft3d_c2c(inbox, outbox, x, -1, 1.0, batch)
    yh = fft3d_c2c(inbox, outbox, y, -1, 1.0, batch)
    return fft3d_c2c(outbox, inbox, xh * yh, +1, scale, batch)


def fft3d_real_c2c(inbox, outbox, x, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(batch * outbox.count(), dtype=np.complex128)
    _check(lib().orc_fft3d_real_c2c(ctypes.byref(inbox), ctypes.byref(outbox), scale, batch, _ptr(x), _ptr(y)))
    return y


def fft3d_r2c(inbox, outbox, r2c_dir, x, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(batch * outbox.count(), dtype=np.complex128)
    _check(lib().orc_fft3d_r2c(ctypes.byref(inbox), ctypes.byref(outbox), r2c_dir, scale, batch, _ptr(x), _ptr(y)))
    return y


def fft3d_c2r(inbox, outbox, r2c_dir, x, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty(batch * inbox.count(), dtype=np.float64)
    _check(lib().orc_fft3d_c2r(ctypes.byref(inbox), ctypes.byref(outbox), r2c_dir, scale, batch, _ptr(x), _ptr(y)))
    return y


def r2r_cos(inbox, outbox, x, sign=-1, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(batch * outbox.count(), dtype=np.float64)
    _check(lib().orc_r2r_cos(ctypes.byref(inbox), ctypes.byref(outbox), sign, scale, batch, _ptr(x), _ptr(y)))
    return y


def r2r_sin(inbox, outbox, x, sign=-1, scale=1.0, batch=1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(batch * outbox.count(), dtype=np.float64)
    _check(lib().orc_r2r_sin(ctypes.byref(inbox), ctypes.byref(outbox), sign, scale, batch, _ptr(x), _ptr(y)))
    return y


def r2r_cos1(inbox, outbox, x, sign=-1, scale=1.0, batch=1):
    """DCT-I (REDFT00) forward, 2 x REDFT00 backward; every extent >= 2."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(batch * outbox.count(), dtype=np.float64)
    _check(lib().orc_r2r_cos1(ctypes.byref(inbox), ctypes.byref(outbox), sign, scale, batch, _ptr(x), _ptr(y)))
    return y


def r2r_1d(x, kind=0, sign=-1):
    """kind 0 = cos (DCT-II fwd / 2 DCT-III bwd), 1 = sin, 2 = cos1 (REDFT00 fwd / 2 REDFT00 bwd)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().orc_r2r_1d(kind, sign, x.size, _ptr(x), _ptr(y))
    return y


def rel_l2(y, ref):
    y = np.asarray(y, dtype=np.complex128 if np.iscomplexobj(y) or np.iscomplexobj(ref) else np.float64)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((y.ravel() - ref.ravel()))
    return float(num / den) if den > 0 else float(num)
