"""lduMatrix of the Vm equation on an arbitrary mesh, solved on the device.

Mirrors what fvMatrix::solve hands to lduMatrix::solve for Vm (monodomainSolver.C:129-135):
diag(), upper() (== lower(), symmetric) and the mesh's lduAddressing (lowerAddr,
upperAddr), with the PCG controls of the fvSolution entry (tolerance, relTol, maxIter) and
OpenFOAM's normalised residual.  The addressing becomes ELL rows on the device once.
"""
import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib


def _arg(a, dtype):
    """(pointer, flags, keep-alive) for a numpy array or a contiguous device tensor"""
    if isinstance(a, np.ndarray):
        a2 = np.ascontiguousarray(a, dtype=dtype)
        return C.c_void_p(a2.ctypes.data), 0, a2
    if a.dtype.itemsize != np.dtype(dtype).itemsize or not a.is_contiguous():
        raise TypeError("device tensors must be contiguous %s" % np.dtype(dtype).name)
    return C.c_void_p(a.data_ptr()), _capi.DEVICE_PTRS if a.is_cuda else 0, a


class LduMatrix:
    def __init__(self, lower_addr, upper_addr, n_cells):
        lo = np.ascontiguousarray(lower_addr, dtype=np.int32)
        up = np.ascontiguousarray(upper_addr, dtype=np.int32)
        if lo.shape != up.shape:
            raise ValueError("lowerAddr and upperAddr differ in length")
        h = C.c_void_p()
        check(lib().cg_ldu_create(int(n_cells), int(lo.size),
                                  lo.ctypes.data_as(C.POINTER(C.c_int32)),
                                  up.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h)))
        self.h = h
        self.n = int(n_cells)
        self.n_faces = int(lo.size)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().cg_ldu_destroy(h)
            except Exception:
                pass
            self.h = None

    @property
    def width(self):
        """ELL slots per row"""
        return lib().cg_ldu_width(self.h)

    def set_coeffs(self, diag, upper):
        pd, fd, kd = _arg(diag, np.float64)
        pu, fu, ku = _arg(upper, np.float64)
        if fd != fu:
            raise TypeError("diag and upper must both be host arrays or both device tensors")
        check(lib().cg_ldu_set_coeffs(self.h, pd, pu, fd))

    def Amul(self, x, y=None):
        if y is None:
            y = np.empty(self.n)
        px, fx, kx = _arg(x, np.float64)
        py, fy, ky = _arg(y, np.float64)
        if fx != fy:
            raise TypeError("x and y must live on the same side")
        check(lib().cg_ldu_amul(self.h, px, py, fx))
        return y

    def amul_kernel_ms(self, x, y):
        """the SpMV kernel alone (device tensors), milliseconds"""
        ms = C.c_float()
        check(lib().cg_ldu_amul_timed(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                      C.byref(ms)))
        return ms.value

    def solve(self, psi, source, tolerance=1e-6, relTol=0.0, maxIter=1000):
        """PCG from psi (updated in place); returns the solver performance as a dict"""
        pp, fp, kp = _arg(psi, np.float64)
        if fp == 0 and kp is not psi:
            raise TypeError("psi must be a contiguous float64 array (it is updated in place)")
        ps, fs, ks = _arg(source, np.float64)
        if fp != fs:
            raise TypeError("psi and source must live on the same side")
        it = C.c_int()
        r0, r1 = C.c_double(), C.c_double()
        check(lib().cg_ldu_solve(self.h, pp, ps, float(tolerance), float(relTol), int(maxIter),
                                 fp, C.byref(it), C.byref(r0), C.byref(r1)))
        return {"nIterations": it.value, "initialResidual": r0.value,
                "finalResidual": r1.value}
