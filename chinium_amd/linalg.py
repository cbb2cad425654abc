"""Dense FP64 linear algebra on the device for the SCF driver around the XC path:
rocBLAS (gemm) and rocSOLVER (symmetric eigensolver, Cholesky factor and solve)
from the ROCm installation the library is built against, called through ctypes
on device buffers, in the same HIP runtime as libchinium_xc. Plumbing for the
SCF loop (orbitals, densities, DIIS residuals); no kernels of this project.
Matrices are column-major."""
import ctypes
import os

import numpy as np

from . import hip

_OP_N, _OP_T = 111, 112          # rocblas_operation_none / _transpose
_FILL_UPPER, _FILL_LOWER = 121, 122
_EVECT_ORIGINAL = 211            # rocblas_evect_original

_rb = _rs = None
_handle = None


def _libs():
    global _rb, _rs, _handle
    if _rb is None:
        root = os.environ.get("ROCM_PATH", "/opt/rocm")
        rb = ctypes.CDLL(os.path.join(root, "lib", "librocblas.so"))
        rs = ctypes.CDLL(os.path.join(root, "lib", "librocsolver.so"))
        vp, i, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        rb.rocblas_create_handle.argtypes = [ctypes.POINTER(vp)]
        rb.rocblas_dgemm.argtypes = [vp, i, i, i, i, i, dp, vp, i, vp, i, dp, vp, i]
        rs.rocsolver_dsyevd.argtypes = [vp, i, i, i, vp, i, vp, vp, vp]
        rs.rocsolver_dpotrf.argtypes = [vp, i, i, vp, i, vp]
        rs.rocsolver_dpotrs.argtypes = [vp, i, i, i, vp, i, vp, i]
        h = vp()
        _ok(rb.rocblas_create_handle(ctypes.byref(h)), "rocblas_create_handle")
        _rb, _rs, _handle = rb, rs, h
    return _rb, _rs, _handle


def _ok(st, what):
    if st != 0:
        raise RuntimeError("%s failed (status %d)" % (what, st))


class DMat:
    """An r x c FP64 matrix in device memory, column-major."""

    def __init__(self, r, c=1):
        self.r, self.c = int(r), int(c)
        self.buf = hip.DeviceArray(self.r * self.c)
        self.ptr = self.buf.ptr

    @classmethod
    def of(cls, a):
        a = np.asarray(a, dtype=np.float64)
        m = cls(a.shape[0], a.shape[1] if a.ndim == 2 else 1)
        m.set(a)
        return m

    def set(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.size == self.r * self.c
        hip._ok(hip.hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * a.size, 1), "hipMemcpy H2D")
        return self

    def get(self):
        out = np.empty((self.r, self.c), dtype=np.float64, order="F")
        hip._ok(hip.hip().hipMemcpy(out.ctypes.data, self.ptr, 8 * out.size, 2), "hipMemcpy D2H")
        return out if self.c > 1 else out[:, 0]


def gemm(A, B, C, ta=False, tb=False, alpha=1.0, beta=0.0, m=None, n=None, k=None):
    """C[:m, :n] = alpha op(A) op(B) + beta C, op = transpose where ta / tb; k defaults to the shared size."""
    rb, _, h = _libs()
    m = m if m is not None else (A.c if ta else A.r)
    n = n if n is not None else (B.r if tb else B.c)
    k = k if k is not None else (A.r if ta else A.c)
    a, b = ctypes.c_double(alpha), ctypes.c_double(beta)
    _ok(rb.rocblas_dgemm(h, _OP_T if ta else _OP_N, _OP_T if tb else _OP_N, m, n, k, ctypes.byref(a), A.ptr, A.r,
                         B.ptr, B.r, ctypes.byref(b), C.ptr, C.r), "rocblas_dgemm")
    return C


def syevd(A, w, work, info):
    """Eigenvalues (ascending) into w and eigenvectors over A (symmetric, lower triangle read)."""
    _, rs, h = _libs()
    _ok(rs.rocsolver_dsyevd(h, _EVECT_ORIGINAL, _FILL_LOWER, A.r, A.ptr, A.r, w.ptr, work.ptr, info.ptr),
        "rocsolver_dsyevd")
    hip.device_synchronize()
    code = int(np.frombuffer(np.asarray(info.get(), dtype=np.float64).tobytes()[:4], dtype=np.int32)[0])
    if code != 0:
        raise RuntimeError("rocsolver_dsyevd: info %d" % code)


def potrf(A, info):
    """Cholesky factor A = L L^T in place (lower triangle)."""
    _, rs, h = _libs()
    _ok(rs.rocsolver_dpotrf(h, _FILL_LOWER, A.r, A.ptr, A.r, info.ptr), "rocsolver_dpotrf")
    hip.device_synchronize()
    code = int(np.frombuffer(np.asarray(info.get(), dtype=np.float64).tobytes()[:4], dtype=np.int32)[0])
    if code != 0:
        raise RuntimeError("rocsolver_dpotrf: info %d (the matrix is not positive definite)" % code)


def potrs(L, B):
    """B <- A^-1 B with the factor of potrf."""
    _, rs, h = _libs()
    _ok(rs.rocsolver_dpotrs(h, _FILL_LOWER, L.r, B.c, L.ptr, L.r, B.ptr, B.r), "rocsolver_dpotrs")
