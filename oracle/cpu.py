"""ctypes binding of oracle/libcpu_fem.so (cpu_fem.c), the multi-threaded CPU
restatement timed as the benchmark's CPU baseline.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

_LIB = None
_P = C.c_void_p
_I64 = C.c_int64


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libcpu_fem.so')
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `make -C oracle`")
        L = C.CDLL(path)
        L.cfem_threads.restype = C.c_int
        L.cfem_csr_pattern.restype = _I64
        L.cfem_csr_pattern.argtypes = [_I64, _I64, _P, C.c_int32, _P, _P]
        L.cfem_assemble_elastic.restype = None
        L.cfem_assemble_elastic.argtypes = [_I64, _P, _P, C.c_double, C.c_double, _P, _P, _P, _P]
        L.cfem_zero_rows.restype = None
        L.cfem_zero_rows.argtypes = [_P, _P, _P, _I64, _P]
        for name in ('cfem_bicgstab', 'cfem_pcg'):
            f = getattr(L, name)
            f.restype = _I64
            f.argtypes = [_I64, _P, _P, _P, _P, _P, C.c_double, C.c_double, _I64, _I64]
        L.cfem_spmv.restype = None
        L.cfem_spmv.argtypes = [_I64, _P, _P, _P, _P, _P]
        _LIB = L
    return _LIB


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class CsrMatrix:
    """Scalar CSR (int64 row pointers, int32 sorted columns, float64 values)."""

    def __init__(self, n, rowptr, col, vals):
        self.n, self.rowptr, self.col, self.vals = n, rowptr, col, vals

    def scipy(self):
        return sp.csr_matrix((self.vals, self.col, self.rowptr), shape=(self.n, self.n))


def elastic_pattern(points, cells, vec=3):
    L = lib()
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    Nn = len(points)
    nnz = L.cfem_csr_pattern(Nn, len(cells), _p(cells), vec, None, None)
    if nnz < 0:
        raise RuntimeError(f"cfem_csr_pattern failed ({nnz})")
    rowptr = np.empty(Nn * vec + 1, dtype=np.int64)
    col = np.empty(nnz, dtype=np.int32)
    L.cfem_csr_pattern(Nn, len(cells), _p(cells), vec, _p(rowptr), _p(col))
    return CsrMatrix(Nn * vec, rowptr, col, np.zeros(nnz))


def assemble_elastic(A, points, cells, lmbda, mu, E=None, bc_dofs=None):
    """A.vals = the elastic tangent (per-cell modulus scale E), then zeroRows(bc_dofs, 1)."""
    L = lib()
    A.vals[:] = 0.
    pts = np.ascontiguousarray(points, dtype=np.float64)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    Ev = None if E is None else np.ascontiguousarray(E, dtype=np.float64)
    L.cfem_assemble_elastic(len(cells), _p(pts), _p(cells), float(lmbda), float(mu), _p(Ev), _p(A.rowptr), _p(A.col),
                            _p(A.vals))
    if bc_dofs is not None and len(bc_dofs):
        rows = np.ascontiguousarray(bc_dofs, dtype=np.int64)
        L.cfem_zero_rows(_p(A.rowptr), _p(A.col), _p(A.vals), len(rows), _p(rows))
    return A


def solve(A, b, x0, method='bicgstab', tol=1e-10, atol=1e-10, maxiter=10000, fixed=0):
    """(x, iterations); fixed > 0 runs exactly that many iterations (a timing block)."""
    L = lib()
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64)
    f = L.cfem_bicgstab if method == 'bicgstab' else L.cfem_pcg
    k = f(A.n, _p(A.rowptr), _p(A.col), _p(A.vals), _p(b), _p(x), tol, atol, maxiter, fixed)
    return x, int(k)


def threads():
    return int(lib().cfem_threads())
