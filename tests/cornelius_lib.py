"""ctypes binding of the cornelius restatement (oracle/liboracle.so) and, when
it has been built from the reference tree, of the reference's own cornelius
(oracle/_ref/libref_cornelius.so).  Test infrastructure."""
import ctypes as C
import os

import numpy as np

import oracle_lib as O

REF_PATH = os.path.join(O.ORACLE_DIR, "_ref", "libref_cornelius.so")
_ref = None


def _bind(fn):
    dp = C.POINTER(C.c_double)
    fn.argtypes = [C.c_int, dp, dp, C.c_double, dp, dp]
    fn.restype = C.c_int
    return fn


def _call(fn, dim, corners, dx, v0):
    v = np.ascontiguousarray(np.asarray(corners, dtype=np.float64).ravel())
    d = np.ascontiguousarray(np.asarray(dx, dtype=np.float64))
    cen = np.zeros(10 * dim)
    nrm = np.zeros(10 * dim)
    n = fn(dim, O._ptr(v), O._ptr(d), v0, O._ptr(cen), O._ptr(nrm))
    return cen[:n * dim].reshape(n, dim), nrm[:n * dim].reshape(n, dim)


def find(dim, corners, dx, v0):
    """Elements of one cell from the restatement: (centroids, normals)."""
    return _call(_bind(O.lib().cfo_find), dim, corners, dx, v0)


def ref_available():
    return os.path.exists(REF_PATH)


def ref_find(dim, corners, dx, v0):
    global _ref
    if _ref is None:
        _ref = _bind(C.CDLL(REF_PATH).ref_cornelius_find)
    return _call(_ref, dim, corners, dx, v0)


def glibc_rand_sequence(n):
    """glibc rand() from the default seed (1), as double in [0,1]
    (rand()/RAND_MAX), the input generator of the reference's
    example_cornelius.cpp."""
    libc = C.CDLL("libc.so.6")
    libc.srand(1)
    libc.rand.restype = C.c_int
    out = np.empty(n)
    for i in range(n):
        out[i] = libc.rand() / 2147483647.0
    return out
