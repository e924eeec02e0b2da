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


def example_inputs():
    """Corner values of example_cornelius.cpp: 10000 squares (2x2), then 10000
    cubes (2x2x2), then 10000 hypercubes (2x2x2x2), all from one rand()
    stream, row-major."""
    seq = glibc_rand_sequence(10000 * (4 + 8 + 16))
    sq = seq[:40000].reshape(10000, 4)
    cu = seq[40000:120000].reshape(10000, 8)
    hc = seq[120000:].reshape(10000, 16)
    return sq, cu, hc


def format_elements(dim, counts, cen, nrm):
    """Elements printed as example_cornelius does (ostream default, 6
    significant digits), one line per element."""
    lines = []
    for i in range(len(counts)):
        for k in range(counts[i]):
            lines.append(" ".join("%g" % x for x in list(cen[i][k]) + list(nrm[i][k])) + " ")
    return lines


def degenerate_cells(dim, n, seed):
    """Cells whose corners sit on a few levels including the iso-value 0.5
    exactly (ties, corners on the surface, ambiguous faces)."""
    rng = np.random.default_rng(seed)
    return rng.choice([0.2, 0.5, 0.8, 0.5, 0.5, 0.35], size=(n, 1 << dim))
