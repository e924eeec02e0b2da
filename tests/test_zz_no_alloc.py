"""gvh_cornelius_cells_noalloc (device buffers, static scratch, no
allocation) against gvh_cornelius_cells on the same cells: identical counts,
centroids and normals, on two calls in a row, the second with fewer cells
than the first (the scratch keeps the first call's contents and its overflow
counter must start again at zero).  Also the size limit and argument errors.
"""
import ctypes as C

import numpy as np
import pytest

import cornelius_lib as CL
import gpuvh

pytestmark = pytest.mark.gpu


def say(*a):
    print("no-alloc:", *a, flush=True)


def cells_for(dim, n, seed):
    """Random and iso-degenerate cells of one dimension."""
    _, cu, hc = CL.example_inputs()
    base = cu if dim == 3 else hc
    rng = np.random.default_rng(seed)
    pick = base[rng.integers(0, len(base), n - n // 4)]
    return np.concatenate([pick, CL.degenerate_cells(dim, n // 4, seed)])


class Hip:
    """The HIP runtime that the library itself uses (libamdhip64.so.7 is
    already loaded with it), for device buffers of the test.  torch's own
    bundled runtime is not used here: in a process where the library's
    runtime started first it finds no device."""

    def __init__(self):
        gpuvh.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        for f in (self.L.hipMalloc, self.L.hipFree, self.L.hipMemcpy, self.L.hipDeviceSynchronize):
            f.restype = C.c_int

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ps):
        for p in ps:
            assert self.L.hipFree(p) == 0


def noalloc(dim, corners, dx, v0, out_fill=0.0):
    hip = Hip()
    n = corners.shape[0]
    counts0 = np.full(n, -1, dtype=np.int32)
    out0 = np.full((n, 10, dim), out_fill, dtype=np.float64)
    c = hip.upload(np.asarray(corners, dtype=np.float64))
    counts = hip.upload(counts0)
    cen = hip.upload(out0)
    nrm = hip.upload(out0)
    d = np.ascontiguousarray(np.asarray(dx, dtype=np.float64))
    rc = gpuvh.lib().gvh_cornelius_cells_noalloc(dim, n, c, d.ctypes.data_as(C.POINTER(C.c_double)),
                                                v0, counts, cen, nrm, None)
    res = (rc, hip.download(counts, counts0), hip.download(cen, out0), hip.download(nrm, out0))
    hip.free(c, counts, cen, nrm)
    return res


def assert_same(got, ref, dim):
    counts, cen, nrm = got
    np.testing.assert_array_equal(counts, ref[0])
    mask = np.arange(10)[None, :] < ref[0][:, None]
    for k, arr in ((1, cen), (2, nrm)):
        a, b = arr[mask], ref[k][mask]
        assert a.shape == b.shape
        diff = float(np.abs(a - b).max()) if a.size else 0.0
        assert diff == 0.0, "max difference %g" % diff


@pytest.mark.parametrize("dim", [3, 4])
def test_noalloc_matches_allocating_entry_point(dim):
    dx = np.array([0.01, 0.1, 0.2, 0.15][:dim])
    for call, n in enumerate((9000, 2500)):
        cells = cells_for(dim, n, 40 + call)
        say("dim", dim, "call", call, "cells", n)
        ref = gpuvh.cornelius_cells(dim, cells, dx, 0.5)
        rc, counts, cen, nrm = noalloc(dim, cells, dx, 0.5, out_fill=np.nan)
        assert rc == 0, gpuvh.lib().gvh_last_error()
        assert_same((counts, cen, nrm), ref, dim)
        # slots past each cell's count are left as the caller had them
        mask = np.arange(10)[None, :] >= counts[:, None]
        assert np.isnan(cen[mask]).all() and np.isnan(nrm[mask]).all()
        assert counts.sum() > 0


def test_noalloc_limit_and_arguments():
    lim = 16384
    n = lim + 1
    say("limit", n)
    dx = np.ones(4)
    # rejected on its size before any device pointer is used
    rc = gpuvh.lib().gvh_cornelius_cells_noalloc(4, n, C.c_void_p(1),
                                                dx.ctypes.data_as(C.POINTER(C.c_double)), 0.5,
                                                C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None)
    assert rc == -1
    assert b"at most" in gpuvh.lib().gvh_last_error()
    rc = gpuvh.lib().gvh_cornelius_cells_noalloc(5, 10, None, None, 0.5, None, None, None, None)
    assert rc == -1
    rc = gpuvh.lib().gvh_cornelius_cells_noalloc(3, 0, None, None, 0.5, None, None, None, None)
    assert rc == 0
