"""simox_cell_sort_noalloc (the allocation-free CELL_SORT): the same cell
occupancy and link order as simox_cell_sort + simox_get_cells, bit for bit,
on two calls in a row (the second with fewer atoms, so stale scratch of the
first would show), through the C-ABI on device buffers."""
import ctypes as C

import numpy as np
import pytest

from simox_amd import MODE_FP64, Simox, api
from simox_amd.systems import config, tiny

pytestmark = pytest.mark.gpu


def context_cells(s):
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    g.set_state(s.ids, s.x, s.v, None, s.q)  # sorts into cells
    npt, link = g.get_cells(link_stride=500)
    g.close()
    return npt, np.concatenate([link[c, :npt[c]] for c in range(len(npt))])


class Hip:
    """The caller's device buffers, from the HIP runtime the library itself
    links (torch bundles another HIP runtime, which cannot share the device
    with this one in one process)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        self.bufs.append(p)
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def noalloc_cells(s, hip):
    nc = [int(b / s.rcut) for b in s.box]
    ncall = nc[0] * nc[1] * nc[2]
    fill = -7
    x = hip.put(np.asarray(s.x, dtype=np.float64))
    npt0 = np.full(ncall, fill, np.int32)
    start0 = np.full(ncall + 1, fill, np.int32)
    link0 = np.full(s.n, fill, np.int32)
    npt, start, link = hip.put(npt0), hip.put(start0), hip.put(link0)
    api.cell_sort_noalloc((x, s.n), s.box, s.rcut, npt, start, link)
    npt, start, link = hip.get(npt, npt0), hip.get(start, start0), hip.get(link, link0)
    assert start[0] == 0 and np.array_equal(np.diff(start), npt) and start[-1] == s.n
    return npt, link


def test_noalloc_cell_sort_equals_context_sort():
    big = config("c2")                     # 77,824 atoms, 9 x 9 x 14 cells
    small = tiny(seed=4, nx=5, ny=5, nz=4)  # a few hundred atoms, fewer cells
    assert small.n < big.n
    hip = Hip()
    try:
        for s in (big, small):
            npt_c, link_c = context_cells(s)
            npt_n, link_n = noalloc_cells(s, hip)
            np.testing.assert_array_equal(npt_n, npt_c)
            np.testing.assert_array_equal(link_n, link_c)
    finally:
        hip.free()


def test_noalloc_cell_sort_rejects_oversize():
    s = tiny(seed=4)
    # n beyond the scratch's 2^21 atoms: refused before any device access
    with pytest.raises(api.SimoxError) as e:
        api.cell_sort_noalloc((1, (1 << 21) + 1), s.box, s.rcut, 1, 1, 1)
    assert e.value.code == 5  # SIMOX_ERR_ARG
