"""GPU: cmx_meson_2pt_noalloc (device arrays, the library's fixed work space, no allocation) against
cmx_meson_2pt on the same propagators, bit for bit, on two calls in a row (the second on a smaller
lattice), and its refusal of a spatial volume above the work space's limit.

The test is the caller: it allocates the device arrays itself through the HIP runtime the library uses.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from chroma_amd import lib as cmx
from tests.util import rel

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime already loaded with the library."""

    def __init__(self):
        cmx.init()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def download(self, p, out):
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def noalloc(hip, L, d1, d2, origin, dm, nmom, dout):
    dims = (ctypes.c_int * 4)(*L)
    org = (ctypes.c_int * 4)(*origin)
    return cmx.lib().cmx_meson_2pt_noalloc(dims, d1, d2, org, int(nmom), dm, dout, None)


@pytest.mark.parametrize("L,origin,same", [((32, 32, 32, 10), (1, 2, 3, 4), True), ((4, 4, 4, 8), (0, 1, 0, 5), False)])
def test_meson_noalloc_equals_allocating_entry_point(L, origin, same):
    """Two calls in a row share the work space: a 32^3x10 lattice (two passes of time slices, the last one
    short), then a 4^3x8 one with two distinct propagators."""
    say("meson noalloc", L, origin, same)
    hip = Hip()
    V = int(np.prod(L))
    r = np.random.default_rng(90 + L[3])
    S1 = r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    S2 = S1 if same else r.standard_normal((V, 12, 12)) + 1j * r.standard_normal((V, 12, 12))
    moms = np.array([(0, 0, 0), (1, 0, 0), (0, -1, 1), (1, 1, 1)], np.int32)
    want = cmx.meson_2pt(L, S1, None if same else S2, origin, moms)
    try:
        d1 = hip.upload(S1)
        d2 = d1 if same else hip.upload(S2)
        dm = hip.upload(moms)
        dout = hip.empty(want.nbytes)
        cmx.check(noalloc(hip, L, d1, d2, origin, dm, len(moms), dout))
        got = hip.download(dout, np.zeros_like(want))
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        if V <= 4096:
            assert rel(got, oracle.mesons(L, S1, None if same else S2, origin, moms)) < 1e-13
    finally:
        hip.free()


def test_meson_noalloc_rejects_large_spatial_volume():
    """A spatial volume above 64^3 sites is refused before anything runs on the device."""
    hip = Hip()
    try:
        small = hip.empty(1 << 20)
        rc = noalloc(hip, (66, 64, 64, 2), small, small, (0, 0, 0, 0), small, 1, small)
        assert rc != 0
        assert b"262144" in cmx.lib().cmx_last_error()
    finally:
        hip.free()
