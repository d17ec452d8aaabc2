"""sphg_direct_find_noalloc: the pair search on caller-owned device arrays
with the library's static scratch instead of any allocation.  Its pairs,
W, grad W, countiac and niac equal those of sphg_direct_find (the context
entry point) bit for bit, on two calls in a row with the second smaller.

The device arrays are allocated through the HIP runtime the library itself
links (the test's own hipMalloc; the entry point allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

from sources_amd import abi, scenarios
from sources_amd.sph import Solver

pytestmark = pytest.mark.gpu


class Dev:
    """Device buffers through the HIP runtime libsphg.so links."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(nbytes), 8)) == 0
        self.bufs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p

    def get(self, p, dtype, count):
        out = np.empty(count, dtype)
        if count:
            assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _context_pairs(sc, phase):
    s = Solver(sc.params, sc.water, sc.soil)
    s.direct_find(phase)
    s.close()
    ph = getattr(sc, phase)
    k = ph.niac
    return ph.pair_i[:k].copy(), ph.pair_j[:k].copy(), ph.w[:k].copy(), ph.dwdx[:k].copy(), ph.countiac[:ph.n].copy()


def _noalloc_pairs(lib, dev, params, ph, max_interaction):
    n = ph.n
    x = dev.put(ph.x[:n])                                   # x(2,n) interleaved
    h = dev.put(ph.hsml[:n])
    fill_i = np.full(max_interaction, -7, np.int32)
    pi, pj = dev.put(fill_i), dev.put(fill_i)
    w = dev.put(np.full(max_interaction, np.nan))
    dw = dev.put(np.full(2 * max_interaction, np.nan))
    cnt = dev.put(np.full(n, -7, np.int32))
    niac = C.c_int32(-1)
    rc = lib.sphg_direct_find_noalloc(C.byref(params), n, x, h, max_interaction, pi, pj, w, dw, cnt,
                                      C.byref(niac), None)
    assert rc == abi.SPHG_OK, (lib.sphg_strerror(rc), "hip error", lib.sphg_last_hip_error(None))
    k = niac.value
    got_pi = dev.get(pi, np.int32, max_interaction)
    return (got_pi[:k], dev.get(pj, np.int32, k), dev.get(w, np.float64, k),
            dev.get(dw, np.float64, 2 * k).reshape(k, 2), dev.get(cnt, np.int32, n), got_pi[k:])


def _column(dx):
    return scenarios.submerged_column(dx=dx, soil_w=0.1, soil_h=0.05, tank_w=0.3, tank_h=0.12,
                                      jitter=0.05, seed=7)


def test_noalloc_direct_find_equals_the_context_entry_point():
    lib = abi.load_library()
    dev = Dev()
    try:
        # two calls in a row, the second on a smaller input: the shared scratch
        # keeps the first call's contents, which must not leak into the second
        for dx, phase in ((0.0025, "water"), (0.005, "soil")):
            sc = _column(dx)
            ph = getattr(sc, phase)
            assert ph.n <= abi.NOALLOC_MAX_N
            print(f"noalloc direct_find dx={dx} {phase} n={ph.n}", flush=True)
            mi = ph.pair_i.shape[0]
            got = _noalloc_pairs(lib, dev, sc.params, ph, mi)
            want = _context_pairs(sc, phase)
            assert len(want[0]) > 0
            for a, b in zip(got[:5], want):
                np.testing.assert_array_equal(a, b)
            assert np.all(got[5] == -7)                  # nothing written past niac
    finally:
        dev.close()


def test_noalloc_direct_find_limits():
    lib = abi.load_library()
    dev = Dev()
    try:
        sc = _column(0.005)
        ph = sc.water
        n = ph.n
        x, h = dev.put(ph.x[:n]), dev.put(ph.hsml[:n])
        pi = dev.put(np.full(8, -7, np.int32))
        w, dw = dev.put(np.zeros(8)), dev.put(np.zeros(16))
        cnt = dev.put(np.zeros(n, np.int32))
        got_n = C.c_int32(-1)
        # too many interactions: the count comes back, no pair is written
        rc = lib.sphg_direct_find_noalloc(C.byref(sc.params), n, x, h, 8, pi, pi, w, dw, cnt, C.byref(got_n), None)
        assert rc == abi.SPHG_ERR_TOO_MANY_INTERACTIONS
        assert got_n.value > 8
        assert np.all(dev.get(pi, np.int32, 8) == -7)
        # beyond the scratch's particle limit: rejected before any device work
        rc = lib.sphg_direct_find_noalloc(C.byref(sc.params), abi.NOALLOC_MAX_N + 1, x, h, 8, pi, pi, w, dw, cnt,
                                          C.byref(got_n), None)
        assert rc == abi.SPHG_ERR_INVALID
    finally:
        dev.close()
