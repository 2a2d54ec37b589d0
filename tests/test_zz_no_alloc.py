"""The allocation-free dense call (storm_hip_allpairs_contig_noalloc) through
the C ABI against the allocating storm_hip_allpairs on the same rows, on two
calls in a row (the second smaller), fp4 and int8, and against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Dev:
    def __init__(self, sh, nbytes):
        self.R = sh.hip_runtime()
        self.p = C.c_void_p()
        assert self.R.hipMalloc(C.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        assert self.R.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert self.R.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p, a.nbytes, D2H) == 0

    def free(self):
        self.R.hipFree(self.p)


def _run_noalloc(sh, ctx, words, ld, bb):
    n, nw = words.shape
    padded = np.zeros((n, ld), np.uint64)
    padded[:, :nw] = words
    dw, dt, ds = Dev(sh, padded.nbytes), Dev(sh, max(1, n * (n - 1) // 2) * 4), Dev(sh, 8)
    try:
        dw.put(padded)
        sh.check(sh.lib().storm_hip_allpairs_contig_noalloc(ctx.h, dw.p, n, nw, ld,
                                                            sh.BB[bb], dt.p, ds.p), "noalloc")
        ctx.sync()
        tri = np.zeros(n * (n - 1) // 2, np.uint32)
        if len(tri):
            dt.get(tri)
        s = np.zeros(1, np.uint64)
        ds.get(s)
        return int(s[0]), tri
    finally:
        dw.free(), dt.free(), ds.free()


@pytest.mark.parametrize("bb", ["mfma_fp4", "mfma", "auto"])
def test_noalloc_matches_allocating_call(sh, ctx, bb):
    # (n_rows, n_words, row stride, p, seed): first call, then a smaller one.
    for n, nw, ld, p, seed in [(700, 300, 304, 0.05, 61), (130, 40, 48, 0.3, 62),
                               (257, 16, 16, 0.5, 63)]:
        print("case noalloc", bb, n, nw, ld, flush=True)
        words = O.bernoulli_rows(seed, p, nw * 64, 0, n)
        with ctx.upload_contig(words) as c:
            st = c.allpairs(out="tri", bb=bb)
            want_tri = c.tri()
        s, tri = _run_noalloc(sh, ctx, words, ld, bb)
        assert s == st["sum"]
        assert np.array_equal(tri, want_tri)
        ws, wt = O.allpairs_words(words)
        assert s == ws and np.array_equal(tri, wt)


def test_noalloc_rejects(sh, ctx):
    d = Dev(sh, 8)
    try:
        L = sh.lib()
        assert L.storm_hip_allpairs_contig_noalloc(ctx.h, d.p, (1 << 20) + 1, 8, 16, 0, None,
                                                   d.p) == -6
        assert L.storm_hip_allpairs_contig_noalloc(ctx.h, d.p, 10, 8, 12, 0, None, d.p) == -1
        assert L.storm_hip_allpairs_contig_noalloc(ctx.h, d.p, 10, 8, 16, 1, None, d.p) == -5
        assert L.storm_hip_allpairs_contig_noalloc(ctx.h, d.p, 10, 8, 16, 0, None, None) == -1
    finally:
        d.free()
