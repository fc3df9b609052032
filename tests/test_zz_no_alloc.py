"""ddx_efld_noalloc: the field of ddx_efld with device pointers and no
allocation (its partial sums live in the library's scratch array).

Called through the C-ABI on device buffers that the test allocates with the
HIP runtime libddx itself is linked against (a second runtime, such as the
one bundled with torch, cannot share a process with it), on a stream of its
own, and checked against ddx_efld (the same field, host pointers, handle
scratch) on the same inputs: twice in a row, the second call smaller, so
the second reads a scratch array that still holds the first call's
partials.  Both must also
agree with the oracle's field (efld.f90:87-116 restated) on a subset of
targets.  Floating-point tolerance: 1e-12 of the largest field component
(the chunk sums run in a different order).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(seed, nsrc, ntrg):
    rng = np.random.default_rng(seed)
    csrc = rng.uniform(-30.0, 30.0, size=(nsrc, 3))
    src = rng.uniform(-0.5, 0.5, size=nsrc)
    ctrg = rng.uniform(-32.0, 32.0, size=(ntrg, 3))
    return src, csrc, ctrg


class _Hip:
    """hipMalloc / hipMemcpy / hipFree / streams of the runtime libddx loaded."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]

    def check(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)))
        self.check(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))   # host to device
        return p

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(p), max(nbytes, 8)))
        return p

    def download(self, p, shape):
        out = np.empty(shape)
        self.check(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2))  # device to host
        return out


def _noalloc(L, hip, stream, src, csrc, ctrg):
    ntrg = ctrg.shape[0]
    bufs = [hip.upload(src), hip.upload(csrc), hip.upload(ctrg)]
    ef = hip.upload(np.full((ntrg, 3), np.nan))
    try:
        rc = L.ddx_efld_noalloc(src.size, bufs[0], bufs[1], ntrg, bufs[2], ef, stream)
        assert rc == 0, L.ddx_last_error()
        hip.check(hip.rt.hipStreamSynchronize(stream))
        return hip.download(ef, (ntrg, 3))
    finally:
        for p in bufs + [ef]:
            hip.rt.hipFree(p)


def test_noalloc_field_matches_efld_twice(oracle_mod):
    from ddpcm_amd import ddx
    from ddpcm_amd.inputs import config_input
    L = ddx.load_library()
    d = config_input("c1")
    h = ddx.DDCosmo.from_input(d)
    hip = _Hip()
    stream = C.c_void_p()
    hip.check(hip.rt.hipStreamCreate(C.byref(stream)))
    # first: many sources, few targets (the chunked path: partials in the
    # scratch array); second, smaller; third: one chunk, no scratch
    for seed, nsrc, ntrg in ((1, 300000, 2500), (2, 9000, 700), (3, 700, 5000)):
        print("ddx_efld_noalloc case seed=%d nsrc=%d ntrg=%d" % (seed, nsrc, ntrg), flush=True)
        src, csrc, ctrg = _case(seed, nsrc, ntrg)
        got = _noalloc(L, hip, stream, src, csrc, ctrg)
        ref = h.efld(src, csrc, ctrg)
        scale = np.max(np.abs(ref))
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - ref)) <= 1e-12 * scale, np.max(np.abs(got - ref)) / scale
        sub = slice(0, min(ntrg, 64))
        o = oracle_mod.efld(src, csrc, ctrg[sub])
        assert np.max(np.abs(got[sub] - o)) <= 1e-12 * scale
    hip.rt.hipStreamDestroy(stream)


def test_noalloc_rejects_bad_arguments():
    from ddpcm_amd import ddx
    L = ddx.load_library()
    assert L.ddx_efld_noalloc(-1, None, None, 1, None, None, None) == ddx.DDX_ERR_ARG
    assert L.ddx_efld_noalloc(4, None, None, 1, None, None, None) == ddx.DDX_ERR_ARG
    assert b"invalid" in L.ddx_last_error()
