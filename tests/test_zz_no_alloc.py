"""ddx_efld_noalloc: the field of ddx_efld with device pointers and no
allocation (its partial sums live in the library's scratch array).

Called through the C-ABI on device buffers that torch allocated, and checked
against ddx_efld (the same field, host pointers, handle scratch) on the same
inputs: twice in a row, the second call smaller, so the second reads a
scratch array that still holds the first call's partials.  Both must also
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


def _noalloc(L, torch, src, csrc, ctrg):
    dev = torch.device("cuda:0")
    q = torch.from_numpy(src).to(dev)
    cs = torch.from_numpy(np.ascontiguousarray(csrc)).to(dev)
    ct = torch.from_numpy(np.ascontiguousarray(ctrg)).to(dev)
    ef = torch.full((ctrg.shape[0], 3), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.ddx_efld_noalloc(src.size, q.data_ptr(), cs.data_ptr(), ctrg.shape[0], ct.data_ptr(),
                            ef.data_ptr(), C.c_void_p(stream))
    assert rc == 0, L.ddx_last_error()
    torch.cuda.synchronize()
    return ef.cpu().numpy()


def test_noalloc_field_matches_efld_twice(oracle_mod):
    import torch
    from ddpcm_amd import ddx
    from ddpcm_amd.inputs import config_input
    L = ddx.load_library()
    d = config_input("c1")
    h = ddx.DDCosmo.from_input(d)
    # first: many sources, few targets (the chunked path: partials in the
    # scratch array); second, smaller; third: one chunk, no scratch
    for seed, nsrc, ntrg in ((1, 300000, 2500), (2, 9000, 700), (3, 700, 5000)):
        print("ddx_efld_noalloc case seed=%d nsrc=%d ntrg=%d" % (seed, nsrc, ntrg), flush=True)
        src, csrc, ctrg = _case(seed, nsrc, ntrg)
        got = _noalloc(L, torch, src, csrc, ctrg)
        ref = h.efld(src, csrc, ctrg)
        scale = np.max(np.abs(ref))
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - ref)) <= 1e-12 * scale, np.max(np.abs(got - ref)) / scale
        sub = slice(0, min(ntrg, 64))
        o = oracle_mod.efld(src, csrc, ctrg[sub])
        assert np.max(np.abs(got[sub] - o)) <= 1e-12 * scale


def test_noalloc_rejects_bad_arguments():
    from ddpcm_amd import ddx
    L = ddx.load_library()
    assert L.ddx_efld_noalloc(-1, None, None, 1, None, None, None) == ddx.DDX_ERR_ARG
    assert L.ddx_efld_noalloc(4, None, None, 1, None, None, None) == ddx.DDX_ERR_ARG
    assert b"invalid" in L.ddx_last_error()
