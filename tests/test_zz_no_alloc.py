"""fgr_record_noalloc (the record that allocates nothing) through the C ABI: the same outputs as fgr_record on a
handle, bit for bit, on two calls in a row (the second with a smaller pair), no change in the device's free memory,
and refusals with a reason for what it cannot run without allocating."""
import ctypes

import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from gpu_util import dev, hip
from test_gpu_formats import encode, to_dev

pytestmark = pytest.mark.gpu


def _free_bytes():
    torch.cuda.synchronize()
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    h = hip()
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    assert h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _handle_output(a, c, t, settings=None, fmt=fgr.FORMAT_RGBA8, **kw):
    t0 = a[0] if isinstance(a, tuple) else a
    H, W = t0.shape[:2]
    fi = fgr.Interpolator(W, H, fmt=fmt, settings=settings, **kw)
    out = (torch.zeros_like(a[0]), torch.zeros_like(a[1])) if isinstance(a, tuple) else torch.zeros_like(a)
    assert fi.record(a, c, out, t)
    torch.cuda.synchronize()
    fi.close()
    return out


def _equal(x, y):
    if isinstance(x, tuple):
        return all(torch.equal(p, q) for p, q in zip(x, y))
    return torch.equal(x, y)


def test_noalloc_matches_record_on_two_calls():
    """1080p at t = 1/4, then a smaller 640x360 pair at t = 1/2 (the array holds the first call's contents): each
    output equal to fgr_record's on a handle; after the set-up call the device's free memory does not change."""
    cases = [(scenes.pairs_1080p(2, n_pairs=2)[0][0], 0.25), (scenes.pair_translate(3), 0.5)]
    dev_pairs = [(dev(a), dev(c), t) for (a, c), t in cases]
    want = [_handle_output(a, c, t) for a, c, t in dev_pairs]
    outs = [torch.full_like(a, 7) for a, _, _ in dev_pairs]
    # the first call on a device is its set-up: the HIP runtime loads the code object that holds the array (placing
    # it on this GPU) and the array's address is read; calls after it allocate nothing
    a, c, _ = dev_pairs[1]
    ok, why = fgr.record_noalloc(a, c, torch.zeros_like(a), 0.5, a.shape[1], a.shape[0])
    assert ok, why
    free0 = _free_bytes()
    for (a, c, t), o in zip(dev_pairs, outs):
        H, W = a.shape[:2]
        print(f"case noalloc {W}x{H} t={t}", flush=True)
        ok, why = fgr.record_noalloc(a, c, o, t, W, H)
        assert ok, why
    torch.cuda.synchronize()
    assert _free_bytes() == free0
    for k, (o, w) in enumerate(zip(outs, want)):
        assert torch.equal(o, w), f"call {k}"


def test_noalloc_p010_and_settings_match_record():
    a, c = scenes.pair_translate(5, w=320, h=180, sprite=40, sprite_pos=(100, 50), shift=(5, -3))
    A, C = to_dev(encode(a, fgr.FORMAT_P010), fgr.FORMAT_P010), to_dev(encode(c, fgr.FORMAT_P010), fgr.FORMAT_P010)
    s = fgr.default_settings(flowScale=1, pyramidMinSide=24, trustLow=8.0, trustHigh=30.0)
    want = _handle_output(A, C, 0.75, settings=s, fmt=fgr.FORMAT_P010, yuv_matrix=fgr.YUV_BT2020)
    out = (torch.zeros_like(A[0]), torch.zeros_like(A[1]))
    ok, why = fgr.record_noalloc(A, C, out, 0.75, 320, 180, fmt=fgr.FORMAT_P010, settings=s,
                                 yuv_matrix=fgr.YUV_BT2020)
    assert ok, why
    torch.cuda.synchronize()
    assert _equal(out, want)


def test_noalloc_refusals_name_the_reason():
    a = torch.zeros((64, 64, 4), dtype=torch.uint8, device="cuda:0")
    cases = [
        (dict(width=3840, height=2160, settings=fgr.default_settings(flowScale=1)), "FGR_NOALLOC_SCRATCH_BYTES"),
        (dict(width=64, height=64, fmt=fgr.FORMAT_RGBAF16), "RgbaF16"),
        (dict(width=64, height=64, settings=fgr.default_settings(seedNet=1)), "seedNet"),
        (dict(width=64, height=64, settings=fgr.default_settings(trustHigh=5.0)), "invalid settings"),
        (dict(width=128, height=64), "pitch"),
    ]
    for kw, reason in cases:
        out = torch.full_like(a, 7)
        ok, why = fgr.record_noalloc(a, a, out, 0.5, **kw)
        assert not ok and reason in why, (kw, why)
        torch.cuda.synchronize()
        assert int(out.min()) == 7 and int(out.max()) == 7
    ok, why = fgr.record_noalloc(a, a, a.clone(), 1.0, 64, 64)
    assert not ok and "t must be" in why
