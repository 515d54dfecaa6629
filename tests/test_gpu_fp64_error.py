"""The synthesized output's error against the FP64 oracle, stated per format: max and mean over every output sample of a
seeded 640x360 pair (noise translated by (+5, -3) px plus a moving 64x64 sprite) at t = 1/4, 1/2, 3/4, through the C ABI.

Bar: within +-1 level (8-bit scale) on >= 99.9 % of samples and +-2 everywhere. RGBA8: the 8-bit output against the
FP64 oracle's rounded output; the other formats: their levels (oracle/formats.py level_error) against the FP64 oracle's
unrounded levels. Each case prints one line (pytest -s shows them).
"""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from oracle import formats as F
from oracle import v1
from test_gpu_formats import CASES, encode, lut_of, matrix_of, run_fmt

pytestmark = pytest.mark.gpu

W, H = 640, 360
TS = (0.25, 0.5, 0.75)


def pair():
    return scenes.pair_translate(1, w=W, h=H)


def report(name, errs):
    e = np.concatenate([x.ravel() for x in errs]).astype(np.float64)
    print(f"\nfp64 error {name}: max {e.max():.4f}  mean {e.mean():.5f}  <=1: {(e <= 1.0).mean() * 100:.4f} %  "
          f"({e.size} samples, t = {TS})", flush=True)
    assert e.max() <= 2.0 and (e <= 1.0).mean() >= 0.999, name


def test_rgba8_error_against_fp64():
    a, c = pair()
    fi = fgr.Interpolator(W, H)
    ta, tc = torch.from_numpy(a).cuda(), torch.from_numpy(c).cuda()
    assert fi.record_estimate(ta, tc)
    est = v1.estimate(a, c)
    errs = []
    for t in TS:
        o = torch.zeros_like(ta)
        assert fi.record_synth(ta, tc, o, t)
        ref64, why, _ = v1.synth(a, c, est, t, exact64=True)
        assert why is None
        errs.append(np.abs(o.cpu().numpy()[..., :3].astype(np.int64) - ref64[..., :3].astype(np.int64)))
    fi.close()
    report("rgba8", errs)


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_format_error_against_fp64(name, fmt, linear):
    a, c = pair()
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    lut = lut_of(fmt, linear)
    yuv = F.yuv_constants(matrix_of(fmt))
    _, outs, _ = run_fmt(A, C, fmt, TS, linear)
    est = F.estimate(A, C, fmt, lut=lut)
    errs = []
    for t, o in zip(TS, outs):
        _, why, i64 = F.synth(A, C, est, t, fmt, exact64=True, lut=lut, yuv=yuv)
        assert why is None
        errs.append(F.level_error(fmt, o, i64, lut=lut))
    report(name, errs)
