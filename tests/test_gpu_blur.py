"""recordSynthBlur (framegen.cpp:3534-3634) on the GPU, bit-exact against the oracle's weighted mean of the synth
frames at the sample times, rounded once; and its collapses to recordSynth(t1) / a real frame."""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from gpu_util import dev
from oracle import formats as F
from oracle import v1
from test_gpu_formats import empty_like, encode, lut_of, matrix_of, same, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,gauss,t0,t1", [(4, 0, 0.1, 0.6), (5, 1, 0.3, 0.9), (16, 0, 0.0, 1.0), (2, 1, 0.45, 0.55)])
def test_blur_rgba8(n, gauss, t0, t1):
    a, c = scenes.pair_translate(31, w=320, h=180, sprite=40, sprite_pos=(120, 60), shift=(4, -2))
    s = fgr.default_settings(blurSamples=n, blurWeights=gauss)
    fi = fgr.Interpolator(320, 180, settings=s)
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    out = torch.zeros_like(ta)
    assert fi.record_synth_blur(ta, tc, out, t0, t1)
    est = v1.estimate(a, c)
    want = v1.synth_blur(a, c, est, t0, t1, n, bool(gauss))
    got = out.cpu().numpy()
    assert np.array_equal(got, want), int((got != want).sum())
    fi.close()


def test_blur_scene_cut_and_collapses():
    a, c = scenes.noise_pair(32, 256, 144)
    fi = fgr.Interpolator(256, 144, settings=fgr.default_settings(blurSamples=4, blurWeights=0))
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    out = torch.zeros_like(ta)
    assert fi.record_synth_blur(ta, tc, out, 0.2, 0.8)   # samples 0.275, 0.425 -> prev; 0.575, 0.725 -> curr
    est = v1.estimate(a, c)
    assert est["cut"]
    assert np.array_equal(out.cpu().numpy(), v1.synth_blur(a, c, est, 0.2, 0.8, 4))
    fi.close()
    # blur off / t0 == t1: recordSynth(t1); t1 = 1 or t1 = 0 without blur: the real frame
    a, c = scenes.pair_translate(33, w=256, h=144)
    for s, t0, t1 in ((fgr.default_settings(), 0.2, 0.7), (fgr.default_settings(blurSamples=4), 0.4, 0.4)):
        fi = fgr.Interpolator(256, 144, settings=s)
        ta, tc = dev(a), dev(c)
        assert fi.record_estimate(ta, tc)
        o1, o2 = torch.zeros_like(ta), torch.zeros_like(ta)
        assert fi.record_synth_blur(ta, tc, o1, t0, t1)
        assert fi.record_synth(ta, tc, o2, t1)
        assert np.array_equal(o1.cpu().numpy(), o2.cpu().numpy())
        assert fi.record_synth_blur(ta, tc, o1, 1.0, 1.0) and np.array_equal(o1.cpu().numpy(), c)
        assert fi.record_synth_blur(ta, tc, o1, 0.0, 0.0) and np.array_equal(o1.cpu().numpy(), a)
        assert not fi.record_synth_blur(ta, tc, o1, 0.7, 0.2)                 # t0 > t1
        assert not fi.record_synth_blur(ta, tc, o1, float("nan"), 0.5)
        fi.close()


@pytest.mark.parametrize("fmt,linear", [(F.RGB10, False), (F.F16, True), (F.NV12, False), (F.P010, False)])
def test_blur_formats(fmt, linear):
    a, c = scenes.pair_translate(34, w=256, h=144, sprite=32, sprite_pos=(90, 40), shift=(3, 2))
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    s = fgr.default_settings(blurSamples=4, blurWeights=1)
    fi = fgr.Interpolator(256, 144, fmt=fmt, settings=s, transfer=fgr.TRANSFER_LINEAR if linear else 0,
                          yuv_matrix=matrix_of(fmt))
    ta, tc = to_dev(A, fmt), to_dev(C, fmt)
    assert fi.record_estimate(ta, tc)
    out = empty_like(ta, fmt)
    assert fi.record_synth_blur(ta, tc, out, 0.25, 0.75)
    lut = lut_of(fmt, linear)
    est = F.estimate(A, C, fmt, lut=lut)
    want = F.synth_blur(A, C, est, 0.25, 0.75, 4, fmt, gauss=True, lut=lut, yuv=F.yuv_constants(matrix_of(fmt)))
    assert same(to_host(out, fmt), want, fmt)
    fi.close()


def test_records_queued_before_a_larger_blur_keep_their_outputs():
    """A blurred record queued with blurSamples 2, then settings raising it to 16 (a larger sub-frame scratch) and a
    second blurred record, synchronised only at the end: both outputs equal the oracle's (the first record's scratch
    outlives the settings change; no record allocates)."""
    a, c = scenes.pair_translate(35, w=320, h=180, sprite=40, sprite_pos=(100, 50), shift=(5, -3))
    fi = fgr.Interpolator(320, 180, settings=fgr.default_settings(blurSamples=2))
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    o1, o2 = torch.zeros_like(ta), torch.zeros_like(ta)
    assert fi.record_synth_blur(ta, tc, o1, 0.1, 0.6)
    assert fi.set_settings(fgr.default_settings(blurSamples=16))
    assert fi.record_synth_blur(ta, tc, o2, 0.1, 0.6)
    torch.cuda.synchronize()
    est = v1.estimate(a, c)
    assert np.array_equal(o1.cpu().numpy(), v1.synth_blur(a, c, est, 0.1, 0.6, 2, gauss=True))
    assert np.array_equal(o2.cpu().numpy(), v1.synth_blur(a, c, est, 0.1, 0.6, 16, gauss=True))
    fi.close()
