"""GPU parity of CrosshairV2 on Rgb10, RgbaF16 (encoded and linear light), Nv12 and P010 frames, through the C ABI.

Bar: the box state and every kept mask bit-exact against oracle/ui.py on the format's integer levels; the protected
frames (tap rejection on the format's raw sums with the 4:2:0 footprint for Nv12 / P010, the luma and chroma paste-back)
bit-exact against the float32 oracle of oracle/formats.py; the estimate fill on the format's fixed-point luma.
"""
import numpy as np
import pytest
import torch

import fgr
from gpu_util import Internals
from oracle import formats as F
from oracle import ui
from test_gpu_formats import CASES, empty_like, encode, first_diff, lut_of, matrix_of, same, to_dev, to_host
from test_gpu_ui import scene, ui_dict, v2_settings
from test_gpu_v1 import check_estimate, oracle_settings

pytestmark = pytest.mark.gpu

W, H = 192, 112


def encoded(frames, fmt, linear):
    return [encode(f, fmt, linear) for f in frames]


def yuv_k(fmt):
    return F.yuv_constants(matrix_of(fmt)) if fmt in F.YUV else None


def oracle_run(A, fmt, lut, s):
    so = ui_dict(s)
    st = ui.V2State(ui.work_rect(so, W, H), ui.box_rect(so, W, H))
    states, masks = [], []
    for k, f in enumerate(A):
        st.update(ui.levels(f, fmt, lut), so, reset=(k == 0), yuv=yuv_k(fmt))
        states.append((st.w0.copy(), st.w1.copy(), st.w2.copy()))
        masks.append(st.mask())
    return states, masks, ui.work_rect(so, W, H)


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_protected_formats_match_the_oracle(name, fmt, linear):
    # 4:2:0: an even pan and a 4 px cross on whole 2x2 blocks (a pixel whose chroma block holds moving background
    # changes with it and is never still)
    rgba = scene(11, seed=21, step=(4, 2), arm=4) if fmt in F.YUV else scene(11, seed=21)
    A = encoded(rgba, fmt, linear)
    lut = lut_of(fmt, linear)
    for fill in (0, 1):
        s = v2_settings(uiEstimateFill=fill)
        states, masks, rect = oracle_run(A, fmt, lut, s)
        assert (masks[-1] == 255).any(), "the scene must hold a detected crosshair"
        fi = fgr.Interpolator(W, H, fmt=fmt, settings=s, yuv_matrix=matrix_of(fmt),
                              transfer=fgr.TRANSFER_LINEAR if linear else fgr.TRANSFER_ENCODED)
        devs = [to_dev(f, fmt) for f in A]
        for k, d in enumerate(devs):
            print(f"{name} fill={fill} observe {k}", flush=True)
            assert fi.record_observe(d)
            st, _ = fi.debug_ui_mask(None)
            w = st.cpu().numpy().view(np.uint32).astype(np.int64)
            for j in range(3):
                assert np.array_equal(w[..., j], states[k][j]), (k, j)
            m, r = fi.debug_ui_mask(d)
            assert r == rect and np.array_equal(m.cpu().numpy(), masks[k]), k
        assert fi.record_estimate(devs[-2], devs[-1])
        a, c = A[-2], A[-1]
        os_ = oracle_settings(s)
        if fill:
            S = s.flowScale
            fw, fh, _ = F.v1.layout(W, H, S, s.pyramidMinSide)
            la, lc = F.luma_level0(a, fmt, S, fw, fh, lut=lut), F.luma_level0(c, fmt, S, fw, fh, lut=lut)
            la = ui.luma_fill(la, None, masks[-2], rect, S, y16=F.luma16(a, fmt, lut=lut))
            lc = ui.luma_fill(lc, None, masks[-1], rect, S, y16=F.luma16(c, fmt, lut=lut))
            est = F.v1.estimate(None, None, os_, luma0=(la, lc), size=(W, H))
        else:
            est = F.estimate(a, c, fmt, s=os_, lut=lut)
        check_estimate(Internals(fi), est)
        for t in (0.5, 0.25):
            o = empty_like(devs[0], fmt)
            assert fi.record_synth(devs[-2], devs[-1], o, t)
            got = to_host(o, fmt)
            ref, why, _ = F.synth(a, c, est, t, fmt, lut=lut, yuv=yuv_k(fmt), ui=(masks[-2], rect, masks[-1], rect))
            ref = ui.patch(ref, c, masks[-1], rect, fmt)
            assert same(got, ref, fmt), f"{name} fill={fill} t={t} ({why}): {first_diff(got, ref, fmt)}"
        fi.close()
