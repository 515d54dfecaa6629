"""Known-shift acceptance of the long-vector settings on 1920x1080 (gpu/shifttest.py's three cases).

A 1/f-noise picture and the same picture panned by a known whole-pixel shift (the content that leaves the frame is
replaced by the picture's next columns / rows). Required, in the region visible in both frames: the per-pixel half-vector
within 1 flow pixel of shift / (2 flowScale) on >= 95 % of the flow pixels, and the t = 0.5 frame against the picture
panned by half the shift at >= 35 dB PSNR. Cases: flowScale 2 and shift (400, 100) at searchRadius 12, searchRadiusTop
12, pyramidMinSide 16 (beyond the default reach); flowScale 1 and shift (600, 300) (beyond +-127 px, so the exact
two-pass wide median decides); the first shift with the default settings must NOT be found (under half the pixels).
The whole-frame fallbacks are off (globalFallback 1, sceneCutSad 255), as in the reference's cases. Bit-exactness of these
settings against the oracle is covered at smaller sizes by test_gpu_v1::test_wide_reach_settings.
"""
import numpy as np
import pytest
import torch

import fgr
from gpu_util import Internals

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
WIDE = dict(searchRadius=12, searchRadiusTop=12, pyramidMinSide=16)
# the strip a pan reveals is more than the default 15 % of the frame: the whole-frame fallbacks are off (not under test)
NOFALLBACK = dict(globalFallback=1.0, sceneCutSad=255.0)


def noise(shape, rng):
    """1/f noise: white noise filtered with an amplitude spectrum 1 / f, plus a little white noise"""
    f = np.fft.fft2(rng.standard_normal(shape))
    fy = np.fft.fftfreq(shape[0])[:, None]
    fx = np.fft.fftfreq(shape[1])[None, :]
    r = np.sqrt(fx * fx + fy * fy)
    r[0, 0] = 1.0
    n = np.fft.ifft2(f / r).real
    n = (n - n.mean()) / n.std()
    return n + 0.15 * rng.standard_normal(shape)


def scene(dx, dy):
    """(prev, curr, truth at t = 0.5): curr(x) = prev(x - d), truth = prev panned by d / 2"""
    rng = np.random.default_rng(12345)
    cw, ch = W + abs(dx) + 16, H + abs(dy) + 16
    can = np.stack([noise((ch, cw), rng) for _ in range(3)], -1)
    can = np.clip(128 + 42 * can, 0, 255).astype(np.uint8)
    ox, oy = max(dx, 0) + 8, max(dy, 0) + 8

    def crop(sx, sy):
        out = np.full((H, W, 4), 255, np.uint8)
        out[..., :3] = can[oy - sy:oy - sy + H, ox - sx:ox - sx + W]
        return out
    return crop(0, 0), crop(dx, dy), crop(dx // 2, dy // 2)


def run(settings, prev, curr):
    fi = fgr.Interpolator(W, H, settings=settings)
    a, c = torch.from_numpy(prev).cuda(), torch.from_numpy(curr).cuda()
    o = torch.zeros_like(a)
    assert fi.record(a, c, o, 0.5)
    I = Internals(fi)
    pf = I.pf()
    out = o.cpu().numpy()
    fi.close()
    return pf, out


def check(pf, out, truth, dx, dy, fs):
    hx, hy = dx / (2.0 * fs), dy / (2.0 * fs)
    fh, fw = pf.shape[:2]
    ys, xs = np.mgrid[0:fh, 0:fw]
    m = 8
    ok = (xs - abs(hx) >= m) & (xs + abs(hx) <= fw - 1 - m) & (ys - abs(hy) >= m) & (ys + abs(hy) <= fh - 1 - m)
    good = (np.abs(pf[..., 0] - hx) <= 1.0) & (np.abs(pf[..., 1] - hy) <= 1.0)
    frac = float(good[ok].mean())
    ys, xs = np.mgrid[0:H, 0:W]
    mm = 16
    vis = (xs - abs(dx) // 2 >= mm) & (xs + abs(dx) // 2 <= W - 1 - mm) & (ys - abs(dy) // 2 >= mm) & \
          (ys + abs(dy) // 2 <= H - 1 - mm)
    e = (out[..., :3].astype(np.float64) - truth[..., :3].astype(np.float64))[vis]
    mse = float((e * e).mean())
    psnr = 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)
    print(f"shift ({dx}, {dy}) flowScale {fs}: {100 * frac:.1f} % within 1 flow px, PSNR {psnr:.1f} dB", flush=True)
    return frac, psnr


@pytest.mark.parametrize("dx,dy,fs", [(400, 100, 2), (600, 300, 1)])
def test_known_shift_is_found_with_the_wide_settings(dx, dy, fs):
    prev, curr, truth = scene(dx, dy)
    pf, out = run(fgr.default_settings(flowScale=fs, **WIDE, **NOFALLBACK), prev, curr)
    frac, psnr = check(pf, out, truth, dx, dy, fs)
    assert frac >= 0.95 and psnr >= 35.0


def test_known_shift_is_beyond_the_default_reach():
    prev, curr, truth = scene(400, 100)
    pf, out = run(fgr.default_settings(**NOFALLBACK), prev, curr)
    frac, psnr = check(pf, out, truth, 400, 100, 2)
    assert frac < 0.5
