"""The format oracle (oracle/formats.py) pinned against closed forms, brute force and invariants."""
import math
from fractions import Fraction

import numpy as np
import pytest

from fgr import scenes
from oracle import formats as F
from oracle import v1


@pytest.fixture(scope="module")
def luts():
    return {False: F.f16_level_table(False), True: F.f16_level_table(True)}


def half(v):
    return int(np.array(v, np.float16).view(np.uint16))


def test_f16_level_table_closed_forms(luts):
    enc, lin = luts[False], luts[True]
    assert enc[half(1.0)] == 4080 and enc[half(0.5)] == 2040 and enc[half(-1.0)] == -4080 and enc[half(0.0)] == 0
    assert enc[half(np.inf)] == 65535 and enc[half(-np.inf)] == -65535 and enc[0x7E00] == 0   # inf, NaN
    assert enc[half(100.0)] == 65535                                                        # clamped (+-4096 levels)
    # linear: 1.0 = 80 nits; 10000 nits (125.0) is PQ 1 = 255 levels; 0 and negatives are PQ(0) ~ 7e-7 -> 0
    assert lin[half(125.0)] == 4080 and lin[half(1000.0)] == 4080 and lin[half(0.0)] == 0 and lin[half(-2.0)] == 0
    # 100 nits is PQ code 0.5081 (SMPTE ST 2084 table value 520/1023 in 10 bits)
    assert abs(F.pq_code(0.01) - 0.50807) < 1e-4
    assert lin[half(1.25)] == math.floor(16 * 255 * F.pq_code(1.25 * 80 / 10000) + 0.5)
    # monotone in the value for finite non-negative halves
    pos = np.arange(0, 0x7C00)
    assert np.all(np.diff(enc[pos]) >= 0) and np.all(np.diff(lin[pos]) >= 0)


def test_luma16_closed_forms(luts):
    codes = np.array([0, 1, 2, 511, 512, 1000, 1023], np.uint32)
    gray = codes | (codes << 10) | (codes << 20)
    y = F.luma16(gray[None, :], F.RGB10)[0]
    for c, v in zip(codes, y):   # 16 * c * 255 / 1023, halves up (the weights sum to 16384)
        assert v == math.floor(Fraction(16 * 255 * int(c), 1023) + Fraction(1, 2))
    # P010 limited range: code 64 is black (0), 940 is white (255 levels = 4080), halves up, below black negative
    yp = np.array([[64 << 6, 940 << 6, 502 << 6, 0, 1023 << 6]], np.uint16)
    got = F.luma16(yp, F.P010)[0]
    want = [math.floor(Fraction((4 * c - 256) * 255, 219) + Fraction(1, 2)) for c in (64, 940, 502, 0, 1023)]
    assert list(got) == want and got[0] == 0 and got[1] == 4080
    assert list(F.luma16(yp, F.P010, full_range=True)[0]) == [256, 3760, 2008, 0, 4092]
    # Nv12: code 16 is black, 235 white (255 levels), 16 code at full range
    y8 = np.array([[16, 235, 126, 0, 255]], np.uint8)
    want8 = [math.floor(Fraction((16 * c - 256) * 255, 219) + Fraction(1, 2)) for c in (16, 235, 126, 0, 255)]
    assert list(F.luma16(y8, F.NV12)[0]) == want8 and want8[:2] == [0, 4080]
    assert list(F.luma16(y8, F.NV12, full_range=True)[0]) == [256, 3760, 2016, 0, 4080]
    # RgbaF16 encoded gray v: luma = 16 * 255 v (weights sum to 16384), rounded
    v = np.array([0.25, 0.5, 1.0], np.float16)
    px = np.stack([v, v, v, np.ones_like(v)], -1).view(np.uint16)[None]
    assert list(F.luma16(px, F.F16, lut=luts[False])[0]) == [1020, 2040, 4080]
    # bgra swaps R and B
    w = np.array([[100 | (200 << 10) | (300 << 20)]], np.uint32)
    n = 300 * 4899 + 200 * 9617 + 100 * 1868
    assert F.luma16(w, F.RGB10, bgra=True)[0, 0] == (n * 255 + 523776) // 1047552


def _scalar_fetch(img, x, y):
    h, w = img.shape[:2]
    return img[min(max(y, 0), h - 1), min(max(x, 0), w - 1)]


def _scalar_bilinear(img, X, Y, unit):
    sh = unit.bit_length() - 1
    ix, iy, fx, fy = X >> sh, Y >> sh, X & (unit - 1), Y & (unit - 1)
    taps = [_scalar_fetch(img, ix, iy), _scalar_fetch(img, ix + 1, iy), _scalar_fetch(img, ix, iy + 1),
            _scalar_fetch(img, ix + 1, iy + 1)]
    ws = [(unit - fx) * (unit - fy), fx * (unit - fy), (unit - fx) * fy, fx * fy]
    return sum(int(w) * np.asarray(t, np.int64) for w, t in zip(ws, taps))


@pytest.mark.parametrize("fmt", [F.RGB10, F.P010, F.NV12, F.F16])
def test_mismatch_matches_a_scalar_restatement(fmt, luts):
    """e of every pixel of a tiny pair with a random field, vectorised oracle vs a per-pixel loop."""
    W, H = 24, 20
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    c = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    enc = {F.RGB10: F.encode_rgb10, F.P010: F.encode_p010, F.NV12: F.encode_nv12, F.F16: F.encode_f16}[fmt]
    A, C = enc(a), enc(c)
    fw, fh, lv = v1.layout(W, H, 2, 16)
    pf = (rng.integers(-40, 41, (fh, fw, 2)) / 8.0).astype(np.float32)
    est = dict(settings=dict(v1.DEFAULTS), pf=pf, cut=False)
    lut = luts[False]
    yuv = F.yuv_constants(2)
    t = 0.25
    _, _, info = F.synth(A, C, est, t, fmt, lut=lut, yuv=yuv)
    (AX, AY), (CX, CY), _, _ = F._maps(est, t, W, H)
    for y in range(H):
        for x in range(W):
            ax, ay, cx, cy = (int(v[y, x]) for v in (AX, AY, CX, CY))
            if fmt == F.RGB10:
                sa = _scalar_bilinear(F.rgb10_channels(A), ax, ay, 32)
                sc = _scalar_bilinear(F.rgb10_channels(C), cx, cy, 32)
                d = max(abs(float(np.float32(sa[k]) * F.K10) - float(np.float32(sc[k]) * F.K10)) for k in range(3))
                e = int(min(d, 255.0))
            elif fmt in F.YUV:
                q = 4 if fmt == F.P010 else 1
                code = (lambda p: (p >> 6).astype(np.int64)) if fmt == F.P010 else (lambda p: p.astype(np.int64))
                ya = _scalar_bilinear(code(A[0]), ax, ay, 32)
                yc = _scalar_bilinear(code(C[0]), cx, cy, 32)
                ca = _scalar_bilinear(code(A[1]), ax - 16, ay - 16, 64)
                cc = _scalar_bilinear(code(C[1]), cx - 16, cy - 16, 64)
                dY = np.float32(int(ya) - int(yc)) * np.float32(1 / (1024 * q))
                dC = [np.float32(int(ca[k]) - int(cc[k])) * np.float32(1 / (4096 * q)) for k in range(2)]
                m = F._yuv_diff(np.array([dY]), np.array([dC[0]]), np.array([dC[1]]), yuv)[0]
                e = int(min(m, np.float32(255)))
            else:
                va = F._f16_warp(F.f16_values(A)[..., :3], np.array([ax]), np.array([ay]))[0]
                vc = F._f16_warp(F.f16_values(C)[..., :3], np.array([cx]), np.array([cy]))[0]
                La = lut[va.astype(np.float16).view(np.uint16)].astype(np.int64)
                Lc = lut[vc.astype(np.float16).view(np.uint16)].astype(np.int64)
                e = min(int(np.abs(La - Lc).max()) >> 4, 255)
            assert info["e"][y, x] == e, (x, y)


@pytest.mark.parametrize("fmt,linear", [(F.RGB10, False), (F.F16, False), (F.F16, True), (F.P010, False),
                                        (F.NV12, False)])
def test_static_pair_reproduces_the_frame(fmt, linear, luts):
    """Zero motion: every warp tap lands on a texel (the P010 chroma of a block at its centre texel), e = 0, and the
    blend returns the frame's own codes / halves."""
    a, _ = scenes.pair_translate(3, w=64, h=48, sprite=16, sprite_pos=(20, 10))
    enc = {F.RGB10: F.encode_rgb10, F.P010: F.encode_p010, F.NV12: F.encode_nv12}.get(fmt,
                                                                                   lambda x: F.encode_f16(x, linear))
    A = enc(a)
    lut = luts[linear] if fmt == F.F16 else None
    est = F.estimate(A, A, fmt, lut=lut)
    assert not est["cut"] and np.all(est["pf"] == 0)
    for t in (0.5, 0.25):
        out, why, info = F.synth(A, A, est, t, fmt, lut=lut, yuv=F.yuv_constants(2))
        assert why is None and info["count"] == 0 and not info["e"].any()
        if fmt in F.YUV:
            assert np.array_equal(out[0], A[0]) and np.array_equal(out[1], A[1])
        else:
            assert np.array_equal(out, A)


def test_encoders_follow_the_reference_scales():
    v = np.array([[[0, 0, 0, 255], [255, 255, 255, 255], [128, 64, 32, 255]]], np.uint8)
    w = F.encode_rgb10(v)
    assert (w[0, 0] & 1023) == round(F.pq_enc(0.0) * 1023) and (w[0, 1] & 1023) == round(F.pq_enc(1000.0) * 1023)
    assert (w >> 30).tolist() == [[3, 3, 3]]
    h = F.encode_f16(v, linear=True).view(np.float16)
    assert float(h[0, 1, 0]) == float(np.float16(1000.0 / 80.0)) and float(h[0, 0, 3]) == 1.0
    y, cbcr = F.encode_p010(np.repeat(np.repeat(v, 2, 0), 2, 1)[:, :4])
    assert y.shape == (2, 4) and cbcr.shape == (1, 2, 2)
    assert (y[0, 0] >> 6) == 64                    # black at limited range
    assert abs(int(cbcr[0, 0, 0] >> 6) - 512) <= 1 and abs(int(cbcr[0, 1, 1] >> 6) - 512) <= 1   # neutral chroma
