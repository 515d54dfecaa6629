"""GPU parity of the V1 path for Rgb10, RgbaF16 (encoded and linear light), Nv12 and P010 frames, through the C ABI.

Bar: every estimate intermediate bit-exact against the oracle (oracle/formats.py pins the fixed-point luma and the
exact taps of each format), the untrusted count and cut / flick decisions exact, the output bit-identical to the
float32 oracle, and within +-1 level (8-bit scale) on >= 99.9 % and +-2 everywhere of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from gpu_util import Internals
from oracle import formats as F
from oracle import v1
from test_gpu_v1 import check_estimate, oracle_settings

pytestmark = pytest.mark.gpu

_LUTS = {}


def lut_of(fmt, linear, white=80.0):
    if fmt != F.F16:
        return None
    key = (linear, white)
    if key not in _LUTS:
        _LUTS[key] = F.f16_level_table(linear, white)
    return _LUTS[key]


def to_dev(frame, fmt):
    if fmt == F.NV12:
        return (torch.from_numpy(frame[0].copy()).cuda(), torch.from_numpy(frame[1].copy()).cuda())
    if fmt == F.P010:
        y, c = frame
        return (torch.from_numpy(y.view(np.int16).copy()).cuda(), torch.from_numpy(c.view(np.int16).copy()).cuda())
    if fmt == F.RGB10:
        return torch.from_numpy(frame.view(np.int32).copy()).cuda()
    return torch.from_numpy(frame.view(np.int16).copy()).cuda()


def to_host(t, fmt):
    if fmt == F.NV12:
        return (t[0].cpu().numpy(), t[1].cpu().numpy())
    if fmt == F.P010:
        return (t[0].cpu().numpy().view(np.uint16), t[1].cpu().numpy().view(np.uint16))
    if fmt == F.RGB10:
        return t.cpu().numpy().view(np.uint32)
    return t.cpu().numpy().view(np.uint16)


def empty_like(t, fmt, fill=0):
    if fmt in F.YUV:
        return (torch.full_like(t[0], fill), torch.full_like(t[1], fill))
    return torch.full_like(t, fill)


def encode(rgba, fmt, linear=False):
    if fmt == F.RGB10:
        return F.encode_rgb10(rgba)
    if fmt == F.F16:
        return F.encode_f16(rgba, linear)
    if fmt == F.NV12:
        return F.encode_nv12(rgba)
    return F.encode_p010(rgba)


def matrix_of(fmt):
    """the YCbCr matrix of the encoders: Nv12 BT.709 (SDR), P010 BT.2020 (HDR10)"""
    return fgr.YUV_BT709 if fmt == F.NV12 else fgr.YUV_BT2020


CASES = [   # (name, format, linear transfer)
    ("rgb10", F.RGB10, False),
    ("f16", F.F16, False),
    ("f16lin", F.F16, True),
    ("nv12", F.NV12, False),
    ("p010", F.P010, False),
]


def run_fmt(A, C, fmt, ts, linear=False, settings=None, yuv_matrix=None, full_range=False, bgra=False):
    H, W = (A[0] if fmt in F.YUV else A).shape[:2]
    yuv_matrix = matrix_of(fmt) if yuv_matrix is None else yuv_matrix
    fi = fgr.Interpolator(W, H, fmt=fmt, settings=settings, bgra=bgra,
                          transfer=fgr.TRANSFER_LINEAR if linear else fgr.TRANSFER_ENCODED,
                          yuv_matrix=yuv_matrix, yuv_full_range=full_range)
    ta, tc = to_dev(A, fmt), to_dev(C, fmt)
    assert fi.record_estimate(ta, tc)
    I = Internals(fi)
    outs, counts = [], []
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    for t in ts:
        o = empty_like(ta, fmt)
        assert fi.record_synth(ta, tc, o, t)
        assert fi.record_debug_counters(cnt)
        outs.append(to_host(o, fmt))
        counts.append(cnt.cpu().numpy().astype(np.uint32).copy())
    fi.close()
    return I, outs, counts


def same(got, want, fmt):
    if fmt in F.YUV:
        return np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    return np.array_equal(got, want)


def first_diff(got, want, fmt):
    if fmt in F.YUV:
        for k in range(2):
            if not np.array_equal(got[k], want[k]):
                idx = np.argwhere(got[k] != want[k])[0]
                return f"plane {k} at {tuple(idx)}: {got[k][tuple(idx)]} vs {want[k][tuple(idx)]}"
        return None
    idx = np.argwhere(got != want)[0]
    return f"at {tuple(idx)}: {got[tuple(idx)]} vs {want[tuple(idx)]}"


def check_fmt_output(out, A, C, est, t, fmt, lut, yuv, cnt):
    ref, why, info = F.synth(A, C, est, t, fmt, lut=lut, yuv=yuv)
    assert same(out, ref, fmt), f"t={t} ({why}): {first_diff(out, ref, fmt)}"
    if why is None:
        _, _, i64 = F.synth(A, C, est, t, fmt, exact64=True, lut=lut, yuv=yuv)
        err = F.level_error(fmt, out, i64, lut=lut)
        assert err.max() <= 2.0 and (err <= 1.0).mean() >= 0.999, f"t={t}: vs FP64 max {err.max()}"
    if why != "cut":
        assert int(cnt[0]) == info["count"], "untrusted count"
    assert int(cnt[1]) == est["sadsum"]
    return why


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_format_pair(name, fmt, linear):
    W, H = 320, 180
    a, c = scenes.pair_translate(11, w=W, h=H, sprite=40, sprite_pos=(100, 50), shift=(5, -3))
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    ts = (0.5, 0.25, 0.75)
    I, outs, counts = run_fmt(A, C, fmt, ts, linear)
    lut = lut_of(fmt, linear)
    yuv = F.yuv_constants(matrix_of(fmt))
    est = F.estimate(A, C, fmt, lut=lut)
    check_estimate(I, est)
    for t, o, cnt in zip(ts, outs, counts):
        assert check_fmt_output(o, A, C, est, t, fmt, lut, yuv, cnt) is None


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_format_sizes_and_settings(name, fmt, linear):
    W, H = (202, 118) if fmt in F.YUV else (203, 117)
    a, c = scenes.pair_translate(12, w=W, h=H, sprite=32, sprite_pos=(W // 3, H // 3), shift=(3, -2))
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    for kw in (dict(flowScale=1, pyramidMinSide=24), dict(flowScale=4, pyramidMinSide=16, subpixel=0)):
        s = fgr.default_settings(**kw)
        I, outs, counts = run_fmt(A, C, fmt, (0.5, 0.3), linear, settings=s)
        lut = lut_of(fmt, linear)
        est = F.estimate(A, C, fmt, s=oracle_settings(s), lut=lut)
        check_estimate(I, est)
        for t, o, cnt in zip((0.5, 0.3), outs, counts):
            check_fmt_output(o, A, C, est, t, fmt, lut, F.yuv_constants(matrix_of(fmt)), cnt)


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_format_scene_cut(name, fmt, linear):
    a, c = scenes.noise_pair(13, 256, 144)
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    I, outs, counts = run_fmt(A, C, fmt, (0.25, 0.75), linear)
    est = F.estimate(A, C, fmt, lut=lut_of(fmt, linear))
    assert est["cut"]
    check_estimate(I, est)
    assert same(outs[0], A, fmt) and same(outs[1], C, fmt)


@pytest.mark.parametrize("name,fmt,linear", CASES)
def test_format_flick(name, fmt, linear):
    # more untrusted pixels than globalFallback allows: the fallback pass copies the nearer frame over the synth's output
    # (4- and 8-byte pixels as 16-byte chunks of a tile's rows with a word tail at an odd width; 4:2:0 pixel by pixel)
    W, H = (202, 118) if fmt in F.YUV else (203, 117)
    a, c = scenes.pair_translate(14, w=W, h=H, sprite=32, sprite_pos=(W // 3, H // 3), shift=(3, -2))
    A, C = encode(a, fmt, linear), encode(c, fmt, linear)
    s = fgr.default_settings(globalFallback=1e-5)
    I, outs, counts = run_fmt(A, C, fmt, (0.5, 0.75), linear, settings=s)
    lut = lut_of(fmt, linear)
    est = F.estimate(A, C, fmt, s=oracle_settings(s), lut=lut)
    for t, o, cnt in zip((0.5, 0.75), outs, counts):
        assert check_fmt_output(o, A, C, est, t, fmt, lut, F.yuv_constants(matrix_of(fmt)), cnt) == "untrusted"
        assert same(o, A if t <= 0.5 else C, fmt)


def test_p010_full_range_bt709_and_rgb10_bgra():
    a, c = scenes.pair_translate(14, w=256, h=144, sprite=32, sprite_pos=(90, 40), shift=(4, 2))
    A, C = F.encode_p010(a), F.encode_p010(c)
    I, outs, counts = run_fmt(A, C, F.P010, (0.5,), yuv_matrix=fgr.YUV_BT709, full_range=True)
    est = F.estimate(A, C, F.P010, full_range=True)
    check_estimate(I, est)
    check_fmt_output(outs[0], A, C, est, 0.5, F.P010, None, F.yuv_constants(1, True), counts[0])
    # A2R10G10B10: R and B swapped in the word; the luma (the only channel-order-dependent step) follows bgra
    A, C = F.encode_rgb10(a[..., [2, 1, 0, 3]]), F.encode_rgb10(c[..., [2, 1, 0, 3]])
    I, outs, counts = run_fmt(A, C, F.RGB10, (0.5,), bgra=True)
    est = F.estimate(A, C, F.RGB10, bgra=True)
    check_estimate(I, est)
    check_fmt_output(outs[0], A, C, est, 0.5, F.RGB10, None, None, counts[0])


@pytest.mark.parametrize("rows", [32, 16, 64])
def test_mixed_format_batch_matches_per_pair(rows):
    """One batch of RGBA8, Rgb10, RgbaF16, Nv12 and P010 pairs of different sizes and multipliers: one launch per pass for
    the three non-RGBA8 formats together; every output equal to the per-pair path (itself checked above), with synth
    tiles of 32 rows and of the tuner's 16 and 64 rows."""
    specs = [(F.RGBA8, 320, 180, 2), (F.RGB10, 256, 144, 3), (F.F16, 288, 160, 4), (F.P010, 320, 176, 4),
             (F.RGB10, 202, 118, 2), (F.P010, 256, 144, 3), (F.NV12, 224, 128, 3)]
    items, refs = [], []
    for i, (fmt, W, H, n) in enumerate(specs):
        a, c = scenes.pair_translate(20 + i, w=W, h=H, sprite=32, sprite_pos=(W // 3, H // 3), shift=(2 + i, -1))
        A, C = (a, c) if fmt == F.RGBA8 else (encode(a, fmt), encode(c, fmt))
        ta = torch.from_numpy(A).cuda() if fmt == F.RGBA8 else to_dev(A, fmt)
        tc = torch.from_numpy(C).cuda() if fmt == F.RGBA8 else to_dev(C, fmt)
        outs = [empty_like(ta, fmt) for _ in range(n - 1)]
        items.append((ta, tc, n, outs, fmt))
        # per-pair reference through the plain C ABI
        fi = fgr.Interpolator(W, H, fmt=fmt, yuv_matrix=fgr.YUV_BT2020)
        assert fi.record_estimate(ta, tc)
        want = []
        for k in range(1, n):
            o = empty_like(ta, fmt)
            assert fi.record_synth(ta, tc, o, k / n)
            want.append(o.cpu().numpy() if fmt == F.RGBA8 else to_host(o, fmt))
        fi.close()
        refs.append(want)
    b = fgr.Batch(items, use_graph=True, yuv_matrix=fgr.YUV_BT2020)
    assert b.set_tuning(rows, 64) and b.get_tuning()[0] == rows
    b.run()
    b.run()
    torch.cuda.synchronize()
    for p, it in enumerate(items):
        fmt = it[4]
        for k, o in enumerate(it[3]):
            got = o.cpu().numpy() if fmt == F.RGBA8 else to_host(o, fmt)
            assert same(got, refs[p][k], fmt if fmt != F.RGBA8 else F.RGB10), f"pair {p} output {k}"
    b.close()
