"""GPU parity of the V1 path against the CPU oracle, through the C ABI (libfgr.so).

Bar: every intermediate of the estimate (luma pyramid, raw and median-filtered block fields per level, global
medians, sub-pixel field, scene-cut SAD sum, per-pixel field) and the cut / flick decisions are bit-exact; the RGBA8
output is bit-exact against the float32 oracle and within +-1 level on >= 99.9 % of pixels and +-2 everywhere of the
FP64 oracle.
"""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from gpu_util import Internals, dev
from oracle import v1

pytestmark = pytest.mark.gpu


def _mismatch(name, got, want):
    got, want = np.asarray(got), np.asarray(want)
    if got.shape != want.shape:
        return f"{name}: shape {got.shape} != {want.shape}"
    bad = got != want
    if bad.any():
        idx = np.argwhere(bad)[0]
        return f"{name}: {int(bad.sum())} of {bad.size} differ, first at {tuple(idx)}: {got[tuple(idx)]} vs {want[tuple(idx)]}"
    return None


def check_estimate(I, est):
    errs = []
    s = est["settings"]
    for i in range(len(est["levels"])):
        errs.append(_mismatch(f"luma a L{i}", I.luma_level(i, "a"), est["luma_a"][i]))
        errs.append(_mismatch(f"luma c L{i}", I.luma_level(i, "c"), est["luma_c"][i]))
        if est["medians"][i] is not None:
            errs.append(_mismatch(f"median L{i}", I.median(i), est["medians"][i]))
        errs.append(_mismatch(f"search L{i}", I.fs(i), est["fields_raw"][i]))
        errs.append(_mismatch(f"median3 L{i}", I.fm(i), est["fields"][i]))
    if s["subpixel"]:
        errs.append(_mismatch("subpix", I.sub(), est["block_field"]))
    errs.append(_mismatch("sadsum", I.sadsum(), est["sadsum"]))
    errs.append(_mismatch("pixfield", I.pf(), est["pf"]))
    errs = [e for e in errs if e]
    assert not errs, "\n".join(errs)


def check_output(out, a, c, est, t):
    ref, why, info = v1.synth(a, c, est, t)
    ref64, _, _ = v1.synth(a, c, est, t, exact64=True)
    d64 = np.abs(out.astype(np.int64) - ref64.astype(np.int64))
    assert d64.max() <= 2 and (d64 <= 1).mean() >= 0.999, f"t={t}: vs FP64 max {d64.max()}"
    e = _mismatch(f"output t={t} ({why})", out, ref)
    assert e is None, e
    return why, info


def run(a, c, ts, settings=None, **kw):
    H, W = a.shape[:2]
    fi = fgr.Interpolator(W, H, settings=settings, **kw)
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    I = Internals(fi)
    outs, counts = [], []
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    for t in ts:
        out = torch.zeros_like(ta)
        assert fi.record_synth(ta, tc, out, t)
        assert fi.record_debug_counters(cnt)
        outs.append(out.cpu().numpy())
        counts.append(cnt.cpu().numpy().astype(np.uint32).copy())
    return fi, I, outs, counts


def oracle_settings(s):
    return v1.settings(**{k: (bool(v) if k == "subpixel" else v) for k, v in s.as_dict().items()
                          if k in v1.DEFAULTS})


@pytest.mark.parametrize("ts", [(0.5,), (0.25, 0.5, 0.75)])
def test_check_pair_640x360(ts):
    a, c = scenes.pair_translate(1)
    fi, I, outs, counts = run(a, c, ts)
    est = v1.estimate(a, c)
    check_estimate(I, est)
    for t, o, cnt in zip(ts, outs, counts):
        why, info = check_output(o, a, c, est, t)
        assert why is None
        assert int(cnt[0]) == info["count"], "untrusted count"
        assert int(cnt[1]) == est["sadsum"]
    fi.close()


@pytest.mark.parametrize("W,H,kw", [
    (203, 117, dict(pyramidMinSide=16)),
    (331, 250, dict(flowScale=1, pyramidMinSide=32)),
    (402, 310, dict(flowScale=4, pyramidMinSide=16)),
    (256, 160, dict(subpixel=0)),
    (300, 200, dict(searchPenalty=0.3, searchRadius=3, searchRadiusTop=8, smoothBonus=0.2, hudBonus=2.0,
                    trustLow=10.0, trustHigh=30.5, pyramidMinSide=24)),
])
def test_sizes_and_settings(W, H, kw):
    a, c = scenes.pair_translate(7, w=W, h=H, sprite=32, sprite_pos=(W // 3, H // 3), shift=(3, -2))
    s = fgr.default_settings(**kw)
    fi, I, outs, _ = run(a, c, (0.5, 0.3), settings=s)
    est = v1.estimate(a, c, oracle_settings(s))
    check_estimate(I, est)
    for t, o in zip((0.5, 0.3), outs):
        check_output(o, a, c, est, t)
    fi.close()


def test_scene_cut_gives_near_frame():
    a, c = scenes.noise_pair(3, 320, 192)
    fi, I, outs, _ = run(a, c, (0.25, 0.5, 0.75))
    est = v1.estimate(a, c)
    assert est["cut"]
    check_estimate(I, est)
    for t, o in zip((0.25, 0.5, 0.75), outs):
        assert np.array_equal(o, a if t <= 0.5 else c)
    fi.close()


@pytest.mark.parametrize("size", [(320, 200), (203, 117)])
def test_flick_gives_near_frame(size):
    w, h = size
    a, c = scenes.pair_translate(4, w=w, h=h, sprite=48, sprite_pos=(w // 3, h // 3))
    s = fgr.default_settings(globalFallback=1e-5)
    fi, I, outs, counts = run(a, c, (0.5, 0.75), settings=s)
    est = v1.estimate(a, c, oracle_settings(s))
    for t, o, cnt in zip((0.5, 0.75), outs, counts):
        ref, why, info = v1.synth(a, c, est, t)
        assert why == "untrusted" and int(cnt[0]) == info["count"]
        assert np.array_equal(o, a if t <= 0.5 else c)
    fi.close()


def test_static_pair_is_exact():
    a, _ = scenes.pair_translate(5, w=256, h=144)
    fi, I, outs, _ = run(a, a.copy(), (0.5, 0.25))
    for o in outs:
        assert np.array_equal(o, a)
    fi.close()


@pytest.mark.parametrize("shift", [(0, 0), (4, -2)])
def test_uniform_fields_match_the_oracle(shift):
    # a static pair and an integer pan: block fields uniform away from the borders, so the interior pixel-field tiles
    # have every candidate's vector equal to the own block's (the kernel reuses the own-block errors and window sums)
    a, c = scenes.pair_shift(11, 320, 192, *shift)
    est = v1.estimate(a, c)
    fi, I, outs, _ = run(a, c, (0.5, 0.25))
    check_estimate(I, est)
    for o, t in zip(outs, (0.5, 0.25)):
        check_output(o, a, c, est, t)
    fi.close()


def test_bgra_order_is_symmetric():
    """README.md:144: RGBA and BGRA give byte-identical results (channels swapped)."""
    a, c = scenes.pair_translate(6, w=320, h=180)
    _, _, o1, _ = run(a, c, (0.5,))
    sw = [2, 1, 0, 3]
    _, _, o2, _ = run(a[..., sw], c[..., sw], (0.5,), bgra=True)
    assert np.array_equal(o1[0][..., sw], o2[0])


def test_refusals_record_nothing():
    a, c = scenes.pair_translate(8, w=128, h=96)
    fi = fgr.Interpolator(128, 96)
    ta, tc = dev(a), dev(c)
    out = torch.full_like(ta, 7)
    assert not fi.record_synth(ta, tc, out, 0.5)             # no estimate yet (framegen.h:547)
    assert fi.record_estimate(ta, tc)
    assert not fi.record_synth(tc, ta, out, 0.5)             # not the estimate's views
    assert not fi.record_synth(ta, tc, out, 0.0)             # t outside (0, 1)
    assert not fi.record_synth(ta, tc, out, 1.0)
    assert not fi.record_synth(ta, tc, out, float("nan"))
    torch.cuda.synchronize()
    assert int(out.min()) == 7 and int(out.max()) == 7       # out untouched
    fi.begin_frame()
    for _ in range(fgr.SET_RING):
        assert fi.record_synth(ta, tc, out, 0.5)
    assert not fi.record_synth(ta, tc, out, 0.5)             # ring full (framegen.h:537-540)
    fi.begin_frame()
    assert fi.record_synth(ta, tc, out, 0.5)
    fi.close()


def test_frames_the_kernels_cannot_address_are_refused():
    """A pitch below a row's bytes, of 2^24 or more, or a null plane: the record returns 0 and enqueues nothing."""
    a, c = scenes.pair_translate(8, w=128, h=96)
    fi = fgr.Interpolator(128, 96)
    ta, tc = dev(a), dev(c)
    out = torch.full_like(ta, 7)
    for pitch in (128 * 4 - 4, 1 << 24):
        bad = fgr.frame_of(tc)
        bad.pitch = pitch
        assert not fi.record_estimate(ta, bad)
        assert not fi.record(ta, bad, out, 0.5)
    null = fgr.frame_of(tc)
    null.ptr = None
    assert not fi.record_estimate(ta, null)
    assert fi.record_estimate(ta, tc)
    o = fgr.frame_of(out)
    o.pitch = 100
    assert not fi.record_synth(ta, tc, o, 0.5)
    assert not fi.record_synth_blur(ta, tc, o, 0.25, 0.75)
    torch.cuda.synchronize()
    assert int(out.min()) == 7 and int(out.max()) == 7
    fi.close()


@pytest.mark.parametrize("k", [0, 7])
def test_1080p_workload_pair(k):
    pairs, _ = scenes.pairs_1080p(2, n_pairs=8)
    a, c = pairs[k]
    ts = (0.25, 0.5, 0.75)
    fi, I, outs, counts = run(a, c, ts)
    est = v1.estimate(a, c)
    check_estimate(I, est)
    assert est["cut"] == (k == 7)
    for t, o in zip(ts, outs):
        check_output(o, a, c, est, t)
    fi.close()


def _read_fgmaps(path):
    """The .fgmaps layout (fgtest.cpp:24-28): 64-byte header, block field, per-pixel field, packed mismatch."""
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:64], np.uint32)
    W, H, fw, fh, bw, bh, fs, es = (int(x) for x in h[2:10])
    o = 64
    block = np.frombuffer(raw, np.float32, bw * bh * 2, o).reshape(bh, bw, 2)
    o += bw * bh * 8
    pf = np.frombuffer(raw, np.float32, fw * fh * 2, o).reshape(fh, fw, 2)
    o += fw * fh * 8
    ew = np.frombuffer(raw, np.uint32, es * H, o).reshape(H, es)
    assert o + es * H * 4 == len(raw)
    e = np.stack([(ew >> (8 * i)) & 255 for i in range(4)], -1).reshape(H, es * 4)[:, :W]
    return h, np.frombuffer(raw[40:56], np.float32), block, pf, e


@pytest.mark.parametrize("t", [0.5, 0.75])
def test_fgmaps_dump_matches_the_oracle(t, tmp_path):
    a, c = scenes.pair_translate(41, w=203, h=117, sprite=32, sprite_pos=(70, 40), shift=(3, -2))
    fi = fgr.Interpolator(203, 117)
    ta, tc = dev(a), dev(c)   # the debug maps read the record's frames: they must still be alive
    out = torch.zeros_like(ta)
    assert fi.record_estimate(ta, tc) and fi.record_synth(ta, tc, out, t)
    path = str(tmp_path / "pair_0000.fgmaps")
    assert fi.dump_fgmaps(path)
    h, hf, block, pf, e = _read_fgmaps(path)
    est = v1.estimate(a, c)
    _, _, info = v1.synth(a, c, est, t)
    assert h[0] == 0x504D4746 and h[1] == 5 and list(h[2:10]) == [203, 117, 102, 58, 13, 8, 2, 51]
    assert list(hf) == [12.0, 40.0, 0.5, 1.0] and h[14] == 0 and h[15] == 0
    assert np.array_equal(block, est["block_field"]) and np.array_equal(pf, est["pf"])
    assert np.array_equal(e, info["e"])
    fi.close()


@pytest.mark.parametrize("kw", [dict(searchRadius=12, searchRadiusTop=12, pyramidMinSide=16),
                                dict(searchRadius=6, searchRadiusTop=12, pyramidMinSide=16)])
def test_wide_reach_settings(kw):
    """Reach settings off their defaults: radius-12 windows (the wide search) and, where the parent level's reach bound
    exceeds 127 flow pixels, the exact two-pass median (medhist.comp FG_WIDE); a large motion (120, -80) px."""
    a, c = scenes.pair_translate(43, w=1280, h=720, sprite=64, sprite_pos=(500, 300), shift=(120, -80))
    s = fgr.default_settings(**kw)
    fi, I, outs, counts = run(a, c, (0.5, 0.25), settings=s)
    est = v1.estimate(a, c, oracle_settings(s))
    n = len(est["levels"])
    assert any(v1.medhist_shift(est["settings"], n, j) >= 0 for j in range(1, n))
    check_estimate(I, est)
    for t, o in zip((0.5, 0.25), outs):
        check_output(o, a, c, est, t)
    fi.close()
