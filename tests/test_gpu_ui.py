"""GPU parity of CrosshairV2 static-UI protection (Rgba8; the other formats: test_gpu_ui_formats.py) through the C ABI.

Bar: the box state words and every kept mask bit-exact against oracle/ui.py after every observed / tracked frame; a
protected synth (tap rejection, scene-cut / flick fallbacks, paste-back) bit-exact against the float32 oracle and within
the FP64 bar; the implicit observes of a host that never observes give the same frames as explicit ones; the refusals
of framegen.h:575-693 (evicted entries, modes not built, batches).
"""
import numpy as np
import pytest
import torch

import fgr
from oracle import ui, v1
from gpu_util import Internals
from test_gpu_v1 import check_estimate, oracle_settings

pytestmark = pytest.mark.gpu

W, H = 192, 112


def scene(n, seed=7, W=W, H=H, step=(3, 1), arm=2, cut_at=None):
    """n frames of a textured background panning by `step` px a frame under a static white cross at the centre"""
    rng = np.random.default_rng(seed)
    pad = 8 + n * max(abs(step[0]), abs(step[1]))
    base = rng.integers(30, 220, (H // 4 + pad, W // 4 + pad, 3)).astype(np.float64)
    big = np.kron(base, np.ones((4, 4, 1)))   # blocky texture: a clear optical flow
    big += rng.normal(0, 6, big.shape)
    big = np.clip(big, 0, 255).astype(np.uint8)
    cx, cy = W // 2, H // 2
    frames = []
    for k in range(n):
        ox, oy = pad + step[0] * k, pad + step[1] * k
        f = np.empty((H, W, 4), np.uint8)
        f[..., :3] = big[oy:oy + H, ox:ox + W] if cut_at is None or k < cut_at else 255 - big[oy:oy + H, ox:ox + W]
        f[..., 3] = 255
        f[cy - arm // 2:cy + arm // 2 + arm % 2, cx - 7:cx + 7, :3] = 250
        f[cy - 7:cy + 7, cx - arm // 2:cx + arm // 2 + arm % 2, :3] = 250
        frames.append(f)
    return frames


def v2_settings(**kw):
    kw.setdefault("uiBoxWidth", 28)
    kw.setdefault("uiBoxHeight", 24)
    return fgr.default_settings(uiProtection=fgr.UI_CROSSHAIR_V2, **kw)


def ui_dict(s):
    return {k: v for k, v in s.as_dict().items() if k.startswith("ui")}


def state_of(fi):
    st, _ = fi.debug_ui_mask(None)
    w = st.cpu().numpy().view(np.uint32).astype(np.int64)
    return w[..., 0], w[..., 1], w[..., 2]


def mask_of(fi, f):
    m = fi.debug_ui_mask(f)
    assert m is not None
    return m[0].cpu().numpy(), m[1]


def test_observe_state_and_masks_match_the_oracle():
    frames = scene(14)
    s = v2_settings(uiBoxOffsetX=5, uiBoxOffsetY=-3)
    so = ui_dict(s)
    fi = fgr.Interpolator(W, H, settings=s)
    work, box = fgr.work_rect(s, W, H), fgr.box_rect(s, W, H)
    assert work == ui.work_rect(so, W, H) and box == ui.box_rect(so, W, H)
    st = ui.V2State(work, box)
    devs = [torch.from_numpy(f).cuda() for f in frames]
    entered = False
    for k, f in enumerate(frames):
        print(f"observe {k}", flush=True)
        assert fi.record_observe(devs[k])
        st.update(f, so, reset=(k == 0))
        w0, w1, w2 = state_of(fi)
        assert np.array_equal(w0, st.w0) and np.array_equal(w1, st.w1) and np.array_equal(w2, st.w2), k
        m, rect = mask_of(fi, devs[k])
        assert rect == work and np.array_equal(m, st.mask()), k
        entered = entered or bool((st.mask() == 255).any())
    assert entered
    fi.close()


def run_protected(frames, s, ts, explicit=True, blur=None):
    """observe every frame (or rely on the implicit observes), estimate the last pair, synthesise at ts"""
    fi = fgr.Interpolator(W, H, settings=s)
    devs = [torch.from_numpy(f).cuda() for f in frames]
    if explicit:
        for d in devs:
            assert fi.record_observe(d)
    else:
        for k in range(len(devs) - 2):
            fi.begin_frame()
            assert fi.record_estimate(devs[k], devs[k + 1])
    fi.begin_frame()
    assert fi.record_estimate(devs[-2], devs[-1])
    outs, counts = [], []
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda")
    for t in ts:
        o = torch.zeros_like(devs[0])
        if blur is None:
            assert fi.record_synth(devs[-2], devs[-1], o, t)
        else:
            assert fi.record_synth_blur(devs[-2], devs[-1], o, t[0], t[1])
        assert fi.record_debug_counters(cnt)
        outs.append(o.cpu().numpy())
        counts.append(cnt.cpu().numpy().copy())
    masks = (mask_of(fi, devs[-2]), mask_of(fi, devs[-1]))
    fi.close()
    return outs, counts, masks


def oracle_masks(frames, s):
    so = ui_dict(s)
    st = ui.V2State(ui.work_rect(so, W, H), ui.box_rect(so, W, H))
    masks = []
    for k, f in enumerate(frames):
        st.update(f, so, reset=(k == 0))
        masks.append(st.mask())
    return masks, ui.work_rect(so, W, H)


def oracle_estimate(a, c, s, ma, mc, rect):
    os_ = oracle_settings(s)
    if not s.uiEstimateFill:
        return v1.estimate(a, c, os_)
    S = s.flowScale
    fw, fh, _ = v1.layout(W, H, S, s.pyramidMinSide)
    la = ui.luma_fill(v1.luma_level0(a, S, fw, fh), a, ma, rect, S)
    lc = ui.luma_fill(v1.luma_level0(c, S, fw, fh), c, mc, rect, S)
    return v1.estimate(a, c, os_, luma0=(la, lc), size=(W, H))


def check_protected(frames, s, ts, outs, counts, masks):
    om, rect = oracle_masks(frames, s)
    ma, mc = om[-2], om[-1]
    assert np.array_equal(masks[0][0], ma) and np.array_equal(masks[1][0], mc)
    assert (mc == 255).any(), "the scene must hold a detected crosshair"
    a, c = frames[-2], frames[-1]
    est = oracle_estimate(a, c, s, ma, mc, rect)
    for t, out, cnt in zip(ts, outs, counts):
        ref, why, info = v1.synth(a, c, est, t, ui=(ma, rect, mc, rect))
        ref = ui.patch(ref, c, mc, rect)
        diff = np.argwhere(out != ref)
        assert diff.size == 0, f"t={t} ({why}): {len(diff)} differ, first {tuple(diff[0])}: {out[tuple(diff[0])]} vs {ref[tuple(diff[0])]}"
        if why is None:
            assert int(cnt[0]) == info["count"]
            r64, _, _ = v1.synth(a, c, est, t, exact64=True, ui=(ma, rect, mc, rect))
            r64 = ui.patch(r64, c, mc, rect)
            err = np.abs(out[..., :3].astype(np.int64) - r64[..., :3].astype(np.int64))
            assert err.max() <= 2 and (err <= 1).mean() >= 0.999


def test_protected_synth_matches_the_oracle():
    frames = scene(11)
    s = v2_settings()
    ts = (0.5, 0.25, 0.75)
    outs, counts, masks = run_protected(frames, s, ts)
    check_protected(frames, s, ts, outs, counts, masks)
    # the protection changed the frame (rejected taps and pasted pixels) against the same pair unprotected
    fu = fgr.Interpolator(W, H)
    a, c = torch.from_numpy(frames[-2]).cuda(), torch.from_numpy(frames[-1]).cuda()
    o = torch.zeros_like(a)
    assert fu.record(a, c, o, 0.5)
    assert (o.cpu().numpy() != outs[0]).any()
    fu.close()


def test_estimate_fill_matches_the_oracle():
    """uiEstimateFill: the UI cells of both frames' level-0 luma re-averaged over their not-UI pixels, every estimate
    intermediate and the frames bit-exact"""
    frames = scene(11, seed=13, step=(2, -1))
    s = v2_settings(uiEstimateFill=1)
    fi = fgr.Interpolator(W, H, settings=s)
    devs = [torch.from_numpy(f).cuda() for f in frames]
    for d in devs:
        assert fi.record_observe(d)
    assert fi.record_estimate(devs[-2], devs[-1])
    I = Internals(fi)
    om, rect = oracle_masks(frames, s)
    est = oracle_estimate(frames[-2], frames[-1], s, om[-2], om[-1], rect)
    check_estimate(I, est)
    plain = v1.estimate(frames[-2], frames[-1], oracle_settings(s))
    assert not np.array_equal(est["luma_a"][0], plain["luma_a"][0])   # the fill changed the crosshair's cells
    fi.close()
    outs, counts, masks = run_protected(frames, s, (0.5, 0.25))
    check_protected(frames, s, (0.5, 0.25), outs, counts, masks)


def test_implicit_observes_match_explicit_ones():
    frames = scene(11, seed=9, step=(-2, 2))
    s = v2_settings()
    ts = (0.5, 0.3)
    exp = run_protected(frames, s, ts, explicit=True)
    imp = run_protected(frames, s, ts, explicit=False)
    for a, b in zip(exp[0], imp[0]):
        assert np.array_equal(a, b)
    check_protected(frames, s, ts, *imp)


def test_protected_scene_cut_and_blur():
    frames = scene(11, seed=4, cut_at=10)   # the last pair is a scene cut: the near frame, then the paste
    s = v2_settings()
    outs, counts, masks = run_protected(frames, s, (0.25, 0.75))
    check_protected(frames, s, (0.25, 0.75), outs, counts, masks)
    frames = scene(11, seed=5)
    s = v2_settings(blurSamples=3)
    outs, _, masks = run_protected(frames, s, [(0.2, 0.8)], blur=True)
    om, rect = oracle_masks(frames, s)
    a, c = frames[-2], frames[-1]
    est = v1.estimate(a, c, oracle_settings(s))
    ref = v1.synth_blur(a, c, est, 0.2, 0.8, 3, gauss=True, ui=(om[-2], rect, om[-1], rect))
    ref = ui.patch(ref, c, om[-1], rect)
    assert np.array_equal(outs[0], ref)


def test_track_advances_the_same_state_as_observe():
    frames = scene(12, seed=11)
    s = v2_settings()
    fa, fb = fgr.Interpolator(W, H, settings=s), fgr.Interpolator(W, H, settings=s)
    devs = [torch.from_numpy(f).cuda() for f in frames]
    for k, d in enumerate(devs):
        assert fa.record_observe(d)
        assert (fb.record_track(d) if k < 10 else fb.record_observe(d))
    for x, y in zip(state_of(fa), state_of(fb)):
        assert np.array_equal(x, y)
    assert np.array_equal(mask_of(fa, devs[-1])[0], mask_of(fb, devs[-1])[0])
    assert fb.debug_ui_mask(devs[5]) is None      # a tracked frame is not kept
    fa.close()
    fb.close()


def test_refusals_and_settings():
    frames = scene(6, seed=2)
    devs = [torch.from_numpy(f).cuda() for f in frames]
    # modes and formats that are not built, the estimate fill
    assert "Off and CrosshairV2" in fgr.settings_error_for_format(fgr.default_settings(uiProtection=1))
    assert "Off and CrosshairV2" in fgr.settings_error_for_format(fgr.default_settings(uiProtection=2))
    assert fgr.settings_error_for_format(v2_settings(), fgr.FORMAT_RGBAF16) is None
    assert fgr.settings_error_for_format(v2_settings(uiEstimateFill=1)) is None
    assert not fgr.validate(v2_settings(uiHistory=2)) and not fgr.validate(v2_settings(uiLeaveFrames=0))
    # Off: observe and track are accepted and record nothing
    fo = fgr.Interpolator(W, H)
    assert fo.record_observe(devs[0]) and fo.record_track(devs[1])
    fo.close()
    # an evicted entry refuses the synths of its pair (uiHistory 3: two more observes evict the pair's prev)
    fi = fgr.Interpolator(W, H, settings=v2_settings())
    for d in devs[:2]:
        assert fi.record_observe(d)
    assert fi.record_estimate(devs[0], devs[1])
    o = torch.zeros_like(devs[0])
    assert fi.record_synth(devs[0], devs[1], o, 0.5)
    assert fi.record_observe(devs[2]) and fi.record_observe(devs[3])
    assert not fi.record_synth(devs[0], devs[1], o, 0.5)
    # explicit observes: a pair with an unobserved frame runs unprotected (the plain V1 frame)
    assert fi.record_estimate(devs[3], devs[4])
    o.zero_()
    assert fi.record_synth(devs[3], devs[4], o, 0.5)
    fu = fgr.Interpolator(W, H)
    o2 = torch.zeros_like(devs[0])
    assert fu.record(devs[3], devs[4], o2, 0.5)
    assert torch.equal(o, o2)
    fu.close()
    # a change of the box drops the estimate
    assert fi.record_observe(devs[4]) and fi.record_estimate(devs[3], devs[4])
    assert fi.set_settings(v2_settings(uiBoxOffsetX=4))
    assert not fi.record_synth(devs[3], devs[4], o, 0.5)
    fi.close()
    # batches run unprotected only
    with pytest.raises(RuntimeError, match="uiProtection Off"):
        fgr.Batch([(devs[0], devs[1], 2, [torch.zeros_like(devs[0])])], settings=v2_settings())
