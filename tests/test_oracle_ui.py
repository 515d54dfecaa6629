"""Pins of the CrosshairV2 oracle (oracle/ui.py) without a GPU.

The reference commits no vectors for CrosshairV2, so each function is pinned by what fixes it independently:
  * the box geometry by the examples framegen.h:247-257 documents and by the library's own host-only C functions;
  * the contrast median by scipy.ndimage.median_filter (nearest = clamped border);
  * the state machine by scenario invariants: a static contrasty crosshair enters the mask after exactly uiStillFrames
    still frames, a flat still area never enters, a masked pixel leaves after uiLeaveFrames changed frames and comes
    back after uiReentryFrames frames of its confirmed colour; the mask is the 8-neighbourhood dilation (scipy) gated
    by the pair;
  * the tap rejection by its purpose: with the background's exact half-vectors, the generated frame equals the
    background-only one wherever not both taps land on the crosshair, while the unprotected frame shows the doubled
    crosshair there.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi

from fgr import scenes
from oracle import ui, v1

import fgr


def s_v2(**kw):
    s = ui.defaults()
    s["uiProtection"] = ui.UI_V2
    s.update(kw)
    return s


def lib_settings(s):
    return fgr.default_settings(**{k: v for k, v in s.items()})


@pytest.mark.parametrize("H,side", [(960, 24), (1080, 27), (1440, 36), (2160, 54)])
def test_box_side_is_the_documented_fraction_of_the_height(H, side):
    s = s_v2()
    W = H * 16 // 9
    assert ui.box_size(s, W, H) == (side, side)
    x, y, w, h = ui.box_rect(s, W, H)
    assert (w, h) == (side, side) and x == (W - side) // 2 and y == (H - side) // 2
    assert ui.work_rect(s, W, H) == (x - 1, y - 1, side + 2, side + 2)


def test_box_legacy_area_explicit_sizes_and_clamps():
    s = s_v2(uiBoxHeightFrac=0.0)
    assert ui.box_size(s, 2560, 1440) == (333, 333)            # framegen.h:254-256 example
    assert ui.box_size(s_v2(uiBoxWidth=2, uiBoxHeight=5000), 640, 360) == (4, 360)
    assert ui.box_rect(s_v2(uiBoxOffsetX=10000, uiBoxOffsetY=-10000), 640, 360) == (640 - 9, 0, 9, 9)
    assert ui.work_rect(s_v2(uiBoxOffsetX=10000, uiBoxOffsetY=-10000), 640, 360) == (640 - 10, 0, 10, 10)
    assert ui.box_margin(s_v2()) == 1 and ui.box_margin(s_v2(uiProtection=ui.UI_CROSSHAIR)) == 2
    assert ui.box_margin(s_v2(uiProtection=ui.UI_OFF)) == 0


@pytest.mark.parametrize("kw,size", [({}, (1920, 1080)), (dict(uiBoxHeightFrac=0.0), (2560, 1440)),
                                     (dict(uiBoxWidth=40, uiBoxHeight=17, uiBoxOffsetX=-300), (640, 360)),
                                     (dict(uiBoxOffsetX=5000, uiBoxOffsetY=7), (333, 201))])
def test_library_box_geometry_matches_the_oracle(kw, size):
    s = s_v2(**kw)
    L = lib_settings(s)
    assert fgr.box_size(L, *size) == ui.box_size(s, *size)
    assert fgr.box_rect(L, *size) == ui.box_rect(s, *size)
    assert fgr.work_rect(L, *size) == ui.work_rect(s, *size)
    assert fgr.box_margin(L) == 1


def test_median7_is_scipys_median_filter():
    rng = np.random.default_rng(1)
    frame = rng.integers(0, 256, (40, 50, 4), dtype=np.uint8)
    rect = (30, 20, 20, 20)   # touching the right and bottom edges: clamped windows
    got = ui.median7(frame, rect)
    for c in range(3):
        want = ndi.median_filter(frame[..., c].astype(np.int64), size=7, mode="nearest")
        assert np.array_equal(got[..., c], want[20:40, 30:50])


def crosshair_scene(n, W=96, H=64, seed=3, arm=2):
    """n frames: a moving noise background (a new random shift every frame) and a static white cross at the centre"""
    rng = np.random.default_rng(seed)
    bg = rng.integers(40, 200, (H + 64, W + 64, 4), dtype=np.uint8)
    bg[..., 3] = 255
    frames = []
    cx, cy = W // 2, H // 2
    for k in range(n):
        dx, dy = int(rng.integers(0, 32)), int(rng.integers(0, 32))
        f = bg[dy:dy + H, dx:dx + W].copy()
        f[cy - arm // 2:cy + arm // 2 + arm % 2, cx - 6:cx + 6, :3] = 255
        f[cy - 6:cy + 6, cx - arm // 2:cx + arm // 2 + arm % 2, :3] = 255
        frames.append(f)
    return frames


def cross_mask(W, H, arm=2):
    m = np.zeros((H, W), bool)
    cx, cy = W // 2, H // 2
    m[cy - arm // 2:cy + arm // 2 + arm % 2, cx - 6:cx + 6] = True
    m[cy - 6:cy + 6, cx - arm // 2:cx + arm // 2 + arm % 2] = True
    return m


def run_states(frames, s):
    H, W = frames[0].shape[:2]
    st = ui.V2State(ui.work_rect(s, W, H), ui.box_rect(s, W, H))
    out = []
    for k, f in enumerate(frames):
        st.update(f, s, reset=(k == 0))
        out.append((((st.w1 >> 24) & 1).astype(bool), st.mask()))
    return st, out


def test_static_crosshair_enters_after_still_frames_and_the_background_never():
    s = s_v2(uiBoxWidth=24, uiBoxHeight=24)
    frames = crosshair_scene(12)
    H, W = frames[0].shape[:2]
    x0, y0, w, h = ui.work_rect(s, W, H)
    cross = cross_mask(W, H)[y0:y0 + h, x0:x0 + w]
    _, seq = run_states(frames, s)
    for k, (inm, mask) in enumerate(seq):
        if k < 8:   # frame 0 resets, frames 1..7 count 1..7 still frames
            assert not inm.any(), k
        else:
            assert np.array_equal(inm, cross), k
            grown = ndi.binary_dilation(cross, np.ones((3, 3), bool))
            # the pair gate: the cross is still (255); the grown ring over the moving background changed (64)
            assert np.array_equal(mask != 0, grown)
            assert (mask[cross] == 255).all()
            ring = grown & ~cross
            assert (mask[ring] == 64).mean() > 0.9


def test_flat_still_area_never_enters():
    s = s_v2(uiBoxWidth=24, uiBoxHeight=24)
    f = np.full((64, 96, 4), 90, np.uint8)
    _, seq = run_states([f] * 20, s)
    assert not seq[-1][0].any() and not seq[-1][1].any()
    s0 = s_v2(uiBoxWidth=24, uiBoxHeight=24, uiContrastThreshold=0.0)   # no contrast gate: everything still enters
    _, seq = run_states([f] * 20, s0)
    inm = seq[-1][0]
    assert inm[1:-1, 1:-1].all()                  # the whole box; the 1 px ring is never detected
    assert not inm[0].any() and not inm[-1].any() and not inm[:, 0].any() and not inm[:, -1].any()


def test_leave_after_changed_frames_and_fast_reentry():
    s = s_v2(uiBoxWidth=24, uiBoxHeight=24)
    frames = crosshair_scene(26)
    H, W = frames[0].shape[:2]
    x0, y0, w, h = ui.work_rect(s, W, H)
    cross = cross_mask(W, H)[y0:y0 + h, x0:x0 + w]
    hidden = []
    for k in range(26):
        f = frames[k].copy()
        if 12 <= k < 18:   # the cross goes dark (a flash) for 6 frames
            f[cross_mask(W, H), :3] = 10
        hidden.append(f)
    _, seq = run_states(hidden, s)
    # frames 12..14: changed once / still dark (still but not matching) -> in the mask until 4 changed frames
    assert np.array_equal(seq[11][0], cross)
    assert np.array_equal(seq[12][0], cross)      # 1 changed frame (leave 1 < 4)
    # frames 13..17 the dark cross is still, so it stays (leave resets to 0 on a still frame), with its colour stored
    assert np.array_equal(seq[17][0], cross)
    # back to white at 18: a change (leave 1); the stored colour is now the dark one
    # a crosshair that changes 4 frames in a row leaves
    s2 = s_v2(uiBoxWidth=24, uiBoxHeight=24, uiLeaveFrames=2)
    flick = []
    for k in range(16):
        f = frames[k].copy()
        if k >= 12:
            f[cross_mask(W, H), :3] = 10 if k % 2 == 0 else 250
        flick.append(f)
    _, seq2 = run_states(flick, s2)
    assert np.array_equal(seq2[12][0], cross)     # one changed frame
    assert not seq2[13][0].any()                  # two in a row: out
    # fast re-entry: white again (its confirmed colour) for uiReentryFrames = 2 contrasty frames
    back = flick[:14] + [frames[14], frames[15]]
    _, seq3 = run_states(back, s2)
    assert not seq3[13][0].any()
    assert not seq3[14][0].any()                  # 1 matching frame
    assert np.array_equal(seq3[15][0], cross)     # 2: back in at once (not after 8 still frames)


def test_footprint_and_hits_closed_forms():
    X = np.array([32 * 5, 32 * 5 + 7, -32, -1])
    Y = np.array([32 * 2, 32 * 2, 0, 31])
    lx, ly, hx, hy = ui.footprint(X, Y)
    assert list(lx) == [5, 5, -1, -1] and list(hx) == [5, 6, -1, 0]
    assert list(ly) == [2, 2, 0, 0] and list(hy) == [2, 2, 0, 1]
    mask = np.zeros((4, 4), np.uint8)
    mask[1, 2] = 64
    rect = (4, 1, 4, 4)   # mask pixel (2, 1) is frame pixel (6, 2)
    got = ui.hits(mask, rect, lx, ly, hx, hy)
    assert list(got) == [False, True, False, False]
    assert not ui.hits(None, (0, 0, 0, 0), lx, ly, hx, hy).any()


def test_rejection_removes_the_doubled_crosshair():
    """Background: an even integer shift, so away from the frame edges the half-vectors are exact and a generated
    pixel's two taps read the same background point. The same estimate (of the background alone) drives three synths:
    of the background, of the background with a static crosshair unprotected, and protected."""
    W, H = 160, 96
    a, c = scenes.pair_translate(5, w=W, h=H, sprite=0, sprite_pos=(0, 0), shift=(8, 0))
    est = v1.estimate(a, c)
    cm = cross_mask(W, H, arm=2)
    A, C = a.copy(), c.copy()
    A[cm, :3] = 255
    C[cm, :3] = 255
    mask = cm.astype(np.uint8) * 255
    rect = (0, 0, W, H)
    t = 0.5
    bg, _, _ = v1.synth(a, c, est, t)
    raw, _, _ = v1.synth(A, C, est, t)
    prot, _, info = v1.synth(A, C, est, t, ui=(mask, rect, mask, rect))
    prot = ui.patch(prot, C, mask, rect)
    # pixels whose taps both land on the cross, or that the patch restores: the cross itself
    interior = np.zeros((H, W), bool)
    interior[8:-8, 16:-16] = True
    differs_raw = (raw[..., :3] != bg[..., :3]).any(-1) & interior & ~cm
    differs_prot = (prot[..., :3] != bg[..., :3]).any(-1) & interior & ~cm
    assert differs_raw.sum() > 20          # the doubled crosshair, without protection
    assert differs_prot.sum() == 0         # gone with the rejection
    assert np.array_equal(prot[cm, :3], C[cm, :3])


def test_luma_fill_closed_forms():
    """ui_v2_lumafill.comp: a cell with a UI pixel = the float32 mean of its other pixels' integer lumas; a cell all UI
    = the mean over the cell widened by S a side; cells without UI unchanged"""
    rng = np.random.default_rng(3)
    f = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    S = 2
    l0 = v1.luma_level0(f, S, 8, 8)
    y8 = v1.luma8(f)
    mask = np.zeros((6, 6), np.uint8)      # the rectangle (4, 4, 6, 6)
    mask[0, 0] = 255                         # frame pixel (4, 4): cell (2, 2) has one UI pixel
    mask[2:4, 2:4] = 64                      # frame pixels (6..7, 6..7): cell (3, 3) is all UI
    out = ui.luma_fill(l0, f, mask, (4, 4, 6, 6), S)
    assert out[2, 2] == np.float32(np.float32(y8[4:6, 4:6].sum() - y8[4, 4]) / np.float32(3))
    win = y8[4:10, 4:10].astype(np.int64)
    ui_px = np.zeros((6, 6), bool)
    ui_px[0, 0] = True
    ui_px[2:4, 2:4] = True
    assert out[3, 3] == np.float32(np.float32(win[~ui_px].sum()) / np.float32((~ui_px).sum()))
    keep = np.ones((8, 8), bool)
    keep[2, 2] = keep[3, 3] = False
    assert np.array_equal(out[keep], l0[keep])


def test_format_levels_are_the_float_forms():
    """ui.levels pins FG_STILL / fgStillYuv exactly: Rgb10 (510 code + 1023) // 2046 equals floor(code / 1023 * 255 +
    1/2) in float32 for every code (no ties); P010 (code + 2) >> 2 equals floor(code / 4 + 1/2); Nv12 codes as they are;
    RgbaF16 the rounded metric of the table ((L16 + 8) >> 4 = floor(L16 / 16 + 1/2))."""
    codes = np.arange(1024, dtype=np.int64)
    rgb10 = (codes | (codes[::-1] << 10) | ((codes * 7 % 1024) << 20)).astype(np.uint32).reshape(1, -1)
    lv = ui.levels(rgb10, 1)
    f = np.floor((codes.astype(np.float32) / np.float32(1023.0) * np.float32(255.0)).astype(np.float32) + np.float32(0.5))
    assert np.array_equal(lv[0, :, 0], f.astype(np.int64))
    y = (codes << 6).astype(np.uint16).reshape(2, 512)
    cc = (np.stack([codes[:256], codes[256:512]], -1) << 6).astype(np.uint16).reshape(1, 256, 2)
    lv = ui.levels((y, cc), 4)
    assert np.array_equal(lv[..., 0].ravel(), np.minimum(np.floor(codes / 4 + 0.5), 255).astype(np.int64))
    y8 = np.arange(256, dtype=np.uint8).reshape(2, 128)
    c8 = np.full((1, 64, 2), 77, np.uint8)
    assert np.array_equal(ui.levels((y8, c8), 3)[..., 0], y8.astype(np.int64))
    lut = np.arange(-100, 65436, dtype=np.int64)   # a table of 1/16-level metrics
    h = np.array([[[0, 50, 100, 0], [108, 116, 5000, 0]]], np.uint16)
    got = ui.levels(h, 2, lut)
    want = np.clip(np.floor(lut[h[..., :3]] / 16.0 + 0.5), 0, 255).astype(np.int64)
    assert np.array_equal(got, want)
