"""CPU oracle of CrosshairV2 static-UI protection: the box geometry, the per-pixel detection state machine of an
observed frame, its mask, the tap rejection of the synthesis, the paste-back and the estimate fill. numpy only.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: only tests/ use it, as the checker.

Follows gpu/framegen.cpp:1041-1075 (boxSize, boxRect, boxMargin, workRect), gpu/shaders/ui_v2_update.comp (the state
machine), ui_v2_mask.comp (the mask and its pair gate), ui_reject.glsl (tap rejection: a warp tap whose bilinear
footprint touches its own frame's UI takes the other frame's tap), ui_patch.comp with `binary` set (the pixels of
mask byte 255 become curr's) and ui_v2_lumafill.comp (uiEstimateFill: the UI cells of the estimate's level-0 luma
re-averaged over their not-UI pixels).

The state machine works on 8-bit-equivalent integer levels of each pixel (format.glsl FG_STILL / fgStillYuv), pinned
here exactly: Rgba8 the three bytes; Rgb10 floor(255 code / 1023 + 1/2) = (510 code + 1023) // 2046 (no ties: the float
form the reference evaluates lands on the same integer for every code); RgbaF16 clamp(floor(metric + 1/2), 0, 255) of
the metric in 1/16 level (the table of oracle/formats.py), i.e. (L16 + 8) >> 4 clamped; Nv12 the (Y, Cb, Cr) codes, the
chroma of the pixel's 2x2 block; P010 (code + 2) >> 2 clamped to 255. The contrast of a 4:2:0 pixel converts the
(Y, Cb, Cr) difference to the R'G'B' one first (fgYuvDiff, the constants of formats.yuv_constants, float32 as the
mismatch metric).

State, per pixel of the work rectangle (the box grown by 1 px, clamped to the frame), three words as the shader keeps
them: w0 = previous original (r | g << 8 | b << 16) | still-and-contrasty count << 24, w1 = last confirmed colour |
flags << 24 (bit 0 in the mask, bit 1 has a stored colour, bit 2 still in this pair), w2 = consecutive changed frames
in the mask | consecutive stored-colour matches out of it << 8.
"""
import math

import numpy as np

UI_OFF, UI_CROSSHAIR, UI_WHOLESCREEN, UI_V2 = 0, 1, 2, 3


def defaults():
    """framegen.h:225-305, the UI fields"""
    return dict(uiProtection=UI_OFF, uiHistory=3, uiBoxHeightFrac=0.025, uiBoxWidth=0, uiBoxHeight=0,
                uiCrosshairArea=0.03, uiBoxOffsetX=0, uiBoxOffsetY=0, uiStillFrames=8, uiStillThreshold=3.5,
                uiInpaintReach=12, uiMaxStillFraction=0.75, uiContrastThreshold=20.0, uiLeaveFrames=4,
                uiReentryFrames=2, uiEstimateFill=0)


def _lround(x):
    """std::lround: half away from zero"""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def box_size(s, w, h):
    """framegen.cpp:1041-1050: explicit sizes win over the height fraction (or the legacy area); clamped to the frame"""
    frac = float(np.float32(s["uiBoxHeightFrac"]))
    if frac > 0.0:
        side = min(max(_lround(h * frac), 8), min(w, h))
    else:
        side = min(max(_lround(math.sqrt(float(np.float32(s["uiCrosshairArea"])) * w * h)), 32), min(w, h))
    bw = min(max(s["uiBoxWidth"], 4), w) if s["uiBoxWidth"] else side
    bh = min(max(s["uiBoxHeight"], 4), h) if s["uiBoxHeight"] else side
    return bw, bh


def box_rect(s, w, h):
    """framegen.cpp:1053-1060: centred on the frame centre plus the offset, clamped into the frame"""
    bw, bh = box_size(s, w, h)
    x = min(max((w - bw) // 2 + s["uiBoxOffsetX"], 0), w - bw)
    y = min(max((h - bh) // 2 + s["uiBoxOffsetY"], 0), h - bh)
    return (x, y, bw, bh)


def box_margin(s):
    """framegen.cpp:1065-1067"""
    return 2 if s["uiProtection"] == UI_CROSSHAIR else (1 if s["uiProtection"] == UI_V2 else 0)


def work_rect(s, w, h):
    """framegen.cpp:1069-1075: the box grown by the margin, clamped to the frame"""
    x, y, bw, bh = box_rect(s, w, h)
    m = box_margin(s)
    x0, y0 = max(x - m, 0), max(y - m, 0)
    x1, y1 = min(x + bw + m, w), min(y + bh + m, h)
    return (x0, y0, x1 - x0, y1 - y0)


def _pack(c):
    return c[..., 0] | (c[..., 1] << 8) | (c[..., 2] << 16)


def _unpack(wd):
    return np.stack([wd & 255, (wd >> 8) & 255, (wd >> 16) & 255], -1)


def _maxabs(d):
    return np.abs(d).max(-1)


def levels(frame, fmt=0, lut=None):
    """the (H, W, 3) integer levels of every pixel the state machine reads (see the module header); fmt as
    oracle/formats.py: 0 Rgba8, 1 Rgb10, 2 RgbaF16, 3 Nv12, 4 P010"""
    if fmt == 0:
        return frame[..., :3].astype(np.int64)
    if fmt == 1:
        w = frame.astype(np.int64)
        codes = np.stack([w & 1023, (w >> 10) & 1023, (w >> 20) & 1023], -1)
        return (510 * codes + 1023) // 2046
    if fmt == 2:
        L16 = lut.astype(np.int64)[frame[..., :3].astype(np.uint16)]
        return np.clip((L16 + 8) >> 4, 0, 255)
    yp, cp = frame
    H, W = yp.shape
    ten = fmt == 4
    y = (yp.astype(np.int64) >> 6) if ten else yp.astype(np.int64)
    c = (cp.astype(np.int64) >> 6) if ten else cp.astype(np.int64)
    c = np.repeat(np.repeat(c, 2, 0), 2, 1)[:H, :W]
    lv = np.concatenate([y[..., None], c], -1)
    return np.clip((lv + 2) >> 2, 0, 255) if ten else lv


def median7(frame, rect):
    """per channel, the median of the 7x7 window (frame-clamped positions) around each pixel of rect: the 25th
    smallest of the 49 values (ui_v2_update.comp contrastAt's binary search finds the smallest m with 25 values <= m).
    frame: RGBA8 pixels or an (H, W, 3) level array."""
    x0, y0, w, h = rect
    H, W = frame.shape[:2]
    ys = np.clip(np.arange(y0 - 3, y0 + h + 3), 0, H - 1)
    xs = np.clip(np.arange(x0 - 3, x0 + w + 3), 0, W - 1)
    reg = frame[ys][:, xs, :3].astype(np.int64)
    win = np.stack([reg[j:j + h, i:i + w] for j in range(7) for i in range(7)], 0)
    return np.sort(win, 0)[24]


class V2State:
    """The box state of one handle (the state buffer), advanced by every observed or tracked frame."""

    def __init__(self, work, box):
        self.work, self.box = work, box
        _, _, w, h = work
        self.w0 = np.zeros((h, w), np.int64)
        self.w1 = np.zeros((h, w), np.int64)
        self.w2 = np.zeros((h, w), np.int64)

    def update(self, frame, s, reset, yuv=None):
        """ui_v2_update.comp main(): one observed / tracked frame (RGBA8 pixels or an (H, W, 3) level array; yuv: the
        YCbCr -> R'G'B' constants of a 4:2:0 format, for the contrast)"""
        x0, y0, w, h = self.work
        bx, by, bw, bh = self.box
        now = frame[y0:y0 + h, x0:x0 + w, :3].astype(np.int64)
        thr, cthr = np.float32(s["uiStillThreshold"]), np.float32(s["uiContrastThreshold"])
        ys, xs = np.meshgrid(np.arange(y0, y0 + h), np.arange(x0, x0 + w), indexing="ij")
        in_box = (xs >= bx) & (ys >= by) & (xs < bx + bw) & (ys < by + bh)
        w0, w1, w2 = self.w0, self.w1, self.w2
        # the margin ring: the pair gate only
        ring_still = (~np.bool_(reset)) & (_maxabs(now - _unpack(w0)) <= thr)
        ring = (np.where(ring_still, 4, 0) << 24, np.zeros_like(w2))
        if reset:
            cnt = np.zeros_like(w0)
            flags = np.zeros_like(w0)
            stored = np.zeros_like(now)
            leave = np.zeros_like(w0)
            match = np.zeros_like(w0)
        else:
            flags = w1 >> 24
            stored = _unpack(w1)
            leave = w2 & 255
            match = (w2 >> 8) & 255
            still = _maxabs(now - _unpack(w0)) <= thr
            in_mask = (flags & 1) != 0
            has_stored = (flags & 2) != 0
            match_s = (~in_mask) & has_stored & (s["uiReentryFrames"] > 0) & (_maxabs(now - stored) <= thr)
            med = median7(frame, self.work)
            if yuv is None:
                contrast = _maxabs(now - med)
            else:
                from . import formats
                d = (now - med).astype(np.float32)
                contrast = formats._yuv_diff(d[..., 0], d[..., 1], d[..., 2], yuv)
            if cthr <= 0:
                contrasty = np.ones_like(still)
            else:
                contrasty = (~in_mask) & (still | match_s) & (contrast >= cthr)
            # in the mask
            leave_m = np.where(still, 0, np.minimum(leave + 1, 255))
            stored_m = np.where(still[..., None], now, stored)
            leaves = leave_m >= s["uiLeaveFrames"]
            # out of the mask
            cnt_o = np.where(still & contrasty, np.minimum((w0 >> 24) + 1, 255), 0)
            match_o = np.where(match_s & contrasty, np.minimum(match + 1, 255), 0)
            enters = (cnt_o >= s["uiStillFrames"]) | ((s["uiReentryFrames"] > 0) & has_stored &
                                                       (match_o >= s["uiReentryFrames"]))
            new_in = np.where(in_mask, ~leaves, enters)
            new_has = has_stored | ((~in_mask) & enters)
            stored = np.where(in_mask[..., None], stored_m, np.where(enters[..., None], now, stored))
            leave = np.where(in_mask, np.where(leaves, 0, leave_m), 0)
            match = np.where(in_mask, np.where(leaves, 0, match), np.where(enters, 0, match_o))
            cnt = np.where(in_mask, 0, cnt_o)
            flags = new_in.astype(np.int64) | (new_has.astype(np.int64) << 1) | (still.astype(np.int64) << 2)
        nw0 = _pack(now) | (cnt << 24)
        nw1 = _pack(stored) | (flags << 24)
        nw2 = leave | (match << 8)
        self.w0 = np.where(in_box, nw0, _pack(now))
        self.w1 = np.where(in_box, nw1, ring[0])
        self.w2 = np.where(in_box, nw2, ring[1])

    def mask(self):
        """ui_v2_mask.comp: 255 = within 1 px of a mask pixel and still in this pair, 64 = within 1 px but changed,
        0 = not UI (bytes over the work rectangle)"""
        h, w = self.w1.shape
        inm = ((self.w1 >> 24) & 1) != 0
        pad = np.zeros((h + 2, w + 2), bool)
        pad[1:-1, 1:-1] = inm
        near = np.zeros((h, w), bool)
        for dy in range(3):
            for dx in range(3):
                near |= pad[dy:dy + h, dx:dx + w]
        still = ((self.w1 >> 24) & 4) != 0
        return np.where(near, np.where(still, 255, 64), 0).astype(np.uint8)


def footprint(X, Y, yuv=False):
    """ui_reject.glsl uiFootprint: RGB the 2x2 texels of a tap at the 1/32 map (X, Y), those of weight 0 dropped; 4:2:0
    the pixels of the 2x2 chroma texels around it, i.e. up to 4x4 pixels (which hold the luma taps)"""
    if yuv:
        cx, cy = X - 16, Y - 16
        ix, iy = cx >> 6, cy >> 6
        nx, ny = (cx & 63) != 0, (cy & 63) != 0
        return 2 * ix, 2 * iy, 2 * (ix + nx) + 1, 2 * (iy + ny) + 1
    lox, loy = X >> 5, Y >> 5
    return lox, loy, lox + ((X & 31) != 0), loy + ((Y & 31) != 0)


def hits(mask, rect, lox, loy, hix, hiy):
    """ui_reject.glsl FG_UI_HIT: does the texel rectangle [lo, hi] touch a non-zero byte of the mask over rect?"""
    out = np.zeros(np.shape(lox), bool)
    if mask is None or rect[2] == 0 or rect[3] == 0:
        return out
    ox, oy, w, h = rect
    lx, ly = np.maximum(lox - ox, 0), np.maximum(loy - oy, 0)
    hx, hy = np.minimum(hix - ox, w - 1), np.minimum(hiy - oy, h - 1)
    span = int(max(np.max(hix - lox, initial=0), np.max(hiy - loy, initial=0))) + 1
    for dy in range(span):
        for dx in range(span):
            x, y = lx + dx, ly + dy
            ok = (x <= hx) & (y <= hy)
            v = mask[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)] != 0
            out |= ok & v
    return out


def reject(wa, wc, maps_a, maps_c, ui, yuv=False):
    """ui_reject.glsl uiReject: (maskA, rectA, maskC, rectC); a tap on its own frame's UI (the other not) takes the
    other frame's tap (wa, wc: arrays with a trailing channel axis)"""
    mask_a, rect_a, mask_c, rect_c = ui
    ha = hits(mask_a, rect_a, *footprint(*maps_a, yuv=yuv))
    hc = hits(mask_c, rect_c, *footprint(*maps_c, yuv=yuv))
    rej = ha != hc
    na = np.where((rej & ha)[..., None], wc, wa)
    nc = np.where((rej & hc)[..., None], wa, wc)
    return na, nc


def patch(out, curr, mask, rect, fmt=0):
    """ui_patch.comp / ui_patch_chroma.comp, binary: the pixels of mask byte 255 take curr's colour (the output's
    alpha stays); 4:2:0: Y per pixel, and a chroma texel whose 2x2 pixels' largest mask byte (inside the rectangle) is
    255 takes curr's chroma"""
    x0, y0, w, h = rect
    sel = mask == 255
    if fmt in (0, 2):
        o = out.copy()
        reg = o[y0:y0 + h, x0:x0 + w]
        reg[..., :3] = np.where(sel[..., None], curr[y0:y0 + h, x0:x0 + w, :3], reg[..., :3])
        return o
    if fmt == 1:
        o = out.copy()
        reg = o[y0:y0 + h, x0:x0 + w]
        cw_ = curr[y0:y0 + h, x0:x0 + w]
        reg[...] = np.where(sel, (cw_ & np.uint32(0x3FFFFFFF)) | (reg & np.uint32(0xC0000000)), reg)
        return o
    yo, co = out[0].copy(), out[1].copy()
    yr = yo[y0:y0 + h, x0:x0 + w]
    yr[...] = np.where(sel, curr[0][y0:y0 + h, x0:x0 + w], yr)
    H, W = out[0].shape
    big = np.zeros((H + 1, W + 1), np.int64)
    big[y0:y0 + h, x0:x0 + w] = mask
    for cy in range(y0 >> 1, (y0 + h - 1) // 2 + 1):
        for cx in range(x0 >> 1, (x0 + w - 1) // 2 + 1):
            if big[2 * cy:2 * cy + 2, 2 * cx:2 * cx + 2].max() == 255:
                co[cy, cx] = curr[1][cy, cx]
    return yo, co


def luma_fill(l0, frame, mask, rect, S, reach=8, bgra=False, y16=None):
    """ui_v2_lumafill.comp: every flow cell (S x S pixels) of level-0 luma holding a UI pixel (mask byte != 0) becomes
    the mean of the integer lumas of the not-UI pixels of its cell, widened by S px a side per step up to `reach` px
    while it has none; float32 sum (exact: integers), one float32 division. Other cells are unchanged. y16: the other
    formats' fixed-point lumas (1/16 level, oracle/formats.py luma16), whose mean is sum / (16 n)."""
    from . import v1
    fh, fw = l0.shape
    if y16 is not None:
        H, W = y16.shape
        y8, unit = y16, 16
    else:
        H, W = frame.shape[:2]
        y8, unit = v1.luma8(frame, bgra), 1
    x0, y0, w, h = rect
    uipx = np.zeros((H, W), bool)
    uipx[y0:y0 + h, x0:x0 + w] = mask != 0
    out = l0.copy()
    for oy in range(y0 // S, min(fh, (y0 + h - 1) // S + 1)):
        for ox in range(x0 // S, min(fw, (x0 + w - 1) // S + 1)):
            px, py = ox * S, oy * S
            if not uipx[py:min(py + S, H), px:min(px + S, W)].any():
                continue
            for k in range(0, reach + 1, S):
                ya, yb = max(py - k, 0), min(py + S + k, H)
                xa, xb = max(px - k, 0), min(px + S + k, W)
                sel = ~uipx[ya:yb, xa:xb]
                n = int(sel.sum())
                if n > 0:
                    out[oy, ox] = np.float32(np.float32(y8[ya:yb, xa:xb][sel].sum()) / np.float32(unit * n))
                    break
    return out
