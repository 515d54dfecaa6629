"""CPU oracle of the V1 path for the frame formats other than RGBA8: Rgb10 (A2B10G10R10), RgbaF16 (scRGB / encoded
float), Nv12 (8-bit 4:2:0 YCbCr, two planes) and P010 (10-bit 4:2:0 YCbCr, two planes). numpy only. A level is an
8-bit-equivalent code: Nv12 codes as they are, 10-bit codes / 4 (q = codes per level: 1, 4).

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: only tests/, the smoke check and the
CPU-baseline leg of bench.py use it, as the checker.

The reference (gpu/shaders/format.glsl:29-106, luma.comp:9-27, motion_v1.glsl:85-144, resolve.comp:48-83) reads
these formats through the GPU's texture filter and leaves the luma unrounded, so its non-RGBA8 arithmetic is not
reproducible bit for bit on another GPU. This build keeps every meaning of the reference (levels = an 8-bit-equivalent
scale; the estimate on luma in levels; the mismatch on the signal converted to levels; blending and warping in the
source's own space) and pins the arithmetic as follows:

  * luma is a fixed-point value in units of 1/16 level (y16, an integer): Rgb10 round((4899 R + 9617 G + 1868 B) *
    255 / (1023 * 16384) * 16) on 10-bit codes; RgbaF16 (4899 L_r + 9617 L_g + 1868 L_b + 8192) >> 14 on the metric
    levels L (1/16 level) of a 65536-entry table built in float64; P010 round(16 (code / 4 - black) * ys), i.e.
    round((4 code - 256) * 255 / 219) at limited range and 4 code at full range (Nv12: 16 code in place of 4 code).
    Halves round up.
    The level-0 luma of an S x S cell is float32(sum of y16) / float32(16 n).
  * the warp map is the RGBA8 one (1/32 px, half-even); taps are exact integer bilinear sums with weights
    (32 - fx)(32 - fy) .. fx fy: Rgb10 keeps the sum S (code * 1024 units), the level is float32(S) * float32(255 /
    1047552); RgbaF16 sums the four float32 products w * v in the order p00, p01, p10, p11 and scales by 1/1024; P010
    (and Nv12) luma S_Y / (1024 q) levels, chroma sampled at the same point of the half-resolution plane, i.e. at
    (X - 16) / 64 chroma texels with weights (64 - fx)(64 - fy) .., S_C / (4096 q) levels (even frame sizes only).
  * mismatch e (0..255): Rgb10 uint(min(max |wa - wc|, 255)) in float32 levels (format.glsl FG_DIFF / FG_PACKE);
    RgbaF16 max |L(half(wa)) - L(half(wc))| >> 4, capped at 255 (the warped value rounded to half precision, then the
    table); Nv12 / P010 uint(min(max |fgYuvDiff(d)|, 255)) with the reference's float32 YCbCr -> R'G'B' constants
    (framegen.cpp:1519-1533) and its operation order.
  * blend (resolve.comp:73-83): m = 0.5 (wa + wc) at t = 1/2, else wA wa + wC wc; o = m (1 - u) + near u in float32.
    Rgb10 out code = floor(clamp(o, 0, 255) * float32(1023 / 255) + 0.5), alpha of the nearer frame; RgbaF16 out =
    half(o) (round to nearest even), alpha of the nearer frame; Nv12 / P010 Y code = clamp(floor(q o + 0.5), 0, top) at every
    pixel, CbCr of each 2x2 block at its even-even pixel, warped at the block centre p + 1/2 (warpChroma
    motion_v1.glsl:134-144) with that pixel's trust.
"""
import math

import numpy as np

from . import v1

f32 = np.float32
f64 = np.float64

RGBA8, RGB10, F16, NV12, P010 = 0, 1, 2, 3, 4
YUV = (NV12, P010)


def yuv_q(fmt):
    """codes per level and the largest code of a two-plane format"""
    return (4, 1023) if fmt == P010 else (1, 255)


def yuv_codes(plane, fmt):
    """Y codes (H, W) or CbCr codes (ch, cw, 2) of an Nv12 / P010 plane, int64."""
    return ((plane >> 6) if fmt == P010 else plane).astype(np.int64)


def yuv_store(codes, fmt):
    return (codes.astype(np.uint16) << 6).astype(np.uint16) if fmt == P010 else codes.astype(np.uint8)
HDR_PEAK, HDR_GAMMA = 1000.0, 2.4          # fgtest.cpp:203: an 8-bit value v is 1000 (v / 255)^2.4 nits
K10 = f32(255.0 / 1047552.0)               # Rgb10: level per unit of a bilinear sum (code * 1024)
K255_1023 = f32(255.0 / 1023.0)
K1023_255 = f32(1023.0 / 255.0)
L16_MAX = 65535                            # RgbaF16 metric levels are clamped to +-4096 levels


# ------------------------------------------------------------------------------------------------ encoders

def pq_enc(nits):
    """SMPTE ST 2084 inverse EOTF (fgtest.cpp:204-207): nits -> PQ code in [0, 1]."""
    p = math.pow(min(max(nits, 0.0) / 10000.0, 1.0), 0.1593017578125)
    return math.pow((0.8359375 + 18.8515625 * p) / (1.0 + 18.6875 * p), 78.84375)


def nits_of(v):
    return HDR_PEAK * math.pow(v / 255.0, HDR_GAMMA)


_PQ8 = None


def _pq8():
    global _PQ8
    if _PQ8 is None:
        _PQ8 = np.array([pq_enc(nits_of(v)) for v in range(256)], f64)
    return _PQ8


def encode_rgb10(rgba):
    """fgtest.cpp:222-229: each 8-bit value v as the PQ code of 1000 (v / 255)^2.4 nits (HDR10), alpha 3.
    uint32 (H, W) words, R in bits 0-9."""
    pq = _pq8()
    codes = np.rint(pq[rgba[..., :3]] * 1023.0).astype(np.uint32)
    return (3 << 30) | codes[..., 0] | (codes[..., 1] << 10) | (codes[..., 2] << 20)


def encode_f16(rgba, linear=False, scale=1.0):
    """fgtest.cpp:230-238: Transfer::Encoded stores the PQ code as a float, Transfer::Linear the light in units of
    80 nits (scRGB); `scale` multiplies the linear light (values up to 12.5 * scale). uint16 (H, W, 4) half bits."""
    if linear:
        lut = np.array([nits_of(v) / 80.0 * scale for v in range(256)], f64)
    else:
        lut = _pq8()
    h = lut[rgba[..., :3]].astype(np.float16)
    out = np.empty(rgba.shape[:2] + (4,), np.float16)
    out[..., :3] = h
    out[..., 3] = np.float16(1.0)
    return out.view(np.uint16)


def encode_nv12(rgba):
    """fgtest.cpp:240-268 (Nv12: BT.709 limited range, the 8-bit values as R'G'B'): Y plane uint8 (H, W), CbCr plane
    uint8 (H/2, W/2, 2), chroma of each 2x2 block averaged."""
    H, W = rgba.shape[:2]
    kr, kb = 0.2126, 0.0722
    kg = 1.0 - kr - kb
    c3 = rgba[..., :3].astype(f64) / 255.0
    Y = kr * c3[..., 0] + kg * c3[..., 1] + kb * c3[..., 2]
    cb = ((c3[..., 2] - Y) / (2 * (1 - kb))).astype(np.float32).astype(f64)
    cr = ((c3[..., 0] - Y) / (2 * (1 - kr))).astype(np.float32).astype(f64)
    y = np.rint(np.clip(16 + 219 * Y, 0.0, 255.0)).astype(np.uint8)
    cw, ch = (W + 1) // 2, (H + 1) // 2
    yi = np.minimum(np.arange(2 * ch), H - 1)
    xi = np.minimum(np.arange(2 * cw), W - 1)
    sb = cb[yi][:, xi].reshape(ch, 2, cw, 2).sum((1, 3))
    sr = cr[yi][:, xi].reshape(ch, 2, cw, 2).sum((1, 3))
    b = np.rint(np.clip(128 + 224 * sb / 4, 0.0, 255.0)).astype(np.uint8)
    r = np.rint(np.clip(128 + 224 * sr / 4, 0.0, 255.0)).astype(np.uint8)
    return y, np.stack([b, r], -1)


def encode_p010(rgba):
    """fgtest.cpp:240-268 (P010: BT.2020 limited range, PQ light of encode_rgb10): Y plane uint16 (H, W), CbCr plane
    uint16 (H/2, W/2, 2), 10-bit codes in the high bits."""
    H, W = rgba.shape[:2]
    pq = _pq8()
    kr, kb = 0.2627, 0.0593
    kg = 1.0 - kr - kb
    c3 = pq[rgba[..., :3]]
    Y = kr * c3[..., 0] + kg * c3[..., 1] + kb * c3[..., 2]
    cb = ((c3[..., 2] - Y) / (2 * (1 - kb))).astype(np.float32).astype(f64)
    cr = ((c3[..., 0] - Y) / (2 * (1 - kr))).astype(np.float32).astype(f64)
    y = np.rint(np.clip((16 + 219 * Y) * 4.0, 0.0, 1023.0)).astype(np.uint16)
    cw, ch = (W + 1) // 2, (H + 1) // 2
    yi = np.minimum(np.arange(2 * ch), H - 1)
    xi = np.minimum(np.arange(2 * cw), W - 1)
    sb = cb[yi][:, xi].reshape(ch, 2, cw, 2).sum((1, 3))
    sr = cr[yi][:, xi].reshape(ch, 2, cw, 2).sum((1, 3))
    b = np.rint(np.clip((128 + 224 * sb / 4) * 4.0, 0.0, 1023.0)).astype(np.uint16)
    r = np.rint(np.clip((128 + 224 * sr / 4) * 4.0, 0.0, 1023.0)).astype(np.uint16)
    return (y << 6).astype(np.uint16), (np.stack([b, r], -1) << 6).astype(np.uint16)


# ------------------------------------------------------------------------------------------------ RgbaF16 levels

def pq_code(y):
    """fgPq format.glsl:65-69 in float64, y already scaled and clamped to [0, 1]."""
    p = math.pow(y, 0.1593017578125)
    return math.pow((0.8359375 + 18.8515625 * p) / (1.0 + 18.6875 * p), 78.84375)


def f16_level_table(linear, white=80.0):
    """The metric level (fgMetric format.glsl:70-73) of every half value in units of 1/16 level, built in float64:
    encoded M = v * 255, linear M = 255 PQ(clamp(v * white / 10000, 0, 1)); L16 = floor(16 M + 0.5) clamped to
    +-65535; NaN -> 0. int32 (65536,)."""
    v = np.arange(65536, dtype=np.uint16).view(np.float16).astype(f64)
    out = np.zeros(65536, np.int64)
    for i in range(65536):
        x = v[i]
        if math.isnan(x):
            continue
        if linear:
            y = min(max(x * (white / 10000.0), 0.0), 1.0)
            m = 255.0 * pq_code(y)
        else:
            m = x * 255.0
        q = math.floor(16.0 * m + 0.5) if math.isfinite(m) else (L16_MAX if m > 0 else -L16_MAX)
        out[i] = min(max(q, -L16_MAX), L16_MAX)
    return out.astype(np.int32)


# ------------------------------------------------------------------------------------------------ luma

def luma16(frame, fmt, bgra=False, lut=None, full_range=False):
    """Per-pixel luma in units of 1/16 level (int64 (H, W))."""
    if fmt == RGB10:
        w = frame.astype(np.int64)
        r, g, b = w & 1023, (w >> 10) & 1023, (w >> 20) & 1023
        if bgra:
            r, b = b, r
        n = r * 4899 + g * 9617 + b * 1868
        return (n * 255 + 523776) // 1047552
    if fmt == F16:
        L = lut.astype(np.int64)[frame[..., :3]]
        n = L[..., 0] * 4899 + L[..., 1] * 9617 + L[..., 2] * 1868
        return (n + 8192) >> 14
    if fmt in YUV:
        yp = frame[0] if isinstance(frame, tuple) else frame
        c = yuv_codes(yp, fmt) * (16 // yuv_q(fmt)[0])   # 16 code / q = 16 x level
        if full_range:
            return c
        return ((c - 256) * 510 + 219) // 438
    raise ValueError(fmt)


def luma_level0(frame, fmt, S, fw, fh, **kw):
    """luma.comp:58-72 on the fixed-point luma: float32(sum of y16 over the valid pixels of the cell) / float32(16 n)."""
    y = luma16(frame, fmt, **kw)
    H, W = y.shape
    Hp, Wp = fh * S, fw * S
    yp = np.zeros((max(Hp, H), max(Wp, W)), np.int64)
    vp = np.zeros_like(yp)
    yp[:H, :W] = y
    vp[:H, :W] = 1
    sums = yp[:Hp, :Wp].reshape(fh, S, fw, S).sum((1, 3))
    n = vp[:Hp, :Wp].reshape(fh, S, fw, S).sum((1, 3))
    return sums.astype(f32) / (16 * n).astype(f32)


def estimate(a, c, fmt, s=None, bgra=False, lut=None, full_range=False, proposals=None):
    """v1.estimate on the format's luma (proposals: as v1.estimate's, learned seeds of the coarsest level)."""
    s = s or v1.DEFAULTS
    H, W = (a[0] if isinstance(a, tuple) else a).shape[:2]
    S = s["flowScale"]
    fw, fh, _ = v1.layout(W, H, S, s["pyramidMinSide"])
    kw = dict(bgra=bgra, lut=lut, full_range=full_range)
    la = luma_level0(a, fmt, S, fw, fh, **kw)
    lc = luma_level0(c, fmt, S, fw, fh, **kw)
    return v1.estimate(None, None, s, luma0=(la, lc), size=(W, H), proposals=proposals)


# ------------------------------------------------------------------------------------------------ warps

def _maps(est, t, W, H):
    """The 1/32 warp maps of a and c at time t (X, Y int64, as warp_rgba8 computes them) and the field."""
    s = est["settings"]
    ts, wA, wC = v1.time_constants(t)
    g = v1.field_at(est["pf"], t)
    hv = v1.half_vec(g, W, H, s["flowScale"])
    py, px = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    pxf, pyf = px.astype(f32), py.astype(f32)
    ka, kc = f32(2.0) * wC, f32(2.0) * wA
    ax = (pxf - (ka * hv[..., 0]).astype(f32)).astype(f32)
    ay = (pyf - (ka * hv[..., 1]).astype(f32)).astype(f32)
    cx = (pxf + (kc * hv[..., 0]).astype(f32)).astype(f32)
    cy = (pyf + (kc * hv[..., 1]).astype(f32)).astype(f32)
    q = lambda v: np.rint((v * f32(32.0)).astype(f32)).astype(np.int64)
    return (q(ax), q(ay)), (q(cx), q(cy)), hv, (ka, kc)


def _taps(img, X, Y, unit=32, sh=5):
    """The four clamped taps and integer weights of a bilinear sample at map (X, Y) in 1/unit texel units."""
    ix, iy, fx, fy = X >> sh, Y >> sh, X & (unit - 1), Y & (unit - 1)
    p00, p01 = v1._fetch(img, ix, iy), v1._fetch(img, ix + 1, iy)
    p10, p11 = v1._fetch(img, ix, iy + 1), v1._fetch(img, ix + 1, iy + 1)
    gx, gy = unit - fx, unit - fy
    return (p00, p01, p10, p11), (gx * gy, fx * gy, gx * fy, fx * fy)


def _isum(img, X, Y, unit=32, sh=5):
    (p00, p01, p10, p11), (w00, w01, w10, w11) = _taps(img.astype(np.int64), X, Y, unit, sh)
    ex = (lambda w: w[..., None]) if img.ndim == 3 else (lambda w: w)
    return ex(w00) * p00 + ex(w01) * p01 + ex(w10) * p10 + ex(w11) * p11


def rgb10_channels(frame):
    w = frame.astype(np.int64)
    return np.stack([w & 1023, (w >> 10) & 1023, (w >> 20) & 1023], -1)


def f16_values(frame):
    return frame.view(np.float16).astype(f32)


def _f16_warp(vals, X, Y):
    (p00, p01, p10, p11), ws = _taps(vals, X, Y)
    w00, w01, w10, w11 = [w.astype(f32)[..., None] for w in ws]
    s = ((w00 * p00).astype(f32) + (w01 * p01).astype(f32)).astype(f32)
    s = (s + (w10 * p10).astype(f32)).astype(f32)
    s = (s + (w11 * p11).astype(f32)).astype(f32)
    return (s * f32(1.0 / 1024.0)).astype(f32)


def yuv_constants(matrix=2, full_range=False):
    """framegen.cpp:1519-1533: (ys, rcr, bcb, gcb, gcr) as float32; matrix 0 BT.601, 1 BT.709, 2 BT.2020."""
    kr, kb = {0: (0.299, 0.114), 1: (0.2126, 0.0722), 2: (0.2627, 0.0593)}[matrix]
    kg = 1.0 - kr - kb
    sY = 1.0 if full_range else 255.0 / 219.0
    sC = 1.0 if full_range else 255.0 / 224.0
    return (f32(sY), f32(sC * 2.0 * (1.0 - kr)), f32(sC * 2.0 * (1.0 - kb)), f32(-sC * 2.0 * (1.0 - kb) * kb / kg),
            f32(-sC * 2.0 * (1.0 - kr) * kr / kg))


def _yuv_diff(dY, dCb, dCr, k):
    ys, rcr, bcb, gcb, gcr = k
    R = ((ys * dY).astype(f32) + (rcr * dCr).astype(f32)).astype(f32)
    G = ((ys * dY).astype(f32) + (gcb * dCb).astype(f32)).astype(f32)
    G = (G + (gcr * dCr).astype(f32)).astype(f32)
    B = ((ys * dY).astype(f32) + (bcb * dCb).astype(f32)).astype(f32)
    return np.maximum(np.abs(R), np.maximum(np.abs(G), np.abs(B)))


def _chroma_sum(cplane, X, Y, fmt=P010):
    """CbCr of the half-resolution plane at luma map (X, Y) (1/32 px): the point (X - 16) / 64 chroma texels."""
    return _isum(yuv_codes(cplane, fmt), X - 16, Y - 16, unit=64, sh=6)


# ------------------------------------------------------------------------------------------------ synth

def synth(a, c, est, t, fmt, exact64=False, lut=None, yuv=None, ui=None):
    """v1.synth for Rgb10 / RgbaF16 / Nv12 / P010 frames (arrays as the encoders return them). Returns (out, reason,
    info); for Nv12 / P010 out is (Y plane, CbCr plane). ui: (maskA, rectA, maskC, rectC) of a CrosshairV2 pair: the
    warp taps are rejected as oracle/ui.py reject() does, on the raw sums (motion_v1.glsl:124, 141; 4:2:0 footprints for
    Nv12 / P010, the luma tap's for Y and CbCr of the mismatch and the blend, the block centre's for the blended CbCr)."""
    from . import ui as uimod
    s = est["settings"]
    yp_a = a[0] if fmt in YUV else a
    H, W = yp_a.shape[:2]
    ts, wA, wC = v1.time_constants(t)
    near_c = wC > f32(0.5)
    near = c if near_c else a
    info = {}
    if est["cut"]:
        return (tuple(x.copy() for x in near) if fmt in YUV else near.copy()), "cut", info
    (AX, AY), (CX, CY), hv, (ka, kc) = _maps(est, t, W, H)
    if fmt == RGB10:
        sa, sc = _isum(rgb10_channels(a), AX, AY), _isum(rgb10_channels(c), CX, CY)
        if ui is not None:
            sa, sc = uimod.reject(sa, sc, (AX, AY), (CX, CY), ui)
        wa, wc = (sa.astype(f32) * K10).astype(f32), (sc.astype(f32) * K10).astype(f32)
        e = np.minimum(np.abs(wa - wc).max(-1), f32(255.0)).astype(np.int64)
    elif fmt == F16:
        wa, wc = _f16_warp(f16_values(a)[..., :3], AX, AY), _f16_warp(f16_values(c)[..., :3], CX, CY)
        if ui is not None:
            wa, wc = uimod.reject(wa, wc, (AX, AY), (CX, CY), ui)
        La = lut.astype(np.int64)[wa.astype(np.float16).view(np.uint16)]
        Lc = lut.astype(np.int64)[wc.astype(np.float16).view(np.uint16)]
        e = np.minimum(np.abs(La - Lc).max(-1) >> 4, 255)
    elif fmt in YUV:
        qq, top = yuv_q(fmt)
        ya, yc = yuv_codes(a[0], fmt), yuv_codes(c[0], fmt)
        sya, syc = _isum(ya, AX, AY), _isum(yc, CX, CY)
        sca, scc = _chroma_sum(a[1], AX, AY, fmt), _chroma_sum(c[1], CX, CY, fmt)
        if ui is not None:
            ra, rc = uimod.reject(np.concatenate([sya[..., None], sca], -1), np.concatenate([syc[..., None], scc], -1),
                                  (AX, AY), (CX, CY), ui, yuv=True)
            sya, sca, syc, scc = ra[..., 0], ra[..., 1:], rc[..., 0], rc[..., 1:]
        ky, kc_ = f32(1.0 / (1024.0 * qq)), f32(1.0 / (4096.0 * qq))
        dY = ((sya - syc).astype(f32) * ky).astype(f32)
        dC = ((sca - scc).astype(f32) * kc_).astype(f32)
        e = np.minimum(_yuv_diff(dY, dC[..., 0], dC[..., 1], yuv), f32(255.0)).astype(np.int64)
    else:
        raise ValueError(fmt)
    S49 = v1.box_int(e, 3)
    lo, hi = f32(s["trustLow"]), f32(s["trustHigh"])
    den = f32(hi - lo)
    e7 = (S49.astype(f64) * (1.0 / 49.0)).astype(f32)
    u32 = np.clip(((e7 - lo).astype(f32) / den).astype(f32), f32(0), f32(1)).astype(f32)
    count = int((u32 > f32(0.5)).sum())
    info.update(e=e, S49=S49, count=count)
    flick = f32(count) > (f32(s["globalFallback"]) * f32(W)) * f32(H)
    if flick:
        return (tuple(x.copy() for x in near) if fmt in YUV else near.copy()), "untrusted", info
    if exact64:
        u = np.clip((S49.astype(f64) / 49.0 - float(lo)) / float(den), 0.0, 1.0)
    else:
        u = u32

    def blend(va, vc, vn, uu):
        if exact64:
            m = 0.5 * (va.astype(f64) + vc) if wC == f32(0.5) else float(wA) * va.astype(f64) + float(wC) * vc
            return m * (1.0 - uu) + vn.astype(f64) * uu
        if wC == f32(0.5):
            m = (f32(0.5) * (va + vc).astype(f32)).astype(f32)
        else:
            m = ((wA * va).astype(f32) + (wC * vc).astype(f32)).astype(f32)
        return ((m * (f32(1.0) - uu)).astype(f32) + (vn.astype(f32) * uu).astype(f32)).astype(f32)

    if fmt == RGB10:
        ncodes = rgb10_channels(near)
        n8 = (ncodes.astype(f32) * K255_1023).astype(f32)
        o = blend(wa, wc, n8, u[..., None])
        if exact64:
            code = np.floor(np.clip(o, 0, 255) * (1023.0 / 255.0) + 0.5).astype(np.int64)
        else:
            code = np.floor((np.clip(o, f32(0), f32(255)) * K1023_255).astype(f32) + f32(0.5)).astype(np.int64)
        out = (near.astype(np.uint32) & np.uint32(0xC0000000)) | code[..., 0].astype(np.uint32) | \
            (code[..., 1].astype(np.uint32) << 10) | (code[..., 2].astype(np.uint32) << 20)
        info.update(levels=o)
        return out.astype(np.uint32), None, info
    if fmt == F16:
        vn = f16_values(near)
        o = blend(wa, wc, vn[..., :3], u[..., None])
        out = np.empty_like(near)
        out[..., :3] = o.astype(np.float16).view(np.uint16)
        out[..., 3] = near[..., 3]
        info.update(values=o)
        return out, None, info
    # Nv12 / P010: luma at every pixel, chroma at the even-even pixel of each 2x2 block
    kn = f32(1.0 / qq)
    wya = (sya.astype(f32) * ky).astype(f32)
    wyc = (syc.astype(f32) * ky).astype(f32)
    ny = (yuv_codes(near[0], fmt).astype(f32) * kn).astype(f32)
    oy = blend(wya, wyc, ny, u)
    yout = yuv_store(np.clip(np.floor(oy * f32(qq) + f32(0.5)), 0, top), fmt)
    # warpChroma: the block centre q = p + 1/2 of every even-even pixel p
    ch, cw = near[1].shape[:2]
    py, px = np.meshgrid(np.arange(ch) * 2, np.arange(cw) * 2, indexing="ij")
    hvb = hv[py, px]
    qx, qy = (px.astype(f32) + f32(0.5)), (py.astype(f32) + f32(0.5))
    q = lambda v: np.rint((v * f32(32.0)).astype(f32)).astype(np.int64)
    BX = q((qx - (ka * hvb[..., 0]).astype(f32)).astype(f32))
    BY = q((qy - (ka * hvb[..., 1]).astype(f32)).astype(f32))
    DX = q((qx + (kc * hvb[..., 0]).astype(f32)).astype(f32))
    DY = q((qy + (kc * hvb[..., 1]).astype(f32)).astype(f32))
    sba, sbc = _chroma_sum(a[1], BX, BY, fmt), _chroma_sum(c[1], DX, DY, fmt)
    if ui is not None:
        sba, sbc = uimod.reject(sba, sbc, (BX, BY), (DX, DY), ui, yuv=True)
    ca = (sba.astype(f32) * kc_).astype(f32)
    cc = (sbc.astype(f32) * kc_).astype(f32)
    nc = (yuv_codes(near[1], fmt).astype(f32) * kn).astype(f32)
    oc = blend(ca, cc, nc, u[py, px][..., None])
    cout = yuv_store(np.clip(np.floor(oc * f32(qq) + f32(0.5)), 0, top), fmt)
    info.update(levels_y=oy, levels_c=oc)
    return (yout, cout), None, info


def synth_blur(a, c, est, t0, t1, n, fmt, gauss=False, lut=None, yuv=None, ui=None):
    """recordSynthBlur for the other formats: the weighted mean of the stored sub-frames (synth() at each sample time,
    the real frames at t <= 0 / t >= 1) in the format's own codes, rounded once (halves: to nearest even)."""
    ts, ws = v1.blur_samples(t0, t1, n, gauss)
    frames = []
    for t in ts:
        if t <= 0:
            frames.append(a)
        elif t >= 1:
            frames.append(c)
        else:
            frames.append(synth(a, c, est, float(t), fmt, lut=lut, yuv=yuv, ui=ui)[0])
    if fmt == RGB10:
        ch = [np.concatenate([rgb10_channels(f), (f >> 30).astype(np.int64)[..., None]], -1) for f in frames]
        r = v1.average_frames(ch[:1] + ch[1:], ws, hi=1023).astype(np.uint32)
        r[..., 3] = np.minimum(r[..., 3], 3)
        return (r[..., 0] | (r[..., 1] << 10) | (r[..., 2] << 20) | (r[..., 3] << 30)).astype(np.uint32)
    if fmt == F16:
        wsum = f32(0.0)
        for w in ws:
            wsum = f32(wsum + w)
        acc = np.zeros(frames[0].shape, f32)
        for f, w in zip(frames, ws):
            acc = (acc + (w * f.view(np.float16).astype(f32)).astype(f32)).astype(f32)
        return (acc / wsum).astype(f32).astype(np.float16).view(np.uint16)
    top = yuv_q(fmt)[1]
    y = v1.average_frames([yuv_codes(f[0], fmt) for f in frames], ws, hi=top)
    cc = v1.average_frames([yuv_codes(f[1], fmt) for f in frames], ws, hi=top)
    return yuv_store(y, fmt), yuv_store(cc, fmt)


def level_error(fmt, got, want64_info, lut=None):
    """The output error in 8-bit levels against the FP64 oracle's unrounded levels (for the +-1 / +-2 bar)."""
    if fmt == RGB10:
        codes = rgb10_channels(got).astype(f64)
        return np.abs(codes * (255.0 / 1023.0) - np.clip(want64_info["levels"], 0, 255))
    if fmt == F16:
        L = lut.astype(f64)[got[..., :3]] / 16.0
        ref = want64_info["values"].astype(np.float16).view(np.uint16)
        return np.abs(L - lut.astype(f64)[ref] / 16.0)
    if fmt in YUV:
        qq, top = yuv_q(fmt)
        y = yuv_codes(got[0], fmt).astype(f64) / qq
        cc = yuv_codes(got[1], fmt).astype(f64) / qq
        ey = np.abs(y - np.clip(want64_info["levels_y"], 0, top / qq))
        ec = np.abs(cc - np.clip(want64_info["levels_c"], 0, top / qq))
        return np.concatenate([ey.ravel(), ec.ravel()])
    raise ValueError(fmt)
