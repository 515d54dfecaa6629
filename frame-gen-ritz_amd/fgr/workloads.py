"""Synthetic benchmark workloads generated on the GPU (torch as device memory and elementwise ops only).

  * mixed_batch -- 64 pairs, 16 each of 1280x720, 1920x1080, 2560x1440 and 3840x2160, formats cycling RGBA8 / Rgb10
    (HDR10 PQ) / RgbaF16 scRGB (linear light, values 0-4) / P010 (BT.2020 PQ, limited range), multipliers 2x / 3x / 4x,
    value-noise content moved by a global integer shift in +-24 px plus a moving sprite.
  * uhd_batch -- 128 pairs at 3840x2160, half RGBA8 and half P010, 4x; 5 % of the pairs a scene cut, 30 % with four
    large occluding sprites moving across the background.

Content: 3-octave value noise (wavelengths 64, 16, 4 px; lattice values uniform, smoothstep interpolation) mapped to
luma 16..239 with low-frequency per-channel offsets, the second frame the first moved by an integer shift with edge
clamp; sprites are value-noise squares pasted at their own positions in each frame. The HDR encodings follow
fgtest.cpp:158-238: an 8-bit value v is 1000 (v / 255)^2.4 nits, stored as its PQ code (Rgb10, P010) or as linear
light in units of 80 nits times a scale (RgbaF16).
"""
import math

import numpy as np
import torch

FMT_RGBA8, FMT_RGB10, FMT_F16, FMT_NV12, FMT_P010 = 0, 1, 2, 3, 4


def _pq(nits):
    p = math.pow(min(max(nits, 0.0) / 10000.0, 1.0), 0.1593017578125)
    return math.pow((0.8359375 + 18.8515625 * p) / (1.0 + 18.6875 * p), 78.84375)


def _nits(v):
    return 1000.0 * math.pow(v / 255.0, 2.4)


def _luts(device):
    pq = torch.tensor([_pq(_nits(v)) for v in range(256)], dtype=torch.float64, device=device)
    lin = torch.tensor([_nits(v) / 80.0 for v in range(256)], dtype=torch.float64, device=device)
    return pq, lin


def noise(gen, h, w, lams=(64, 16, 4), amps=(1.0, 0.5, 0.25), device="cuda:0"):
    """value noise in [0, 1), float32 (h, w)"""
    ys = torch.arange(h, device=device, dtype=torch.float32)
    xs = torch.arange(w, device=device, dtype=torch.float32)
    acc = torch.zeros((h, w), device=device)
    for lam, amp in zip(lams, amps):
        gh, gw = int(h / lam) + 3, int(w / lam) + 3
        lat = torch.rand((gh, gw), generator=gen, device=device)
        fy, fx = ys / lam, xs / lam
        iy, ix = fy.floor().long().clamp(0, gh - 2), fx.floor().long().clamp(0, gw - 2)
        sy, sx = fy - fy.floor(), fx - fx.floor()
        sy, sx = (sy * sy * (3 - 2 * sy))[:, None], (sx * sx * (3 - 2 * sx))[None, :]
        a = lat[iy][:, ix]
        b = lat[iy][:, ix + 1]
        c = lat[iy + 1][:, ix]
        d = lat[iy + 1][:, ix + 1]
        top = a + (b - a) * sx
        acc += amp * (top + (c + (d - c) * sx - top) * sy)
    return acc / sum(amps)


def _rgba(lum3):
    out = torch.empty(lum3.shape[:2] + (4,), dtype=torch.uint8, device=lum3.device)
    out[..., :3] = lum3.round().clamp(0, 255).to(torch.uint8)
    out[..., 3] = 255
    return out


def _shift(img, dx, dy):
    h, w = img.shape[:2]
    ys = (torch.arange(h, device=img.device) - dy).clamp(0, h - 1)
    xs = (torch.arange(w, device=img.device) - dx).clamp(0, w - 1)
    return img[ys][:, xs]


def pair(gen, w, h, shift, n_sprites=1, sprite=96, cut=False, device="cuda:0"):
    """(prev, curr) uint8 RGBA (h, w, 4) tensors."""
    lum = 16.0 + 223.0 * noise(gen, h, w, device=device)
    chroma = torch.stack([12.0 * (noise(gen, h, w, (128,), (1.0,), device) - 0.5) for _ in range(3)], -1)
    a = lum[..., None] + chroma
    c = _shift(a, int(shift[0]), int(shift[1]))
    a, c = a.clone(), c.clone()
    for _ in range(n_sprites):
        tex = 16.0 + 223.0 * noise(gen, sprite, sprite, (24, 8, 3), (1.0, 0.5, 0.25), device)
        tex = tex[..., None] + (torch.rand(3, generator=gen, device=device) * 80.0 - 40.0)
        r = torch.rand(4, generator=gen, device=device).tolist()
        x0 = int(32 + r[0] * (w - sprite - 64))
        y0 = int(32 + r[1] * (h - sprite - 64))
        vx, vy = int(r[2] * 48) - 24, int(r[3] * 48) - 24
        x1, y1 = min(max(x0 + vx, 0), w - sprite), min(max(y0 + vy, 0), h - sprite)
        a[y0:y0 + sprite, x0:x0 + sprite] = tex
        c[y1:y1 + sprite, x1:x1 + sprite] = tex
    prev = _rgba(a)
    if cut:
        bits = torch.rand((h // 2 + 1, w // 2 + 1), generator=gen, device=device) < 0.5
        v = torch.where(bits, 235.0, 20.0).repeat_interleave(2, 0).repeat_interleave(2, 1)[:h, :w]
        return prev, _rgba(v[..., None].expand(h, w, 3))
    return prev, _rgba(c)


def encode(rgba, fmt, luts, f16_scale=0.32):
    """RGBA8 tensor -> the frame of `fmt` (a tensor, or (Y, CbCr) for P010), fgtest.cpp:216-268 encodings."""
    pq, lin = luts
    v = rgba[..., :3].long()
    if fmt == FMT_RGBA8:
        return rgba
    if fmt == FMT_RGB10:
        codes = (pq[v] * 1023.0).round().to(torch.int32)
        return (codes[..., 0] | (codes[..., 1] << 10) | (codes[..., 2] << 20) | (3 << 30)).to(torch.int32)
    if fmt == FMT_F16:
        out = torch.empty(rgba.shape[:2] + (4,), dtype=torch.float16, device=rgba.device)
        out[..., :3] = (lin[v] * f16_scale).to(torch.float16)
        out[..., 3] = 1.0
        return out
    if fmt == FMT_P010:
        kr, kb = 0.2627, 0.0593
        kg = 1.0 - kr - kb
        c3 = pq[v]
        Y = kr * c3[..., 0] + kg * c3[..., 1] + kb * c3[..., 2]
        cb = (c3[..., 2] - Y) / (2 * (1 - kb))
        cr = (c3[..., 0] - Y) / (2 * (1 - kr))
        y = ((16 + 219 * Y) * 4.0).clamp(0, 1023).round().to(torch.int32)
        h, w = Y.shape
        sb = cb.reshape(h // 2, 2, w // 2, 2).sum((1, 3))
        sr = cr.reshape(h // 2, 2, w // 2, 2).sum((1, 3))
        b = ((128 + 224 * sb / 4) * 4.0).clamp(0, 1023).round().to(torch.int32)
        r = ((128 + 224 * sr / 4) * 4.0).clamp(0, 1023).round().to(torch.int32)
        return ((y << 6).to(torch.int16), torch.stack([(b << 6), (r << 6)], -1).to(torch.int16))
    raise ValueError(fmt)


def empty_like_frame(f):
    if isinstance(f, tuple):
        return tuple(torch.empty_like(x) for x in f)
    return torch.empty_like(f)


def frame_bytes(w, h, fmt):
    return w * h * {FMT_RGBA8: 4, FMT_RGB10: 4, FMT_F16: 8, FMT_NV12: 1.5, FMT_P010: 3}[fmt]


def pair_bytes(w, h, fmt, n, cut=False, S=2):
    """B = 2F + P + (N - 1)(3F + P), a cut pair's synths copies (2F)"""
    F = frame_bytes(w, h, fmt)
    P = round(w / S) * round(h / S) * 8
    return 2 * F + P + (n - 1) * (2 * F if cut else 3 * F + P)


def mixed_batch(seed=3, device="cuda:0"):
    """64 items (prev, curr, multiplier, outs, fmt) and their algorithmic bytes."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    luts = _luts(device)
    sizes = [(1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]
    fmts = [FMT_RGBA8, FMT_RGB10, FMT_F16, FMT_P010]
    items, total = [], 0
    for i in range(64):
        w, h = sizes[i // 16]
        fmt = fmts[i % 4]
        n = 2 + i % 3
        r = torch.rand(2, generator=gen, device=device).tolist()
        a, c = pair(gen, w, h, (int(r[0] * 49) - 24, int(r[1] * 49) - 24), device=device)
        A, C = encode(a, fmt, luts), encode(c, fmt, luts)
        items.append((A, C, n, [empty_like_frame(A) for _ in range(n - 1)], fmt))
        total += pair_bytes(w, h, fmt, n)
    return items, total


def uhd_batch(seed=5, n_pairs=128, device="cuda:0"):
    """128 items at 3840x2160 (RGBA8 and P010, 4x), their algorithmic bytes and cut flags."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    luts = _luts(device)
    w, h = 3840, 2160
    items, total, cuts = [], 0, []
    for i in range(n_pairs):
        fmt = FMT_RGBA8 if i % 2 == 0 else FMT_P010
        cut = i % 20 == 7                       # 5 %
        occl = i % 10 < 3                       # 30 %: four large occluding sprites
        r = torch.rand(2, generator=gen, device=device).tolist()
        a, c = pair(gen, w, h, (int(r[0] * 49) - 24, int(r[1] * 49) - 24), n_sprites=4 if occl else 1,
                    sprite=384 if occl else 96, cut=cut, device=device)
        A, C = encode(a, fmt, luts), encode(c, fmt, luts)
        items.append((A, C, 4, [empty_like_frame(A) for _ in range(3)], fmt))
        total += pair_bytes(w, h, fmt, 4, cut)
        cuts.append(cut)
    return items, total, cuts
