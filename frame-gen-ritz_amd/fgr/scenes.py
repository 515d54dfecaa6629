"""Seeded synthetic frame pairs for the V1 workloads (RGBA8, row 0 = top, bytes R, G, B, A).

Every generator takes a ``numpy.random.Generator`` (or a seed) and returns uint8 arrays of shape (H, W, 4).
The same seed gives the same bytes on every machine: only IEEE +, -, *, / and floor are used in float64,
then one round-half-even to uint8.

Workloads:
  * ``pair_translate`` -- the 640x360 check pair: 3-octave value noise, the whole frame translated by an
    integer (+5, -3) px with edge clamp, plus a 64x64 uniform-noise sprite moving (+12, +7) px.
  * ``pairs_1080p`` -- the 1920x1080 4x workload: value-noise background shifted by a global sub-pixel
    motion (uniform in +-16 px, quarter-pel steps, bilinear resample), four moving 96x96 sprites, a static
    300x60 binary-noise HUD, and the last pair's second frame replaced by unrelated binary noise (a scene cut).
"""
import numpy as np

__all__ = ["value_noise", "to_rgba", "pair_translate", "pair_shift", "pairs_1080p", "noise_pair", "cut_frame",
           "rng_of"]


def rng_of(seed):
    return seed if isinstance(seed, np.random.Generator) else np.random.default_rng(seed)


def _octave(rng, h, w, lam, oy=0.0, ox=0.0):
    """One octave of value noise sampled at pixel centres (x + ox, y + oy): lattice values uniform [0, 1),
    smoothstep interpolation between lattice points spaced `lam` px apart."""
    gh, gw = int(np.ceil((h + abs(oy) + 2) / lam)) + 3, int(np.ceil((w + abs(ox) + 2) / lam)) + 3
    lat = rng.random((gh, gw))
    ys = (np.arange(h, dtype=np.float64) + oy) / lam + 1.0
    xs = (np.arange(w, dtype=np.float64) + ox) / lam + 1.0
    iy, ix = np.floor(ys).astype(np.int64), np.floor(xs).astype(np.int64)
    fy, fx = ys - iy, xs - ix
    sy, sx = fy * fy * (3.0 - 2.0 * fy), fx * fx * (3.0 - 2.0 * fx)
    iy = np.clip(iy, 0, gh - 2)
    ix = np.clip(ix, 0, gw - 2)
    a = lat[iy][:, ix]
    b = lat[iy][:, ix + 1]
    c = lat[iy + 1][:, ix]
    d = lat[iy + 1][:, ix + 1]
    top = a + (b - a) * sx[None, :]
    bot = c + (d - c) * sx[None, :]
    return top + (bot - top) * sy[:, None]


def value_noise(rng, h, w, lams=(64, 16, 4), amps=(1.0, 0.5, 0.25)):
    """3-octave value noise in [0, 1), float64 (h, w)."""
    acc = np.zeros((h, w))
    for lam, amp in zip(lams, amps):
        acc += amp * _octave(rng, h, w, lam)
    return acc / sum(amps)


def _chroma(rng, h, w):
    """Low-frequency per-channel offsets (levels) that make RGB luma-correlated but not grey."""
    return np.stack([12.0 * (value_noise(rng, h, w, (128,), (1.0,)) - 0.5) for _ in range(3)], -1)


def to_rgba(lum, chroma=None):
    """float64 luma levels (h, w) (+ per-channel offsets (h, w, 3)) -> uint8 RGBA, alpha 255."""
    rgb = lum[..., None] + (0.0 if chroma is None else chroma)
    out = np.empty(lum.shape + (4,), np.uint8)
    out[..., :3] = np.clip(np.rint(rgb), 0, 255).astype(np.uint8)
    out[..., 3] = 255
    return out


def _background(rng, h, w):
    return 16.0 + 223.0 * value_noise(rng, h, w)


def _shift_clamp(img, dx, dy):
    """Content moves by (+dx, +dy): out[y, x] = img[clamp(y - dy), clamp(x - dx)]."""
    h, w = img.shape[:2]
    ys = np.clip(np.arange(h) - dy, 0, h - 1)
    xs = np.clip(np.arange(w) - dx, 0, w - 1)
    return img[ys][:, xs]


def pair_translate(seed=1, w=640, h=360, shift=(5, -3), sprite=64, sprite_pos=(200, 120), sprite_move=(12, 7)):
    """The 640x360 check pair (prev, curr): noise translated by an integer shift, plus a moving sprite."""
    rng = rng_of(seed)
    prev = to_rgba(_background(rng, h, w), _chroma(rng, h, w))
    curr = _shift_clamp(prev, *shift).copy()
    spr = np.empty((sprite, sprite, 4), np.uint8)
    spr[..., :3] = rng.integers(0, 256, (sprite, sprite, 3), dtype=np.uint8)
    spr[..., 3] = 255
    x0, y0 = sprite_pos
    _paste(prev, spr, x0, y0)
    _paste(curr, spr, x0 + sprite_move[0], y0 + sprite_move[1])
    return prev, curr


def _paste(img, spr, x, y):
    """Draw spr with its top-left corner at (x, y), clipped to the image."""
    h, w = img.shape[:2]
    sh, sw = spr.shape[:2]
    x0, y0, x1, y1 = max(x, 0), max(y, 0), min(x + sw, w), min(y + sh, h)
    if x1 > x0 and y1 > y0:
        img[y0:y1, x0:x1] = spr[y0 - y:y1 - y, x0 - x:x1 - x]


def pair_shift(seed, w, h, dx, dy):
    """A plain integer translation pair (edge clamp), for reach and exact-vector checks."""
    rng = rng_of(seed)
    prev = to_rgba(_background(rng, h, w), _chroma(rng, h, w))
    return prev, _shift_clamp(prev, dx, dy).copy()


def cut_frame(rng, h, w):
    """An unrelated frame of binary noise in 2x2-pixel cells (grey 0 or 255): the flow-resolution luma keeps its full
    contrast, no vector explains it, and the mean block SAD against any other picture is far above 30 levels."""
    cells = np.where(rng.random(((h + 1) // 2, (w + 1) // 2)) < 0.5, 0, 255).astype(np.uint8)
    g = np.repeat(np.repeat(cells, 2, 0), 2, 1)[:h, :w]
    out = np.empty((h, w, 4), np.uint8)
    out[..., :3] = g[..., None]
    out[..., 3] = 255
    return out


def noise_pair(seed, w, h):
    """A value-noise frame followed by an unrelated binary-noise frame (a scene cut)."""
    rng = rng_of(seed)
    a = to_rgba(_background(rng, h, w), _chroma(rng, h, w))
    return a, cut_frame(rng, h, w)


def _bilinear_shift(canvas, pad, h, w, sx, sy):
    """canvas (h + 2 pad, w + 2 pad[, C]) sampled at (x + pad - sx, y + pad - sy): the content moves by (sx, sy)."""
    ys = np.arange(h, dtype=np.float64) + pad - sy
    xs = np.arange(w, dtype=np.float64) + pad - sx
    iy, ix = np.floor(ys).astype(np.int64), np.floor(xs).astype(np.int64)
    fy, fx = (ys - iy), (xs - ix)
    if canvas.ndim == 3:
        fy, fx = fy[:, None, None], fx[None, :, None]
    else:
        fy, fx = fy[:, None], fx[None, :]
    a = canvas[iy][:, ix]
    b = canvas[iy][:, ix + 1]
    c = canvas[iy + 1][:, ix]
    d = canvas[iy + 1][:, ix + 1]
    top = a + (b - a) * fx
    bot = c + (d - c) * fx
    return top + (bot - top) * fy


def pairs_1080p(seed=2, n_pairs=8, w=1920, h=1080, max_shift=16.0, sprite=96, n_sprites=4, hud=(300, 60),
                cut_pair=7):
    """The 1080p 4x workload: a list of (prev, curr) uint8 RGBA pairs and their global shifts."""
    rng = rng_of(seed)
    pad = int(max_shift) + 4
    hud_rng = np.random.default_rng(rng.integers(1 << 62))
    hud_w, hud_h = hud
    hud_bits = np.where(hud_rng.random((hud_h // 6, hud_w // 4)) < 0.35, 255, 24).astype(np.float64)
    hud_img = np.repeat(np.repeat(hud_bits, 6, 0), 4, 1)[:hud_h, :hud_w]
    hud_pos = (w - hud_w - 40, h - hud_h - 30)
    pairs, shifts = [], []
    for k in range(n_pairs):
        lum = _background(rng, h + 2 * pad, w + 2 * pad)
        chroma = _chroma(rng, h + 2 * pad, w + 2 * pad)
        canvas = lum[..., None] + chroma
        sx = rng.integers(-4 * max_shift, 4 * max_shift + 1) / 4.0
        sy = rng.integers(-4 * max_shift, 4 * max_shift + 1) / 4.0
        a = canvas[pad:pad + h, pad:pad + w]
        c = _bilinear_shift(canvas, pad, h, w, sx, sy)
        a, c = a.copy(), c.copy()
        for _ in range(n_sprites):
            tex = 16.0 + 223.0 * value_noise(rng, sprite, sprite, (24, 8, 3))
            tex = tex[..., None] + np.array([rng.uniform(-40, 40) for _ in range(3)])
            x0 = int(rng.integers(64, w - sprite - 64))
            y0 = int(rng.integers(64, h - sprite - 64))
            vx, vy = int(rng.integers(-24, 25)), int(rng.integers(-24, 25))
            a[y0:y0 + sprite, x0:x0 + sprite] = tex
            x1, y1 = x0 + vx, y0 + vy
            c[y1:y1 + sprite, x1:x1 + sprite] = tex
        for img in (a, c):
            img[hud_pos[1]:hud_pos[1] + hud_h, hud_pos[0]:hud_pos[0] + hud_w] = hud_img[..., None]
        prev = to_rgba(np.zeros((h, w)), a)
        if k == cut_pair:
            curr = cut_frame(rng, h, w)
        else:
            curr = to_rgba(np.zeros((h, w)), c)
        pairs.append((prev, curr))
        shifts.append((sx, sy))
    return pairs, shifts
