"""CPU oracle of frame-gen-ritz MotionVersion::V1 (estimate + synth), numpy only.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: only tests/, the smoke check and the
CPU-baseline leg of bench.py use it, as the checker.

It restates proto/interp.py (the algorithm) with the arithmetic of the reference's compute shaders
(gpu/shaders/*.comp, motion_v1.glsl) where interp.py delegates to OpenCV. Every float32 operation is written
out in the order the shaders (or cv2) perform it, so a kernel that follows the same order is bit-exact.

Pinned choices where interp.py and the shaders differ (each also stated where it is used):
  * luma is the C++ 14-bit integer form (luma.comp:23-26), not cv2 5.x's 15-bit one;
  * block SADs: seed candidates reduce with the shader's 64-lane tree (search.comp:41-49), search candidates
    sum y-outer / x-inner sequentially (search.comp:71-74), sub-pixel candidates row by row (subpix.comp:42-55);
  * the per-pixel 5x5 window sum is exact (float64: horizontal 5 then vertical 5), then * (1/25) in float64
    and rounded to float32 -- cv2.blur's CV_32F path (sum in double, scaled in double); the shader's float32
    running sum agrees with it below 4096 levels (pixfield.comp:53-57);
  * the 7x7 mismatch box is an exact integer sum, e7 = float32(S * (1/49)), trust = (e7 - lo) / (hi - lo) in
    float32 (interp.py:236-237); the untrusted count is trust > 0.5 (resolve.comp:44);
  * the scene-cut mean is the C++ fixed-point sum: sum of roundEven(256 * blockSAD) (subpix.comp:68), then
    float(sum) * (1/256) / float(blocks) > sceneCutSad (fallback.comp:10-11);
  * the flick test is float(count) > globalFallback * float(W) * float(H) (fallback.comp:12);
  * the RGBA8 warp is cv2.remap on uint8: map rounded to 1/32 px (round half even), integer weights,
    (sum + 512) >> 10 (round half up of the exact value) (interp.py:214; LAYOUT.md 5.3 "fgSampleRGB8");
  * GLSL mix(x, y, a) is x * (1 - a) + y * a (two products, one sum, no fused multiply-add).
"""
import numpy as np

f32 = np.float32
f64 = np.float64

# interp.py:43-61 constants == framegen.h Settings defaults (framegen.h:173-206, 403-408)
DEFAULTS = dict(trustLow=12.0, trustHigh=40.0, globalFallback=0.15, sceneCutSad=30.0, hudBonus=1.0,
                smoothBonus=0.5, searchPenalty=0.5, searchRadius=2, searchRadiusTop=6, subpixel=True,
                flowScale=2, pyramidMinSide=48)
BLOCK = 8


def settings(**kw):
    s = dict(DEFAULTS)
    for k, v in kw.items():
        if k not in s:
            raise KeyError(k)
        s[k] = v
    return s


def rhe(v):
    """round half to even of a double (framegen.cpp:657 rintHalfEven == cv2 saturate_cast<int>)."""
    return int(round(v))


def layout(W, H, S=2, min_side=48):
    """framegen.cpp:658-685 sizeLayout: flow size and the pyramid levels (w, h, bw, bh), finest first."""
    fw, fh = rhe(W * (1.0 / S)), rhe(H * (1.0 / S))
    lv = []
    w, h = fw, fh
    while True:
        lv.append((w, h, (w + 7) // 8, (h + 7) // 8))
        if min(w, h) // 2 < min_side:
            break
        w, h = rhe(w * 0.5), rhe(h * 0.5)
    return fw, fh, lv


# ----------------------------------------------------------------------------------------------- luma + pyramid

def luma8(frame, bgra=False):
    """luma.comp:23-26 (cv2 BGR2GRAY, 14-bit fixed point): (4899 R + 9617 G + 1868 B + 8192) >> 14."""
    f = frame.astype(np.int64)
    r, g, b = (f[..., 2], f[..., 1], f[..., 0]) if bgra else (f[..., 0], f[..., 1], f[..., 2])
    return (r * 4899 + g * 9617 + b * 1868 + 8192) >> 14


def luma_level0(frame, S, fw, fh, bgra=False):
    """luma.comp:58-72 (interp.py:150 INTER_AREA 1/S): mean of the valid integer lumas of each SxS cell.
    Sums of integers are exact, so the order does not matter; the mean is one float32 division."""
    y = luma8(frame, bgra)
    H, W = y.shape
    Hp, Wp = fh * S, fw * S
    yp = np.zeros((max(Hp, H), max(Wp, W)), np.int64)
    vp = np.zeros_like(yp)
    yp[:H, :W] = y
    vp[:H, :W] = 1
    sums = yp[:Hp, :Wp].reshape(fh, S, fw, S).sum((1, 3))
    n = vp[:Hp, :Wp].reshape(fh, S, fw, S).sum((1, 3))
    return sums.astype(f32) / n.astype(f32)


def down(src, w, h):
    """down.comp:9-26: 2x2 mean of the valid pixels, summed j-outer / i-inner in float32, times 0.25 or
    float32(1 / n)."""
    ph, pw = src.shape
    acc = np.zeros((h, w), f32)
    n = np.zeros((h, w), np.int64)
    for j in range(2):
        for i in range(2):
            ys = 2 * np.arange(h) + j
            xs = 2 * np.arange(w) + i
            valid = (ys < ph)[:, None] & (xs < pw)[None, :]
            vals = src[np.minimum(ys, ph - 1)][:, np.minimum(xs, pw - 1)]
            acc = np.where(valid, acc + vals, acc)
            n += valid
    s = np.where(n == 4, f32(0.25), f32(1.0) / n.astype(f32)).astype(f32)
    return (acc * s).astype(f32)


def pyramid(frame_a, frame_c, S, lv, bgra=False):
    """interp.py:147-152 flow() luma + pyramid(): lists of float32 level images, finest first."""
    fw, fh = lv[0][0], lv[0][1]
    pa = [luma_level0(frame_a, S, fw, fh, bgra)]
    pc = [luma_level0(frame_c, S, fw, fh, bgra)]
    for (w, h, _, _) in lv[1:]:
        pa.append(down(pa[-1], w, h))
        pc.append(down(pc[-1], w, h))
    return pa, pc


# ----------------------------------------------------------------------------------------------- block search

def _fetch(img, x, y):
    """L() motion_v1.glsl:42-44: clamp to edge (== BORDER_REPLICATE == pad_to_block)."""
    h, w = img.shape[:2]
    return img[np.clip(y, 0, h - 1), np.clip(x, 0, w - 1)]


def _block_grid(bw, bh):
    by, bx = np.meshgrid(np.arange(bh) * BLOCK, np.arange(bw) * BLOCK, indexing="ij")
    return bx, by


def block_sad_seq(la, lc, vx, vy):
    """search.comp:71-74: raw SAD sum of each 8x8 block for integer half-vector (vx, vy), float32, y outer,
    x inner, sequential. vx, vy: int arrays (bh, bw) (broadcastable)."""
    bh, bw = np.broadcast(vx, vy).shape
    bx, by = _block_grid(bw, bh)
    acc = np.zeros((bh, bw), f32)
    for y in range(BLOCK):
        for x in range(BLOCK):
            qx, qy = bx + x, by + y
            acc = acc + np.abs(_fetch(la, qx - vx, qy - vy) - _fetch(lc, qx + vx, qy + vy))
    return acc


def block_sad_tree(la, lc, vx, vy):
    """search.comp:41-49: per-pixel |a - c| in lane order lid = 8 y + x, then the shared-memory tree
    red[lid] += red[lid + s], s = 32 .. 1."""
    bh, bw = np.broadcast(vx, vy).shape
    bx, by = _block_grid(bw, bh)
    d = np.empty((bh, bw, 64), f32)
    for lid in range(64):
        qx, qy = bx + (lid & 7), by + (lid >> 3)
        d[..., lid] = np.abs(_fetch(la, qx - vx, qy - vy) - _fetch(lc, qx + vx, qy + vy))
    s = 32
    while s >= 1:
        d = d[..., :s] + d[..., s:2 * s]
        s //= 2
    return d[..., 0]


def search(la, lc, seed, R, lam):
    """search.comp:62-86 (interp.py:116-124): integer +-R window around the seed, dy-major order, cost =
    SAD/64 + lambda * (|dx| + |dy|), strict < (first wins). Returns (vectors int (bh, bw, 2), winning SAD/64)."""
    D = 2 * R + 1
    best = bi = best_sad = None
    for ci in range(D * D):
        dx, dy = ci % D - R, ci // D - R
        s = block_sad_seq(la, lc, seed[..., 0] + dx, seed[..., 1] + dy)
        sadc = (s * f32(1.0 / 64.0)).astype(f32)
        cost = (sadc + f32(lam) * f32(abs(dx) + abs(dy))).astype(f32)
        if ci == 0:
            best, bi, best_sad = cost, np.zeros(cost.shape, np.int64), sadc
        else:
            m = cost < best
            best = np.where(m, cost, best)
            bi = np.where(m, ci, bi)
            best_sad = np.where(m, sadc, best_sad)
    v = np.stack([seed[..., 0] + bi % D - R, seed[..., 1] + bi // D - R], -1)
    return v, best_sad


def global_median2(parent, bw, bh, shift=-1):
    """medhist.comp:73-106 (interp.py:142 np.median of the own-parent vectors): exact per-component median of
    roundEven(parent(min(ix/2, pw-1), min(iy/2, ph-1))) over all fine blocks, bins clamped to [-128, 127];
    returns float32 2 * 0.5 * (k1-th + k2-th), k1 = (n-1)//2, k2 = n//2.
    shift >= 0 (medhist.comp:9-71, FG_WIDE): first-pass bins b = clamp((v >> shift) + 128, 0, 255); the k-th element's
    bin, then the k-th among the low `shift` bits (v & mask) of the vectors in that bin, scanned up to mask - 1."""
    ph, pw = parent.shape[:2]
    iy, ix = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
    v = np.rint(parent[np.minimum(iy // 2, ph - 1), np.minimum(ix // 2, pw - 1)]).astype(np.int64)
    n = bw * bh
    k1, k2 = (n - 1) // 2, n // 2
    out = []
    for c in range(2):
        if shift < 0:
            vals = np.sort(np.clip(v[..., c], -128, 127).ravel())
            m = f32(0.5) * (f32(vals[k1]) + f32(vals[k2]))
        else:
            x = v[..., c].ravel()
            bins = np.clip((x >> shift) + 128, 0, 255)
            mask = (1 << shift) - 1
            ks = []
            for k in (k1, k2):
                hist = np.bincount(bins, minlength=256)
                cum = np.cumsum(hist)
                b = int(np.argmax(cum > k))
                acc = int(cum[b] - hist[b])
                fine = np.bincount(x[bins == b] & mask, minlength=max(mask + 1, 1))
                f = 0
                while f < mask:
                    acc += int(fine[f])
                    if acc > k:
                        break
                    f += 1
                ks.append((b - 128) * (1 << shift) + f)
            m = f32(0.5) * (f32(ks[0]) + f32(ks[1]))
        out.append(f32(2.0) * m)
    return np.array(out, f32)


def medhist_shift(s, n, parent):
    """framegen.cpp:429-451: -1 (the clamped histogram) unless a reach setting is off its default and the reach
    bound Rt 2^(n-1-j) + R (2^(n-1-j) - 1) of the parent level j exceeds 127; then the smallest shift holding it."""
    a = 1 << (n - 1 - parent)
    b = s["searchRadiusTop"] * a + s["searchRadius"] * (a - 1)
    non_default = s["searchRadius"] != 2 or s["searchRadiusTop"] != 6 or s["pyramidMinSide"] != 48
    sh = 0
    while (b >> sh) > 127:
        sh += 1
    return sh if non_default and b > 127 else -1


def seed_from(la, lc, parent, med2, bw, bh, lam, extra=None):
    """search.comp:28-59 (interp.py:131-144): 6 candidates (own parent x2, vertical, horizontal and diagonal
    neighbour parents x2, zero, global median), cost = SAD/64 (+ lambda for all but own), strict <.
    extra: a 7th candidate per block (float32 (bh, bw, 2), rounded half-even), the learned proposal."""
    ph, pw = parent.shape[:2]
    by, bx = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
    px, py = np.minimum(bx // 2, pw - 1), np.minimum(by // 2, ph - 1)
    sy = np.where(by % 2 == 0, -1, 1)
    sx = np.where(bx % 2 == 0, -1, 1)
    pr = np.rint(parent).astype(np.int64)
    cands = []
    for (ox, oy) in ((0, 0), (0, 1), (1, 0), (1, 1)):
        cx = np.clip(px + ox * sx, 0, pw - 1)
        cy = np.clip(py + oy * sy, 0, ph - 1)
        cands.append(2 * pr[cy, cx])
    cands.append(np.zeros((bh, bw, 2), np.int64))
    m = np.rint(med2).astype(np.int64)
    cands.append(np.broadcast_to(m, (bh, bw, 2)).copy())
    if extra is not None:
        cands.append(np.rint(extra).astype(np.int64))
    return _pick_seed(la, lc, cands, lam)


def seed_coarse(la, lc, extra, lam):
    """The coarsest level with a learned proposal: candidates zero (the plain seed, no lambda) and the proposal
    (rounded half-even, + lambda), chosen as seed_from chooses."""
    return _pick_seed(la, lc, [np.zeros(extra.shape, np.int64), np.rint(extra).astype(np.int64)], lam)


def _pick_seed(la, lc, cands, lam):
    best = bi = None
    for k, cv in enumerate(cands):
        s = block_sad_tree(la, lc, cv[..., 0], cv[..., 1])
        c = (s * f32(1.0 / 64.0)).astype(f32)
        if k == 0:
            best, bi = c, np.zeros(c.shape, np.int64)
        else:
            c = (c + f32(lam)).astype(f32)
            msk = c < best
            best = np.where(msk, c, best)
            bi = np.where(msk, k, bi)
    allc = np.stack(cands, 0)
    return np.take_along_axis(allc, bi[None, ..., None].repeat(2, -1), 0)[0]


def median3(v):
    """median3.comp:8-28 (interp.py:127-128 cv2.medianBlur 3): per-component 3x3 median, replicate border."""
    pad = np.pad(v, ((1, 1), (1, 1), (0, 0)), mode="edge")
    bh, bw = v.shape[:2]
    stack = np.stack([pad[j:j + bh, i:i + bw] for j in range(3) for i in range(3)], -1)
    return np.sort(stack, -1)[..., 4].astype(v.dtype)


def _tile_tap(img, gx0, gy0, px, py):
    """subpix.comp:15-24 tap(): bilinear at tile coords p (>= 0) of the tile whose (0, 0) is global (gx0, gy0),
    tile texels clamped to the image; weights (1-fy)(1-fx), (1-fy)fx, fy(1-fx), fy fx summed in that order."""
    flx, fly = np.floor(px).astype(f32), np.floor(py).astype(f32)
    fx, fy = (px - flx).astype(f32), (py - fly).astype(f32)
    ix, iy = gx0 + flx.astype(np.int64), gy0 + fly.astype(np.int64)
    p00, p01 = _fetch(img, ix, iy), _fetch(img, ix + 1, iy)
    p10, p11 = _fetch(img, ix, iy + 1), _fetch(img, ix + 1, iy + 1)
    one = f32(1.0)
    return (p00 * ((one - fy) * (one - fx)) + p01 * ((one - fy) * fx)
            + p10 * (fy * (one - fx)) + p11 * (fy * fx)).astype(f32)


def subpix(la, lc, fieldM, lam):
    """subpix.comp:26-73 (interp.py:162-163): 3x3 at 0.5 px then 3x3 at 0.25 px around the integer vector,
    cost = SAD/64 + (lambda * step) * |d|_1, strict <. Returns (float32 field (bh, bw, 2), int SAD*256 per block)."""
    bh, bw = fieldM.shape[:2]
    s = np.rint(fieldM).astype(np.int64)
    bx, by = _block_grid(bw, bh)
    O = 2
    ax0, ay0 = bx - s[..., 0] - O, by - s[..., 1] - O
    cx0, cy0 = bx + s[..., 0] - O, by + s[..., 1] - O
    base = np.zeros((bh, bw, 2), f32)
    for p, step in enumerate((f32(0.5), f32(0.25))):
        costs, sads = [], []
        for k in range(9):
            d = (f32(k % 3 - 1), f32(k // 3 - 1))
            ox = (base[..., 0] + d[0] * step).astype(f32)
            oy = (base[..., 1] + d[1] * step).astype(f32)
            total = np.zeros((bh, bw), f32)
            for y in range(8):
                row = np.zeros((bh, bw), f32)
                for x in range(8):
                    ta = _tile_tap(la, ax0, ay0, (f32(x + O) - ox).astype(f32), (f32(y + O) - oy).astype(f32))
                    tc = _tile_tap(lc, cx0, cy0, (f32(x + O) + ox).astype(f32), (f32(y + O) + oy).astype(f32))
                    row = (row + np.abs(ta - tc)).astype(f32)
                total = (total + row).astype(f32)
            sad = (total * f32(1.0 / 64.0)).astype(f32)
            sads.append(sad)
            pen = (f32(lam) * step) * f32(abs(d[0]) + abs(d[1]))
            costs.append((sad + f32(pen)).astype(f32))
        best, bi = costs[0], np.zeros((bh, bw), np.int64)
        for k in range(1, 9):
            m = costs[k] < best
            best = np.where(m, costs[k], best)
            bi = np.where(m, k, bi)
        dxb = (bi % 3 - 1).astype(f32)
        dyb = (bi // 3 - 1).astype(f32)
        base = np.stack([(base[..., 0] + dxb * step).astype(f32), (base[..., 1] + dyb * step).astype(f32)], -1)
        best_sad = np.take_along_axis(np.stack(sads, 0), bi[None], 0)[0]
    field = (s.astype(f32) + base).astype(f32)
    sad256 = np.rint((best_sad * f32(256.0)).astype(f32)).astype(np.int64)
    return field, sad256


# ----------------------------------------------------------------------------------------------- per-pixel field

def mix(x, y, a):
    """GLSL mix as pinned here: x * (1 - a) + y * a (float32, no fused multiply-add)."""
    one = f32(1.0)
    return (x * (one - a) + y * a).astype(f32)


def sample_l(img, px, py):
    """sampleL() motion_v1.glsl:48-56 (cv2.remap INTER_LINEAR, BORDER_REPLICATE on float32): map rounded to
    1/32 px (round half even), four clamped taps, weights summed in the order p00, p01, p10, p11."""
    X = np.rint((px * f32(32.0)).astype(f32)).astype(np.int64)
    Y = np.rint((py * f32(32.0)).astype(f32)).astype(np.int64)
    ix, iy = X >> 5, Y >> 5
    fx = ((X & 31).astype(f32) * f32(1.0 / 32.0)).astype(f32)
    fy = ((Y & 31).astype(f32) * f32(1.0 / 32.0)).astype(f32)
    one = f32(1.0)
    p00, p01 = _fetch(img, ix, iy), _fetch(img, ix + 1, iy)
    p10, p11 = _fetch(img, ix, iy + 1), _fetch(img, ix + 1, iy + 1)
    if img.ndim == 3:
        fx, fy = fx[..., None], fy[..., None]
    return (p00 * ((one - fy) * (one - fx)) + p01 * ((one - fy) * fx)
            + p10 * (fy * (one - fx)) + p11 * (fy * fx)).astype(f32)


def _block_v(bf, bx, by):
    bh, bw = bf.shape[:2]
    return bf[np.clip(by, 0, bh - 1), np.clip(bx, 0, bw - 1)]


def pix_candidates(bf, w, h):
    """pixfield.comp:19-35 cand() (interp.py:172-176): the 7 candidate vectors per flow pixel, float32
    (7, h, w, 2): bilinear block field (cv2.resize INTER_LINEAR, s = (q + 0.5)/8 - 0.5), own block, above,
    below, left, right, zero."""
    qy, qx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    sx = ((qx.astype(f32) + f32(0.5)) / f32(8.0) - f32(0.5)).astype(f32)
    sy = ((qy.astype(f32) + f32(0.5)) / f32(8.0) - f32(0.5)).astype(f32)
    flx, fly = np.floor(sx), np.floor(sy)
    fx, fy = (sx - flx).astype(f32)[..., None], (sy - fly).astype(f32)[..., None]
    ix, iy = flx.astype(np.int64), fly.astype(np.int64)
    top = mix(_block_v(bf, ix, iy), _block_v(bf, ix + 1, iy), fx)
    bot = mix(_block_v(bf, ix, iy + 1), _block_v(bf, ix + 1, iy + 1), fx)
    c0 = mix(top, bot, fy)
    bx, by = qx // BLOCK, qy // BLOCK
    cands = [c0, _block_v(bf, bx, by), _block_v(bf, bx, by - 1), _block_v(bf, bx, by + 1),
             _block_v(bf, bx - 1, by), _block_v(bf, bx + 1, by), np.zeros((h, w, 2), f32)]
    return np.stack(cands, 0).astype(f32)


def box_f64(img, r):
    """Exact window sum (2r+1)^2 with replicate border, float64: horizontal sums left to right, then vertical
    sums top to bottom (the order the kernels use)."""
    h, w = img.shape
    d = img.astype(f64)
    xs = np.arange(w)
    hs = np.zeros((h, w), f64)
    for k in range(-r, r + 1):
        hs = hs + d[:, np.clip(xs + k, 0, w - 1)]
    ys = np.arange(h)
    vs = np.zeros((h, w), f64)
    for k in range(-r, r + 1):
        vs = vs + hs[np.clip(ys + k, 0, h - 1)]
    return vs


def pixel_field(la, lc, bf, smooth, zero):
    """pixfield.comp:37-63 (interp.py:167-183): per flow pixel, the candidate whose 5x5 window of
    |a(q - v) - c(q + v)| (each window pixel with its own candidate, replicate border) costs least;
    cost = float32(exact sum * (1/25)) (- smoothBonus for the bilinear, - hudBonus for zero), strict <."""
    h, w = la.shape
    cands = pix_candidates(bf, w, h)
    qy, qx = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    qxf, qyf = qx.astype(f32), qy.astype(f32)
    best = bk = None
    for k in range(7):
        vx, vy = cands[k, ..., 0], cands[k, ..., 1]
        ea = sample_l(la, (qxf - vx).astype(f32), (qyf - vy).astype(f32))
        ec = sample_l(lc, (qxf + vx).astype(f32), (qyf + vy).astype(f32))
        err = np.abs(ea - ec).astype(f32)
        c = (box_f64(err, 2) * (1.0 / 25.0)).astype(f32)
        if k == 0:
            c = (c - f32(smooth)).astype(f32)
        if k == 6:
            c = (c - f32(zero)).astype(f32)
        if k == 0:
            best, bk = c, np.zeros((h, w), np.int64)
        else:
            m = c < best
            best = np.where(m, c, best)
            bk = np.where(m, k, bk)
    return np.take_along_axis(cands, bk[None, ..., None].repeat(2, -1), 0)[0]


# ----------------------------------------------------------------------------------------------- estimate

def estimate(a, c, s=None, bgra=False, luma0=None, size=None, proposals=None):
    """interp.py:186-191 estimate() with the GPU's schedule (framegen.cpp:2932-3153): everything that depends
    only on the pair. The per-pixel field is computed even for a scene cut (LAYOUT.md 3.3); the cut is applied
    in the synth. Returns a dict with every intermediate. luma0 = (level-0 luma of a, of c) and size = (W, H)
    replace the RGBA8 frames for the other formats (oracle/formats.py). proposals: learned seed vectors of the
    coarsest level's blocks (float32 (bh, bw, 2)), an extra seed candidate of that level's search."""
    s = s or DEFAULTS
    W, H = size if size is not None else (a.shape[1], a.shape[0])
    S = s["flowScale"]
    fw, fh, lv = layout(W, H, S, s["pyramidMinSide"])
    if luma0 is None:
        pa, pc = pyramid(a, c, S, lv, bgra)
    else:
        pa, pc = [luma0[0]], [luma0[1]]
        for (w, h, _, _) in lv[1:]:
            pa.append(down(pa[-1], w, h))
            pc.append(down(pc[-1], w, h))
    lam = s["searchPenalty"]
    n = len(lv)
    fields_raw, fields, meds = [None] * n, [None] * n, [None] * n
    sad256 = None
    v = None
    for i in range(n - 1, -1, -1):
        w, h, bw, bh = lv[i]
        if i == n - 1:
            if proposals is None:
                seed = np.zeros((bh, bw, 2), np.int64)
            else:
                seed = seed_coarse(pa[i], pc[i], proposals, lam)
            R = s["searchRadiusTop"]
        else:
            meds[i] = global_median2(v, bw, bh, medhist_shift(s, n, i + 1))
            seed = seed_from(pa[i], pc[i], v, meds[i], bw, bh, lam)
            R = s["searchRadius"]
        vi, sadv = search(pa[i], pc[i], seed, R, lam)
        fields_raw[i] = vi.astype(f32)
        v = median3(fields_raw[i])
        fields[i] = v
        if i == 0 and not s["subpixel"]:
            sad256 = np.rint((sadv * f32(256.0)).astype(f32)).astype(np.int64)
    if s["subpixel"]:
        bfield, sad256 = subpix(pa[0], pc[0], fields[0], lam)
    else:
        bfield = fields[0]
    pf = pixel_field(pa[0], pc[0], bfield, s["smoothBonus"], s["hudBonus"])
    sadsum = int(sad256.sum()) & 0xFFFFFFFF
    nblocks = lv[0][2] * lv[0][3]
    return dict(W=W, H=H, fw=fw, fh=fh, levels=lv, luma_a=pa, luma_c=pc, fields_raw=fields_raw, fields=fields,
                medians=meds, block_field=bfield, sad256=sad256, sadsum=sadsum, nblocks=nblocks, pf=pf,
                cut=bool(cut_decision(sadsum, nblocks, s["sceneCutSad"])), settings=dict(s))


def cut_decision(sadsum, nblocks, cut):
    """fallback.comp:10-11: float(sum) * (1/256) / float(blocks) > sceneCutSad, in float32."""
    mean = (f32(sadsum) * f32(1.0 / 256.0)) / f32(nblocks)
    return bool(f32(mean) > f32(cut))


# ----------------------------------------------------------------------------------------------- synth

def time_constants(t):
    """framegen.cpp:3490-3497: fold s = 2t - 1, wA = 1 - t, wC = t in double from the float t, round to float."""
    td = float(f32(t))
    return f32(2.0 * td - 1.0), f32(1.0 - td), f32(td)


def field_at(pf, t):
    """lookup.comp:10-19 (interp.py:194-205): g <- pf(q - s g) twice (LOOKUP_ITERS), s = float32(2t - 1),
    the product then the difference in float32, pf sampled with sampleF (1/32 map, replicate)."""
    s, _, _ = time_constants(t)
    if f32(t) == f32(0.5):
        return pf
    fh, fw = pf.shape[:2]
    qy, qx = np.meshgrid(np.arange(fh), np.arange(fw), indexing="ij")
    qxf, qyf = qx.astype(f32), qy.astype(f32)
    g = pf
    for _ in range(2):
        px = (qxf - (s * g[..., 0]).astype(f32)).astype(f32)
        py = (qyf - (s * g[..., 1]).astype(f32)).astype(f32)
        g = sample_l(pf, px, py)
    return g


def half_vec(field, W, H, S):
    """halfVec() motion_v1.glsl:73-84 (interp.py:229-231): cv2.resize(field, x S, INTER_LINEAR) + edge pad,
    times S. float32 (H, W, 2)."""
    fh, fw = field.shape[:2]
    py, px = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    px = np.minimum(px, S * fw - 1)
    py = np.minimum(py, S * fh - 1)
    inv = f32(1.0 / S)
    sx = ((px.astype(f32) + f32(0.5)) * inv - f32(0.5)).astype(f32)
    sy = ((py.astype(f32) + f32(0.5)) * inv - f32(0.5)).astype(f32)
    flx, fly = np.floor(sx), np.floor(sy)
    fx, fy = (sx - flx).astype(f32)[..., None], (sy - fly).astype(f32)[..., None]
    i0x = np.clip(flx.astype(np.int64), 0, fw - 1)
    i1x = np.clip(flx.astype(np.int64) + 1, 0, fw - 1)
    i0y = np.clip(fly.astype(np.int64), 0, fh - 1)
    i1y = np.clip(fly.astype(np.int64) + 1, 0, fh - 1)
    top = mix(field[i0y, i0x], field[i0y, i1x], fx)
    bot = mix(field[i1y, i0x], field[i1y, i1x], fx)
    return (mix(top, bot, fy) * f32(S)).astype(f32)


def maps32(px, py):
    """the warp's map in 1/32 px: 32 p rounded half even (int64)"""
    X = np.rint((px * f32(32.0)).astype(f32)).astype(np.int64)
    Y = np.rint((py * f32(32.0)).astype(f32)).astype(np.int64)
    return X, Y


def warp_rgba8(frame, px, py):
    """cv2.remap INTER_LINEAR / BORDER_REPLICATE on uint8 (interp.py:214): map rounded to 1/32 px (half even),
    integer weights (32 - fx)(32 - fy) .. fx fy, result (sum + 512) >> 10. int64 (H, W, 4)."""
    X, Y = maps32(px, py)
    ix, iy, fx, fy = X >> 5, Y >> 5, X & 31, Y & 31
    img = frame.astype(np.int64)
    p00, p01 = _fetch(img, ix, iy), _fetch(img, ix + 1, iy)
    p10, p11 = _fetch(img, ix, iy + 1), _fetch(img, ix + 1, iy + 1)
    w00 = ((32 - fx) * (32 - fy))[..., None]
    w01 = (fx * (32 - fy))[..., None]
    w10 = ((32 - fx) * fy)[..., None]
    w11 = (fx * fy)[..., None]
    return (p00 * w00 + p01 * w01 + p10 * w10 + p11 * w11 + 512) >> 10


def box_int(img, r):
    """Exact integer window sum with replicate border."""
    h, w = img.shape
    xs, ys = np.arange(w), np.arange(h)
    hs = np.zeros((h, w), np.int64)
    for k in range(-r, r + 1):
        hs += img[:, np.clip(xs + k, 0, w - 1)]
    vs = np.zeros((h, w), np.int64)
    for k in range(-r, r + 1):
        vs += hs[np.clip(ys + k, 0, h - 1)]
    return vs


def synth(a, c, est, t, exact64=False, bgra=False, ui=None):
    """interp.py:219-246 synth() with the GPU's arithmetic (lookup.comp, error.comp, resolve.comp,
    fallback.comp). Returns (uint8 frame (H, W, 4), reason None / "cut" / "untrusted", dict of intermediates).
    exact64: the blend (trust ramp, time mix, resolve) in float64 -- the FP64 reference of the output; the
    sampling geometry (vectors, 1/32 maps) and the decisions stay the pinned float32 ones.
    ui: (maskA, rectA, maskC, rectC) of a CrosshairV2 pair: the warp taps are rejected as ui.reject does
    (motion_v1.glsl:124, before the mismatch and the blend); the paste-back is ui.patch, after this."""
    s = est["settings"]
    H, W = a.shape[:2]
    S = s["flowScale"]
    ts, wA, wC = time_constants(t)
    near_c = wC > f32(0.5)
    near = c if near_c else a
    info = {}
    if est["cut"]:
        return near.copy(), "cut", info
    g = field_at(est["pf"], t)
    hv = half_vec(g, W, H, S)
    py, px = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    pxf, pyf = px.astype(f32), py.astype(f32)
    ka, kc = f32(2.0) * wC, f32(2.0) * wA
    pax, pay = (pxf - (ka * hv[..., 0]).astype(f32)).astype(f32), (pyf - (ka * hv[..., 1]).astype(f32)).astype(f32)
    pcx, pcy = (pxf + (kc * hv[..., 0]).astype(f32)).astype(f32), (pyf + (kc * hv[..., 1]).astype(f32)).astype(f32)
    wa = warp_rgba8(a, pax, pay)
    wc = warp_rgba8(c, pcx, pcy)
    if ui is not None:
        from . import ui as uimod
        wa, wc = uimod.reject(wa, wc, maps32(pax, pay), maps32(pcx, pcy), ui)
    e = np.abs(wa[..., :3] - wc[..., :3]).max(-1)
    S49 = box_int(e, 3)
    lo, hi = f32(s["trustLow"]), f32(s["trustHigh"])
    den = f32(hi - lo)
    e7 = (S49.astype(f64) * (1.0 / 49.0)).astype(f32)
    u32 = np.clip(((e7 - lo).astype(f32) / den).astype(f32), f32(0), f32(1)).astype(f32)
    count = int((u32 > f32(0.5)).sum())
    info.update(g=g, hv=hv, e=e, S49=S49, count=count)
    flick = f32(count) > (f32(s["globalFallback"]) * f32(W)) * f32(H)
    if flick:
        return near.copy(), "untrusted", info
    n8 = near[..., :3].astype(np.int64)
    wa3, wc3 = wa[..., :3], wc[..., :3]
    if exact64:
        u = np.clip((S49.astype(f64) / 49.0 - float(lo)) / float(den), 0.0, 1.0)[..., None]
        if wC == f32(0.5):
            m = 0.5 * (wa3 + wc3)
        else:
            m = float(wA) * wa3 + float(wC) * wc3
        o = m * (1.0 - u) + n8 * u
    else:
        u = u32[..., None]
        if wC == f32(0.5):
            m = (f32(0.5) * (wa3 + wc3).astype(f32)).astype(f32)
        else:
            m = ((wA * wa3.astype(f32)).astype(f32) + (wC * wc3.astype(f32)).astype(f32)).astype(f32)
        o = ((m * (f32(1.0) - u)).astype(f32) + (n8.astype(f32) * u).astype(f32)).astype(f32)
    out = np.empty_like(a)
    out[..., :3] = np.clip(np.floor(o + 0.5), 0, 255).astype(np.uint8)
    out[..., 3] = near[..., 3]
    return out, None, info


def blur_samples(t0, t1, n, gauss):
    """framegen.cpp:3549-3555 + blur_resolve.comp:128-135: the midpoints t_k of n equal slices of [t0, t1] (double,
    then float32, clamped to [t0, t1]) and the weights (1, or exp(-2 x^2), x = (2k + 1 - n) / n, double -> float32)."""
    t0, t1 = float(f32(t0)), float(f32(t1))
    ts, ws = [], []
    for k in range(n):
        t = f32(t0 + (t1 - t0) * (k + 0.5) / n)
        ts.append(min(max(t, f32(t0)), f32(t1)))
        x = (2.0 * k + 1.0 - n) / n
        ws.append(f32(np.exp(-2.0 * x * x)) if gauss else f32(1.0))
    return ts, ws


def average_frames(frames, ws, hi=255, alpha_hi=255):
    """The blur's rounding: acc = sum_k w_k v_k (float32, in order), out = clamp(floor(acc / wsum + 0.5), 0, hi)."""
    wsum = f32(0.0)
    for w in ws:
        wsum = f32(wsum + w)
    acc = np.zeros(frames[0].shape, f32)
    for v, w in zip(frames, ws):
        acc = (acc + (w * v.astype(f32)).astype(f32)).astype(f32)
    return np.clip(np.floor((acc / wsum).astype(f32) + f32(0.5)), 0, hi)


def synth_blur(a, c, est, t0, t1, n, gauss=False, ui=None):
    """recordSynthBlur (framegen.cpp:3534-3634) with blurSamples = n >= 2, t0 < t1: the weighted mean of the
    RGBA8 frames synth() gives at the sample times (prev / curr at t <= 0 / t >= 1), rounded once. ui: as synth()."""
    ts, ws = blur_samples(t0, t1, n, gauss)
    frames = []
    for t in ts:
        if t <= 0:
            frames.append(a)
        elif t >= 1:
            frames.append(c)
        else:
            frames.append(synth(a, c, est, float(t), ui=ui)[0])
    return average_frames(frames, ws).astype(np.uint8)


def interpolate_multi(a, c, n, s=None):
    """interp.py:254-258: one estimate, then the n - 1 frames at t = k / n."""
    est = estimate(a, c, s)
    return est, [synth(a, c, est, k / n) for k in range(1, n)]
