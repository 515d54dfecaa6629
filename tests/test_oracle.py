"""Pins of the CPU oracle (oracle/v1.py) without a GPU.

The reference holds no golden vectors for V1 (its goldens are regenerated from interp.py + OpenCV at test time, and
neither runs here), so each oracle function is pinned by at least one of:
  * the scalar restatement in tests/bruteforce.py (written per block / pixel from the shaders), bit for bit;
  * a library routine or closed form (scipy.ndimage median / convolution, np.median, exact rationals, BT.601 luma);
  * invariants of the algorithm (exact vectors and output for even integer shifts, static pairs, scene cuts, the
    interp.py self-check thresholds, interp.py:261-291).
"""
import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.signal as sig

import bruteforce as bf
from fgr import scenes
from oracle import v1

f32 = np.float32


def psnr(x, y):
    d = np.mean((x.astype(np.float64) - y.astype(np.float64)) ** 2)
    return 99.0 if d == 0 else 10 * math.log10(255.0 ** 2 / d)


# --------------------------------------------------------------------------------------------- layout and luma

def test_layout_matches_the_reference_examples():
    # LAYOUT.md:91 (S=1: L=5, S=2: L=4, S=4: L=3 at 1920x1080) and SURVEY block grids
    assert v1.layout(1920, 1080, 2)[2] == [(960, 540, 120, 68), (480, 270, 60, 34), (240, 135, 30, 17), (120, 68, 15, 9)]
    assert len(v1.layout(1920, 1080, 1)[2]) == 5 and len(v1.layout(1920, 1080, 4)[2]) == 3
    assert v1.layout(640, 360, 2)[2] == [(320, 180, 40, 23), (160, 90, 20, 12)]
    assert v1.layout(3840, 2160, 2)[2][0] == (1920, 1080, 240, 135) and len(v1.layout(3840, 2160, 2)[2]) == 5
    assert v1.layout(1280, 720, 2)[2][-1][:2] == (160, 90) and len(v1.layout(1280, 720, 2)[2]) == 3
    # round half even: 135 / 2 = 67.5 -> 68, 65 / 2 = 32.5 -> 32
    assert v1.rhe(67.5) == 68 and v1.rhe(32.5) == 32


def test_luma_matches_bt601_and_integer_formula():
    px = np.array([[[255, 0, 0, 255], [0, 255, 0, 255], [0, 0, 255, 255], [255, 255, 255, 255]]], np.uint8)
    assert v1.luma8(px).tolist() == [[76, 150, 29, 255]]          # round(0.299 / 0.587 / 0.114 * 255)
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (50, 70, 4), dtype=np.uint8)
    y = v1.luma8(img)
    ref = 0.299 * img[..., 0] + 0.587 * img[..., 1] + 0.114 * img[..., 2]
    assert np.abs(y - ref).max() <= 1                               # the 14-bit form is BT.601 within 1 level
    assert np.array_equal(v1.luma8(img[..., [2, 1, 0, 3]], bgra=True), y)


@pytest.mark.parametrize("W,H,S", [(37, 23, 2), (40, 24, 4), (41, 29, 4), (19, 17, 1)])
def test_luma_level0_is_the_cell_mean(W, H, S):
    rng = np.random.default_rng(W)
    img = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    fw, fh, lv = v1.layout(W, H, S, 16)
    got = v1.luma_level0(img, S, fw, fh)
    y = v1.luma8(img)
    for oy in range(fh):
        for ox in range(fw):
            cell = y[S * oy:S * oy + S, S * ox:S * ox + S]
            assert got[oy, ox] == f32(cell.sum()) / f32(cell.size)


def test_down_is_the_valid_pixel_mean():
    rng = np.random.default_rng(1)
    src = (rng.integers(0, 1021, (9, 13)) / 4.0).astype(f32)       # level-0 values at flow scale 2
    d = v1.down(src, v1.rhe(13 * 0.5), v1.rhe(9 * 0.5))
    assert d.shape == (4, 6)   # 4.5 -> 4, 6.5 -> 6
    for oy in range(4):
        for ox in range(6):
            cell = src[2 * oy:2 * oy + 2, 2 * ox:2 * ox + 2].astype(np.float64)
            assert d[oy, ox] == cell.mean()                            # exact dyadic mean


# --------------------------------------------------------------------------------------------- library reductions

def test_median3_is_scipy_median_filter():
    rng = np.random.default_rng(2)
    v = rng.integers(-9, 10, (11, 17, 2)).astype(f32)
    got = v1.median3(v)
    for c in range(2):
        assert np.array_equal(got[..., c], ndi.median_filter(v[..., c], size=3, mode="nearest"))


def test_box_sums_are_convolutions():
    rng = np.random.default_rng(3)
    e = rng.integers(0, 256, (23, 31))
    pad = np.pad(e, 3, mode="edge")
    assert np.array_equal(v1.box_int(e, 3), sig.convolve2d(pad, np.ones((7, 7), np.int64), mode="valid"))
    x = (rng.integers(0, 1 << 20, (19, 21)) / 4096.0).astype(f32)
    padx = np.pad(x.astype(np.float64), 2, mode="edge")
    assert np.array_equal(v1.box_f64(x, 2), sig.convolve2d(padx, np.ones((5, 5)), mode="valid"))


def test_global_median_is_np_median_of_own_parents():
    rng = np.random.default_rng(4)
    for bw, bh, pw, ph in ((15, 9, 8, 5), (30, 17, 15, 9), (7, 7, 4, 4), (17, 12, 9, 6)):
        parent = rng.integers(-20, 21, (ph, pw, 2)).astype(f32)
        iy, ix = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
        own = 2 * parent[np.minimum(iy // 2, ph - 1), np.minimum(ix // 2, pw - 1)]
        want = np.median(own.reshape(-1, 2), 0)                     # interp.py:142
        assert np.array_equal(v1.global_median2(parent, bw, bh), want.astype(f32))


def test_samplers_reduce_to_texels_and_means():
    rng = np.random.default_rng(5)
    img = (rng.integers(0, 1021, (12, 14)) / 4.0).astype(f32)
    ys, xs = np.meshgrid(np.arange(12), np.arange(14), indexing="ij")
    assert np.array_equal(v1.sample_l(img, xs.astype(f32), ys.astype(f32)), img)     # integer map = texel
    half = v1.sample_l(img, (xs[:-1, :-1] + 0.5).astype(f32), ys[:-1, :-1].astype(f32))
    assert np.array_equal(half, ((img[:-1, :-1].astype(np.float64) + img[:-1, 1:]) / 2).astype(f32))
    rgba = rng.integers(0, 256, (12, 14, 4), dtype=np.uint8)
    w = v1.warp_rgba8(rgba, xs.astype(f32), ys.astype(f32))
    assert np.array_equal(w, rgba.astype(np.int64))
    # 1/32 map: the exact bilinear value rounded half up (cv2's fixed-point remap)
    px = (xs[:-1, :-1] + 0.40625).astype(f32)                       # 13 / 32
    py = (ys[:-1, :-1] + 0.78125).astype(f32)                       # 25 / 32
    w = v1.warp_rgba8(rgba, px, py)[..., :3]
    a = rgba.astype(np.float64)
    fx, fy = 13 / 32, 25 / 32
    exact = ((1 - fx) * (1 - fy) * a[:-1, :-1] + fx * (1 - fy) * a[:-1, 1:] + (1 - fx) * fy * a[1:, :-1]
             + fx * fy * a[1:, 1:])[..., :3]
    assert np.array_equal(w, np.floor(exact + 0.5).astype(np.int64))


def test_half_vec_and_lookup_of_constant_fields():
    field = np.full((10, 12, 2), f32(1.25), f32)
    field[..., 1] = f32(-0.75)
    hv = v1.half_vec(field, 24, 20, 2)
    assert np.all(hv[..., 0] == 2.5) and np.all(hv[..., 1] == -1.5)
    for t in (0.25, 0.75, 0.3):
        assert np.array_equal(v1.field_at(field, t), field)


def test_pix_candidates_closed_forms():
    rng = np.random.default_rng(6)
    bfield = (rng.integers(-16, 17, (4, 5, 2)) / 4.0).astype(f32)
    c = v1.pix_candidates(bfield, 40, 32)
    # the bilinear candidate is cv2.resize INTER_LINEAR: at q, s = (q + 0.5)/8 - 0.5 (multiples of 1/16), the exact
    # bilinear of the clamped block vectors (all values dyadic, so float32 holds it exactly)
    bh, bw = bfield.shape[:2]
    B = bfield.astype(np.float64)
    for qy in range(32):
        for qx in range(40):
            sx, sy = (qx + 0.5) / 8 - 0.5, (qy + 0.5) / 8 - 0.5
            ix, iy = math.floor(sx), math.floor(sy)
            fx, fy = sx - ix, sy - iy
            at = lambda x, y: B[min(max(y, 0), bh - 1), min(max(x, 0), bw - 1)]
            want = ((1 - fy) * ((1 - fx) * at(ix, iy) + fx * at(ix + 1, iy))
                    + fy * ((1 - fx) * at(ix, iy + 1) + fx * at(ix + 1, iy + 1)))
            assert np.array_equal(c[0, qy, qx].astype(np.float64), want)
    assert np.array_equal(c[1, 9, 17], bfield[1, 2]) and np.array_equal(c[2, 9, 17], bfield[0, 2])
    assert np.array_equal(c[3, 9, 17], bfield[2, 2]) and np.array_equal(c[4, 9, 17], bfield[1, 1])
    assert np.array_equal(c[5, 9, 17], bfield[1, 3]) and not c[6].any()


# --------------------------------------------------------------------------------------------- t != 1/2

def _to_f32(q):
    """The rational q rounded to the nearest float32 (ties to even; normal range), as a Fraction."""
    if q == 0:
        return Fraction(0)
    sign, a = (-1, -q) if q < 0 else (1, q)
    e = 0
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    return sign * round(a / ulp) * ulp


def _affine_lookup(A, b, fw, fh, s, qx, qy, iters=2, sign=1):
    """lookup.comp:10-19 on the affine field pf(x, y) = A (x, y) + b, in exact rationals: the float32 product s g and
    difference q - s g are rounded with _to_f32, the map to 1/32 px half even, and the bilinear sample of an affine
    field at (X / 32, Y / 32) is the field's value there (bilinear interpolation reproduces affine functions), as long
    as the four taps lie inside the field. Returns the final g, or None when a tap would be clamped."""
    def pf(x, y):
        return (A[0][0] * x + A[0][1] * y + b[0], A[1][0] * x + A[1][1] * y + b[1])
    g = pf(Fraction(qx), Fraction(qy))
    for _ in range(iters):
        px = _to_f32(qx - sign * _to_f32(s * g[0]))
        py = _to_f32(qy - sign * _to_f32(s * g[1]))
        X, Y = round(px * 32), round(py * 32)
        if not (0 <= X // 32 <= fw - 2 and 0 <= Y // 32 <= fh - 2):
            return None
        g = pf(Fraction(X, 32), Fraction(Y, 32))
    return g


@pytest.mark.parametrize("t", [0.25, 0.75, 0.3])
def test_field_at_of_an_affine_field_is_the_two_step_iteration(t):
    """field_at against lookup.comp's g <- pf(q - s g) twice, s = float32(2t - 1), computed in exact rationals on an
    affine dyadic field (where the bilinear sampler is exact). One iteration, a flipped sign of s or the time
    constant of the other t give different fields on most cells, so each of those mistakes fails here."""
    fw, fh = 40, 32
    A = ((Fraction(3, 32), Fraction(-1, 16)), (Fraction(1, 32), Fraction(5, 32)))
    b = (Fraction(-3, 4), Fraction(5, 4))
    qy, qx = np.meshgrid(np.arange(fh), np.arange(fw), indexing="ij")
    pf = np.stack([float(A[k][0]) * qx + float(A[k][1]) * qy + float(b[k]) for k in range(2)], -1).astype(f32)
    got = v1.field_at(pf, t)
    s = Fraction(float(f32(2.0 * float(f32(t)) - 1.0)))
    s_other = Fraction(float(f32(2.0 * float(f32(1.0 - t)) - 1.0)))
    checked = differ_one = differ_sign = differ_t = 0
    for y in range(fh):
        for x in range(fw):
            want = _affine_lookup(A, b, fw, fh, s, x, y)
            if want is None:
                continue
            checked += 1
            assert (Fraction(float(got[y, x, 0])), Fraction(float(got[y, x, 1]))) == want, (x, y)
            differ_one += _affine_lookup(A, b, fw, fh, s, x, y, iters=1) != want
            differ_sign += _affine_lookup(A, b, fw, fh, s, x, y, sign=-1) not in (None, want)
            differ_t += _affine_lookup(A, b, fw, fh, s_other, x, y) not in (None, want)
    assert checked >= 0.6 * fw * fh
    assert min(differ_one, differ_sign, differ_t) >= 0.5 * checked


@pytest.mark.parametrize("t", [0.25, 0.75, 0.3])
def test_time_blend_of_a_zero_motion_pair(t):
    """resolve.comp:73-82 with framegen.cpp:3490-3497's constants: a zero field, frames that differ by at most 5
    levels (every pixel trusted: u = 0), so each output sample is floor(wA a + wC c + 1/2) with wA = float32(1 - t),
    wC = float32(t): (3a + c + 2) >> 2 at t = 1/4, (a + 3c + 2) >> 2 at t = 3/4 (swapped weights fail), the exact
    rational value at t = 0.3 away from rounding ties. Alpha is the nearer frame's."""
    rng = np.random.default_rng(31)
    a = rng.integers(0, 256, (24, 40, 4), dtype=np.uint8)
    c = np.clip(a.astype(np.int64) + rng.integers(-5, 6, a.shape), 0, 255).astype(np.uint8)
    est = dict(pf=np.zeros((12, 20, 2), f32), cut=False, settings=v1.settings())
    out, why, info = v1.synth(a, c, est, t)
    assert why is None and info["count"] == 0
    ai, ci = a[..., :3].astype(np.int64), c[..., :3].astype(np.int64)
    if t == 0.25:
        want = (3 * ai + ci + 2) >> 2
    elif t == 0.75:
        want = (ai + 3 * ci + 2) >> 2
    else:
        wA, wC = Fraction(float(f32(1.0 - float(f32(t))))), Fraction(float(f32(float(f32(t)))))
        exact = [[[wA * int(ai[y, x, k]) + wC * int(ci[y, x, k]) + Fraction(1, 2) for k in range(3)]
                  for x in range(a.shape[1])] for y in range(a.shape[0])]
        want = np.array([[[math.floor(v) for v in px] for px in row] for row in exact])
        far = np.array([[[abs(v - round(v)) > Fraction(1, 2 ** 14) for v in px] for px in row] for row in exact])
        assert far.mean() > 0.7
        got = out[..., :3].astype(np.int64)
        assert np.array_equal(got[far], want[far])
        assert np.abs(got - want).max() <= 1
        return
    assert (want != ((ai + 3 * ci + 2) >> 2 if t == 0.25 else (3 * ai + ci + 2) >> 2)).mean() > 0.5
    assert np.array_equal(out[..., :3], want)
    assert np.array_equal(out[..., 3], (a if t <= 0.5 else c)[..., 3])


@pytest.mark.parametrize("t", [0.25, 0.75, 0.3])
def test_synth_at_any_t_matches_scalar_restatement(t):
    """The vectorised synth (lookup, halfVec, both warps, trust, time blend) equals the scalar per-cell / per-pixel
    restatement at t != 1/2 on a pair with a moving sprite (so trust and the lookup's sampling both matter)."""
    W, H = 80, 56
    a, c = scenes.pair_translate(12, w=W, h=H, sprite=16, sprite_pos=(24, 16), sprite_move=(6, 4), shift=(3, -2))
    s = v1.settings(pyramidMinSide=16)
    est = v1.estimate(a, c, s)
    out, why, info = v1.synth(a, c, est, t)
    rout, rcount = bf.synth_t(a, c, est["pf"], s["flowScale"], s, t)
    assert info["count"] == rcount and rcount > 0
    assert why is None
    assert np.array_equal(out, rout)
    assert np.array_equal(info["g"], bf.lookup_t(est["pf"], t))


# --------------------------------------------------------------------------------------------- arithmetic claims

def test_untrusted_threshold_at_defaults_is_S_above_1274():
    S = np.arange(0, 49 * 255 + 1)
    e7 = (S.astype(np.float64) * (1.0 / 49.0)).astype(f32)
    u = np.clip(((e7 - f32(12)).astype(f32) / f32(28)).astype(f32), 0, 1)
    assert np.array_equal(u > f32(0.5), S > 1274)


@pytest.mark.parametrize("lo,hi", [(12.0, 40.0), (10.0, 30.5), (0.0, 255.0), (3.3, 7.1), (100.25, 101.25),
                                   (1.0, 2.5), (17.7, 200.9)])
def test_trust_quotient_as_double_product_is_the_float_division(lo, hi):
    """The synth kernel forms (e7 - lo) / (hi - lo) as float(double(e7 - lo) * (1 / double(hi - lo))); over every
    7x7 sum this equals numpy's correctly rounded float32 division."""
    S = np.arange(0, 49 * 255 + 1)
    e7 = (S.astype(np.float64) * (1.0 / 49.0)).astype(f32)
    lo32, hi32 = f32(lo), f32(hi)
    den = f32(hi32 - lo32)
    x = (e7 - lo32).astype(f32)
    want = (x / den).astype(f32)
    got = (x.astype(np.float64) * (1.0 / float(den))).astype(f32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("lo,hi", [(12.0, 40.0), (10.0, 30.5), (0.0, 255.0), (3.3, 7.1), (1.0, 2.5),
                                   (17.7, 200.9), (20.0, 21.0), (5.0, 60.0), (12.0, 13.0)])
def test_trust_division_and_resolve_product_take_the_same_untrusted_decision(lo, hi):
    """The oracle's trust is interp.py's float32 quotient (e7 - lo) / (hi - lo); resolve.comp:39 multiplies by the
    reciprocal rounded on the CPU. Over every 7x7 sum the two agree on u > 0.5 (the counted, flick-deciding bit)."""
    S = np.arange(0, 49 * 255 + 1)
    e7 = (S.astype(np.float64) * (1.0 / 49.0)).astype(f32)
    lo32, hi32 = f32(lo), f32(hi)
    x = (e7 - lo32).astype(f32)
    quot = np.clip((x / f32(hi32 - lo32)).astype(f32), 0, 1)
    prod = np.clip((x * f32(1.0 / (float(hi32) - float(lo32)))).astype(f32), 0, 1)
    assert np.array_equal(quot > f32(0.5), prod > f32(0.5))


def test_sums_are_exact_at_the_workload_sizes():
    """Level-k luma at flow scale 2 has 2 + 2k fractional bits; a 64-pixel SAD stays below 2^24 units up to level 4
    (the 4K pyramid), so float32 sums of it are exact in any order."""
    for k in range(5):
        assert 64 * 255 * 2 ** (2 + 2 * k) <= 2 ** 24 * 4 ** max(0, k - 3)
    assert 64 * 255 * 2 ** 10 < 2 ** 24


# --------------------------------------------------------------------------------------------- brute force

@pytest.mark.parametrize("W,H,kw", [
    (96, 64, dict(pyramidMinSide=16)),
    (78, 54, dict(pyramidMinSide=16, searchPenalty=0.3, smoothBonus=0.2, hudBonus=2.0)),
])
def test_estimate_matches_scalar_restatement(W, H, kw):
    a, c = scenes.pair_translate(11, w=W, h=H, sprite=16, sprite_pos=(30, 20), sprite_move=(5, 3), shift=(3, -2))
    s = v1.settings(**kw)
    fw, fh, lv = v1.layout(W, H, s["flowScale"], s["pyramidMinSide"])
    est = v1.estimate(a, c, s)
    ref = bf.estimate(a, c, s, lv)
    for i in range(len(lv)):
        assert np.array_equal(est["luma_a"][i], ref["luma_a"][i])
        assert np.array_equal(est["fields_raw"][i], ref["fields_raw"][i]), f"search L{i}"
        assert np.array_equal(est["fields"][i], ref["fields"][i]), f"median3 L{i}"
        if ref["medians"][i] is not None:
            assert np.array_equal(est["medians"][i], ref["medians"][i])
    assert np.array_equal(est["block_field"], ref["block_field"])
    assert est["sadsum"] == ref["sadsum"]
    assert np.array_equal(est["pf"], ref["pf"])
    out, why, info = v1.synth(a, c, est, 0.5)
    rout, rcount = bf.synth_t05(a, c, est["pf"], s["flowScale"], s)
    assert info["count"] == rcount
    if why is None:
        assert np.array_equal(out[..., :3], rout[..., :3])


# --------------------------------------------------------------------------------------------- invariants

def test_even_integer_shift_gives_exact_vectors_and_frame():
    """Content shifted by (+8, -4) px: the half-vector is (2, -1) flow px (S = 2) on every interior block and pixel,
    and the t = 0.5 frame equals the content shifted by (+4, -2) away from the borders."""
    a, c = scenes.pair_shift(21, 256, 160, 8, -4)
    est = v1.estimate(a, c)
    bfld = est["block_field"]
    assert np.all(bfld[2:-2, 2:-2] == np.array([2.0, -1.0], f32))
    assert np.all(est["pf"][16:-16, 16:-16] == np.array([2.0, -1.0], f32))
    out, why, _ = v1.synth(a, c, est, 0.5)
    mid = scenes._shift_clamp(a, 4, -2)
    assert why is None
    assert np.array_equal(out[24:-24, 24:-24], mid[24:-24, 24:-24])


def test_static_pair_and_scene_cut():
    a, _ = scenes.pair_translate(22, w=128, h=96)
    est = v1.estimate(a, a)
    assert not est["pf"].any() and est["sadsum"] == 0
    for t in (0.25, 0.5, 0.75):
        assert np.array_equal(v1.synth(a, a, est, t)[0], a)
    x, y = scenes.noise_pair(23, 128, 96)
    est = v1.estimate(x, y)
    assert est["cut"]
    for t in (0.25, 0.5, 0.75):
        o, why, _ = v1.synth(x, y, est, t)
        assert why == "cut" and np.array_equal(o, x if t <= 0.5 else y)


def test_interp_selfcheck_thresholds():
    """interp.py:261-291: a textured image panning (24, -12): interpolation beats a plain blend by > 10 dB at 2x and
    at every t of 3x and 4x; a static HUD over the pan stays > 30 dB."""
    rng = np.random.default_rng(0)
    lum = 16.0 + 223.0 * scenes.value_noise(rng, 400, 640, (32, 8, 2), (1.0, 0.5, 0.25))
    tex = scenes.to_rgba(lum)
    shift = lambda img, dx, dy: np.roll(img, (dy, dx), (0, 1))
    a, c = tex, shift(tex, 24, -12)
    m = (slice(40, -40), slice(40, -40))
    est = v1.estimate(a, c)
    for n in (2, 3, 4):
        for k in range(1, n):
            t = k / n
            ref = shift(tex, 24 * k // n, -12 * k // n)
            blend = np.clip(np.floor((1 - t) * a.astype(np.float64) + t * c + 0.5), 0, 255).astype(np.uint8)
            o, why, _ = v1.synth(a, c, est, t)
            assert why is None
            assert psnr(o[m][..., :3], ref[m][..., :3]) > psnr(blend[m][..., :3], ref[m][..., :3]) + 10, (n, k)
    hud = (slice(300, 340), slice(60, 260))
    b = shift(tex, 12, -6)
    a2, b2, c2 = a.copy(), b.copy(), c.copy()
    for f in (a2, b2, c2):
        f[hud] = 255
        f[310:330, 80:240] = 20
        f[..., 3] = 255
    o, _, _ = v1.synth(a2, c2, v1.estimate(a2, c2), 0.5)
    assert psnr(o[hud][..., :3], b2[hud][..., :3]) > 30


def test_wide_median_is_the_exact_median():
    """medhist.comp FG_WIDE: with first-pass bins wide enough for every vector, the two-pass histogram median is the
    exact median (np.median's order statistics); at shift 0 and small vectors it equals the clamped default."""
    rng = np.random.default_rng(3)
    for shift, lim in ((3, 900), (1, 200), (0, 100)):
        parent = rng.integers(-lim, lim + 1, (9, 13, 2)).astype(np.float32)
        bw, bh = 25, 17
        got = v1.global_median2(parent, bw, bh, shift)
        iy, ix = np.meshgrid(np.arange(bh), np.arange(bw), indexing="ij")
        v = parent[np.minimum(iy // 2, 8), np.minimum(ix // 2, 12)].reshape(-1, 2)
        want = np.array([2.0 * np.median(v[:, 0]), 2.0 * np.median(v[:, 1])], np.float32)
        assert np.array_equal(got, want), (shift, got, want)
        if lim <= 127:
            assert np.array_equal(got, v1.global_median2(parent, bw, bh))
    s = v1.settings(searchRadius=12, searchRadiusTop=12, pyramidMinSide=16)
    assert v1.medhist_shift(s, 5, 1) == 1 and v1.medhist_shift(s, 5, 4) == -1 and v1.medhist_shift(v1.DEFAULTS, 6, 1) == -1


# --------------------------------------------------------------------------------------------- blur (recordSynthBlur)

def test_blur_sample_times_are_slice_midpoints_and_weights_a_gaussian():
    """framegen.cpp:3549-3555: the midpoints of n equal slices; blur_resolve.comp:128-135: weights of a Gaussian whose
    window [t0, t1] spans +-2 sigma, sampled at the midpoints (scipy's normal density, normalised to 1 at the centre)."""
    from scipy.stats import norm
    ts, ws = v1.blur_samples(0.25, 0.75, 4, False)
    assert [float(t) for t in ts] == [0.3125, 0.4375, 0.5625, 0.6875]
    assert [float(w) for w in ws] == [1.0] * 4
    for n in (2, 3, 4, 7, 16):
        ts, ws = v1.blur_samples(0.0, 1.0, n, True)
        mids = (np.arange(n) + 0.5) / n
        assert np.array_equal(np.array(ts, f32), mids.astype(f32))
        z = (mids - 0.5) / 0.25                          # sigma = a quarter of the window
        want = (norm.pdf(z) / norm.pdf(0.0)).astype(f32)
        assert np.allclose(np.array(ws, f32), want, rtol=2e-7, atol=0)
        assert ws == ws[::-1]


def test_blur_average_rounds_half_up_once():
    """Two equally weighted frames average to (a + b + 1) >> 1 over every byte pair; any weights of equal frames give
    the frame back; a static pair's blurred frame is the frame."""
    a, b = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    got = v1.average_frames([a.astype(np.uint8), b.astype(np.uint8)], [f32(1.0), f32(1.0)])
    assert np.array_equal(got, (a + b + 1) >> 1)
    _, ws = v1.blur_samples(0.0, 1.0, 5, True)
    frame = np.arange(256, dtype=np.uint8).reshape(16, 16)
    assert np.array_equal(v1.average_frames([frame] * 5, ws), frame)
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (40, 56, 4), dtype=np.uint8)
    img[..., 3] = 255
    est = v1.estimate(img, img)
    assert np.array_equal(v1.synth_blur(img, img, est, 0.2, 0.8, 4, gauss=True), img)
