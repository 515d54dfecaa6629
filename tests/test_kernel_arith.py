"""The integer regroupings and the median network that the kernels use, checked exhaustively (or over every weight) on
the CPU against the formulas they replace: the cv2 fixed-point remap tap of interp.py:214, the trusted blends of
resolve.comp:73-82 at t = 1/4, 3/4, the max-channel mismatch of error.comp:9-21, and the median of median3.comp:8-28.
The GPU tests compare the kernels' outputs with the oracle; these pin the identities over all inputs, including values
the test images may never produce."""
import numpy as np


def lerp_u8(a, b, r):
    """v_lerp_u8 on one byte lane: (a + b + r) >> 1, r the lane's rounding bit"""
    return (a + b + r) >> 1


def test_quarter_blend_is_two_byte_lerps():
    a, c = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    u = lerp_u8(a, c, 0)
    assert np.array_equal(lerp_u8(a, u, 1), (3 * a + c + 2) >> 2)   # t = 1/4: kA = 3, kC = 1
    assert np.array_equal(lerp_u8(c, u, 1), (a + 3 * c + 2) >> 2)   # t = 3/4: kA = 1, kC = 3


def test_mismatch_is_max_of_max_and_minus_min_difference():
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, size=(200000, 3))
    c = rng.integers(0, 256, size=(200000, 3))
    d = a - c
    assert np.array_equal(np.maximum(d.max(axis=1), -d.min(axis=1)), np.abs(d).max(axis=1))


def test_run_warp_regrouping_is_the_cv2_tap():
    """warpRun4: V = (32 - fy) top + fy bottom per texel column (a byte dot product), then
    64 (32 - fx) V0 + 64 fx V1 + 32768 (a 16-bit pair dot product), byte 2 of which is cv2's
    ((32-fx)(32-fy) p00 + fx(32-fy) p01 + (32-fx)fy p10 + fx fy p11 + 512) >> 10"""
    rng = np.random.default_rng(7)
    fx, fy = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    fx, fy = fx.ravel()[:, None], fy.ravel()[:, None]
    p = rng.integers(0, 256, size=(4, 1, 4096))
    p[:, :, :4] = 255                                                # the largest texels too
    p00, p01, p10, p11 = p
    gx, gy = 32 - fx, 32 - fy
    cv2 = (gx * gy * p00 + fx * gy * p01 + gx * fy * p10 + fx * fy * p11 + 512) >> 10
    v0, v1 = gy * p00 + fy * p10, gy * p01 + fy * p11
    assert v0.max() < 1 << 16 and v1.max() < 1 << 16                # each fits a 16-bit lane
    s = 64 * gx * v0 + 64 * fx * v1 + 32768
    assert s.max() < 1 << 24
    assert np.array_equal((s >> 16) & 255, cv2)


# med9 in csrc/v1_estimate.hip: the exchanges (min to the first slot, max to the second) in order
MED9_NET = [(1, 2), (4, 5), (7, 8), (0, 1), (3, 4), (6, 7), (1, 2), (4, 5), (7, 8), (0, 3),
            (5, 8), (4, 7), (3, 6), (1, 4), (2, 5), (4, 7), (4, 2), (6, 4), (4, 2)]


def med9_network(v):
    v = np.array(v, copy=True)
    for a, b in MED9_NET:
        lo, hi = np.minimum(v[..., a], v[..., b]), np.maximum(v[..., a], v[..., b])
        v[..., a], v[..., b] = lo, hi
    return v[..., 4]


def test_median_network_is_the_fifth_smallest():
    """median3.comp's partial selection sort returns the 5th smallest value; the network must too. All 2^9 inputs of
    0s and 1s (the 0-1 principle: a comparator network right on those is right on every input), then random values
    with many ties and the float values of quarter-pel vectors."""
    bits = ((np.arange(512)[:, None] >> np.arange(9)) & 1).astype(np.float32)
    assert np.array_equal(med9_network(bits), np.sort(bits, axis=1)[:, 4])
    rng = np.random.default_rng(3)
    for v in (rng.integers(0, 4, size=(100000, 9)).astype(np.float32),
              (rng.integers(-64, 65, size=(100000, 9)) / 4.0).astype(np.float32)):
        assert np.array_equal(med9_network(v), np.sort(v, axis=1)[:, 4])
