"""The seed network's CPU oracle (oracle/seednet.py) pinned by mathematics: im2col / col2im are adjoint, and the
analytic backward equals central differences of the forward in float64."""
import numpy as np

from oracle import seednet as N


def test_param_count_and_layout():
    assert N.param_count() == 16 * 18 + 16 + 32 * 144 + 32 + 64 * 288 + 64 + 64 * 576 + 64 + 2 * 576 + 2
    p = np.arange(N.param_count(), dtype=np.float64)
    (W0, b0) = N.split(p)[0]
    assert W0.shape == (16, 18) and b0[0] == 16 * 18


def test_im2col_and_col2im_are_adjoint():
    rng = np.random.default_rng(0)
    for s, (h, w) in ((2, (9, 7)), (1, (5, 6)), (2, (8, 8))):
        x = rng.standard_normal((2, h, w, 3))
        A = N.im2col(x, s)
        y = rng.standard_normal(A.shape)
        assert abs((A * y).sum() - (x * N.col2im(y, x.shape, s)).sum()) < 1e-9
        # the centre tap of output pixel (0, 0) is input pixel (0, 0); the top-left tap is padding
        assert np.array_equal(A[:, 0, 0, 4 * 3:5 * 3], x[:, 0, 0]) and not A[:, 0, 0, :3].any()


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    h, w = 16, 24
    la = rng.uniform(16, 240, (1, h, w))
    lc = rng.uniform(16, 240, (1, h, w))
    p = np.zeros(N.param_count())
    for (W, b), (ci, co, _, _) in zip(N.split(p), N.LAYERS):
        W[...] = rng.standard_normal(W.shape) * np.sqrt(2.0 / (9 * ci))
        b[...] = rng.standard_normal(b.shape) * 0.1
    target = rng.standard_normal((1, 2, 3, 2))
    g, loss = N.backward(la, lc, p, target, emulate=False, dtype=np.float64)
    f = lambda q: N.mse(N.forward(la, lc, q, emulate=False, dtype=np.float64)[0][-1], target)[0]
    assert abs(f(p) - loss) < 1e-12
    o = 0
    for ci, co, _, _ in N.LAYERS:
        n = co * 9 * ci + co
        for i in rng.choice(n, 12, replace=False):
            e = np.zeros_like(p)
            e[o + i] = 1e-6
            num = (f(p + e) - f(p - e)) / 2e-6
            assert abs(num - g[o + i]) <= 1e-5 * max(1.0, abs(num)), (ci, co, i, num, g[o + i])
        o += n


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = N.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_e4m3_rounding_and_range():
    v = np.array([0.0, 1.0, 1.0625, 1.1875, 448.0, 470.0, 1e6, -3.3, 2.0 ** -9, 3 * 2.0 ** -10, 2.0 ** -11,
                  0.01708984375], np.float32)
    q = N.e4m3(v)
    # 1.0625 is halfway between 1 and 1.125 -> even (1); 1.1875 halfway between 1.125 and 1.25 -> 1.25; saturation at
    # 448; -3.3 -> -3.25 (step 0.25 in [2, 4)); subnormals are multiples of 2^-9 (3 * 2^-10 is halfway -> 2^-8 even);
    # 2^-11 -> 0; 0.01708984375 (= 2^-6 * 1.09375) -> 2^-6 * 1.125
    assert list(q) == [0.0, 1.0, 1.0, 1.25, 448.0, 448.0, 448.0, -3.25, 2.0 ** -9, 2.0 ** -8, 0.0, 2.0 ** -6 * 1.125]
    # every representable positive value maps to itself
    vals = np.array([m * 2.0 ** -9 for m in range(8)] + [(1 + m / 8) * 2.0 ** e for e in range(-6, 9) for m in range(8)
                                                         if (1 + m / 8) * 2.0 ** e <= 448], np.float32)
    assert np.array_equal(N.e4m3(vals), vals)
