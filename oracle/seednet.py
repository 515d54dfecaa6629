"""CPU oracle of the learned seed network (the library's seednet: forward, mean squared error, backward), numpy.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: only tests/ use it, as the checker.

The network (no counterpart in the reference; the library's own addition): 3x3 convolutions, padding 1, channels
2 -> 16 -> 32 -> 64 -> 64 -> 2, stride 2 in layers 1-3, ReLU after layers 1-4; input the luma pair normalised
(v - 128) / 64; parameters per layer W [Cout][9 Cin] (k = (ky * 3 + kx) * Cin + ci) then b [Cout], layers in order.

With emulate=True the products take bf16 operands as the GPU's matrix cores do (each operand rounded to bf16, round
to nearest even) and are summed in float64; the GPU sums in float32 in its own order, so the comparison is to a
tolerance. The analytic backward is pinned by central differences in float64 (tests/test_oracle_seednet.py).
"""
import numpy as np

LAYERS = [(2, 16, 2, True), (16, 32, 2, True), (32, 64, 2, True), (64, 64, 1, True), (64, 2, 1, False)]


def param_count():
    return sum(co * 9 * ci + co for ci, co, _, _ in LAYERS)


def split(params):
    out, o = [], 0
    for ci, co, _, _ in LAYERS:
        W = params[o:o + co * 9 * ci].reshape(co, 9 * ci)
        o += co * 9 * ci
        b = params[o:o + co]
        o += co
        out.append((W, b))
    return out


def bf16(x):
    """float32 -> nearest-even bf16, as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def out_size(v, s):
    return (v + s - 1) // s


def im2col(x, stride):
    """x (n, h, w, C) -> (n, oh, ow, 9 C), k = (ky * 3 + kx) * C + c, zero padding 1."""
    n, h, w, C = x.shape
    oh, ow = out_size(h, stride), out_size(w, stride)
    xp = np.zeros((n, h + 2 + stride, w + 2 + stride, C), x.dtype)
    xp[:, 1:h + 1, 1:w + 1] = x
    cols = []
    for ky in range(3):
        for kx in range(3):
            cols.append(xp[:, ky:ky + stride * oh:stride, kx:kx + stride * ow:stride][:, :oh, :ow])
    return np.concatenate(cols, -1)


def col2im(dA, shape, stride):
    """The adjoint of im2col: dA (n, oh, ow, 9 C) -> dx of shape (n, h, w, C)."""
    n, h, w, C = shape
    oh, ow = dA.shape[1:3]
    dxp = np.zeros((n, h + 2 + stride, w + 2 + stride, C), dA.dtype)
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            dxp[:, ky:ky + stride * oh:stride, kx:kx + stride * ow:stride][:, :oh, :ow] += dA[..., t * C:(t + 1) * C]
    return dxp[:, 1:h + 1, 1:w + 1]


def net_input(la, lc):
    return np.stack([(la - np.float32(128.0)) * np.float32(1 / 64), (lc - np.float32(128.0)) * np.float32(1 / 64)],
                    -1).astype(np.float32)


def forward(la, lc, params, emulate=True, dtype=np.float32):
    """Returns (activations of layers 1..5 as `dtype`, their im2col inputs)."""
    x = net_input(la, lc).astype(dtype)
    acts, cols = [], []
    for (W, b), (ci, co, s, relu) in zip(split(params), LAYERS):
        A = im2col(x, s)
        Ae, We = (bf16(A), bf16(W)) if emulate else (A, W)
        z = Ae.astype(np.float64) @ We.astype(np.float64).T + b.astype(np.float64)
        a = np.maximum(z, 0.0) if relu else z
        x = a.astype(dtype)
        acts.append(x)
        cols.append(A)
    return acts, cols


def mse(pred, target):
    nb = pred.size // 2
    d = pred.astype(np.float64) - target
    return float((d * d).sum() / nb), (2.0 * d / nb)


def backward(la, lc, params, target, emulate=True, dtype=np.float32):
    """Gradients of mse(forward(...)[-1], target) with respect to the parameters (flat, like params), and the loss."""
    acts, cols = forward(la, lc, params, emulate, dtype)
    loss, g = mse(acts[-1], target)
    ps = split(params)
    grads = [None] * len(LAYERS)
    x_shapes = [net_input(la, lc).shape] + [a.shape for a in acts[:-1]]
    for l in range(len(LAYERS) - 1, -1, -1):
        ci, co, s, relu = LAYERS[l]
        W, b = ps[l]
        dY = g * (acts[l] > 0) if relu else g
        dYe, Ae, We = (bf16(dY.astype(np.float32)), bf16(cols[l]), bf16(W)) if emulate else (dY, cols[l], W)
        n_, oh, ow = dY.shape[:3]
        dW = dYe.reshape(-1, co).astype(np.float64).T @ Ae.reshape(-1, 9 * ci).astype(np.float64)
        db = dY.reshape(-1, co).astype(np.float64).sum(0)
        grads[l] = np.concatenate([dW.ravel(), db])
        if l > 0:
            dA = dYe.astype(np.float64) @ We.astype(np.float64)
            g = col2im(dA, x_shapes[l], s)
            if emulate:
                g = g.astype(np.float32)
    return np.concatenate(grads), loss


def e4m3(v):
    """float32 -> OCP FP8 E4M3 value (bias 7, largest 448), round to nearest even, saturating (as float32)."""
    v = np.asarray(v, np.float32)
    a = np.abs(v).astype(np.float32)
    out = np.empty_like(a)
    sub = a < np.float32(2.0 ** -6)
    out[sub] = np.rint(a[sub] * np.float32(512.0)) / np.float32(512.0)
    nrm = ~sub
    e = np.frexp(a[nrm])[1] - 1
    out[nrm] = np.ldexp(np.rint(np.ldexp(a[nrm], 3 - e)), e - 3).astype(np.float32)
    out = np.minimum(out, np.float32(448.0))
    out[~(a < np.float32(448.0))] = np.float32(448.0)
    return np.where(v < 0, -out, out).astype(np.float32)


def fp8_tables(params, sa_all):
    """The FP8 weights of every layer as the library builds them from the input scales sa: W' = W sa[ci] (float32),
    sw[co] = max |W'[co]| / 448, wq = e4m3(W' / sw). Returns [(wq values (Cout, 9 Cin), sw (Cout))]."""
    out, o = [], 0
    for (W, b), (ci, co, _, _) in zip(split(params), LAYERS):
        sa = sa_all[o:o + ci].astype(np.float32)
        o += ci
        Wp = (W.astype(np.float32) * np.tile(sa, 9)[None, :]).astype(np.float32)
        m = np.abs(Wp).max(1)
        sw = np.where(m > 0, (m / np.float32(448.0)).astype(np.float32), np.float32(1.0)).astype(np.float32)
        out.append((e4m3((Wp / sw[:, None]).astype(np.float32)), sw))
    return out


def forward_fp8(la, lc, params, sa_all):
    """FP8 forward: each layer's input x quantised per channel as e4m3(x / sa[ci]), times the FP8 weights, summed
    (float64 here), times sw[co], plus the bias; ReLU after layers 1-4."""
    x = net_input(la, lc)
    o = 0
    tabs = fp8_tables(params, sa_all)
    for (W, b), (ci, co, s, relu), (wq, sw) in zip(split(params), LAYERS, tabs):
        sa = sa_all[o:o + ci].astype(np.float32)
        o += ci
        A = e4m3((im2col(x, s) / np.tile(sa, 9)).astype(np.float32))
        z = (A.astype(np.float64) @ wq.astype(np.float64).T).astype(np.float32) * sw + b.astype(np.float32)
        x = (np.maximum(z, 0) if relu else z).astype(np.float32)
    return x, tabs
