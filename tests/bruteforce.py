"""Scalar restatement of V1 for tiny frames (test infrastructure): one loop per block / pixel, written from the
reference's shaders (gpu/shaders/*.comp) with numpy float32 scalars, independently of the vectorised oracle
(oracle/v1.py), which it cross-checks bit for bit. Frames must be small: this is pure Python."""
import numpy as np

f32 = np.float32


def clampi(v, lo, hi):
    return lo if v < lo else (hi if v > hi else v)


def L(img, x, y):
    h, w = img.shape[:2]
    return img[clampi(y, 0, h - 1), clampi(x, 0, w - 1)]


def luma_levels(a, S, lv):
    """luma.comp main() + down.comp main()."""
    H, W = a.shape[:2]
    fw, fh = lv[0][0], lv[0][1]
    out = np.zeros((fh, fw), f32)
    for oy in range(fh):
        for ox in range(fw):
            s, n = 0, 0
            for j in range(S):
                for i in range(S):
                    px, py = S * ox + i, S * oy + j
                    if px < W and py < H:
                        r, g, b = int(a[py, px, 0]), int(a[py, px, 1]), int(a[py, px, 2])
                        s += (r * 4899 + g * 9617 + b * 1868 + 8192) >> 14
                        n += 1
            out[oy, ox] = f32(s) / f32(n)
    levels = [out]
    for (w, h, _, _) in lv[1:]:
        src = levels[-1]
        ph, pw = src.shape
        d = np.zeros((h, w), f32)
        for oy in range(h):
            for ox in range(w):
                s, n = f32(0), 0
                for j in range(2):
                    for i in range(2):
                        px, py = 2 * ox + i, 2 * oy + j
                        if px < pw and py < ph:
                            s = f32(s + src[py, px])
                            n += 1
                sc = f32(0.25) if n == 4 else f32(f32(1.0) / f32(n))
                d[oy, ox] = f32(s * sc)
        levels.append(d)
    return levels


def search_block(la, lc, bx, by, seed, R, lam):
    """search.comp:62-86 for one block."""
    D = 2 * R + 1
    best_ci, best = 0, None
    for ci in range(D * D):
        dx, dy = ci % D - R, ci // D - R
        vx, vy = seed[0] + dx, seed[1] + dy
        s = f32(0)
        for y in range(8):
            for x in range(8):
                qx, qy = bx * 8 + x, by * 8 + y
                s = f32(s + abs(f32(L(la, qx - vx, qy - vy) - L(lc, qx + vx, qy + vy))))
        sadc = f32(s * f32(1.0 / 64.0))
        cost = f32(sadc + f32(f32(lam) * f32(abs(dx) + abs(dy))))
        if ci == 0 or cost < best:
            best, best_ci, best_sad = cost, ci, sadc
    return (seed[0] + best_ci % D - R, seed[1] + best_ci // D - R), best_sad


def seed_block(la, lc, parent, med2, bx, by, lam):
    """search.comp:28-59 for one block (64-lane tree reduction)."""
    ph, pw = parent.shape[:2]
    px, py = min(bx // 2, pw - 1), min(by // 2, ph - 1)
    sy = -1 if by % 2 == 0 else 1
    sx = -1 if bx % 2 == 0 else 1
    pp = [(0, 0), (0, sy), (sx, 0), (sx, sy)]
    cands = []
    for ox, oy in pp:
        cx, cy = clampi(px + ox, 0, pw - 1), clampi(py + oy, 0, ph - 1)
        v = parent[cy, cx]
        cands.append((2 * int(np.rint(v[0])), 2 * int(np.rint(v[1]))))
    cands.append((0, 0))
    cands.append((int(np.rint(med2[0])), int(np.rint(med2[1]))))
    best, bi = None, 0
    for k, (vx, vy) in enumerate(cands):
        red = []
        for lid in range(64):
            qx, qy = bx * 8 + (lid & 7), by * 8 + (lid >> 3)
            red.append(abs(f32(L(la, qx - vx, qy - vy) - L(lc, qx + vx, qy + vy))))
        s = 32
        while s >= 1:
            for lid in range(s):
                red[lid] = f32(red[lid] + red[lid + s])
            s //= 2
        c = f32(red[0] * f32(1.0 / 64.0))
        if k == 0:
            best = c
        else:
            c = f32(c + f32(lam))
            if c < best:
                best, bi = c, k
    return cands[bi]


def median_of_parents(parent, bw, bh):
    """medhist.comp:73-106: histogram over every fine block's own parent."""
    ph, pw = parent.shape[:2]
    hx, hy = [0] * 256, [0] * 256
    for iy in range(bh):
        for ix in range(bw):
            v = parent[min(iy // 2, ph - 1), min(ix // 2, pw - 1)]
            hx[clampi(int(np.rint(v[0])) + 128, 0, 255)] += 1
            hy[clampi(int(np.rint(v[1])) + 128, 0, 255)] += 1
    n = bw * bh

    def kth(h, k):
        acc = 0
        for b in range(256):
            acc += h[b]
            if acc > k:
                return f32(b - 128)
        return f32(127)

    out = []
    for h in (hx, hy):
        m = f32(f32(0.5) * f32(kth(h, (n - 1) // 2) + kth(h, n // 2)))
        out.append(f32(f32(2.0) * m))
    return np.array(out, f32)


def median3(v):
    bh, bw = v.shape[:2]
    out = np.zeros_like(v)
    for by in range(bh):
        for bx in range(bw):
            for c in range(2):
                vals = [v[clampi(by + j, 0, bh - 1), clampi(bx + i, 0, bw - 1), c] for j in (-1, 0, 1) for i in (-1, 0, 1)]
                for i in range(5):
                    for j in range(i + 1, 9):
                        if vals[j] < vals[i]:
                            vals[i], vals[j] = vals[j], vals[i]
                out[by, bx, c] = vals[4]
    return out


def tap(img, gx0, gy0, px, py):
    flx, fly = np.floor(px), np.floor(py)
    fx, fy = f32(px - flx), f32(py - fly)
    ix, iy = gx0 + int(flx), gy0 + int(fly)
    one = f32(1)
    return f32(f32(f32(L(img, ix, iy) * f32((one - fy) * (one - fx))) + f32(L(img, ix + 1, iy) * f32((one - fy) * fx)))
               + f32(L(img, ix, iy + 1) * f32(fy * (one - fx)))) + f32(L(img, ix + 1, iy + 1) * f32(fy * fx))


def subpix_block(la, lc, bx, by, v, lam):
    """subpix.comp:26-73 for one block."""
    sx, sy = int(np.rint(v[0])), int(np.rint(v[1]))
    O = 2
    base = [f32(0), f32(0)]
    for step in (f32(0.5), f32(0.25)):
        costs, sads = [], []
        for k in range(9):
            dx, dy = f32(k % 3 - 1), f32(k // 3 - 1)
            ox, oy = f32(base[0] + f32(dx * step)), f32(base[1] + f32(dy * step))
            total = f32(0)
            for y in range(8):
                row = f32(0)
                for x in range(8):
                    a = tap(la, bx * 8 - sx - O, by * 8 - sy - O, f32(f32(x + O) - ox), f32(f32(y + O) - oy))
                    c = tap(lc, bx * 8 + sx - O, by * 8 + sy - O, f32(f32(x + O) + ox), f32(f32(y + O) + oy))
                    row = f32(row + abs(f32(a - c)))
                total = f32(total + row)
            sad = f32(total * f32(1.0 / 64.0))
            sads.append(sad)
            costs.append(f32(sad + f32(f32(f32(lam) * step) * f32(abs(dx) + abs(dy)))))
        bi = 0
        for k in range(1, 9):
            if costs[k] < costs[bi]:
                bi = k
        base = [f32(base[0] + f32(f32(bi % 3 - 1) * step)), f32(base[1] + f32(f32(bi // 3 - 1) * step))]
        best_sad = sads[bi]
    return (f32(f32(sx) + base[0]), f32(f32(sy) + base[1])), int(np.rint(f32(best_sad * f32(256))))


def sample(img, px, py):
    X, Y = int(np.rint(f32(px * f32(32)))), int(np.rint(f32(py * f32(32))))
    ix, iy = X >> 5, Y >> 5
    fx, fy = f32(f32(X & 31) * f32(1 / 32)), f32(f32(Y & 31) * f32(1 / 32))
    one = f32(1)
    return f32(f32(f32(L(img, ix, iy) * f32((one - fy) * (one - fx))) + f32(L(img, ix + 1, iy) * f32((one - fy) * fx)))
               + f32(L(img, ix, iy + 1) * f32(fy * (one - fx)))) + f32(L(img, ix + 1, iy + 1) * f32(fy * fx))


def mixs(x, y, a):
    one = f32(1)
    return f32(f32(x * f32(one - a)) + f32(y * a))


def pix_cand(k, qx, qy, bf):
    bh, bw = bf.shape[:2]

    def bv(x, y):
        return bf[clampi(y, 0, bh - 1), clampi(x, 0, bw - 1)]
    bx, by = qx // 8, qy // 8
    if k == 0:
        sx = f32(f32(f32(qx) + f32(0.5)) / f32(8) - f32(0.5))
        sy = f32(f32(f32(qy) + f32(0.5)) / f32(8) - f32(0.5))
        fx, fy = f32(sx - np.floor(sx)), f32(sy - np.floor(sy))
        ix, iy = int(np.floor(sx)), int(np.floor(sy))
        r = []
        for c in range(2):
            top = mixs(bv(ix, iy)[c], bv(ix + 1, iy)[c], fx)
            bot = mixs(bv(ix, iy + 1)[c], bv(ix + 1, iy + 1)[c], fx)
            r.append(mixs(top, bot, fy))
        return r
    off = {1: (0, 0), 2: (0, -1), 3: (0, 1), 4: (-1, 0), 5: (1, 0)}
    if k in off:
        v = bv(bx + off[k][0], by + off[k][1])
        return [v[0], v[1]]
    return [f32(0), f32(0)]


def pixel_field(la, lc, bf, smooth, zero):
    """pixfield.comp:37-63, with the window sum in float64 (horizontal rows, then the column) and float32(sum/25)."""
    h, w = la.shape
    err = np.zeros((7, h, w), f32)
    for k in range(7):
        for qy in range(h):
            for qx in range(w):
                v = pix_cand(k, qx, qy, bf)
                a = sample(la, f32(f32(qx) - v[0]), f32(f32(qy) - v[1]))
                c = sample(lc, f32(f32(qx) + v[0]), f32(f32(qy) + v[1]))
                err[k, qy, qx] = abs(f32(a - c))
    pf = np.zeros((h, w, 2), f32)
    for qy in range(h):
        for qx in range(w):
            best, bk = None, 0
            for k in range(7):
                vs = 0.0
                for y in range(-2, 3):
                    hs = 0.0
                    for x in range(-2, 3):
                        hs += float(err[k, clampi(qy + y, 0, h - 1), clampi(qx + x, 0, w - 1)])
                    vs += hs
                c = f32(vs * (1.0 / 25.0))
                if k == 0:
                    c = f32(c - f32(smooth))
                if k == 6:
                    c = f32(c - f32(zero))
                if k == 0 or c < best:
                    best, bk = c, k
            pf[qy, qx] = pix_cand(bk, qx, qy, bf)
    return pf


def estimate(a, c, s, lv):
    S = s["flowScale"]
    pa, pc = luma_levels(a, S, lv), luma_levels(c, S, lv)
    n = len(lv)
    lam = s["searchPenalty"]
    v = None
    res = {"luma_a": pa, "luma_c": pc, "fields_raw": [None] * n, "fields": [None] * n, "medians": [None] * n}
    for i in range(n - 1, -1, -1):
        w, h, bw, bh = lv[i]
        raw = np.zeros((bh, bw, 2), f32)
        sads = np.zeros((bh, bw), f32)
        if i < n - 1:
            med = median_of_parents(v, bw, bh)
            res["medians"][i] = med
        for by in range(bh):
            for bx in range(bw):
                seed = (0, 0) if i == n - 1 else seed_block(pa[i], pc[i], v, med, bx, by, lam)
                R = s["searchRadiusTop"] if i == n - 1 else s["searchRadius"]
                vec, sad = search_block(pa[i], pc[i], bx, by, seed, R, lam)
                raw[by, bx] = vec
                sads[by, bx] = sad
        res["fields_raw"][i] = raw
        v = median3(raw)
        res["fields"][i] = v
    bh, bw = lv[0][3], lv[0][2]
    bf = np.zeros((bh, bw, 2), f32)
    sadsum = 0
    for by in range(bh):
        for bx in range(bw):
            vec, s256 = subpix_block(pa[0], pc[0], bx, by, v[by, bx], lam)
            bf[by, bx] = vec
            sadsum += s256
    res["block_field"] = bf
    res["sadsum"] = sadsum
    res["pf"] = pixel_field(pa[0], pc[0], bf, s["smoothBonus"], s["hudBonus"])
    return res


def lookup_t(pf, t):
    """lookup.comp:10-19 per flow cell: s = float32(2 t - 1) folded in double from the float t (framegen.cpp:3490-3493),
    then twice g = sampleF(pf, q - s g), starting from g = pf(q). Returns the field at time t."""
    td = float(f32(t))
    s = f32(2.0 * td - 1.0)
    fh, fw = pf.shape[:2]
    out = np.zeros_like(pf)
    for qy in range(fh):
        for qx in range(fw):
            g = [pf[qy, qx, 0], pf[qy, qx, 1]]
            for _ in range(2):
                px = f32(f32(qx) - f32(s * g[0]))
                py = f32(f32(qy) - f32(s * g[1]))
                g = [sample(pf[..., 0], px, py), sample(pf[..., 1], px, py)]
            out[qy, qx] = g
    return out


def synth_t05(a, c, pf, S, s):
    """error.comp + resolve.comp at t = 0.5 (no lookup), scalar. Returns (out, untrusted count)."""
    return synth_t(a, c, pf, S, s, 0.5)


def synth_t(a, c, pf, S, s, t):
    """lookup.comp (t != 0.5) + error.comp + resolve.comp at any t in (0, 1), scalar, with the constants of
    framegen.cpp:3490-3497 (wA = float32(1 - t), wC = float32(t) from the double t; warp factors 2 wC, 2 wA) and the
    blend of resolve.comp:73-82 (0.5 (wa + wc) at t = 0.5, else wA wa + wC wc; near = prev for t <= 0.5, else curr).
    Returns (out, untrusted count)."""
    H, W = a.shape[:2]
    td = float(f32(t))
    wA, wC = f32(1.0 - td), f32(td)
    mid = wC == f32(0.5)
    if not mid:
        pf = lookup_t(pf, t)
    ka, kc = f32(f32(2) * wC), f32(f32(2) * wA)
    near = c if wC > f32(0.5) else a
    fh, fw = pf.shape[:2]

    def hv(x, y):
        x, y = min(x, S * fw - 1), min(y, S * fh - 1)
        sx = f32(f32(f32(x) + f32(0.5)) * f32(1.0 / S) - f32(0.5))
        sy = f32(f32(f32(y) + f32(0.5)) * f32(1.0 / S) - f32(0.5))
        fx, fy = f32(sx - np.floor(sx)), f32(sy - np.floor(sy))
        i0x, i1x = clampi(int(np.floor(sx)), 0, fw - 1), clampi(int(np.floor(sx)) + 1, 0, fw - 1)
        i0y, i1y = clampi(int(np.floor(sy)), 0, fh - 1), clampi(int(np.floor(sy)) + 1, 0, fh - 1)
        r = []
        for k in range(2):
            top = mixs(pf[i0y, i0x, k], pf[i0y, i1x, k], fx)
            bot = mixs(pf[i1y, i0x, k], pf[i1y, i1x, k], fx)
            r.append(f32(mixs(top, bot, fy) * f32(S)))
        return r

    def warp(img, px, py):
        X, Y = int(np.rint(f32(px * f32(32)))), int(np.rint(f32(py * f32(32))))
        ix, iy, fx, fy = X >> 5, Y >> 5, X & 31, Y & 31
        out = []
        for ch in range(3):
            v = ((32 - fx) * (32 - fy) * int(L(img, ix, iy)[ch]) + fx * (32 - fy) * int(L(img, ix + 1, iy)[ch])
                 + (32 - fx) * fy * int(L(img, ix, iy + 1)[ch]) + fx * fy * int(L(img, ix + 1, iy + 1)[ch]) + 512) >> 10
            out.append(v)
        return out

    wa = np.zeros((H, W, 3), np.int64)
    wc = np.zeros((H, W, 3), np.int64)
    e = np.zeros((H, W), np.int64)
    for y in range(H):
        for x in range(W):
            h = hv(x, y)
            wa[y, x] = warp(a, f32(f32(x) - f32(ka * h[0])), f32(f32(y) - f32(ka * h[1])))
            wc[y, x] = warp(c, f32(f32(x) + f32(kc * h[0])), f32(f32(y) + f32(kc * h[1])))
            e[y, x] = max(abs(int(wa[y, x, k]) - int(wc[y, x, k])) for k in range(3))
    out = near.copy()
    count = 0
    lo, hi = f32(s["trustLow"]), f32(s["trustHigh"])
    for y in range(H):
        for x in range(W):
            S49 = sum(int(e[clampi(y + j, 0, H - 1), clampi(x + i, 0, W - 1)]) for j in range(-3, 4) for i in range(-3, 4))
            e7 = f32(S49 * (1.0 / 49.0))
            u = f32(f32(e7 - lo) / f32(hi - lo))
            u = min(max(u, f32(0)), f32(1))
            if u > f32(0.5):
                count += 1
            for k in range(3):
                if mid:
                    m = f32(f32(0.5) * f32(wa[y, x, k] + wc[y, x, k]))
                else:
                    m = f32(f32(wA * f32(wa[y, x, k])) + f32(wC * f32(wc[y, x, k])))
                o = f32(f32(m * f32(f32(1) - u)) + f32(f32(near[y, x, k]) * u))
                out[y, x, k] = clampi(int(np.floor(f32(o + f32(0.5)))), 0, 255)
    return out, count
