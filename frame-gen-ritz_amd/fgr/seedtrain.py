"""Training data and training loop of the learned seed network (fgr.SeedNet), on the GPU.

Synthetic pairs at the coarsest pyramid level's resolution: `a` is 3-octave value noise (lattice values, smoothstep
interpolation; wavelengths 12, 5 and 2 level pixels, amplitudes 1, 0.5, 0.25, luma 16..239), `c` the same noise seen
through a random affine motion about the frame centre (shift up to +-shift level pixels, rotation up to +-3 degrees,
zoom 0.95..1.05), both sampled exactly from the continuous noise (no resampling). The training target is the plain
search's field of the pair (fgr.search_field with seed zero, radius 6: what the estimate's coarsest step computes);
the true bilateral half-vector of every block centre is returned too, to measure endpoint errors.

Torch is used only as device memory and for generating the synthetic frames; every network operation is the library's.
"""
import math

import torch

import fgr

OCTAVES = ((12.0, 1.0), (5.0, 0.5), (2.0, 0.25))


def _noise(lattices, x, y):
    """Value noise at float coordinates x, y (n, h, w): sum of octaves, smoothstep bilinear over lattice values."""
    out = torch.zeros_like(x)
    for (lam, amp), L in zip(OCTAVES, lattices):
        n, lh, lw = L.shape
        u, v = x / lam + 8.0, y / lam + 8.0
        i0, j0 = torch.floor(u), torch.floor(v)
        fu, fv = u - i0, v - j0
        fu, fv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
        i0 = i0.long().clamp(0, lw - 2)
        j0 = j0.long().clamp(0, lh - 2)
        Lf = L.reshape(n, -1)
        idx = lambda jj, ii: torch.gather(Lf, 1, (jj * lw + ii).reshape(n, -1)).reshape(x.shape)
        top = idx(j0, i0) * (1 - fu) + idx(j0, i0 + 1) * fu
        bot = idx(j0 + 1, i0) * (1 - fu) + idx(j0 + 1, i0 + 1) * fu
        out = out + amp * (top * (1 - fv) + bot * fv)
    return out


def make_pairs(n, w, h, gen, shift=4.0, device="cuda:0"):
    """n pairs (la, lc) float32 (n, h, w) in luma levels, and the true half-vectors of the 8x8 block centres
    (n, ceil(h/8), ceil(w/8), 2): the d with M(q - d) = q + d for the affine motion M of a onto c."""
    lattices = []
    for lam, _ in OCTAVES:
        lh, lw = int((h + 64) / lam) + 20, int((w + 64) / lam) + 20
        lattices.append(torch.rand((n, lh, lw), generator=gen, device=device))
    ys, xs = torch.meshgrid(torch.arange(h, device=device, dtype=torch.float32),
                            torch.arange(w, device=device, dtype=torch.float32), indexing="ij")
    xs, ys = xs.expand(n, h, w), ys.expand(n, h, w)
    r = lambda lo, hi: lo + (hi - lo) * torch.rand((n, 1, 1), generator=gen, device=device)
    ang = r(-3.0, 3.0) * (math.pi / 180.0)
    zoom = r(0.95, 1.05)
    tx, ty = r(-shift, shift), r(-shift, shift)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    # M(p) = c0 + Z R (p - c0) + t; c(p) = a(M^-1(p))
    ca, sa = torch.cos(ang) * zoom, torch.sin(ang) * zoom
    det = ca * ca + sa * sa
    px, py = xs - cx - tx, ys - cy - ty
    ux, uy = (ca * px + sa * py) / det + cx, (-sa * px + ca * py) / det + cy
    scale = 223.0 / 1.75
    la = 16.0 + scale * _noise(lattices, xs, ys)
    lc = 16.0 + scale * _noise(lattices, ux, uy)
    # half-vectors at block centres q: (A + I) d = (A - I)(q - c0) + t
    bh, bw = (h + 7) // 8, (w + 7) // 8
    qy, qx = torch.meshgrid(torch.arange(bh, device=device, dtype=torch.float32) * 8 + 3.5,
                            torch.arange(bw, device=device, dtype=torch.float32) * 8 + 3.5, indexing="ij")
    qx, qy = qx - cx, qy - cy
    rx = (ca - 1) * qx - sa * qy + tx
    ry = sa * qx + (ca - 1) * qy + ty
    m00, m01, m10, m11 = ca + 1, -sa, sa, ca + 1
    dd = m00 * m11 - m01 * m10
    dx = (m11 * rx - m01 * ry) / dd
    dy = (-m10 * rx + m00 * ry) / dd
    return la.contiguous(), lc.contiguous(), torch.stack([dx, dy], -1).contiguous()


def plain_targets(la, lc, R=6, lam=0.5):
    n, h, w = la.shape
    bh, bw = (h + 7) // 8, (w + 7) // 8
    raw = torch.empty((n, bh, bw, 2), device=la.device)
    out = torch.empty_like(raw)
    assert fgr.search_field(la, lc, R, lam, raw, out)
    return out


def train(net, steps, w, h, batch=32, lr=2e-3, seed=0, shift=4.0, log_every=0, log=print):
    """Adam steps on fresh synthetic batches against the plain field; returns the list of losses (one per step)."""
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    loss = torch.zeros(1, device="cuda:0")
    losses = []
    for i in range(steps):
        la, lc, _ = make_pairs(batch, w, h, gen, shift)
        target = plain_targets(la, lc)
        step_lr = lr * (0.5 * (1 + math.cos(math.pi * i / steps)))   # cosine decay
        assert net.train_step(la, lc, target, max(step_lr, 1e-6), loss)
        losses.append(float(loss.item()))
        if log_every and (i % log_every == 0 or i == steps - 1):
            log(f"seednet step {i} loss {losses[-1]:.4f}")
    return losses


def block_sad(la, lc, v):
    """Mean |a(q - v) - c(q + v)| over every 8x8 block (integer vectors, clamped taps), (n, bh, bw)."""
    n, h, w = la.shape
    bh, bw = v.shape[1:3]
    ys, xs = torch.meshgrid(torch.arange(bh * 8, device=la.device), torch.arange(bw * 8, device=la.device),
                            indexing="ij")
    vv = v.round().long().repeat_interleave(8, 1).repeat_interleave(8, 2)
    ax = (xs - vv[..., 0]).clamp(0, w - 1)
    ay = (ys - vv[..., 1]).clamp(0, h - 1)
    cxx = (xs + vv[..., 0]).clamp(0, w - 1)
    cyy = (ys + vv[..., 1]).clamp(0, h - 1)
    bi = torch.arange(n, device=la.device)[:, None, None]
    d = (la[bi, ay, ax] - lc[bi, cyy, cxx]).abs()
    return d.reshape(n, bh, 8, bw, 8).mean((2, 4))


def evaluate(net, n, w, h, R=2, seed=1, shift=4.0, lam=0.5):
    """Plain seeds (zero) vs learned proposals, both followed by the radius-R search and median3, on n held-out pairs:
    endpoint error against the true half-vectors, block SAD, and the seed-hit rates (blocks whose true vector lies
    within the radius-R window around the seed: zero for the plain seeds, the rounded proposal for the learned)."""
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    la, lc, true = make_pairs(n, w, h, gen, shift)
    bh, bw = true.shape[1:3]
    raw = torch.empty((n, bh, bw, 2), device="cuda:0")
    plain = torch.empty_like(raw)
    learned = torch.empty_like(raw)
    props = torch.empty_like(raw)
    assert fgr.search_field(la, lc, R, lam, raw, plain)
    assert net.forward(la, lc, props)
    assert fgr.search_field(la, lc, R, lam, raw, learned, seeds=props)
    epe = lambda f: float((f - true).norm(dim=-1).mean())
    hit = lambda seed: float(((true - seed).abs().amax(-1) <= R + 0.5).float().mean())
    return dict(epe_plain=epe(plain), epe_learned=epe(learned), epe_proposal=epe(props),
                sad_plain=float(block_sad(la, lc, plain).mean()), sad_learned=float(block_sad(la, lc, learned).mean()),
                seed_hit_plain=hit(torch.zeros_like(true)), seed_hit_learned=hit(props.round()))
