"""GPU parity at the sizes and mixes of the benchmark workloads, through the C ABI.

  * 2560x1440 and 3840x2160 pairs in RGBA8, Rgb10 (HDR10 PQ), RgbaF16 (scRGB linear light) and P010 (BT.2020 PQ):
    the 5-level pyramid of 4K and its 240x135 level-0 block grid, every estimate intermediate bit-exact against the
    oracle, one synth at t != 1/2 bit-exact with its untrusted count;
  * the 64-pair batch of 720p-4K frames in four formats and three multipliers, every output equal to the per-pair path;
  * the seed network at its training shape (the 240x136 coarsest level of 1080p, batch 32): forward and gradients
    against the CPU oracle;
  * 4K pairs with FP8 learned seeds inside a batch: the proposals read back and fed to the oracle reproduce every
    intermediate and the outputs bit for bit.
Content comes from the benchmark's own generators (fgr/workloads.py, seeded), the oracle runs on the host.
"""
import numpy as np
import pytest
import torch

import fgr
from fgr import seedtrain
from fgr import workloads as WL
from gpu_util import Internals
from oracle import formats as F
from oracle import seednet as N
from oracle import v1
from test_gpu_formats import first_diff, lut_of, same, to_host
from test_gpu_seednet import he_params
from test_gpu_v1 import _mismatch, check_estimate, oracle_settings

pytestmark = pytest.mark.gpu

FMTS = [("rgba8", WL.FMT_RGBA8), ("rgb10", WL.FMT_RGB10), ("f16lin", WL.FMT_F16), ("p010", WL.FMT_P010)]


def _host(x, fmt):
    return x.cpu().numpy() if fmt == WL.FMT_RGBA8 else to_host(x, fmt)


def _lut(fmt):
    return lut_of(F.F16, True, 80.0) if fmt == WL.FMT_F16 else None


def _oracle_estimate(A, C, fmt, s=None, proposals=None):
    if fmt == WL.FMT_RGBA8:
        return v1.estimate(A, C, s, proposals=proposals)
    return F.estimate(A, C, fmt, s=s, lut=_lut(fmt), proposals=proposals)


def _check_synth(out, A, C, est, t, fmt, count=None):
    """the output bit-identical to the float32 oracle (and the untrusted count exact when given)"""
    if fmt == WL.FMT_RGBA8:
        ref, why, info = v1.synth(A, C, est, t)
        e = _mismatch(f"output t={t} ({why})", out, ref)
        assert e is None, e
    else:
        ref, why, info = F.synth(A, C, est, t, fmt, lut=_lut(fmt), yuv=F.yuv_constants(fgr.YUV_BT2020))
        assert same(out, ref, fmt), f"t={t} ({why}): {first_diff(out, ref, fmt)}"
    if count is not None and why != "cut":
        assert count == info["count"], "untrusted count"
    return why


def _interpolator(w, h, fmt, settings=None):
    return fgr.Interpolator(w, h, fmt=fmt, settings=settings, yuv_matrix=fgr.YUV_BT2020,
                            transfer=fgr.TRANSFER_LINEAR if fmt == WL.FMT_F16 else fgr.TRANSFER_ENCODED)


@pytest.mark.parametrize("size,t", [((2560, 1440), 0.75), ((3840, 2160), 0.25)], ids=["1440p", "4k"])
@pytest.mark.parametrize("name,fmt", FMTS)
def test_large_pair_per_format(size, t, name, fmt):
    w, h = size
    print(f"case {name} {w}x{h} t={t}", flush=True)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(100 + 10 * fmt + (w > 3000))
    A, C = (WL.encode(x, fmt, WL._luts("cuda:0")) for x in WL.pair(gen, w, h, (17, -11)))
    fi = _interpolator(w, h, fmt)
    assert fi.record_estimate(A, C)
    I = Internals(fi)
    out = WL.empty_like_frame(A)
    assert fi.record_synth(A, C, out, t)
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    assert fi.record_debug_counters(cnt)
    got, count = _host(out, fmt), int(cnt[0].item())
    fi.close()
    Ah, Ch = _host(A, fmt), _host(C, fmt)
    est = _oracle_estimate(Ah, Ch, fmt)
    assert len(est["levels"]) == (5 if w > 3000 else 4)
    check_estimate(I, est)
    assert not est["cut"]
    assert _check_synth(got, Ah, Ch, est, t, fmt, count) is None


def _same_dev(x, y):
    if isinstance(x, tuple):
        return all(torch.equal(a, b) for a, b in zip(x, y))
    return torch.equal(x, y)


def test_mixed_64_pair_batch_matches_per_pair():
    """The 64-pair batch (16 each of 720p, 1080p, 1440p, 4K; RGBA8 / Rgb10 / RgbaF16 linear / P010; 2x / 3x / 4x),
    replayed twice from its graph: every output equal to the per-pair C ABI path's."""
    items, _ = WL.mixed_batch()
    b = fgr.Batch(items, use_graph=True, yuv_matrix=fgr.YUV_BT2020, transfer=fgr.TRANSFER_LINEAR)
    b.run()
    b.run()
    torch.cuda.synchronize()
    for p, (A, C, n, outs, fmt) in enumerate(items):
        t0 = A[0] if isinstance(A, tuple) else A
        h, w = t0.shape[:2]
        fi = _interpolator(w, h, fmt)
        assert fi.record_estimate(A, C)
        for k in range(1, n):
            o = WL.empty_like_frame(A)
            assert fi.record_synth(A, C, o, k / n)
            torch.cuda.synchronize()
            assert _same_dev(o, outs[k - 1]), f"pair {p} ({w}x{h} format {fmt}) output {k}"
        fi.close()
    b.close()


def test_seednet_at_the_training_shape():
    """Forward and gradients of the seed network on 32 pairs of 240x136 coarse luma (the coarsest level of 1080p at
    flowScale 2 with 3 levels) against the CPU oracle's BF16 emulation."""
    net = fgr.SeedNet(0)
    p = he_params(21)
    net.set_weights(p)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(22)
    la, lc, _ = seedtrain.make_pairs(32, 240, 136, gen, shift=10.0)
    out = torch.empty((32, 17, 30, 2), device="cuda:0")
    assert net.forward(la, lc, out)
    lah, lch = la.cpu().numpy(), lc.cpu().numpy()
    acts, _ = N.forward(lah, lch, p)
    ref = acts[-1]
    got = out.cpu().numpy()
    assert got.shape == ref.shape
    err = np.abs(got - ref).max()
    assert err <= 1e-2 * (1.0 + np.abs(ref).max()), err
    target = torch.randn((32, 17, 30, 2), device="cuda:0")
    grad = torch.empty(N.param_count(), device="cuda:0")
    loss = torch.empty(1, device="cuda:0")
    assert net.gradients(la, lc, target, grad, loss)
    g_ref, loss_ref = N.backward(lah, lch, p, target.cpu().numpy().astype(np.float64))
    assert abs(float(loss.item()) - loss_ref) <= 1e-3 * max(1.0, loss_ref)
    g = grad.cpu().numpy().astype(np.float64)
    o = 0
    for ci, co, _, _ in N.LAYERS:
        n = co * 9 * ci + co
        rel = np.linalg.norm(g[o:o + n] - g_ref[o:o + n]) / max(np.linalg.norm(g_ref[o:o + n]), 1e-12)
        assert rel <= 2e-2, (ci, co, rel)
        o += n
    net.close()


def test_uhd_batch_with_fp8_seeds_is_bit_exact():
    """Two 3840x2160 4x pairs (RGBA8 with four 384x384 occluders, P010) in one batch with FP8 learned seeds on the
    240x135 coarsest level (pyramidMinSide 72, top radius 2): each pair's proposals, fed to the oracle, reproduce
    every estimate intermediate and its outputs at t = 1/4 and 3/4 bit for bit."""
    w, h = 3840, 2160
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(55)
    luts = WL._luts("cuda:0")
    specs = [(WL.FMT_RGBA8, 4, 384, (21, -13)), (WL.FMT_P010, 1, 96, (-19, 9))]
    items = []
    for fmt, ns, sp, shift in specs:
        A, C = (WL.encode(x, fmt, luts) for x in WL.pair(gen, w, h, shift, n_sprites=ns, sprite=sp))
        items.append((A, C, 4, [WL.empty_like_frame(A) for _ in range(3)], fmt))
    net = fgr.SeedNet(1)
    net.set_weights(he_params(23))
    la, lc, _ = seedtrain.make_pairs(16, 240, 136, gen, 10.0)
    assert net.calibrate_fp8(la, lc) and net.set_fp8(True)
    s = fgr.default_settings(seedNet=1, pyramidMinSide=72, searchRadiusTop=2)
    b = fgr.Batch(items, settings=s, use_graph=True, yuv_matrix=fgr.YUV_BT2020, seed_net=net)
    b.run()
    torch.cuda.synchronize()
    for p, (A, C, n, outs, fmt) in enumerate(items):
        print(f"case pair {p} format {fmt}", flush=True)
        I = Internals(b.debug_internal(p))
        d = I.d
        L = d.levels - 1
        assert (d.w[L], d.h[L]) == (240, 135)
        props = I.field[d.fieldSeed:d.fieldSeed + d.bw[L] * d.bh[L]].reshape(d.bh[L], d.bw[L], 2)
        assert np.abs(props).max() > 0
        Ah, Ch = _host(A, fmt), _host(C, fmt)
        est = _oracle_estimate(Ah, Ch, fmt, s=oracle_settings(s), proposals=props)
        check_estimate(I, est)
        k = 0 if p == 0 else 2
        assert _check_synth(_host(outs[k], fmt), Ah, Ch, est, (k + 1) / 4, fmt) is None
    b.close()
    net.close()
