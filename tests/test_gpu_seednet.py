"""The learned seed network on the GPU: its BF16-MFMA forward and gradients against the CPU oracle, training on the
GPU, what the trained network gains over the plain seeds, and the estimate with proposals bit-exact against the
oracle fed the same proposals (the proposals only seed the exact search)."""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes, seedtrain
from gpu_util import Internals, dev, read_device
from oracle import seednet as N
from oracle import v1
from test_gpu_v1 import check_estimate, check_output, oracle_settings

pytestmark = pytest.mark.gpu


def he_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(N.param_count(), np.float32)
    for (W, b), (ci, co, _, _) in zip(N.split(p), N.LAYERS):
        W[...] = rng.standard_normal(W.shape) * np.sqrt(2.0 / (9 * ci))
        b[...] = rng.standard_normal(b.shape) * 0.05
    return p


def small_batch(n, w, h, seed):
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(seed)
    return seedtrain.make_pairs(n, w, h, gen, shift=4.0)


def test_forward_matches_oracle():
    net = fgr.SeedNet(0)
    p = he_params(3)
    net.set_weights(p)
    la, lc, _ = small_batch(3, 44, 26, 5)
    out = torch.empty((3, 4, 6, 2), device="cuda:0")
    assert net.forward(la, lc, out)
    acts, _ = N.forward(la.cpu().numpy(), lc.cpu().numpy(), p)
    ref = acts[-1]
    got = out.cpu().numpy()
    assert got.shape == ref.shape
    err = np.abs(got - ref).max()
    assert err <= 1e-2 * (1.0 + np.abs(ref).max()), err
    net.close()


def _e4m3_value(b):
    """OCP FP8 E4M3 bytes -> float64 values (0x7F / 0xFF NaN)"""
    b = np.asarray(b, np.int64)
    s = np.where(b & 0x80, -1.0, 1.0)
    E, M = (b >> 3) & 15, b & 7
    v = np.where(E == 0, M * 2.0 ** -9, (1 + M / 8.0) * 2.0 ** (E - 7.0))
    return np.where((b & 0x7F) == 0x7F, np.nan, s * v)


def test_fp8_conversion_matches_oracle():
    """The FP8 forward's conversion (the conversion instruction after a clamp to +-448) against the oracle's e4m3 on
    every finite E4M3 value, the midpoints between neighbours (ties to even) and points just off them, subnormals,
    values past 448 (saturation) and random floats over the whole range."""
    codes = np.array([c for c in range(256) if (c & 0x7F) != 0x7F])
    vals = np.unique(_e4m3_value(codes))
    mids = (vals[:-1] + vals[1:]) / 2
    near = np.concatenate([np.nextafter(mids.astype(np.float32), np.float32(np.inf)),
                           np.nextafter(mids.astype(np.float32), np.float32(-np.inf))])
    rng = np.random.default_rng(9)
    rand = (rng.standard_normal(20000) * np.exp2(rng.uniform(-14, 10, 20000))).astype(np.float32)
    big = np.array([448.0, 449.0, 463.9, 464.0, 480.0, 1e4, 3e38, -448.0, -464.0, -1e6], np.float32)
    tiny = np.array([2.0 ** -10, 3 * 2.0 ** -11, 2.0 ** -11, 1e-30, 0.0], np.float32)
    x = np.concatenate([vals, mids, near, rand, big, tiny, -tiny]).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    out = torch.empty(x.size, dtype=torch.uint8, device="cuda:0")
    assert fgr.debug_fp8_e4m3(xd, out)
    got = _e4m3_value(out.cpu().numpy())
    want = N.e4m3(x).astype(np.float64)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(x[i]), float(got[i]), float(want[i])) for i in bad[:8]]


def test_fp8_forward_matches_oracle():
    """FP8 (E4M3, per-channel scales) forward on v_mfma_f32_16x16x32_fp8_fp8 against the oracle's emulation with the
    library's calibrated input scales: the weight scales are identical, the outputs equal to a tolerance."""
    net = fgr.SeedNet(0)
    p = he_params(5)
    net.set_weights(p)
    la, lc, _ = small_batch(4, 48, 32, 8)
    assert net.calibrate_fp8(la, lc)
    assert net.set_fp8(True)
    out = torch.empty((4, 4, 6, 2), device="cuda:0")
    assert net.forward(la, lc, out)
    sa, sw = net.fp8_scales()
    assert np.all(np.frexp(sa)[0] == 0.5), "the input scales are powers of two"
    ref, tabs = N.forward_fp8(la.cpu().numpy(), lc.cpu().numpy(), p, sa)
    assert np.array_equal(np.concatenate([t[1] for t in tabs]), sw)
    got = out.cpu().numpy()
    err = np.abs(got - ref).max()
    assert err <= 2e-2 * (1.0 + np.abs(ref).max()), err
    # and close to the BF16 forward of the same network
    assert net.set_fp8(False)
    out16 = torch.empty_like(out)
    assert net.forward(la, lc, out16)
    d = np.abs(out16.cpu().numpy() - got).max()
    assert d <= 0.1 * (1.0 + np.abs(got).max()), d
    net.close()


def test_gradients_match_oracle():
    net = fgr.SeedNet(0)
    p = he_params(4)
    net.set_weights(p)
    la, lc, _ = small_batch(2, 40, 24, 6)
    target = torch.randn((2, 3, 5, 2), device="cuda:0")
    grad = torch.empty(N.param_count(), device="cuda:0")
    loss = torch.empty(1, device="cuda:0")
    assert net.gradients(la, lc, target, grad, loss)
    g_ref, loss_ref = N.backward(la.cpu().numpy(), lc.cpu().numpy(), p, target.cpu().numpy().astype(np.float64))
    assert abs(float(loss.item()) - loss_ref) <= 1e-3 * max(1.0, loss_ref)
    g = grad.cpu().numpy().astype(np.float64)
    o = 0
    for ci, co, _, _ in N.LAYERS:
        n = co * 9 * ci + co
        a, b = g[o:o + n], g_ref[o:o + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel <= 2e-2, (ci, co, rel)
        o += n
    net.close()


def test_training_lowers_the_loss_and_beats_the_plain_seeds():
    """Train on the GPU against plain radius-6 fields, then seed a radius-2 search with the proposals: the endpoint
    error and the SAD against held-out pairs must drop well below the plain (zero-seed) radius-2 search's."""
    torch.manual_seed(0)
    net = fgr.SeedNet(7)
    w, h = 64, 48
    losses = seedtrain.train(net, steps=800, w=w, h=h, batch=32, lr=2e-3, seed=11, shift=10.0)
    first, last = np.mean(losses[:20]), np.mean(losses[-20:])
    r = seedtrain.evaluate(net, 256, w, h, R=2, seed=12, shift=10.0)
    print(f"seednet loss {first:.3f} -> {last:.3f}; evaluation {r}")
    assert last < 0.4 * first
    assert r["epe_learned"] < 0.6 * r["epe_plain"]
    assert r["sad_learned"] < r["sad_plain"]
    assert r["seed_hit_learned"] > r["seed_hit_plain"]
    # FP8 inference of the trained network, calibrated on training-like pairs: nearly the BF16 gain
    la, lc, _ = small_batch(64, w, h, 13)
    assert net.calibrate_fp8(la, lc) and net.set_fp8(True)
    r8 = seedtrain.evaluate(net, 256, w, h, R=2, seed=12, shift=10.0)
    print(f"seednet FP8 evaluation {r8}")
    assert r8["epe_learned"] < 0.7 * r["epe_plain"] and r8["epe_learned"] < 1.25 * r["epe_learned"]
    net.close()


def _large_motion_pair():
    return scenes.pair_translate(21, w=640, h=360, sprite=48, sprite_pos=(300, 150), shift=(22, -14))


@pytest.mark.parametrize("fp8", [False, True])
def test_estimate_with_proposals_is_bit_exact(fp8):
    """seedNet = 1 with a network attached: the proposals (read back from the library) fed to the oracle reproduce
    every intermediate of the estimate and the output bit for bit; the per-pair and the batched paths agree."""
    a, c = _large_motion_pair()
    s = fgr.default_settings(seedNet=1, pyramidMinSide=24, searchRadiusTop=2)
    net = fgr.SeedNet(1)
    net.set_weights(he_params(9))
    if fp8:
        la, lc, _ = small_batch(8, 80, 45, 14)
        assert net.calibrate_fp8(la, lc) and net.set_fp8(True)
    fi = fgr.Interpolator(640, 360, settings=s)
    assert fi.set_seed_net(net)
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    out = torch.zeros_like(ta)
    assert fi.record_synth(ta, tc, out, 0.25)
    I = Internals(fi)
    d = I.d
    L = d.levels - 1
    props = I.field[d.fieldSeed:d.fieldSeed + d.bw[L] * d.bh[L]].reshape(d.bh[L], d.bw[L], 2)
    est = v1.estimate(a, c, oracle_settings(s), proposals=props)
    check_estimate(I, est)
    check_output(out.cpu().numpy(), a, c, est, 0.25)
    # batched, with the network attached to the pair
    outs = [torch.zeros_like(ta) for _ in range(3)]
    b = fgr.Batch([(ta, tc, 4, outs)], settings=s, seed_net=net)
    b.run()
    torch.cuda.synchronize()
    check_estimate(Internals(b.debug_internal(0)), est)
    assert np.array_equal(outs[0].cpu().numpy(), out.cpu().numpy())
    b.close()
    fi.close()
    net.close()


def test_seed_setting_without_network_is_the_plain_path():
    a, c = _large_motion_pair()
    ta, tc = dev(a), dev(c)
    res = []
    for seed_net in (0, 1):
        fi = fgr.Interpolator(640, 360, settings=fgr.default_settings(seedNet=seed_net, pyramidMinSide=24))
        assert fi.record_estimate(ta, tc)
        o = torch.zeros_like(ta)
        assert fi.record_synth(ta, tc, o, 0.5)
        res.append(o.cpu().numpy())
        fi.close()
    assert np.array_equal(res[0], res[1])
