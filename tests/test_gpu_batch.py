"""Batched V1 (fgr_batch_*): one launch per pass over many pairs of mixed sizes and multipliers, replayed as a
hipGraph. Bar: every output bit-identical to the per-pair path (itself bit-exact against the oracle), the estimate's
intermediates bit-exact against the oracle for pairs of every size, and each graph replay identical to the eager run.
"""
import numpy as np
import pytest
import torch

import fgr
from fgr import scenes
from gpu_util import Internals, dev
from oracle import v1
from test_gpu_v1 import check_estimate, check_output

pytestmark = pytest.mark.gpu


def per_pair(a, c, n):
    H, W = a.shape[:2]
    fi = fgr.Interpolator(W, H)
    ta, tc = dev(a), dev(c)
    assert fi.record_estimate(ta, tc)
    outs = []
    for k in range(1, n):
        o = torch.zeros_like(ta)
        assert fi.record_synth(ta, tc, o, k / n)
        outs.append(o.cpu().numpy())
    fi.close()
    return outs


def make_batch(pairs, mults, use_graph=True):
    items, frames = [], []
    for (a, c), n in zip(pairs, mults):
        ta, tc = dev(a), dev(c)
        outs = [torch.zeros_like(ta) for _ in range(n - 1)]
        items.append((ta, tc, n, outs))
    return fgr.Batch(items, use_graph=use_graph), items


def mixed_pairs():
    sizes = [(320, 180), (203, 117), (640, 360), (256, 144), (402, 226)]
    pairs = []
    for i, (w, h) in enumerate(sizes):
        pairs.append(scenes.pair_translate(30 + i, w=w, h=h, sprite=min(48, h // 3), sprite_pos=(w // 3, h // 3),
                                           shift=(3 + i, -2 + i % 3)))
    pairs.append(scenes.noise_pair(40, 288, 160))         # a scene cut inside the batch
    return pairs, [2, 3, 4, 4, 3, 4]


@pytest.mark.parametrize("use_graph", [False, True])
def test_mixed_batch_matches_per_pair_and_oracle(use_graph):
    pairs, mults = mixed_pairs()
    b, items = make_batch(pairs, mults, use_graph)
    b.run()
    torch.cuda.synchronize()
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda:0")
    for p, ((a, c), n) in enumerate(zip(pairs, mults)):
        got = [o.cpu().numpy() for o in items[p][3]]
        want = per_pair(a, c, n)
        for k in range(n - 1):
            assert np.array_equal(got[k], want[k]), f"pair {p} output {k}"
        est = v1.estimate(a, c)
        check_estimate(Internals(b.debug_internal(p)), est)
        for k in range(n - 1):
            why, info = check_output(got[k], a, c, est, (k + 1) / n)
            assert b.debug_counters(p, k, cnt)
            cv = cnt.cpu().numpy().astype(np.uint32)
            assert int(cv[1]) == est["sadsum"]
            if why != "cut":
                assert int(cv[0]) == info["count"], (p, k)
    b.close()


def test_uniform_pairs_batch_matches_the_oracle():
    # the batched pixel field on uniform block fields (a static pair, an integer pan): own-block reuse over whole tiles
    pairs = [scenes.pair_shift(11, 320, 192, 0, 0), scenes.pair_shift(12, 320, 192, 4, -2)]
    b, items = make_batch(pairs, [2, 2])
    b.run()
    torch.cuda.synchronize()
    for p, (a, c) in enumerate(pairs):
        est = v1.estimate(a, c)
        check_estimate(Internals(b.debug_internal(p)), est)
        check_output(items[p][3][0].cpu().numpy(), a, c, est, 0.5)
    b.close()


def test_graph_replays_match_eager():
    """Replay the captured graph several times with eager work and synchronisations in between; every replay must
    reproduce the eager outputs exactly (the graph holds no memset nodes: the counters are cleared by a kernel)."""
    pairs, mults = mixed_pairs()
    be, items_e = make_batch(pairs, mults, use_graph=False)
    be.run()
    torch.cuda.synchronize()
    ref = [[o.cpu().numpy() for o in it[3]] for it in items_e]
    bg, items_g = make_batch(pairs, mults, use_graph=True)
    for rep in range(3):
        for it in items_g:
            for o in it[3]:
                o.zero_()
        torch.cuda.synchronize()
        bg.run()
        x = torch.randn(256, 256, device="cuda:0")   # unrelated eager work between replays
        (x @ x).sum().item()
        torch.cuda.synchronize()
        for p, it in enumerate(items_g):
            for k, o in enumerate(it[3]):
                assert np.array_equal(o.cpu().numpy(), ref[p][k]), f"replay {rep} pair {p} output {k}"
    be.close()
    bg.close()


def test_set_frames_rebinds_pointers():
    pairs, mults = mixed_pairs()
    b, items = make_batch(pairs, mults, use_graph=True)
    b.run()
    torch.cuda.synchronize()
    # swap prev / curr of every pair (same sizes): outputs must equal the per-pair path of the swapped pairs
    swapped = [(it[1], it[0], it[2], it[3]) for it in items]
    assert b.set_frames(swapped)
    b.run()
    torch.cuda.synchronize()
    for p, ((a, c), n) in enumerate(zip(pairs, mults)):
        want = per_pair(c, a, n)
        for k in range(n - 1):
            assert np.array_equal(swapped[p][3][k].cpu().numpy(), want[k]), f"pair {p} output {k}"
    b.close()


def test_1080p_workload_batch_matches_per_pair():
    pairs, _ = scenes.pairs_1080p(2, n_pairs=8)
    b, items = make_batch(pairs, [4] * 8, use_graph=True)
    b.run()
    b.run()
    torch.cuda.synchronize()
    for p, (a, c) in enumerate(pairs):
        want = per_pair(a, c, 4)
        for k in range(3):
            assert np.array_equal(items[p][3][k].cpu().numpy(), want[k]), f"pair {p} output {k}"
    b.close()


@pytest.mark.parametrize("rows,chunk,lanes,form", [(16, 16, 1, 0), (32, 256, 2, 1), (64, 64, 3, 2), (32, 64, 4, 3),
                                                    (32, 64, 1, 3)])
def test_tile_configurations_are_bit_identical(rows, chunk, lanes, form):
    """Every configuration of the tuner (synth tile rows x XCD chunk x lanes x estimate forms: the pixel field from
    LDS windows, the radius-2 search one block a wavefront) gives the per-pair path's bytes."""
    pairs, mults = mixed_pairs()
    b, items = make_batch(pairs, mults, use_graph=True)
    assert b.set_tuning(rows, chunk, lanes) and b.get_tuning() == (rows, chunk, lanes)
    assert b.set_form(form) and b.get_form() == form
    b.run()
    torch.cuda.synchronize()
    for p, ((a, c), n) in enumerate(zip(pairs, mults)):
        want = per_pair(a, c, n)
        for k in range(n - 1):
            assert np.array_equal(items[p][3][k].cpu().numpy(), want[k]), f"pair {p} output {k}"
    b.close()


def test_autotune_picks_the_fastest_and_caches_it(tmp_path):
    pairs, _ = scenes.pairs_1080p(2, n_pairs=4)
    b, items = make_batch(pairs, [4] * 4, use_graph=True)
    times, chosen = b.autotune(reps=3)
    assert len(times) == 17 and all(t > 0 for t in times.values())
    tiles = {k: v for k, v in times.items() if isinstance(k[0], int)}
    assert tiles[chosen[:2]] == min(tiles.values())
    lanes = {k[1]: v for k, v in times.items() if k[0] == "lanes"}
    assert lanes[chosen[2]] == min(lanes.values())
    form = b.get_form()
    for knob, key, on in ((fgr.Batch.FORM_PIXFIELD_WINDOWS, "pixfield", "windows"),
                          (fgr.Batch.FORM_SEARCH_NARROW, "search", "4x1")):
        t_on, t_off = times[(key, on)], min(v for (k, x), v in times.items() if k == key and x != on)
        if t_on != t_off:
            assert bool(form & knob) == (t_on < t_off), (key, t_on, t_off)
    b.run()
    torch.cuda.synchronize()
    want = per_pair(*pairs[1], 4)
    for k in range(3):
        assert np.array_equal(items[1][3][k].cpu().numpy(), want[k])
    # a second batch of the same pairs starts with the cached choice; the cache survives a save / load
    b2, _ = make_batch(pairs, [4] * 4, use_graph=True)
    assert b2.get_tuning() == chosen and b2.get_form() == form
    path = tmp_path / "tuning.txt"
    assert fgr.tuning_save(path) and fgr.tuning_load(path) >= 1
    b.close()
    b2.close()


def test_set_frames_refuses_frames_it_cannot_address():
    """fgr_batch_set_frames checks every new frame as fgr_batch_create does (null planes, pitches), before any
    upload: the batch keeps its old frames and still matches the per-pair path."""
    a, c = scenes.pair_translate(41, w=192, h=128)
    b, items = make_batch([(a, c)], [2])
    ta, tc, n, outs = items[0]
    bad = fgr.frame_of(tc)
    bad.pitch = 192 * 4 - 4
    assert not b.set_frames([(ta, bad, n, outs)])
    null = fgr.frame_of(outs[0])
    null.ptr = None
    assert not b.set_frames([(ta, tc, n, [null])])
    b.run()
    torch.cuda.synchronize()
    assert np.array_equal(outs[0].cpu().numpy(), per_pair(a, c, 2)[0])
    b.close()
