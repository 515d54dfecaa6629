#!/usr/bin/env python3
"""Headline benchmark: 1080p 4x V1 interpolated pairs/s on one MI355X.

Workload: 8 seeded 1920x1080 RGBA8 pairs (value-noise background with a global quarter-pel shift, four moving
96x96 sprites, a static 300x60 HUD, the last pair a scene cut), 4x frame generation: per pair one motion estimate and
three synths at t = 1/4, 1/2, 3/4. Frames are resident in HBM before the timed region. One step = the whole batch of
8 pairs. Two paths are timed:
  * plain:   the C ABI per pair (fgr_record_estimate + 3 fgr_record_synth on one stream);
  * batched: one launch per pass for all pairs (fgr_batch_*), replayed as a hipGraph.
`value` is the batched rate when the batched path is built, else the plain rate (the line says which).

Extra fields:
  * `roofline` of the dominant kernel, the batched synth (one launch = the three synths of all 8 pairs): algorithmic
    bytes per launch (3F + P per synth, 2F for the scene-cut pair's copies) over its average launch duration, from HIP
    events around each launch on the launch stream; `traffic` = the HBM bytes per launch that rocprofv3's FETCH_SIZE
    and WRITE_SIZE counters read (from the committed summary under profiles/), or null; `valu_issue_frac` = the
    fraction of the GPU's VALU issue slots the launch's wavefronts filled (SQ_INSTS_VALU over 256 CUs x cycles, from
    the committed counter summary), the bound this exact integer arithmetic meets before HBM does;
  * `copy_probe`: a device-to-device copy of known size, the bandwidth a plain streaming kernel reaches;
  * `cpu_baseline`: the numpy oracle on 4 pairs of the workload, one core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "frame-gen-ritz_amd"))

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, published
W, H, MULT = 1920, 1080, 4


def frame_bytes(w, h, bpp=4):
    return w * h * bpp


def pf_bytes(w, h, S=2):
    return round(w / S) * round(h / S) * 8


def synth_bytes(w, h, cut=False):
    """Algorithmic bytes of one synth: read both frames, read the per-pixel field, write the output (a cut: copy)."""
    F = frame_bytes(w, h)
    return 2 * F if cut else 3 * F + pf_bytes(w, h)


def pair_bytes(w, h, n, cut=False):
    """B = 2F + P + (N - 1)(3F + P); a cut pair's synths are copies (2F)."""
    F, P = frame_bytes(w, h), pf_bytes(w, h)
    return 2 * F + P + (n - 1) * synth_bytes(w, h, cut)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter
    scaled by what it reads for a device copy of known size), or None when there is no such summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_1080p_4x_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def measured_valu_issue(kernel):
    """VALU issue fraction of `kernel`'s launch from the committed rocprofv3 SQ counter summary, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_1080p_4x_valu.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("valu_issue_frac:"):
            return float(line.split(":")[1])
    return None


def tuning_block(batch, tuning, tune_times):
    """The autotuner's choice for a batch and the time per step of every configuration it tried: synth tile rows x XCD
    chunk (every format's synth), the pixel field's form, the radius-2 search's blocks a wavefront, the lanes."""
    import fgr
    return {"synth_tile_rows": tuning[0], "xcd_chunk": tuning[1], "lanes": tuning[2],
            "pixfield": "windows" if batch.get_form() & fgr.Batch.FORM_PIXFIELD_WINDOWS else "gather",
            "search_r2": "4x1" if batch.get_form() & fgr.Batch.FORM_SEARCH_NARROW else "2x8",
            "us_per_step": {(f"{r}x{c}" if isinstance(r, int) else f"{r}_{c}"): round(t / 1e3, 1)
                            for (r, c), t in tune_times.items()}}


def timed_run(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def plain_runner(items, settings=None, net=None, yuv_matrix=2, transfer=1):
    """The per-pair C ABI path over batch items (prev, curr, n, outs, fmt): one handle per pair."""
    import fgr
    his = []
    for prev, curr, n, outs, fmt in items:
        t = prev[0] if isinstance(prev, tuple) else prev
        fi = fgr.Interpolator(int(t.shape[1]), int(t.shape[0]), fmt=fmt, settings=settings, yuv_matrix=yuv_matrix,
                              transfer=transfer)
        if net is not None:
            fi.set_seed_net(net)
        his.append(fi)

    def step():
        for fi, (prev, curr, n, outs, fmt) in zip(his, items):
            fi.begin_frame()
            assert fi.record_estimate(prev, curr)
            for k in range(1, n):
                assert fi.record_synth(prev, curr, outs[k - 1], k / n)
    return step, his


def config_mixed(args):
    """64 pairs of 1280x720 .. 3840x2160, RGBA8 / Rgb10 / RgbaF16 scRGB / P010, 2x / 3x / 4x: one batch, one launch per
    pass for all formats, replayed as a hipGraph; the per-pair path timed beside it."""
    import fgr
    from fgr import workloads
    import torch
    items, total = workloads.mixed_batch()
    torch.cuda.synchronize()
    batch = fgr.Batch(items, use_graph=True, yuv_matrix=fgr.YUV_BT2020, transfer=fgr.TRANSFER_LINEAR)
    tune_times, tuning = batch.autotune(reps=2)
    bs = timed_run(batch.run, args.steps, args.warmup)
    step, _ = plain_runner(items)
    ps = timed_run(step, max(2, args.steps // 4), 1)
    rate = len(items) * args.steps / bs
    plain = len(items) * max(2, args.steps // 4) / ps
    print(json.dumps({
        "metric": "mixed-size, mixed-format batched V1 pairs/sec on 1 GPU (64 pairs per batch)", "value": round(rate, 2),
        "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * bs / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/u10/f16", "data": "synthetic",
        "config": {"workload": "64 pairs: 16 each of 1280x720, 1920x1080, 2560x1440, 3840x2160; formats RGBA8 / Rgb10 "
                               "PQ / RgbaF16 scRGB (0-4) / P010 BT.2020 PQ; multipliers 2x/3x/4x; motion +-24 px",
                   "launches_per_step": batch.launches()},
        "tuning": tuning_block(batch, tuning, tune_times),
        "plain_pairs_per_s": round(plain, 2), "batched_over_plain": round(rate / plain, 3),
        "whole_path_hbm_frac": round(total * args.steps / bs / 1e9 / HBM_PEAK_GBS, 4),
        "algorithmic_bytes_per_step": int(total)}))
    return 0


def config_seednet(args):
    """The learned seed network: training on the GPU (20k synthetic pairs at the coarsest level of 1080p, flowScale 2,
    3 pyramid levels: 240x136, batch 32), then the gain of its proposals over the plain seeds for a radius-2 search,
    in BF16 and FP8, and the forward latency."""
    import torch
    import fgr
    from fgr import seedtrain
    w, h, batch = 240, 136, 32
    steps = 20000 // batch
    net = fgr.SeedNet(3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = seedtrain.train(net, steps, w, h, batch=batch, lr=2e-3, seed=21, shift=10.0)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    r16 = seedtrain.evaluate(net, 512, w, h, R=2, seed=22, shift=10.0)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(23)
    la, lc, _ = seedtrain.make_pairs(64, w, h, gen, 10.0)
    assert net.calibrate_fp8(la, lc) and net.set_fp8(True)
    r8 = seedtrain.evaluate(net, 512, w, h, R=2, seed=22, shift=10.0)
    out = torch.empty((64, (h + 7) // 8, (w + 7) // 8, 2), device="cuda:0")
    lat = {}
    for name, fp8 in (("bf16", False), ("fp8", True)):
        net.set_fp8(fp8)
        s = timed_run(lambda: net.forward(la, lc, out), 20, 3)
        lat[name] = round(1e6 * s / 20 / 64, 3)
    print(json.dumps({
        "metric": "learned seeds: training pairs/sec on 1 GPU (BF16 MFMA forward + backward + Adam)",
        "value": round(steps * batch / train_s, 1), "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1e3 * train_s / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulate); fp8 e4m3 inference", "data": "synthetic",
        "config": {"workload": "5-layer 3x3 CNN 2-16-32-64-64-2 on 240x136 coarse luma pairs (value noise, affine "
                               "motion: shift +-10 px, rotation +-3 deg, zoom 0.95-1.05), batch 32, 20k pairs, Adam",
                   "width": w, "height": h, "batch": batch},
        "loss_first_last": [round(float(sum(losses[:20]) / 20), 4), round(float(sum(losses[-20:]) / 20), 4)],
        "eval_radius2_bf16": {k: round(v, 4) for k, v in r16.items()},
        "eval_radius2_fp8": {k: round(v, 4) for k, v in r8.items()},
        "forward_us_per_pair": lat}))
    return 0


def _read_dev(ptr, n, dtype):
    """n elements at a device address as a new CUDA tensor (a device-to-device copy)."""
    import ctypes
    import torch
    t = torch.empty(n, dtype=dtype, device="cuda:0")
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    torch.cuda.synchronize()
    assert hip.hipMemcpy(t.data_ptr(), ptr, t.numel() * t.element_size(), 3) == 0
    return t


def config_uhd(args):
    """128 pairs at 3840x2160 (RGBA8 and P010, 4x, 5 % scene cuts, 30 % occluding sprites) in one batch: the plain V1
    path against the learned FP8 seeds (coarsest level 240x135 of a 4-level pyramid, radius-2 top search), with the
    proposals' seed-hit rate, the block SAD and the endpoint error of the final field against the plain one."""
    import torch
    import fgr
    from fgr import seedtrain, workloads
    items, total, cuts = workloads.uhd_batch()
    torch.cuda.synchronize()
    net = fgr.SeedNet(3)
    seedtrain.train(net, 20000 // 32, 240, 136, batch=32, lr=2e-3, seed=21, shift=10.0)
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(23)
    la, lc, _ = seedtrain.make_pairs(64, 240, 136, gen, 10.0)
    assert net.calibrate_fp8(la, lc) and net.set_fp8(True)
    plain_set = fgr.default_settings()
    seed_set = fgr.default_settings(seedNet=1, pyramidMinSide=72, searchRadiusTop=2)
    rates, batches = {}, {}
    for name, s, n in (("plain", plain_set, None), ("learned_fp8", seed_set, net)):
        b = fgr.Batch(items, settings=s, use_graph=True, yuv_matrix=fgr.YUV_BT2020, seed_net=n)
        t = timed_run(b.run, args.steps, args.warmup)
        rates[name] = len(items) * args.steps / t
        batches[name] = b
    # quality: per pair, the final (sub-pixel) block fields and the block SAD of both paths; the proposals' hits
    epe, sad, hits, blocks = [], {"plain": [], "learned_fp8": []}, 0, 0
    for p, cut in enumerate(cuts):
        if cut:
            continue
        d = {k: b.debug_internal(p) for k, b in batches.items()}
        f = {k: _read_dev(x.field + 8 * x.fieldSub, 2 * x.bw[0] * x.bh[0], torch.float32).reshape(-1, 2)
             for k, x in d.items()}
        epe.append(float((f["learned_fp8"] - f["plain"]).norm(dim=-1).mean()))
        for k, x in d.items():
            u = _read_dev(x.counters, 128, torch.int32)[64:128]
            sad[k].append(float(u.sum()) / 256.0 / (x.bw[0] * x.bh[0]))
        x = d["learned_fp8"]
        L = x.levels - 1
        n = x.w[L] * x.h[L]
        la_ = _read_dev(x.luma + 4 * x.lumaOffA[L], n, torch.float32).reshape(1, x.h[L], x.w[L])
        lc_ = _read_dev(x.luma + 4 * x.lumaOffC[L], n, torch.float32).reshape(1, x.h[L], x.w[L])
        raw = torch.empty((1, x.bh[L], x.bw[L], 2), device="cuda:0")
        ref = torch.empty_like(raw)
        assert fgr.search_field(la_, lc_, 6, 0.5, raw, ref)
        prop = _read_dev(x.field + 8 * x.fieldSeed, 2 * x.bw[L] * x.bh[L], torch.float32).reshape(1, x.bh[L], x.bw[L], 2)
        hits += int(((prop.round() - ref).abs().amax(-1) <= 2.5).sum())
        blocks += x.bw[L] * x.bh[L]
    print(json.dumps({
        "metric": "4K 4x batched V1 pairs/sec on 1 GPU, FP8 learned seeds vs plain seeds (128 pairs per batch)",
        "value": round(rates["learned_fp8"], 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * len(items) / rates["learned_fp8"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8/u10; fp8 e4m3 seeds",
        "data": "synthetic",
        "config": {"workload": "128 pairs 3840x2160, RGBA8 and P010 (BT.2020 PQ), 4x, 5 % scene cuts, 30 % with four "
                               "384x384 occluding sprites, motion +-24 px",
                   "learned": "seedNet=1, FP8 E4M3, pyramidMinSide 72, searchRadiusTop 2",
                   "plain": "default settings (pyramidMinSide 48, searchRadiusTop 6)"},
        "plain_pairs_per_s": round(rates["plain"], 2), "learned_over_plain": round(rates["learned_fp8"] / rates["plain"], 3),
        "seed_hit_rate": round(hits / max(blocks, 1), 4),
        "endpoint_error_vs_plain_field": round(sum(epe) / len(epe), 4),
        "mean_block_sad": {k: round(sum(v) / len(v), 4) for k, v in sad.items()},
        "whole_path_hbm_frac": round(total * rates["learned_fp8"] / len(items) / 1e9 / HBM_PEAK_GBS, 4)}))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--path", choices=["auto", "plain", "batched"], default="auto")
    ap.add_argument("--copy-mib", type=int, default=512, help="size of the device copy probe")
    ap.add_argument("--config", type=int, choices=[2, 3, 4, 5], default=2,
                    help="2: the headline 1080p 4x workload (default); 3: 64 mixed sizes / formats / multipliers; "
                         "4: learned seeds, training and gains; 5: 128 4K pairs, plain vs FP8 learned seeds")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every pair is independent, so more GPUs would shard "
                                   "the pair list with no communication; multi-GPU runs are not part of this build"}))
        return 2
    if args.config != 2:
        return {3: config_mixed, 4: config_seednet, 5: config_uhd}[args.config](args)

    import numpy as np
    import torch
    import fgr
    from fgr import scenes

    dev = torch.device("cuda:0")
    pairs, _ = scenes.pairs_1080p(2, n_pairs=args.pairs)
    cut_index = 7
    fr = [(torch.from_numpy(a).to(dev), torch.from_numpy(c).to(dev)) for a, c in pairs]
    outs = [[torch.empty_like(fr[0][0]) for _ in range(MULT - 1)] for _ in fr]
    ts = [k / MULT for k in range(1, MULT)]
    stream = torch.cuda.current_stream()

    fi = fgr.Interpolator(W, H)

    def plain_step():
        for p, (a, c) in enumerate(fr):
            fi.begin_frame()
            assert fi.record_estimate(a, c)
            for k, t in enumerate(ts):
                assert fi.record_synth(a, c, outs[p][k], t)

    batch = None
    if args.path in ("auto", "batched") and hasattr(fgr, "Batch"):
        bouts = [[torch.empty_like(fr[0][0]) for _ in range(MULT - 1)] for _ in fr]
        batch = fgr.Batch([(a, c, MULT, o) for (a, c), o in zip(fr, bouts)], use_graph=True)
        tune_times, tuning = batch.autotune(reps=3)   # tile rows x XCD chunk, timed on this batch

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    plain_s = timed(plain_step, args.steps, args.warmup)
    plain_rate = len(fr) * args.steps / plain_s
    if batch is not None:
        batch_s = timed(batch.run, args.steps, args.warmup)
        batch_rate = len(fr) * args.steps / batch_s
        value, ms = batch_rate, 1e3 * batch_s / args.steps
        path = "batched"
    else:
        batch_rate = None
        value, ms = plain_rate, 1e3 * plain_s / args.steps
        path = "plain"

    # roofline of the dominant kernel, the batched synth (one launch = every synth of every pair): the batch's
    # launches recorded eagerly between HIP events on the launch stream, `rounds` records
    rounds = max(2, args.steps // 4)
    synth_ns, synth_bytes_l = [], []
    rb = batch if batch is not None else None
    if rb is not None:
        rb.set_lanes(1)   # one launch per pass covers every pair (the bytes below); the tuned lanes are restored after
        rb.set_profiling(True)
        for _ in range(rounds):
            rb.run()
            for stage, pas, ns in rb.read_timings():
                if stage == 500:   # the batch's one synth launch: every synth of every pair
                    synth_ns.append(ns)
                    synth_bytes_l.append(sum((MULT - 1) * synth_bytes(W, H, cut=(p == cut_index))
                                             for p in range(len(fr))))
        rb.set_profiling(False)
        rb.set_lanes(tuning[2])
        kernel = {16: "synth_rgba8_s2_t16_batch_kernel", 32: "synth_rgba8_s2_batch_kernel",
                  64: "synth_rgba8_s2_t64_batch_kernel"}[tuning[0]]
    # the plain path's per-synth launch, for comparison (the library's profile: events around each synth)
    fi.set_profiling(True)
    plain_ns = []
    for p, (a, c) in enumerate(fr):
        fi.begin_frame()
        assert fi.record_estimate(a, c)
        for k, t in enumerate(ts):
            assert fi.record_synth(a, c, outs[p][k], t)
        plain_ns += [ns for name, ns in fi.read_timings() if name == "synth" and p != cut_index]
    fi.set_profiling(False)
    if rb is None:
        synth_ns = plain_ns
        synth_bytes_l = [synth_bytes(W, H)] * len(plain_ns)
        kernel = "synth_rgba8_s2_kernel"
    avg_ns = sum(synth_ns) / len(synth_ns)
    bytes_per_launch = sum(synth_bytes_l) / len(synth_bytes_l)
    achieved = bytes_per_launch / (avg_ns * 1e-9) / 1e9
    traffic = measured_traffic(kernel)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "valu_issue_frac": measured_valu_issue(kernel),
                "kernel": kernel, "launches": len(synth_ns), "avg_launch_us": round(avg_ns / 1e3, 3),
                "bytes_per_launch": int(bytes_per_launch),
                "plain_synth_avg_us": round(sum(plain_ns) / len(plain_ns) / 1e3, 3)}

    # a plain device-to-device copy of known size: the bandwidth a streaming kernel reaches on this GPU
    copy_bytes = args.copy_mib << 20
    src = torch.empty(copy_bytes, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * copy_bytes * 10 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import v1
        sample = [0, 1, 2, 3]
        t0 = time.perf_counter()
        for p in sample:
            a, c = pairs[p]
            est = v1.estimate(a, c)
            for t in ts:
                v1.synth(a, c, est, t)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(len(sample) / cpu_s, 4), "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": f"pairs {sample} of the workload (1080p RGBA8, estimate + 3 synths each) through the numpy "
                         "oracle, single-threaded numpy", "seconds": round(cpu_s, 2), "host_cpus": os.cpu_count()}

    total_bytes = sum(pair_bytes(W, H, MULT, cut=(p == cut_index)) for p in range(len(fr)))
    line = {
        "metric": "1080p 4x interp pairs/sec on 1 GPU, batched vs plain V1; % of HBM roofline",
        "value": round(value, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": "1080p RGBA8 4x V1, 8 seeded pairs (value noise + quarter-pel global shift, 4 sprites, "
                               "static HUD, 1 scene cut)", "width": W, "height": H, "multiplier": MULT,
                   "pairs_per_step": len(fr), "path": path},
        "plain_pairs_per_s": round(plain_rate, 2),
        "batched_pairs_per_s": None if batch_rate is None else round(batch_rate, 2),
        "batched_over_plain": None if batch_rate is None else round(batch_rate / plain_rate, 3),
        "whole_path_hbm_frac": round(total_bytes * args.steps / (ms * 1e-3 * args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "tuning": None if batch is None else tuning_block(batch, tuning, tune_times),
        "copy_probe": {"bytes": copy_bytes, "gbs": round(copy_gbs, 1), "frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
