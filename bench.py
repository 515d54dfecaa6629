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
  * `roofline` of the dominant kernel, the batched synth (one launch = the k-th synth of all 8 pairs): algorithmic
    bytes per launch (3F + P per pair, 2F for the scene-cut pair's copy) over its average launch duration, from HIP
    events around each launch on the launch stream; `traffic` = the HBM bytes per launch that rocprofv3's FETCH_SIZE
    and WRITE_SIZE counters read (from the committed summary under profiles/), or null;
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--path", choices=["auto", "plain", "batched"], default="auto")
    ap.add_argument("--copy-mib", type=int, default=512, help="size of the device copy probe")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every pair is independent, so more GPUs would shard "
                                   "the pair list with no communication; multi-GPU runs are not part of this build"}))
        return 2

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

    # roofline of the dominant kernel, the batched synth (one launch = the k-th synth of every pair): the batch's
    # launches recorded eagerly between HIP events on the launch stream, `rounds` records
    rounds = max(2, args.steps // 4)
    synth_ns, synth_bytes_l = [], []
    rb = batch if batch is not None else None
    if rb is not None:
        rb.set_profiling(True)
        for _ in range(rounds):
            rb.run()
            for stage, pas, ns in rb.read_timings():
                if 500 <= stage < 700 and stage % 10 == 0:
                    synth_ns.append(ns)
                    synth_bytes_l.append(sum(synth_bytes(W, H, cut=(p == cut_index)) for p in range(len(fr))))
        rb.set_profiling(False)
        kernel = "synth_rgba8_s2_batch_kernel"
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
        "copy_probe": {"bytes": copy_bytes, "gbs": round(copy_gbs, 1), "frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
