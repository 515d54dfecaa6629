#!/usr/bin/env python3
"""Headline benchmark: 1080p 4x V1 interpolated pairs/s on one MI355X.

Workload: 8 seeded 1920x1080 RGBA8 pairs (value-noise background with a global quarter-pel shift, four moving
96x96 sprites, a static 300x60 HUD, the last pair a scene cut), 4x frame generation: per pair one motion estimate and
three synths at t = 1/4, 1/2, 3/4. Frames are resident in HBM before the timed region. One step = the whole batch of
8 pairs. Two paths are timed:
  * plain:   the C ABI per pair (fgr_record_estimate + 3 fgr_record_synth on one stream);
  * batched: one launch per pass for all pairs (fgr_batch_*), replayed as a hipGraph.
`value` is the batched rate when the batched path is built, else the plain rate (the line says which).

Extra fields: `roofline` of the dominant kernel (the synth kernel: algorithmic bytes per launch over its average
launch duration, from HIP events on the launch stream), `cpu_baseline` (the numpy oracle on one pair, rank 0 only).
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--path", choices=["auto", "plain", "batched"], default="auto")
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

    # roofline of the synth kernel: HIP events around each synth launch on its stream (the library's profile)
    fi.set_profiling(True)
    durs, nbytes = [], []
    for _ in range(max(2, args.steps // 4)):
        for p, (a, c) in enumerate(fr):
            fi.begin_frame()
            assert fi.record_estimate(a, c)
            for k, t in enumerate(ts):
                assert fi.record_synth(a, c, outs[p][k], t)
            tim = fi.read_timings()
            for name, ns in tim:
                if name == "synth":
                    durs.append(ns)
            nbytes += [synth_bytes(W, H, cut=(p == cut_index))] * (MULT - 1)
    fi.set_profiling(False)
    achieved = sum(nbytes) / (sum(durs) * 1e-9) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "synth_rgba8_kernel", "launches": len(durs),
                "avg_launch_us": round(sum(durs) / len(durs) / 1e3, 3),
                "bytes_per_launch": int(sum(nbytes) / len(nbytes))}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, ROOT)
        from oracle import v1
        a, c = pairs[0]
        t0 = time.perf_counter()
        est = v1.estimate(a, c)
        for t in ts:
            v1.synth(a, c, est, t)
        cpu_s = time.perf_counter() - t0
        cpu = {"value": round(1.0 / cpu_s, 4), "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": "pair 0 of the workload (1080p RGBA8, estimate + 3 synths) through the numpy oracle, "
                         "single-threaded numpy", "seconds": round(cpu_s, 2), "host_cpus": os.cpu_count()}

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
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
