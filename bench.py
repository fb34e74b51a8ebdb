#!/usr/bin/env python3
"""Frames/s and Mtri/s of the oibvh collision pipeline on one MI355X.

Default workload (the deforming pair): two UV spheres of 1000 x 501
(1,000,000 triangles each) 0.9 radii apart; every vertex moves along its rest
normal by 0.05 sin(omega t + k . p). One step is one frame: both trees refit
(rebuilt from scratch every 50 frames), then one broad + narrow phase. All
frames' vertex positions are generated on the host and uploaded to HBM before
the timed region; the trees read them in place.

Prints one JSON line. `value` is frames/s over the timed frames, measured
with HIP events on the library's stream; `roofline` is computed for the kernel
that takes most of the frame, from algorithmic bytes per launch over that
kernel's average duration (events around each launch in the timed region);
`cpu_baseline` times the oracle (the CPU restatement of the reference CPU
path, one thread) on the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "frames/s & Mtri/s (build+refit+broad+narrow), batched vs plain, 0 pair diffs"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--workload", default="deforming-uv-pair", choices=["deforming-uv-pair", "icosphere-pair"])
    p.add_argument("--entry-level", type=int, default=None,
                   help="BVTT entry level; default: picked by the autotuner during warm-up")
    p.add_argument("--expand-levels", type=int, default=None,
                   help="levels descended per BVTT step; default: picked by the autotuner")
    p.add_argument("--rebuild-every", type=int, default=50)
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-verify", action="store_true")
    return p.parse_args()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


class EventPairs:
    """Start/stop HIP events recorded on the library's stream around launches."""

    def __init__(self, H, stream, n):
        self.H, self.stream = H, C.c_void_p(stream)
        self.ev = []
        for _ in range(2 * n):
            e = C.c_void_p()
            H.hipEventCreate(C.byref(e))
            self.ev.append(e)
        self.i = 0

    def start(self):
        self.H.hipEventRecord(self.ev[2 * self.i], self.stream)

    def stop(self):
        self.H.hipEventRecord(self.ev[2 * self.i + 1], self.stream)
        self.i += 1

    def total_ms(self):
        t = 0.0
        v = C.c_float()
        for k in range(self.i):
            self.H.hipEventSynchronize(self.ev[2 * k + 1])
            self.H.hipEventElapsedTime(C.byref(v), self.ev[2 * k], self.ev[2 * k + 1])
            t += v.value
        return t

    def close(self):
        for e in self.ev:
            self.H.hipEventDestroy(e)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this pipeline runs on a single "
              "MI355X; multi-GPU sharding is described in DESIGN.md and not built)", file=sys.stderr)
        sys.exit(2)
    from oibvh_amd import _lib as L
    from oibvh_amd import api
    from oibvh_amd import meshgen as M

    K, W = args.steps, args.warmup
    if args.workload == "deforming-uv-pair":
        meshes, waves = M.deforming_uv_pair()
        n_frames = 200
        workload = ("deforming-uv-pair: two UV spheres 1000x501 (1,000,000 triangles each) 0.9 radii apart, "
                    "vertices displaced along the normal by 0.05 sin(wt + k.p); refit every frame, rebuild every "
                    f"{args.rebuild_every}, broad + narrow phase every frame")
    else:
        meshes = M.interpenetrating_icospheres(subdiv=6)
        waves = None
        n_frames = 1
        workload = "icosphere-pair: two noisy icospheres (81,920 triangles each) 1.6 radii apart, build + detect"
    V = meshes[0][0].shape[0]
    Fm = [m[1].shape[0] for m in meshes]
    tri_per_frame = sum(Fm)
    frames_used = min(n_frames, max(K, W))
    t_gen = time.time()
    if waves is not None:
        frames = [[w.frame(f) for f in range(frames_used)] for w in waves]
    else:
        frames = [[m[0]] for m in meshes]
    t_gen = time.time() - t_gen

    ctx = api.Context(0)
    H = L.hip()
    bufs = []
    for m in range(len(meshes)):
        b = api.DeviceBuffer(len(frames[m]) * V * 12)
        for f, x in enumerate(frames[m]):
            b.upload(x, f * V * 12)
        bufs.append(b)
    trees = []
    for m, (x, f) in enumerate(meshes):
        t = api.OibvhTree(ctx, x, f)
        t.set_positions_device(bufs[m].at(0))
        t.build()
        trees.append(t)
    scene = api.Scene(ctx)
    for t in trees:
        scene.addOibvhTree(t)

    rebuild = args.rebuild_every if waves is not None else 1
    ev_refit = EventPairs(H, ctx.stream, K * len(trees) + 1)
    ev_build = EventPairs(H, ctx.stream, K * len(trees) + 1)
    ev_detect = EventPairs(H, ctx.stream, K + 1)

    def step(fi, timed):
        f = fi % len(frames[0])
        for m, t in enumerate(trees):
            t.set_positions_device(bufs[m].at(f * V * 12))
            if fi % rebuild == 0:
                if timed:
                    ev_build.start()
                t.build()
                if timed:
                    ev_build.stop()
            else:
                if timed:
                    ev_refit.start()
                t.refit()
                if timed:
                    ev_refit.stop()
        if timed:
            ev_detect.start()
        scene.detectCollision(shape[0], shape[1], asynchronous=True)
        if timed:
            ev_detect.stop()

    shape = [args.entry_level if args.entry_level is not None else 4,
             args.expand_levels if args.expand_levels is not None else 1]
    # warm-up: synchronous detects size the device buffers; then the autotuner
    # picks entry level, levels per BVTT step, workgroups per launch and the
    # reduce subtree size on the live scene (none of them changes results)
    for fi in range(W):
        step(fi, False)
        scene.getIntTriPairCount()
    tuned = None
    if args.entry_level is None or args.expand_levels is None:
        from oibvh_amd import autotune

        def tune_frame(i):
            f = (W + i) % len(frames[0])
            for m, t in enumerate(trees):
                t.set_positions_device(bufs[m].at(f * V * 12))
                t.refit()

        tuned = autotune.tune(ctx, scene, tune_frame)
        if args.entry_level is None:
            args.entry_level = tuned["entry_level"]
        if args.expand_levels is None:
            args.expand_levels = tuned["expand_levels"]
        ctx.set_tuning(traverse_expand=args.expand_levels)
        shape[:] = [args.entry_level, args.expand_levels]
        scene.detectCollision(args.entry_level, args.expand_levels)
        scene.getIntTriPairCount()
    timer = api.Timer(ctx)
    ctx.synchronize()
    timer.start()
    for fi in range(K):
        step(fi, True)
    timer.stop()
    ctx.synchronize()
    total_ms = timer.ms()
    st = scene.stats()
    pairs_last = scene.getIntTriPairs()
    ms_step = total_ms / K
    fps = 1000.0 / ms_step

    refit_ms = ev_refit.total_ms() / max(1, ev_refit.i)
    build_ms = ev_build.total_ms() / max(1, ev_build.i)
    detect_ms = ev_detect.total_ms() / max(1, ev_detect.i)
    N = [t.node_count for t in trees]
    H_pairs = len(pairs_last)
    # algorithmic bytes per launch (12V + 12F + 24N per refit; 24N per mesh for
    # the broad phase, 12V + 12F per mesh for the narrow phase, 16 B per hit)
    refit_bytes = sum(12 * V + 12 * F + 24 * n for F, n in zip(Fm, N)) / len(trees)
    detect_bytes = sum(24 * n + 12 * V + 12 * F for F, n in zip(Fm, N)) + 16 * H_pairs
    share = {"refit": refit_ms * ev_refit.i, "detect": detect_ms * ev_detect.i, "build": build_ms * ev_build.i}
    kernels = {
        "refit": {"kernel": "k_reduce (+ k_reduce_apex above 2^11 leaves per workgroup)", "avg_ms": refit_ms, "launches": ev_refit.i,
                  "alg_bytes": refit_bytes, "gbps": refit_bytes / (refit_ms * 1e-3) / 1e9 if refit_ms else None},
        "detect": {"kernel": "k_seed, k_bvtt_level x levels, k_narrow", "avg_ms": detect_ms, "launches": ev_detect.i,
                   "alg_bytes": detect_bytes,
                   "gbps": detect_bytes / (detect_ms * 1e-3) / 1e9 if detect_ms else None},
        "build": {"kernels": "k_mesh_box, k_morton, 4 x k_sort_pass, k_reduce", "avg_ms": build_ms,
                  "launches": ev_build.i},
    }
    dom = "detect" if share["detect"] >= share["refit"] else "refit"
    dk = kernels[dom]
    roofline = {"bound": "hbm", "kernel": dk["kernel"], "achieved": dk["gbps"], "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": (dk["gbps"] * 1e9 / HBM_PEAK) if dk["gbps"] else None,
                "traffic": None, "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": dk["avg_ms"]}

    # ---------------------------------------------------------------- parity
    diffs = None
    if not args.no_verify:
        import oracle

        f_last = (K - 1) % len(frames[0])
        otrees = [oracle.Tree(frames[m][f_last], meshes[m][1], order=trees[m].perm()) for m in range(len(meshes))]
        want = oracle.detect(otrees)
        got = pairs_last
        if got.shape == want.shape:
            diffs = int((got != want).any(axis=1).sum())
        else:
            a = set(map(tuple, got.tolist()))
            b = set(map(tuple, want.tolist()))
            diffs = len(a ^ b)

    # ---------------------------------------------------------------- CPU
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        t0 = time.time()
        ot = [oracle.Tree(frames[m][0], meshes[m][1], order="morton") for m in range(len(meshes))]
        t_build = (time.time() - t0) / len(meshes)
        n = 0
        t_frames = 0.0
        budget = args.cpu_seconds
        while n < max(1, frames_used - 1) and (t_frames < budget or n == 0):
            f = 1 + n if frames_used > 1 else 0
            t0 = time.time()
            if waves is not None:
                for m, t in enumerate(ot):
                    t.refit(frames[m][f])
            oracle.detect(ot)
            t_frames += time.time() - t0
            n += 1
        per_frame = t_frames / n + (t_build * len(meshes)) / rebuild
        cpu = {"value": 1.0 / per_frame, "unit": "frames/s", "cores": 1, "kind": "port",
               "mtri_per_s": tri_per_frame / per_frame / 1e6,
               "sample": (f"oracle (CPU restatement of src/cpu, 1 thread of {os.cpu_count()} on '{cpu_model()}'): "
                          f"Morton build of both meshes ({t_build:.3f} s each, amortised over {rebuild} frames) + "
                          f"{n} frame(s) of refit + detect ({t_frames / n:.3f} s each)")}

    out = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "form": "plain", "triangles_per_frame": tri_per_frame,
                   "entry_level": args.entry_level, "expand_levels": args.expand_levels,
                   "tuning": ctx.get_tuning()},
        "autotune": tuned,
        "bvtt_levels_last_frame": scene.level_counts().tolist(),
        "mtri_per_s": tri_per_frame * fps / 1e6,
        "pair_diffs": diffs,
        "pairs_last_frame": H_pairs,
        "async_errors": st["async_errors"],
        "bvtt_tests_last_frame": st["bvtt_tests"],
        "candidates_last_frame": st["candidates"],
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_gen_s": t_gen,
    }
    print(json.dumps(out), flush=True)
    for e in (ev_refit, ev_build, ev_detect):
        e.close()


if __name__ == "__main__":
    main()
