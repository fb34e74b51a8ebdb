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

The same line carries the added forms, each on its own workload:
  batched          512 scenes of 2-16 rigid meshes (noisy icospheres and
                   boxes, 1k-200k triangles each, log-uniform): the batched
                   FP32 + FP64-fallback frame (one hipGraph per frame) against
                   the plain per-tree / per-scene API on the same frames
  precision        64 scenes of two ~250k-triangle planes or boxes placed
                   coplanar or edge-touching: FP64 re-check rate and pair
                   differences against the FP64 oracle
  self_collision   8 folded cloth sheets of 2048 x 2048 vertices: refit +
                   self-BVTT query per frame
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
# HBM bytes of one scene refit of the two 1M-triangle trees (k_reduce), as
# measured by two rocprofv3 counter runs of this bench (FETCH_SIZE, then
# WRITE_SIZE, per dispatch), each scaled by the counter's reading for a 256 MiB
# device-to-device copy (--copy-probe-mb 256): 37.0 MB fetched + 96.0 MB
# written. Re-measure after changing the refit kernel.
TRAFFIC_REFIT = 133.0e6
METRIC = "frames/s & Mtri/s (build+refit+broad+narrow), batched vs plain, 0 pair diffs"


def apex_nodes(F: int, kernel: str) -> int:
    """Real nodes above the stage-1 subtree roots of a tree of F faces (the
    ones k_reduce_apex writes): levels 0 .. L - h - 1 of the implicit layout,
    h = 10 for k_reduce_w."""
    L = max(0, (F - 1).bit_length())
    vl = (1 << L) - F
    M = 4 if kernel == "k_reduce_w" else int(kernel[kernel.index("<") + 1:-1]) if kernel[-2:-1].isdigit() else 4
    h = min(L, 8 + {1: 0, 2: 1, 4: 2, 8: 3}.get(M, 2))
    return sum((1 << l) - (vl >> (L - l)) for l in range(0, L - h))


def reduce_kernel(lpt: int, faces: int, blocks: int = 0, items: int = 0) -> str:
    """Name of the stage-1 reduce kernel the library runs for a launch over
    `faces` faces with tuning reduce_leaves_per_thread = lpt (0 = by size)
    and refit_blocks = blocks, the launch holding `items` subtrees of 1024
    leaves."""
    if lpt == 16 or (lpt == 0 and faces >= (1 << 18)):
        return "k_reduce_wl" if blocks and ((blocks + 7) & ~7) < items else "k_reduce_w"
    return f"k_reduce<{lpt}>" if lpt else "k_reduce<M>"


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
    p.add_argument("--traverse-local", type=int, default=None,
                   help="with --entry-level/--expand-levels: initial pairs per workgroup of the local walk (0: steps)")
    p.add_argument("--traverse-local-levels", type=int, default=0,
                   help="steps the local walk replaces (0: all)")
    p.add_argument("--traverse-tail", type=int, default=None,
                   help="with --entry-level/--expand-levels: front pairs per workgroup pass of the tail walk (0: none)")
    p.add_argument("--traverse-tail-steps", type=int, default=0,
                   help="steps the tail walk replaces, counted from the last (0: all after the head walk)")
    p.add_argument("--narrow-separate", type=int, default=None,
                   help="with --entry-level/--expand-levels: 1 tests leaf pairs in one k_narrow launch")
    p.add_argument("--reduce-leaves", type=int, default=None,
                   help="with --entry-level/--expand-levels: leaves per thread of the build/refit reduce "
                        "(16: k_reduce_w, 1/2/4/8: k_reduce<M>, 0: by size)")
    p.add_argument("--refit-blocks", type=int, default=None,
                   help="at most this many workgroups in the scene refit's k_reduce_w (0: one per 1024 leaves)")
    p.add_argument("--rebuild-every", type=int, default=50)
    p.add_argument("--no-overlap", action="store_true",
                   help="plain form: frames one after another (default: a frame's refit overlaps the previous "
                        "frame's detect, OIBVH_FRAME_OVERLAP)")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-verify", action="store_true")
    p.add_argument("--sections", default="plain,batched,precision,self",
                   help="comma list of: plain (the headline), batched, precision, self")
    p.add_argument("--batched-frames", type=int, default=100)
    p.add_argument("--plain-frames", type=int, default=5, help="frames of the plain form on the batched workload")
    p.add_argument("--batched-scenes", type=int, default=512)
    p.add_argument("--precision-frames", type=int, default=20)
    p.add_argument("--self-frames", type=int, default=100)
    p.add_argument("--cloth-nv", type=int, default=2048)
    p.add_argument("--copy-probe-mb", type=int, default=0,
                   help="first copy this many MiB device to device (a kernel of known traffic, for counter scaling)")
    return p.parse_args()


def progress(msg: str) -> None:
    """Progress to stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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


def run_plain(args, ctx):
    from oibvh_amd import _lib as L
    from oibvh_amd import api
    from oibvh_amd import meshgen as M

    K, W = args.steps, args.warmup
    if args.traverse_local is not None:
        ctx.set_tuning(traverse_local=args.traverse_local, traverse_local_levels=args.traverse_local_levels)
    if args.traverse_tail is not None:
        ctx.set_tuning(traverse_tail=args.traverse_tail, traverse_tail_steps=args.traverse_tail_steps)
    if args.narrow_separate is not None:
        ctx.set_tuning(narrow_separate=args.narrow_separate)
    if args.reduce_leaves is not None:
        ctx.set_tuning(reduce_leaves_per_thread=args.reduce_leaves)
    if args.refit_blocks is not None:
        ctx.set_tuning(refit_blocks=args.refit_blocks)
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
    ev_build = EventPairs(H, ctx.stream, K * len(trees) + 1)
    pos = [[bufs[m].at(f * V * 12) for m in range(len(meshes))] for f in range(len(frames[0]))]
    pos_arr = [api.position_array(p) for p in pos]  # built once, outside the timed region

    def step(fi, timed):
        f = fi % len(frames[0])
        if fi % rebuild == 0:
            for m, t in enumerate(trees):
                t.set_positions_device(pos[f][m])
            for t in trees:
                if timed:
                    ev_build.start()
                t.build()
                if timed:
                    ev_build.stop()
            scene.detectCollision(shape[0], shape[1], asynchronous=True)
        else:
            # the frame in one call: borrow frame f's vertices, refit both
            # trees (one launch pair), detect (oibvh_scene_frame); the
            # positions are known ahead, so the refit may overlap the
            # previous frame's detect
            scene.frame(pos_arr[f], shape[0], shape[1], asynchronous=True, overlap=not args.no_overlap)

    shape = [args.entry_level if args.entry_level is not None else 4,
             args.expand_levels if args.expand_levels is not None else 1]
    progress("plain pair: warm-up")
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
            scene.refit()

        tuned = autotune.tune(ctx, scene, tune_frame)
        if args.entry_level is None:
            args.entry_level = tuned["entry_level"]
        if args.expand_levels is None:
            args.expand_levels = tuned["expand_levels"]
        ctx.set_tuning(traverse_expand=args.expand_levels)
        shape[:] = [args.entry_level, args.expand_levels]
        scene.detectCollision(args.entry_level, args.expand_levels)
        scene.getIntTriPairCount()
    tuned_rb = None
    if not args.no_overlap and args.refit_blocks is None:
        from oibvh_amd import autotune

        def overlapped(i0, n):
            for i in range(n):
                f = (W + i0 + i) % len(frames[0])
                scene.frame(pos_arr[f], shape[0], shape[1], asynchronous=True, overlap=True)

        tuned_rb = autotune.tune_refit_blocks(ctx, scene, overlapped)
    timer = api.Timer(ctx)
    ctx.synchronize()
    # the library's own HIP events time the stage-1 launch of every 8th
    # refit (recorded by the launch itself as it starts and completes) and
    # bracket every 8th detect inside the timed region (an event pair around
    # a detect drains the GPU pipeline for a few microseconds: a sample keeps
    # the timed frames close to untimed ones)
    scene.timing()
    scene.set_timing(8)
    timer.start()
    for fi in range(K):
        step(fi, True)
    # the last detects may still run on the library's side stream: wait for
    # them before the stop event (the wait's host round trip, a few
    # microseconds, is counted)
    ctx.synchronize()
    timer.stop()
    total_ms = timer.ms()
    st = scene.stats()
    pairs_last = scene.getIntTriPairs()
    ms_step = total_ms / K
    fps = 1000.0 / ms_step
    tm = scene.timing()
    scene.set_timing(0)

    refit_ms = tm["refit_ms"] / max(1, tm["refits"])
    build_ms = ev_build.total_ms() / max(1, ev_build.i)
    detect_ms = tm["detect_ms"] / max(1, tm["detects"])
    N = [t.node_count for t in trees]
    H_pairs = len(pairs_last)
    tn = ctx.get_tuning()
    rk = reduce_kernel(tn["reduce_leaves_per_thread"], sum(Fm), tn["refit_blocks"], sum(-(-F // 1024) for F in Fm))
    # algorithmic bytes per launch: the refit's stage 1 reads the positions
    # and sorted faces once and writes every node below the apex (12 V + 12 F
    # + 24 N' per tree, N' = N minus the nodes k_reduce_apex writes), both
    # trees in one launch; a detect reads 24 N per mesh (broad phase) and
    # 12 V + 12 F per mesh (narrow phase) and writes 16 B per hit
    refit_bytes = sum(12 * V + 12 * F + 24 * (n - apex_nodes(F, rk)) for F, n in zip(Fm, N))
    detect_bytes = sum(24 * n + 12 * V + 12 * F for F, n in zip(Fm, N)) + 16 * H_pairs
    share = {"refit": refit_ms * tm["refits"], "detect": detect_ms * tm["detects"], "build": build_ms * ev_build.i}
    kernels = {
        "refit": {"kernel": f"{rk}: stage 1 of oibvh_scene_refit over both trees (k_reduce_apex follows, outside "
                            f"the events)", "avg_ms": refit_ms, "launches": tm["refits"],
                  "alg_bytes": refit_bytes, "gbps": refit_bytes / (refit_ms * 1e-3) / 1e9 if refit_ms else None},
        "detect": {"kernel": "k_seed, k_bvtt_local (first steps, when tuned), k_bvtt_step x steps (leaf pairs "
                             "tested inside the steps)",
                   "avg_ms": detect_ms, "launches": tm["detects"],
                   "alg_bytes": detect_bytes,
                   "gbps": detect_bytes / (detect_ms * 1e-3) / 1e9 if detect_ms else None},
        "build": {"kernels": "k_mesh_box, k_morton, 4 x k_sort_pass, k_reduce", "avg_ms": build_ms,
                  "launches": ev_build.i},
    }
    # the roofline kernel is the refit's stage-1 reduce, the single kernel
    # with the largest share of the frame; the library's events bracket that
    # launch alone
    dk = kernels["refit"]
    roofline = {"bound": "hbm", "kernel": f"{rk} (stage 1 of the scene refit, both trees in one launch)",
                "achieved": dk["gbps"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": (dk["gbps"] * 1e9 / HBM_PEAK) if dk["gbps"] else None, "traffic": TRAFFIC_REFIT,
                "alg_bytes_per_launch": dk["alg_bytes"], "avg_launch_ms": dk["avg_ms"],
                "alg_bytes_unit": "12 V + 12 F + 24 N' per tree (positions, sorted faces, nodes below the apex), "
                                  "both trees"}
    share_ms = {k: v for k, v in share.items()}

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
                   "frame_overlap": not args.no_overlap, "tuning": ctx.get_tuning()},
        "autotune": tuned,
        "autotune_refit_blocks": tuned_rb,
        "bvtt_levels_last_frame": scene.level_counts().tolist(),
        "mtri_per_s": tri_per_frame * fps / 1e6,
        "pair_diffs": diffs,
        "pairs_last_frame": H_pairs,
        "async_errors": st["async_errors"], "retries_last_frame": st["retries"],
        "bvtt_tests_last_frame": st["bvtt_tests"],
        "candidates_last_frame": st["candidates"],
        "kernels": kernels,
        "time_share_ms": share_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_gen_s": t_gen,
    }
    ev_build.close()
    for t in trees:
        t.close()
    scene.close()
    for b in bufs:
        b.free()
    return out

def _colmajor(mats: np.ndarray) -> np.ndarray:
    """(..., 4, 4) matrices -> column-major float32 (glm::mat4 layout)."""
    return np.ascontiguousarray(np.swapaxes(np.asarray(mats, np.float32), -1, -2))


def _pair_diff(a: np.ndarray, b: np.ndarray) -> int:
    sa = set(map(tuple, np.asarray(a).tolist()))
    sb = set(map(tuple, np.asarray(b).tolist()))
    return len(sa ^ sb)


def run_batched(args, ctx):
    """Batched FP32 + FP64-fallback frames against the plain API on the same
    512-scene workload."""
    import ctypes as C

    from oibvh_amd import _lib as L
    from oibvh_amd import api, autotune
    from oibvh_amd import meshgen as M

    progress("batched: generating scenes")
    t0 = time.time()
    rs = M.RigidScenes(n_scenes=args.batched_scenes)
    n = len(rs.instances)
    nf = max(2, args.batched_frames)
    mats = np.stack([_colmajor(rs.transforms(f)) for f in range(nf)])  # (nf, n, 4, 4) column-major
    t_gen = time.time() - t0
    xbuf = api.DeviceBuffer(mats.nbytes)
    xbuf.upload(mats)
    stride = n * 64

    # ---------------- batched
    b = api.Batch(ctx)
    first = {}
    for scene, lib in rs.instances:
        if lib in first:
            b.add_instance(scene, first[lib])
        else:
            xyz, f = rs.library[lib]
            first[lib] = b.add_mesh(scene, xyz, f)
    b.set_transforms_device(xbuf.at(0))
    t1 = time.time()
    b.build()
    ctx.synchronize()
    t_build = time.time() - t1

    def stage(i):
        b.set_transforms_device(xbuf.at((i % nf) * stride))

    progress(f"batched: built {n} meshes in {t_build:.2f} s; autotuning")
    tuned = autotune.tune_batch(ctx, b, stage, flags={"fp64_fallback": True})
    progress(f"batched: tuned {tuned}")
    entry, expand = tuned["entry_level"], tuned["expand_levels"]
    # warm-up, then the timed frames: frame f = transforms of frame f, refit, detect
    for i in range(3):
        stage(i)
        b.frame(entry, expand, fp64_fallback=True)
    timer = api.Timer(ctx)
    ctx.synchronize()
    timer.start()
    for f in range(nf):
        stage(f)
        b.frame(entry, expand, fp64_fallback=True, asynchronous=True)
    timer.stop()
    ms_b = timer.ms() / nf
    st_b = b.stats()
    pairs_b64 = b.pairs()
    lc = b.level_counts().tolist()
    # the last frame again in plain FP32 for the bit-exact comparison with the plain API
    stage(nf - 1)
    b.refit()
    b.detect(entry, expand)
    pairs_b32 = b.pairs()

    progress(f"batched: {1000.0 / ms_b:.1f} frames/s; setting up the plain form")
    # ---------------- plain: one tree per mesh, one scene per scene
    lib = L.load()
    t1 = time.time()
    trees = []
    for i, (scene, li) in enumerate(rs.instances):
        xyz, f = rs.library[li]
        t = api.OibvhTree(ctx, xyz, f)
        trees.append(t)
    scenes = [api.Scene(ctx) for _ in range(rs.n_scenes)]
    fptr = C.POINTER(C.c_float)
    for i, t in enumerate(trees):
        lib.oibvh_tree_set_transform(t.h, mats[0, i].ctypes.data_as(fptr))
        t.build()
    for i, (scene, _) in enumerate(rs.instances):
        scenes[scene].addOibvhTree(trees[i])
    ctx.synchronize()
    t_plain_setup = time.time() - t1
    hs = [t.h for t in trees]
    shs = [s.h for s in scenes if len(s.trees) > 1]
    lone = [s.trees[0].h for s in scenes if len(s.trees) == 1]
    async_flags = L.DETECT_ASYNC

    def plain_frame(f):
        m = mats[f % nf]
        for i, h in enumerate(hs):
            lib.oibvh_tree_set_transform(h, m[i].ctypes.data_as(fptr))
        for sh in shs:
            lib.oibvh_scene_refit(sh)
            lib.oibvh_scene_detect(sh, entry, expand, async_flags)
        for h in lone:
            lib.oibvh_tree_refit(h, None, 0)

    plain_frame(0)
    for s_ in scenes:
        s_.getIntTriPairCount()
    kp = max(1, args.plain_frames)
    ctx.synchronize()
    tw = time.time()
    timer.start()
    for f in range(nf - kp, nf):
        plain_frame(f)
    timer.stop()
    ms_p_gpu = timer.ms() / kp
    ms_p = (time.time() - tw) * 1000.0 / kp
    rows = []
    for si, s_ in enumerate(scenes):
        if len(s_.trees) < 2:
            continue
        p = s_.getIntTriPairs()
        if len(p):
            ids = np.asarray([i for i, (sc, _) in enumerate(rs.instances) if sc == si], np.uint32)
            p = p.copy()
            p[:, 0] = ids[p[:, 0]]
            p[:, 1] = ids[p[:, 1]]
            rows.append(p)
    pairs_plain = np.concatenate(rows) if rows else np.zeros((0, 4), np.uint32)
    diff_plain = _pair_diff(pairs_b32, pairs_plain)

    progress(f"plain: {1000.0 / ms_p:.2f} frames/s; oracle check")
    # ---------------- FP64 oracle on a sample of scenes (last frame)
    diffs64 = None
    sample = []
    if not args.no_verify:
        import oracle

        last = mats[nf - 1]
        sample = list(range(0, rs.n_scenes, max(1, rs.n_scenes // 8)))[:8]
        diffs64 = 0
        for si in sample:
            ids = [i for i, (sc, _) in enumerate(rs.instances) if sc == si]
            ot = []
            for i in ids:
                xyz, f = rs.library[rs.instances[i][1]]
                ot.append(oracle.Tree(oracle.xform(last[i].T, xyz), f, order=b.perm(i)))
            want = oracle.detect(ot, fp64=True)
            if len(want):
                want = want.copy()
                want[:, 0] = np.asarray(ids, np.uint32)[want[:, 0]]
                want[:, 1] = np.asarray(ids, np.uint32)[want[:, 1]]
            got = pairs_b64[np.isin(pairs_b64[:, 0], ids)] if len(pairs_b64) else pairs_b64
            diffs64 += _pair_diff(got, want)

    # overlap statistic: mesh pairs of a scene with at least one hit
    hit_pairs = len(set(map(tuple, pairs_b64[:, :2].tolist()))) if len(pairs_b64) else 0
    mesh_pairs = 0
    per_scene = np.bincount([s for s, _ in rs.instances], minlength=rs.n_scenes)
    mesh_pairs = int(sum(k * (k - 1) // 2 for k in per_scene))
    tri = rs.triangles
    res = {
        "workload": (f"{rs.n_scenes} scenes, {n} rigid meshes (2-16 per scene; noisy icospheres and boxes, "
                     f"log-uniform 1k-200k triangles each; {len(rs.library)} base shapes instanced), "
                     f"{tri} triangles in all, radius 0.25, translations in the unit cube, spin + drift per "
                     f"frame; refit + broad + narrow every frame, build at frame 0"),
        "triangles": tri,
        "meshes": n,
        "frames": nf,
        "batched_fps": 1000.0 / ms_b,
        "batched_ms_per_frame": ms_b,
        "batched_mtri_per_s": tri / (ms_b * 1e-3) / 1e6,
        "batched_form": "FP32 + FP64 fallback, one hipGraph per frame (refit + detect), model matrices staged "
                        "device-to-device from HBM before each frame",
        "plain_fps": 1000.0 / ms_p,
        "plain_ms_per_frame": ms_p,
        "plain_gpu_ms_per_frame": ms_p_gpu,
        "plain_frames": kp,
        "plain_form": "one oibvh_tree per mesh (set_transform) and one oibvh_scene per scene (scene refit + "
                      "asynchronous detect), FP32, driven through the C ABI from Python",
        "speedup_batched_vs_plain": ms_p / ms_b,
        "pairs_last_frame": int(len(pairs_b64)),
        "pair_diffs_batched_fp32_vs_plain": diff_plain,
        "pair_diffs_vs_fp64_oracle_sampled": diffs64,
        "oracle_sample_scenes": sample,
        "fp64_rechecks_last_frame": st_b["fp64_rechecks"],
        # an overflow in a timed frame (which runs without a host check)
        # would have left that frame incomplete: both must be 0
        "async_errors": st_b["async_errors"], "retries_last_frame": st_b["retries"],
        "candidates_last_frame": st_b["candidates"],
        "bvtt_tests_last_frame": st_b["bvtt_tests"],
        "bvtt_levels_last_frame": lc,
        "mesh_pairs_overlapping": hit_pairs,
        "mesh_pairs_total": mesh_pairs,
        "autotune": tuned,
        "build_s": t_build,
        "plain_setup_s": t_plain_setup,
        "host_gen_s": t_gen,
    }
    for t in trees:
        t.close()
    for s_ in scenes:
        s_.close()
    b.close()
    xbuf.free()
    return res


def run_precision(args, ctx):
    """Coplanar / edge-touching scenes: FP32 filter + FP64 re-check."""
    from oibvh_amd import api
    from oibvh_amd import meshgen as M

    progress("precision: generating scenes")
    t0 = time.time()
    scenes = M.precision_scenes()
    t_gen = time.time() - t0
    b = api.Batch(ctx)
    for si, ms in enumerate(scenes):
        for x, f in ms:
            b.add_mesh(si, x, f)
    b.build()
    tri = int(sum(f.shape[0] for ms in scenes for _, f in ms))
    entry, expand = 4, 2
    b.frame(entry, expand, fp64_fallback=True)
    timer = api.Timer(ctx)
    k = max(1, args.precision_frames)
    ctx.synchronize()
    timer.start()
    for _ in range(k):
        b.frame(entry, expand, fp64_fallback=True, asynchronous=True)
    timer.stop()
    ms = timer.ms() / k
    st = b.stats()
    p64 = b.pairs()
    b.detect(entry, expand)
    p32 = b.pairs()
    res = {"workload": (f"{len(scenes)} scenes of two meshes (planes 2*354^2 = 250,632 or boxes 12*144^2 = "
                        f"248,832 triangles) coplanar or edge-touching, perturbations 10^U[-7,-5]; {tri} triangles"),
           "triangles": tri, "frames": k, "fps": 1000.0 / ms, "ms_per_frame": ms,
           "pairs_fp64_fallback": int(len(p64)), "pairs_fp32": int(len(p32)),
           "candidates": st["candidates"], "fp64_axis_rechecks": st["fp64_rechecks"],
           "fp64_recheck_rate_per_axis_test": st["fp64_rechecks"] / max(1, 17 * st["candidates"]),
           "async_errors": st["async_errors"], "retries_last_frame": st["retries"]}
    progress(f"precision: {1000.0 / ms:.1f} frames/s; oracle check")
    if not args.no_verify:
        import oracle

        sample = [0, 1, 2, 3, 4, 5, 6, 7]
        d64 = miss = spur = 0
        for si in sample:
            ot = [oracle.Tree(x, f, order=b.perm(2 * si + j)) for j, (x, f) in enumerate(scenes[si])]
            want = oracle.detect(ot, fp64=True)
            if len(want):
                want = want.copy()
                want[:, :2] += 2 * si
            sel64 = p64[(p64[:, 0] // 2) == si] if len(p64) else p64
            sel32 = p32[(p32[:, 0] // 2) == si] if len(p32) else p32
            d64 += _pair_diff(sel64, want)
            a32 = set(map(tuple, sel32.tolist()))
            a64 = set(map(tuple, want.tolist()))
            miss += len(a64 - a32)
            spur += len(a32 - a64)
        res.update({"oracle_sample_scenes": sample, "pair_diffs_vs_fp64_oracle": d64,
                    "fp32_path_missed_vs_fp64": miss, "fp32_path_spurious_vs_fp64": spur})
    res["host_gen_s"] = t_gen
    b.close()
    return res


def run_self(args, ctx):
    """Folded cloth sheets: refit + self query per frame, 8 instances per launch."""
    from oibvh_amd import api
    from oibvh_amd import meshgen as M

    nv = args.cloth_nv
    progress("self: generating cloth frames")
    t0 = time.time()
    n_inst, n_distinct, dphase = 8, 8, 0.02
    base = M.cloth_instances(n_inst, nv)
    t_gen = time.time() - t0
    V = nv * nv
    # deformation frames: the height field's phase advances dphase per frame
    bufs = []
    for d in range(n_distinct):
        buf = api.DeviceBuffer(n_inst * V * 12)
        for i in range(n_inst):
            buf.upload(M.cloth_positions(nv, 0.2, 0.05, M.CFG5_SEED + i, phase=dphase * d), i * V * 12)
        bufs.append(buf)
    t_gen = time.time() - t0
    b = api.Batch(ctx)
    for i, (x, f) in enumerate(base):
        b.add_mesh(i, x, f)
    b.build()

    def stage(k):
        buf = bufs[k % n_distinct]
        for i in range(n_inst):
            b.set_positions_device(i, buf.at(i * V * 12))

    from oibvh_amd import autotune

    progress("self: built; autotuning")
    ctx.set_tuning(front_capacity=1 << 29)
    # steps only: a sheet's fronts outgrow any workgroup's LDS, and a local
    # walk that overflows on some frames costs them a re-run
    tuned = autotune.tune_batch(ctx, b, stage, entries=(4, 6), expands=(2, 3), blocks=(0,), leaves=(0,),
                                local=(), tail=(), flags={"self_collide": True}, reps=4)
    progress(f"self: tuned {tuned}")
    entry, expand = tuned["entry_level"], tuned["expand_levels"]
    stage(0)
    b.frame(entry, expand, self_collide=True)
    timer = api.Timer(ctx)
    k = max(1, args.self_frames)
    ctx.synchronize()
    timer.start()
    for j in range(k):
        stage(j)
        b.frame(entry, expand, self_collide=True, asynchronous=True)
    timer.stop()
    ms = timer.ms() / k
    st = b.stats()
    pairs = b.pairs()
    tri = int(sum(f.shape[0] for _, f in base))
    # full-size cross-check of the last frame (the oracle is too slow at this
    # size): every sheet as its own plain tree and scene, self mode, eagerly,
    # against the batched frame's pairs (same positions: frame k - 1)
    plain_pairs = None
    if not args.no_verify:
        last = bufs[(k - 1) % n_distinct]
        rows = []
        for i, (x, f) in enumerate(base):
            t = api.OibvhTree(ctx, x, f)
            t.set_positions_device(last.at(i * V * 12))
            t.build()
            sc = api.Scene(ctx)
            sc.addOibvhTree(t)
            sc.detectCollision(entry, expand, self_collide=True)
            p = sc.getIntTriPairs().copy()
            p[:, 0] = i
            p[:, 1] = i
            rows.append(p)
            sc.close()
            t.close()
        plain_pairs = np.concatenate(rows)
        plain_pairs = plain_pairs[np.lexsort((plain_pairs[:, 3], plain_pairs[:, 2], plain_pairs[:, 1],
                                              plain_pairs[:, 0]))]
    res = {"workload": (f"{n_inst} folded cloth sheets of {nv} x {nv} vertices ({base[0][1].shape[0]} triangles "
                        f"each, {tri} in all), amplitude 0.2, wavelength 0.05, fold 1.5; the height field's phase "
                        f"advances {dphase} rad per frame ({n_distinct} distinct frames cycled, resident in HBM); "
                        f"refit + self-BVTT query (adjacent pairs excluded) per frame"),
           "triangles": tri, "frames": k, "fps": 1000.0 / ms, "ms_per_frame": ms,
           "mtri_per_s": tri / (ms * 1e-3) / 1e6, "pairs_last_frame": int(len(pairs)),
           "adjacent_skipped": st["adjacent"], "candidates": st["candidates"], "bvtt_tests": st["bvtt_tests"],
           "async_errors": st["async_errors"], "retries_last_frame": st["retries"],
           "bvtt_levels_last_frame": b.level_counts().tolist(), "autotune": tuned, "host_gen_s": t_gen}
    if plain_pairs is not None:
        res["pair_diffs_batched_vs_plain_full_size"] = _pair_diff(pairs, plain_pairs)
    if not args.no_verify:
        import oracle

        # parity spot check at reduced size (the full sheet is covered by the
        # GPU tests' smaller sheets and the batch's per-mesh equivalence)
        x, f = M.cloth(128, 0.2, 0.05, M.CFG5_SEED)
        c = api.Batch(ctx)
        c.add_mesh(0, x, f)
        c.build()
        c.detect(2, 2, self_collide=True)
        want = oracle.detect([oracle.Tree(x, f, order="morton")], self_collide=True)
        res["spot_check"] = {"cloth_nv": 128, "pairs": int(len(want)), "pair_diffs": _pair_diff(c.pairs(), want)}
        c.close()
    b.close()
    for buf in bufs:
        buf.free()
    return res



def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the library runs each scene "
              "or batch on one MI355X; sharding scenes or trees over several GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    from oibvh_amd import api

    sections = set(args.sections.split(","))
    ctx = api.Context(0)
    if args.copy_probe_mb:
        from oibvh_amd import _lib as L

        nb = args.copy_probe_mb << 20
        a, b = api.DeviceBuffer(nb), api.DeviceBuffer(nb)
        H = L.hip()
        for _ in range(3):
            L.hip_check(H.hipMemcpy(b.at(0), a.at(0), nb, 3), "hipMemcpy D2D")
        a.free()
        b.free()
    out = run_plain(args, ctx) if "plain" in sections else {"metric": METRIC, "value": None}
    extra = [("batched", run_batched), ("precision", run_precision), ("self", run_self)]
    for name, fn in extra:
        if name in sections:
            base = ctx.get_tuning()
            t0 = time.time()
            res = fn(args, ctx)
            res["section_wall_s"] = time.time() - t0
            out["self_collision" if name == "self" else name] = res
            ctx.set_tuning(**base)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
