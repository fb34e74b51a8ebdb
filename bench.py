#!/usr/bin/env python3
"""Frames/s and Mtri/s of the oibvh collision pipeline on one MI355X.

Headline workload (batched scenes): 512 independent scenes of 2-16 rigid
meshes each (noisy icospheres and boxes, log-uniform 1k-200k triangles per
mesh), every mesh with geometry of its own; seeded rigid motion (random
rotation, translation in the unit cube, spin and drift). One step is one
frame of every scene: new model matrices, refit of every tree, broad and
narrow phase. The batched form runs a frame as two captured hipGraphs (refit
on the library's stream; BVTT walk + FP32 SAT with FP64 re-check on its side
stream), frame f's refit beside frame f - 1's detect. All frames' matrices are
resident in HBM before the timed region.

Prints one JSON line. `value` is the batched form's frames/s over the timed
frames (HIP events on the library's stream; the last detect is waited for
inside the region). Beside it the same line carries:

  plain            the plain per-scene form on the same meshes and frames: one
                   oibvh_scene per scene, each frame per scene its matrices
                   (read in place) and one oibvh_scene_frame, driven from C++
                   (libbench_plain.so); the ratio batched / plain
  roofline         the refit's stage-1 kernel (the largest single launch of a
                   frame): algorithmic bytes per launch over its event-timed
                   duration in the timed region
  frame_roofline   refit and detect algorithmic bytes over the frame time
  kernels          the refit's and the detect's event-timed durations, the
                   detect's (the walk's) algorithmic bytes and roofline
                   fraction
  cpu_baseline     the oracle (CPU restatement of the reference CPU path, one
                   thread) on a stated sample of scenes, extrapolated
  instanced        the same scenes as rigid instances of 250 base shapes
  deforming_pair   two 1M-triangle UV spheres, 200 frames of per-vertex
                   deformation, refit every frame, rebuild every 50
  precision        64 coplanar / edge-touching scenes: FP64 re-check rate and
                   pair differences against the FP64 oracle on every scene
  self_collision   8 folded cloth sheets of 2048 x 2048 vertices
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
# HBM bytes per launch of the deforming pair's refit stage 1 (two 1M-triangle
# trees), from two rocprofv3 counter runs (FETCH_SIZE, then WRITE_SIZE, per
# dispatch), each scaled by the counter's reading for a 256 MiB
# device-to-device copy: 37.0 MB fetched + 96.0 MB written.
TRAFFIC_REFIT_PAIR = 133.0e6
# the same for the batched workload's refit stage 1 (k_reduce_w over all
# 4633 meshes, each with its own geometry; median dispatch): 3.097 GB fetched
# + 8.158 GB written
TRAFFIC_REFIT_BATCHED = 11.255e9
METRIC = "frames/s & Mtri/s (build+refit+broad+narrow), batched vs plain, 0 pair diffs"


def apex_nodes(F: int, kernel: str) -> int:
    """Real nodes above the stage-1 subtree roots of a tree of F faces (the
    ones k_reduce_apex writes): levels 0 .. L - h - 1 of the implicit layout,
    h = 10 for the wave-subtree kernels."""
    L = max(0, (F - 1).bit_length())
    vl = (1 << L) - F
    M = {"1": 1, "2": 2, "4": 4, "8": 8}.get(kernel[kernel.find("<") + 1:-1], 4) if "<" in kernel else 4
    h = min(L, 8 + {1: 0, 2: 1, 4: 2, 8: 3}.get(M, 2))
    return sum((1 << l) - (vl >> (L - l)) for l in range(0, L - h))


def node_count(F: int) -> int:
    return 2 * F + bin((1 << max(0, (F - 1).bit_length())) - F).count("1") - 1


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
    p.add_argument("--steps", type=int, default=20, help="timed batched frames")
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--sections", default="batched,instanced,deforming,precision,self",
                   help="comma list of: batched (the headline, with the plain form and the CPU baseline), "
                        "instanced, deforming, precision, self")
    p.add_argument("--batched-scenes", type=int, default=512)
    p.add_argument("--frame-cycle", type=int, default=100, help="distinct frames of matrices resident in HBM")
    p.add_argument("--plain-frames", type=int, default=5, help="timed frames of the plain form")
    p.add_argument("--no-overlap", action="store_true", help="batched frames one after another")
    p.add_argument("--oracle-scenes", type=int, default=64,
                   help="batched scenes checked against the FP64 oracle (every scene whose FP32 and FP64 pair "
                        "sets differ, plus an even sample up to this count)")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-verify", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--tune-reps", type=int, default=3)
    p.add_argument("--rebuild-every", type=int, default=50)
    p.add_argument("--rebuild-reps", type=int, default=21, help="timed rebuilds after the sequence (median)")
    p.add_argument("--precision-frames", type=int, default=20)
    p.add_argument("--self-frames", type=int, default=100)
    p.add_argument("--cloth-nv", type=int, default=2048)
    p.add_argument("--deforming-tuning", default="",
                   help="JSON object: fixed entry_level, expand_levels and tuning knobs for the deforming section "
                        "instead of the autotuner (counter runs)")
    p.add_argument("--self-tuning", default="",
                   help="JSON object: fixed entry_level, expand_levels and tuning knobs for the self-collision "
                        "section instead of the autotuner (counter runs)")
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


def _colmajor(mats: np.ndarray) -> np.ndarray:
    """(..., 4, 4) matrices -> column-major float32 (glm::mat4 layout)."""
    return np.ascontiguousarray(np.swapaxes(np.asarray(mats, np.float32), -1, -2))


def _pair_set(a) -> set:
    return set(map(tuple, np.asarray(a).tolist()))


def _pair_diff(a: np.ndarray, b: np.ndarray) -> int:
    return len(_pair_set(a) ^ _pair_set(b))


def _rows_of_scene(p: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """rows of a sorted pair list whose meshA lies in [lo, hi)"""
    if not len(p):
        return p
    a, b = np.searchsorted(p[:, 0], lo), np.searchsorted(p[:, 0], hi)
    return p[a:b]


# ====================================================================== batched
class BatchedWorkload:
    """The 512-scene workload, its matrices for `nf` frames in HBM, and the
    per-scene mesh ranges (instances are grouped by scene)."""

    def __init__(self, args, api, M):
        t0 = time.time()
        self.rs = M.RigidScenes(n_scenes=args.batched_scenes)
        rs = self.rs
        self.n = len(rs.instances)
        self.nf = max(2, args.frame_cycle)
        self.mats = np.stack([rs.transforms(f) for f in range(self.nf)])  # (nf, n, 4, 4)
        self.xbuf = api.DeviceBuffer(self.mats.nbytes)
        self.xbuf.upload(_colmajor(self.mats))
        self.stride = self.n * 64
        scene_of = np.asarray([s for s, _ in rs.instances])
        self.first = np.searchsorted(scene_of, np.arange(rs.n_scenes + 1))  # scene s: meshes [first[s], first[s+1])
        self.faces = np.asarray([rs.library[b][1].shape[0] for _, b in rs.instances], np.int64)
        self.tri = int(self.faces.sum())
        self.t_gen = time.time() - t0

    def geometry(self, i: int, unique: bool):
        return self.rs.unique_mesh(i) if unique else self.rs.library[self.rs.instances[i][1]]

    def scene_tri(self, s: int) -> int:
        return int(self.faces[self.first[s]:self.first[s + 1]].sum())


def _make_batch(api, ctx, W, unique: bool, geoms=None):
    b = api.Batch(ctx)
    first = {}
    for i, (scene, lib) in enumerate(W.rs.instances):
        if unique:
            xyz, f = geoms[i] if geoms is not None else W.geometry(i, True)
            b.add_mesh(scene, xyz, f)
        elif lib in first:
            b.add_instance(scene, first[lib])
        else:
            xyz, f = W.rs.library[lib]
            first[lib] = b.add_mesh(scene, xyz, f)
    b.set_transforms_device(W.xbuf.at(0))
    b.build()
    return b


def _time_batched(args, api, ctx, b, W, entry, expand, K, warm, overlap, timing_every=8):
    """warm + K frames (frames continue modulo the cycle); returns ms per
    timed frame, the timing sample and the stats of the last frame."""
    def stage(f):
        b.set_transforms_device(W.xbuf.at((f % W.nf) * W.stride))

    for f in range(warm):
        stage(f)
        b.frame(entry, expand, fp64_fallback=True, asynchronous=True, overlap=overlap)
    b.stats()
    timer = api.Timer(ctx)
    ctx.synchronize()
    # event samples start at the second timed frame: the first follows the
    # synchronisation above, so with overlap its refit runs alone
    if K == 1:
        b.set_timing(timing_every)
    timer.start()
    for k in range(K):
        stage(warm + k)
        b.frame(entry, expand, fp64_fallback=True, asynchronous=True, overlap=overlap)
        if k == 0 and K > 1:
            b.set_timing(timing_every)
    # the last frame's detect runs on the library's side stream: wait for it
    # before the stop event (the host round trip, microseconds, is counted)
    ctx.synchronize()
    timer.stop()
    ms = timer.ms() / K
    tm = b.timing()
    b.set_timing(0)
    st = b.stats()
    return ms, tm, st, warm + K - 1


def _refit_bytes(W, geoms_v, rk):
    """algorithmic bytes of one refit stage-1 launch over every mesh: 12 V +
    12 F + 24 N' (positions, sorted faces, nodes below the apex) + 64 (the
    model matrix) per mesh"""
    tot = 0
    for i in range(W.n):
        F = int(W.faces[i])
        tot += 12 * geoms_v[i] + 12 * F + 24 * (node_count(F) - apex_nodes(F, rk)) + 64
    return tot


def run_batched(args, ctx, out):
    from oibvh_amd import _lib as L
    from oibvh_amd import api, autotune
    from oibvh_amd import meshgen as M

    K, Wm = max(1, args.steps), max(0, args.warmup)
    progress("batched: generating scenes")
    W = BatchedWorkload(args, api, M)
    t0 = time.time()
    geoms = [W.geometry(i, True) for i in range(W.n)]
    t_geo = time.time() - t0
    verts = [g[0].shape[0] for g in geoms]
    progress(f"batched: {W.n} meshes, {W.tri} triangles (own geometry, {t_geo:.1f} s); building")
    t1 = time.time()
    b = _make_batch(api, ctx, W, True, geoms)
    ctx.synchronize()
    t_build = time.time() - t1
    overlap = not args.no_overlap

    def stage(i):
        b.set_transforms_device(W.xbuf.at((i % W.nf) * W.stride))

    progress(f"batched: built in {t_build:.2f} s; autotuning")
    tuned = autotune.tune_batch(ctx, b, stage, entries=(0, 1, 2, 3, 4), flags={"fp64_fallback": True},
                                reps=args.tune_reps, narrow=(0, 1))
    progress(f"batched: tuned {tuned}")
    entry, expand = tuned["entry_level"], tuned["expand_levels"]
    # frames one after another, then overlapped (refit of frame f beside the
    # detect of frame f - 1, the refit's workgroups capped as the tuner
    # picks): the faster is the headline, the other is reported beside it
    modes = {}
    ms_b, tm, st_b, f_last = _time_batched(args, api, ctx, b, W, entry, expand, K, Wm, False)
    modes["sequential"] = {"fps": 1000.0 / ms_b, "refit_stage1_ms": tm["refit_ms"] / max(1, tm["refits"]),
                           "detect_ms": tm["detect_ms"] / max(1, tm["detects"]), "tuning": ctx.get_tuning()}
    progress(f"batched: {1000.0 / ms_b:.1f} frames/s one after another")
    chosen = ("sequential", ms_b, tm, st_b, f_last, ctx.get_tuning())
    modes["sequential"]["autotune"] = tuned
    if overlap:
        seq_tuning = ctx.get_tuning()
        # tuned again with overlapped frames: a detect that leaves room on the
        # GPU for the refit beside it can win there while losing alone
        depth = max((int(F) - 1).bit_length() for F in W.faces)
        ov = autotune.tune_batch(ctx, b, stage, entries=(0, 1, 2, 3, 4),
                                 flags={"fp64_fallback": True, "overlap": True},
                                 blocks=(0, 2048, 4096), refit_blocks=(0, 512, 1024, 2048),
                                 reps=args.tune_reps, narrow=(0, 1), depth=depth)
        progress(f"batched: tuned for overlapped frames {ov}")
        e_o, x_o = ov["entry_level"], ov["expand_levels"]
        ms_o, tm_o, st_o, fl_o = _time_batched(args, api, ctx, b, W, e_o, x_o, K, Wm, True)
        modes["overlapped"] = {"fps": 1000.0 / ms_o, "refit_stage1_ms": tm_o["refit_ms"] / max(1, tm_o["refits"]),
                               "detect_ms": tm_o["detect_ms"] / max(1, tm_o["detects"]),
                               "refit_blocks": ov["refit_blocks"], "tuning": ctx.get_tuning(), "autotune": ov}
        progress(f"batched: {1000.0 / ms_o:.1f} frames/s overlapped")
        if ms_o < ms_b:
            chosen = ("overlapped", ms_o, tm_o, st_o, fl_o, ctx.get_tuning())
            entry, expand, tuned = e_o, x_o, ov
        else:
            ctx.set_tuning(**seq_tuning)
            # the comparisons below read the last frame: run it again in the
            # chosen mode
            stage(f_last)
            b.frame(entry, expand, fp64_fallback=True)
    mode, ms_b, tm, st_b, f_last, _ = chosen
    overlap = mode == "overlapped"
    pairs_b64 = b.pairs()
    lc = b.level_counts().tolist()
    mats_last = W.mats[f_last % W.nf]
    # the last frame again in plain FP32 for the comparisons below
    stage(f_last)
    b.refit()
    b.detect(entry, expand)
    pairs_b32 = b.pairs()
    perms = None

    tn = ctx.get_tuning()
    nitems = sum(-(-int(F) // 1024) for F in W.faces)
    rk = reduce_kernel(tn["reduce_leaves_per_thread"], W.tri, tn["refit_blocks"] if overlap else 0, nitems)
    refit_ms = tm["refit_ms"] / max(1, tm["refits"])
    detect_ms = tm["detect_ms"] / max(1, tm["detects"])
    refit_bytes = _refit_bytes(W, verts, rk)
    detect_bytes = sum(24 * node_count(int(F)) + 12 * v + 12 * int(F) for F, v in zip(W.faces, verts)) \
        + 16 * len(pairs_b64)
    roofline = {"bound": "hbm", "kernel": f"{rk} (refit stage 1 over all {W.n} meshes, one launch per frame)",
                "achieved": refit_bytes / (refit_ms * 1e-3) / 1e9 if refit_ms else None,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": refit_bytes / (refit_ms * 1e-3) / HBM_PEAK if refit_ms else None,
                "traffic": TRAFFIC_REFIT_BATCHED, "alg_bytes_per_launch": refit_bytes,
                "avg_launch_ms": refit_ms, "launches_timed": tm["refits"],
                "alg_bytes_unit": "12 V + 12 F + 24 N' + 64 per mesh (positions, sorted faces, nodes below the "
                                  "apex, model matrix)"}
    kernels = {"refit_stage1": {"kernel": rk, "avg_ms": refit_ms, "launches": tm["refits"],
                                "alg_bytes": refit_bytes},
               "detect": {"kernels": "BVTT walk (k_bvtt_local / k_bvtt_step) + k_narrow<FP64 fallback>, one graph "
                                     "on a side stream", "avg_ms": detect_ms, "launches": tm["detects"],
                          "alg_bytes": detect_bytes,
                          "gbps": detect_bytes / (detect_ms * 1e-3) / 1e9 if detect_ms else None,
                          "frac": detect_bytes / (detect_ms * 1e-3) / HBM_PEAK if detect_ms else None}}

    # ------------------------------------------------------------ parity
    verify = {}
    if not args.no_verify:
        import oracle

        progress("batched: FP64 oracle check")
        changed = []
        for s in range(W.rs.n_scenes):
            lo, hi = W.first[s], W.first[s + 1]
            if _pair_set(_rows_of_scene(pairs_b32, lo, hi)) != _pair_set(_rows_of_scene(pairs_b64, lo, hi)):
                changed.append(s)
        even = list(range(0, W.rs.n_scenes, max(1, W.rs.n_scenes // max(1, args.oracle_scenes))))
        sample = sorted(set(changed) | set(even[:max(0, args.oracle_scenes)]))
        d64 = miss32 = spur32 = 0
        t_or = time.time()
        for s in sample:
            ids = list(range(W.first[s], W.first[s + 1]))
            ot = [oracle.Tree(oracle.xform(mats_last[i], geoms[i][0]), geoms[i][1], order=b.perm(i)) for i in ids]
            want = oracle.detect(ot, fp64=True)
            if len(want):
                want = want.copy()
                want[:, 0] = np.asarray(ids, np.uint32)[want[:, 0]]
                want[:, 1] = np.asarray(ids, np.uint32)[want[:, 1]]
            lo, hi = W.first[s], W.first[s + 1]
            d64 += _pair_diff(_rows_of_scene(pairs_b64, lo, hi), want)
            a32, a64 = _pair_set(_rows_of_scene(pairs_b32, lo, hi)), _pair_set(want)
            miss32 += len(a64 - a32)
            spur32 += len(a32 - a64)
        verify = {"pair_diffs_vs_fp64_oracle": d64, "oracle_scenes": len(sample),
                  "oracle_scene_triangles": int(sum(W.scene_tri(s) for s in sample)),
                  "scenes_where_fp64_changed_the_pairs": changed,
                  "fp32_only_missed_vs_fp64_oracle": miss32, "fp32_only_spurious_vs_fp64_oracle": spur32,
                  "oracle_s": time.time() - t_or}

    # ------------------------------------------------------------ plain
    plain = None
    if not args.no_plain:
        progress("batched: setting up the plain form")
        plain = run_plain_form(args, api, ctx, W, geoms, f_last, pairs_b32)
        progress(f"plain: {plain['fps']:.2f} frames/s")

    # ------------------------------------------------------------ CPU
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_batched(args, W, geoms, b)

    fps = 1000.0 / ms_b
    hit_pairs = len(set(map(tuple, pairs_b64[:, :2].tolist()))) if len(pairs_b64) else 0
    per_scene = np.diff(W.first)
    out.update({
        "value": fps, "unit": "frames/s", "steps": K, "warmup": Wm, "ms_per_step": ms_b,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"batched scenes: {W.rs.n_scenes} scenes, {W.n} rigid meshes (2-16 per scene, noisy "
                                f"icospheres and boxes, log-uniform 1k-200k triangles, each mesh with its own "
                                f"geometry), {W.tri} triangles, radius 0.25, translations in the unit cube, spin + "
                                f"drift per frame; per frame new model matrices, refit, broad + narrow phase"),
                   "form": ("batched FP32 + FP64 fallback, " + ("overlapped frames (refit of frame f beside the "
                            "detect of frame f - 1), one hipGraph per half" if overlap else
                            "frames one after another, one hipGraph per frame")),
                   "scenes": W.rs.n_scenes, "meshes": W.n, "triangles_per_frame": W.tri,
                   "entry_level": entry, "expand_levels": expand, "tuning": ctx.get_tuning()},
        "mtri_per_s": W.tri * fps / 1e6,
        "frame_modes": modes,
        "plain": plain,
        "speedup_batched_vs_plain": (fps / plain["fps"]) if plain else None,
        "pair_diffs_batched_fp32_vs_plain": plain["pair_diffs_vs_batched_fp32"] if plain else None,
        **verify,
        "pairs_last_frame": int(len(pairs_b64)),
        "fp64_candidates_last_frame": st_b["fp64_candidates"],
        "fp64_axis_evaluations_last_frame": st_b["fp64_rechecks"],
        "candidates_last_frame": st_b["candidates"],
        "bvtt_tests_last_frame": st_b["bvtt_tests"],
        "async_errors": st_b["async_errors"], "retries_last_frame": st_b["retries"],
        "bvtt_levels_last_frame": lc,
        "mesh_pairs_overlapping": hit_pairs, "mesh_pairs_total": int(sum(k * (k - 1) // 2 for k in per_scene)),
        "kernels": kernels,
        "roofline": roofline,
        # the frame as a whole: refit and detect algorithmic bytes over the
        # frame time (with overlapped frames both halves share HBM, so each
        # kernel's own duration stretches while the frame shortens)
        "frame_roofline": {"bound": "hbm", "alg_bytes_per_frame": refit_bytes + detect_bytes,
                           "achieved": (refit_bytes + detect_bytes) / (ms_b * 1e-3) / 1e9,
                           "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": (refit_bytes + detect_bytes) / (ms_b * 1e-3) / HBM_PEAK},
        "cpu_baseline": cpu,
        "autotune": tuned,
        "build_s": t_build, "host_gen_s": W.t_gen + t_geo,
    })
    b.close()
    del geoms
    return W


def run_plain_form(args, api, ctx, W, geoms, f_last, pairs_b32):
    """One oibvh_tree per mesh and one oibvh_scene per scene, driven from C++
    (bench_plain.cpp): per frame and scene, the scene's matrices read in
    place (oibvh_scene_set_transforms, borrowed) and one oibvh_scene_frame.
    FP32, as the reference."""
    from oibvh_amd import _lib as L

    lib = C.CDLL(os.path.join(ROOT, "libbench_plain.so"))
    lib.plain_frames.restype = C.c_int
    lib.plain_frames.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p, C.POINTER(C.c_uint64), C.c_uint64,
                                 C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.POINTER(C.c_uint32)]
    t0 = time.time()
    trees = [api.OibvhTree(ctx, x, f) for x, f in geoms]
    scenes = [api.Scene(ctx) for _ in range(W.rs.n_scenes)]
    # build every tree under its frame-0 matrix (as the batch), then add it
    # to its scene
    for s, sc in enumerate(scenes):
        for i in range(W.first[s], W.first[s + 1]):
            trees[i].set_transform(W.mats[0, i])
            trees[i].build()
            sc.addOibvhTree(trees[i])
    ctx.synchronize()
    t_setup = time.time() - t0
    ns = W.rs.n_scenes
    handles = (C.c_void_p * ns)(*[sc.h.value for sc in scenes])
    mat_off = (C.c_uint64 * ns)(*[int(W.first[s]) * 16 for s in range(ns)])
    failed = C.c_uint32()

    def frames(f0, k, entry, expand):
        st = lib.plain_frames(handles, ns, C.c_void_p(W.xbuf.at(0)), mat_off, W.n * 16, f0, k, W.nf, entry, expand,
                              0, C.byref(failed))
        if st != 0:
            raise RuntimeError(f"plain_frames: status {st} at scene {failed.value}: "
                               f"{L.load().oibvh_last_error(ctx.h)}")

    timer = api.Timer(ctx)
    # a small tuning pass over (entry level, levels per step), none of which
    # changes results
    cands = [(0, 3), (2, 2), (4, 2), (4, 3)]
    best = None
    frames(0, 1, 0, 3)
    for s in scenes:
        s.getIntTriPairCount()  # sizes every scene's buffers
    for e, x in cands:
        frames(1, 1, e, x)
        ctx.synchronize()
        timer.start()
        frames(2, 2, e, x)
        ctx.synchronize()
        timer.stop()
        ms = timer.ms() / 2
        for s in scenes:
            s.getIntTriPairCount()
        if best is None or ms < best[0]:
            best = (ms, e, x)
    _, entry, expand = best
    kp = max(1, args.plain_frames)
    f0 = f_last - kp + 1
    frames(f0 - 1, 1, entry, expand)  # warm
    for s in scenes:
        s.getIntTriPairCount()
    ctx.synchronize()
    tw = time.time()
    timer.start()
    frames(f0, kp, entry, expand)
    ctx.synchronize()
    timer.stop()
    ms_gpu = timer.ms() / kp
    ms_wall = (time.time() - tw) * 1000.0 / kp
    rows = []
    retries = async_errors = 0
    for s, sc in enumerate(scenes):
        if W.first[s + 1] - W.first[s] < 2:
            continue
        p = sc.getIntTriPairs()
        stt = sc.stats()
        retries += stt["retries"]
        async_errors |= stt["async_errors"]
        if len(p):
            p = p.copy()
            p[:, :2] += np.uint32(W.first[s])
            rows.append(p)
    pairs = np.concatenate(rows) if rows else np.zeros((0, 4), np.uint32)
    res = {"fps": 1000.0 / ms_gpu, "ms_per_frame": ms_gpu, "wall_ms_per_frame": ms_wall, "frames": kp,
           "form": ("one oibvh_tree per mesh, one oibvh_scene per scene; per frame and scene "
                    "oibvh_scene_set_transforms (matrices read in place from HBM) + oibvh_scene_frame (refit of "
                    "every tree + asynchronous detect), FP32, driven from C++ (bench_plain.cpp)"),
           "entry_level": entry, "expand_levels": expand,
           "pair_diffs_vs_batched_fp32": _pair_diff(pairs, pairs_b32), "pairs": int(len(pairs)),
           "retries_last_frame": retries, "async_errors": async_errors, "setup_s": t_setup}
    for sc in scenes:
        sc.close()
    for t in trees:
        t.close()
    return res


def cpu_baseline_batched(args, W, geoms, b):
    """The oracle on every 16th scene, frames until the time budget: per frame
    and sampled scene, every mesh's positions transformed and refit, then the
    scene's detect. Extrapolated to all scenes by triangle count."""
    import oracle

    progress("batched: CPU baseline")
    sample = list(range(0, W.rs.n_scenes, 16))
    trees = {}
    for s in sample:
        trees[s] = [oracle.Tree(oracle.xform(W.mats[0][i], geoms[i][0]), geoms[i][1], order=b.perm(i))
                    for i in range(W.first[s], W.first[s + 1])]
    tri_sample = sum(W.scene_tri(s) for s in sample)
    n = 0
    t = 0.0
    while n == 0 or (t < args.cpu_seconds and n < W.nf - 1):
        f = 1 + n
        t0 = time.time()
        for s in sample:
            ids = range(W.first[s], W.first[s + 1])
            for k, i in enumerate(ids):
                trees[s][k].refit(oracle.xform(W.mats[f][i], geoms[i][0]))
            if len(ids) > 1:
                oracle.detect(trees[s])
        t += time.time() - t0
        n += 1
    per_frame_sample = t / n
    per_frame = per_frame_sample * W.tri / tri_sample
    return {"value": 1.0 / per_frame, "unit": "frames/s", "cores": 1, "kind": "port",
            "mtri_per_s": W.tri / per_frame / 1e6,
            "sample": (f"oracle (CPU restatement of src/cpu, 1 thread of {os.cpu_count()} on '{cpu_model()}'): "
                       f"{len(sample)} of {W.rs.n_scenes} scenes (every 16th, {tri_sample} triangles), {n} frame(s) "
                       f"of transform + refit + detect at {per_frame_sample:.3f} s each, extrapolated to all "
                       f"{W.tri} triangles by triangle count")}


def run_instanced(args, ctx, W):
    """The same scenes with every mesh a rigid instance of one of the base
    shapes (shared vertex and face buffers): batched form only, overlapped
    frames (unless --no-overlap), tuned for them."""
    from oibvh_amd import api, autotune

    progress("instanced: building")
    b = _make_batch(api, ctx, W, False)

    def stage(i):
        b.set_transforms_device(W.xbuf.at((i % W.nf) * W.stride))

    base = ctx.get_tuning()
    overlap = not args.no_overlap
    tuned = autotune.tune_batch(ctx, b, stage, entries=(0, 1, 2, 3, 4),
                                flags={"fp64_fallback": True, "overlap": overlap},
                                blocks=(0, 2048, 4096) if overlap else (0, 4096, 8192),
                                refit_blocks=(0, 512, 1024, 2048) if overlap else (0,), reps=args.tune_reps,
                                narrow=(0, 1))
    ms, tm, st, _ = _time_batched(args, api, ctx, b, W, tuned["entry_level"], tuned["expand_levels"],
                                  max(1, args.steps), max(0, args.warmup), overlap)
    res = {"workload": f"as the headline, the {W.n} meshes rigid instances of {len(W.rs.library)} base shapes",
           "frames": "overlapped" if overlap else "one after another",
           "fps": 1000.0 / ms, "ms_per_frame": ms, "mtri_per_s": W.tri / (ms * 1e-3) / 1e6,
           "refit_stage1_ms": tm["refit_ms"] / max(1, tm["refits"]),
           "detect_ms": tm["detect_ms"] / max(1, tm["detects"]),
           "pairs_last_frame": st["pairs"], "async_errors": st["async_errors"], "autotune": tuned}
    b.close()
    ctx.set_tuning(**base)
    progress(f"instanced: {res['fps']:.1f} frames/s")
    return res


# ====================================================================== deforming pair
def run_deforming(args, ctx):
    """Two UV spheres of 1000 x 501 (1,000,000 triangles each) 0.9 radii
    apart; every vertex moves along its rest normal by 0.05 sin(wt + k.p).
    The timed region is the whole 200-frame sequence: refit every frame,
    rebuild both trees at frames 0, 50, 100, 150, broad + narrow phase every
    frame; refit frames overlap the previous frame's detect."""
    from oibvh_amd import api, autotune
    from oibvh_amd import meshgen as M

    meshes, waves = M.deforming_uv_pair()
    n_frames, rebuild = 200, args.rebuild_every
    V = meshes[0][0].shape[0]
    Fm = [m[1].shape[0] for m in meshes]
    t_gen = time.time()
    frames = [[w.frame(f) for f in range(n_frames)] for w in waves]
    t_gen = time.time() - t_gen
    bufs = []
    for m in range(len(meshes)):
        bb = api.DeviceBuffer(n_frames * V * 12)
        for f, x in enumerate(frames[m]):
            bb.upload(x, f * V * 12)
        bufs.append(bb)
    trees = []
    for m, (x, f) in enumerate(meshes):
        t = api.OibvhTree(ctx, x, f)
        t.set_positions_device(bufs[m].at(0))
        t.build()
        trees.append(t)
    scene = api.Scene(ctx)
    for t in trees:
        scene.addOibvhTree(t)
    pos = [[bufs[m].at(f * V * 12) for m in range(len(meshes))] for f in range(n_frames)]
    pos_arr = [api.position_array(p) for p in pos]
    def tune_frame(i):
        for m, t in enumerate(trees):
            t.set_positions_device(pos[(1 + i) % n_frames][m])
        scene.refit()

    def overlapped(i0, n):
        for i in range(n):
            scene.frame(pos_arr[(1 + i0 + i) % n_frames], entry, expand, asynchronous=True, overlap=True)

    def overlapped_at(i0, n, e, x):
        for i in range(n):
            scene.frame(pos_arr[(1 + i0 + i) % n_frames], e, x, asynchronous=True, overlap=True)

    base = ctx.get_tuning()
    if args.deforming_tuning:
        # fixed settings (counter runs, whose passes must see the same kernels)
        tuned = json.loads(args.deforming_tuning)
        entry, expand = tuned.pop("entry_level"), tuned.pop("expand_levels")
        ctx.set_tuning(**{**tuned, "traverse_expand": expand})
        tuned = {"fixed": True, "entry_level": entry, "expand_levels": expand, **tuned}
        tuned_rb = tuned_alone = None
        scene.detectCollision(entry, expand)
        scene.getIntTriPairCount()
    else:
        progress("deforming pair: autotuning")
        tuned = tuned_alone = autotune.tune(ctx, scene, tune_frame)
        # then the walk's shape again on overlapped frames, around that pick
        keys = ("traverse_blocks", "reduce_leaves_per_thread")
        cands = [tuned] + [{"entry_level": tuned["entry_level"], "expand_levels": x,
                            **{k: tuned[k] for k in keys}, "traverse_local": lc, "traverse_local_levels": ll,
                            "traverse_tail": tc, "traverse_tail_steps": 0, "narrow_separate": ns}
                           for x in (2, 3) for lc, ll in ((0, 0), (128, 3), (256, 3), (256, 2))
                           for tc in (0, 32, 64) for ns in (0, 1)]
        tuned = autotune.tune_overlapped(ctx, scene, overlapped_at, cands)
        entry, expand = tuned["entry_level"], tuned["expand_levels"]
        ctx.set_tuning(traverse_expand=expand)
        scene.detectCollision(entry, expand)
        scene.getIntTriPairCount()
        tuned_rb = autotune.tune_refit_blocks(ctx, scene, overlapped)

    def step(f):
        if f % rebuild == 0:
            for m, t in enumerate(trees):
                t.set_positions_device(pos[f][m])
            scene.build()  # both trees in every launch of one build chain
            scene.detectCollision(entry, expand, asynchronous=True)
        else:
            scene.frame(pos_arr[f], entry, expand, asynchronous=True, overlap=True)

    for f in range(1, 6):  # warm-up
        step(f)
    scene.getIntTriPairCount()
    timer = api.Timer(ctx)
    ctx.synchronize()
    scene.timing()
    scene.set_timing(8)
    timer.start()
    for f in range(n_frames):
        step(f)
    ctx.synchronize()
    timer.stop()
    total_ms = timer.ms()
    st = scene.stats()
    pairs_last = scene.getIntTriPairs()
    tm = scene.timing()
    scene.set_timing(0)
    # the rebuild's own cost, event-timed apart (four per sequence): both
    # trees (oibvh_scene_build), and one tree after the other; medians over
    # --rebuild-reps calls, each started on an idle GPU and timed to its last
    # launch's end (the stop event follows the launches in the stream; the
    # host has queued them long before the GPU ends them); then the same
    # calls queued back to back, as the sequence's rebuilds run behind frames
    tb = api.Timer(ctx)
    build_v, build_seq_v = [], []
    for _ in range(max(1, args.rebuild_reps)):
        ctx.synchronize()
        tb.start()
        scene.build()
        tb.stop()
        build_v.append(tb.ms())
    ctx.synchronize()
    tb.start()
    for _ in range(max(1, args.rebuild_reps)):
        scene.build()
    tb.stop()
    build_b2b_ms = tb.ms() / max(1, args.rebuild_reps)
    for _ in range(max(1, args.rebuild_reps)):
        ctx.synchronize()
        tb.start()
        for t in trees:
            t.build()
        tb.stop()
        build_seq_v.append(tb.ms())
    build_ms = float(np.median(build_v))
    build_seq_ms = float(np.median(build_seq_v))
    ms_step = total_ms / n_frames
    fps = 1000.0 / ms_step
    refit_ms = tm["refit_ms"] / max(1, tm["refits"])
    detect_ms = tm["detect_ms"] / max(1, tm["detects"])
    tn = ctx.get_tuning()
    N = [t.node_count for t in trees]
    rk = reduce_kernel(tn["reduce_leaves_per_thread"], sum(Fm), tn["refit_blocks"], sum(-(-F // 1024) for F in Fm))
    refit_bytes = sum(12 * V + 12 * F + 24 * (n - apex_nodes(F, rk)) for F, n in zip(Fm, N))
    detect_bytes = sum(24 * n + 12 * V + 12 * F for F, n in zip(Fm, N)) + 16 * len(pairs_last)
    diffs = None
    if not args.no_verify:
        import oracle

        f_last = n_frames - 1
        otrees = [oracle.Tree(frames[m][f_last], meshes[m][1], order=trees[m].perm()) for m in range(len(meshes))]
        diffs = _pair_diff(pairs_last, oracle.detect(otrees))
    cpu = None
    if not args.no_cpu_baseline:
        import oracle

        t0 = time.time()
        ot = [oracle.Tree(frames[m][0], meshes[m][1], order="morton") for m in range(len(meshes))]
        t_build = (time.time() - t0) / len(meshes)
        n, tf = 0, 0.0
        while n < n_frames - 1 and (tf < args.cpu_seconds or n == 0):
            t0 = time.time()
            for m, t in enumerate(ot):
                t.refit(frames[m][1 + n])
            oracle.detect(ot)
            tf += time.time() - t0
            n += 1
        per_frame = tf / n + (t_build * len(meshes)) / rebuild
        cpu = {"value": 1.0 / per_frame, "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": (f"oracle, 1 thread of {os.cpu_count()} on '{cpu_model()}': Morton build of both meshes "
                          f"({t_build:.3f} s each, amortised over {rebuild} frames) + {n} frame(s) of refit + detect "
                          f"({tf / n:.3f} s each)")}
    res = {"workload": ("two UV spheres 1000x501 (1,000,000 triangles each) 0.9 radii apart, vertices displaced "
                        "along the normal by 0.05 sin(wt + k.p); the 200-frame sequence timed whole: refit every "
                        f"frame, rebuild of both trees every {rebuild} frames, broad + narrow phase every frame"),
           "fps": fps, "ms_per_frame": ms_step, "mtri_per_s": sum(Fm) * fps / 1e6, "frames": n_frames,
           "rebuild_ms_both_trees": build_ms, "rebuild_ms_back_to_back": build_b2b_ms,
           "rebuild_ms_one_after_another": build_seq_ms,
           "rebuild_reps": len(build_v), "rebuild_ms_range": [float(min(build_v)), float(max(build_v))],
           "frame_overlap": True,
           "entry_level": entry, "expand_levels": expand, "autotune": tuned, "autotune_alone": tuned_alone,
           "autotune_refit_blocks": tuned_rb,
           "pair_diffs_last_frame_vs_oracle": diffs, "pairs_last_frame": int(len(pairs_last)),
           "async_errors": st["async_errors"], "retries_last_frame": st["retries"],
           "kernels": {"refit_stage1": {"kernel": rk, "avg_ms": refit_ms, "launches": tm["refits"],
                                        "alg_bytes": refit_bytes},
                       "detect": {"avg_ms": detect_ms, "launches": tm["detects"], "alg_bytes": detect_bytes,
                                  "alg_bytes_unit": "24 N per tree (broad) + 12 V + 12 F per tree (narrow) + 16 B "
                                                    "per hit",
                                  "gbps": detect_bytes / (detect_ms * 1e-3) / 1e9 if detect_ms else None,
                                  "frac": detect_bytes / (detect_ms * 1e-3) / HBM_PEAK if detect_ms else None}},
           "roofline": {"bound": "hbm", "kernel": f"{rk} (refit stage 1, both trees in one launch)",
                        "achieved": refit_bytes / (refit_ms * 1e-3) / 1e9 if refit_ms else None,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": refit_bytes / (refit_ms * 1e-3) / HBM_PEAK if refit_ms else None,
                        "traffic": TRAFFIC_REFIT_PAIR, "alg_bytes_per_launch": refit_bytes, "avg_launch_ms": refit_ms},
           "cpu_baseline": cpu, "host_gen_s": t_gen}
    for t in trees:
        t.close()
    scene.close()
    for bb in bufs:
        bb.free()
    ctx.set_tuning(**base)
    return res


# ====================================================================== precision
def run_precision(args, ctx):
    """Coplanar / edge-touching scenes: FP32 filter + FP64 re-check; every
    scene checked against the FP64 oracle."""
    from oibvh_amd import api, autotune
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
    # static geometry: every frame refits and detects the same positions
    overlap = not args.no_overlap
    base = ctx.get_tuning()
    tuned = autotune.tune_batch(ctx, b, lambda i: None, entries=(0, 1, 2, 3, 4),
                                flags={"fp64_fallback": True, "overlap": overlap},
                                blocks=(0, 2048, 4096), refit_blocks=(0, 512, 1024) if overlap else (0,),
                                reps=args.tune_reps, narrow=(0, 1))
    entry, expand = tuned["entry_level"], tuned["expand_levels"]
    b.frame(entry, expand, fp64_fallback=True, overlap=overlap)
    timer = api.Timer(ctx)
    k = max(1, args.precision_frames)
    ctx.synchronize()
    timer.start()
    for _ in range(k):
        b.frame(entry, expand, fp64_fallback=True, asynchronous=True, overlap=overlap)
    ctx.synchronize()  # an overlapped frame's detect runs on a side stream
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
           "candidates": st["candidates"], "fp64_candidates": st["fp64_candidates"],
           "fp64_candidate_rate": st["fp64_candidates"] / max(1, st["candidates"]),
           "fp64_axis_evaluations": st["fp64_rechecks"],
           "async_errors": st["async_errors"], "retries_last_frame": st["retries"],
           "frames_overlapped": overlap, "autotune": tuned}
    progress(f"precision: {1000.0 / ms:.1f} frames/s; FP64 oracle on every scene")
    if not args.no_verify:
        import oracle

        d64 = miss = spur = 0
        changed = []
        t0 = time.time()
        for si in range(len(scenes)):
            ot = [oracle.Tree(x, f, order=b.perm(2 * si + j)) for j, (x, f) in enumerate(scenes[si])]
            want = oracle.detect(ot, fp64=True)
            if len(want):
                want = want.copy()
                want[:, :2] += 2 * si
            sel64 = _rows_of_scene(p64, 2 * si, 2 * si + 2)
            sel32 = _rows_of_scene(p32, 2 * si, 2 * si + 2)
            d64 += _pair_diff(sel64, want)
            a32, a64 = _pair_set(sel32), _pair_set(want)
            miss += len(a64 - a32)
            spur += len(a32 - a64)
            if a32 != a64:
                changed.append(si)
        res.update({"oracle_scenes": len(scenes), "pair_diffs_vs_fp64_oracle": d64,
                    "fp32_path_missed_vs_fp64": miss, "fp32_path_spurious_vs_fp64": spur,
                    "scenes_where_fp32_and_fp64_differ": changed, "oracle_s": time.time() - t0})
    res["host_gen_s"] = t_gen
    b.close()
    ctx.set_tuning(**base)
    return res


# ====================================================================== self collision
def run_self(args, ctx):
    """Folded cloth sheets: refit + self query per frame, 8 instances per
    launch, frames overlapped."""
    from oibvh_amd import api, autotune
    from oibvh_amd import meshgen as M

    nv = args.cloth_nv
    progress("self: generating cloth frames")
    t0 = time.time()
    n_inst, n_distinct, dphase = 8, 8, 0.02
    base = M.cloth_instances(n_inst, nv)
    V = nv * nv
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

    progress("self: built; autotuning")
    tb = ctx.get_tuning()
    ctx.set_tuning(front_capacity=1 << 29)
    # no head walk: a sheet's early fronts outgrow any workgroup's LDS. Tail
    # walks of at most 128 front pairs per workgroup over the last one or two
    # steps cannot overflow (two levels of 4 x 4 children: 16 x 128 = 2048
    # pairs, the LDS queue) and keep the largest fronts out of HBM
    depth = max((int(f.shape[0]) - 1).bit_length() for _, f in base)
    if args.self_tuning:
        # fixed settings (counter runs, whose passes must see the same kernels)
        tuned = json.loads(args.self_tuning)
        ctx.set_tuning(**{k: v for k, v in tuned.items() if k not in ("entry_level", "expand_levels")},
                       traverse_expand=tuned["expand_levels"])
        tuned = {"fixed": True, **tuned}
    else:
        tuned = autotune.tune_batch(ctx, b, stage, entries=(3, 4, 5, 6), expands=(2, 3),
                                    blocks=(0, 2048, 4096, 8192),
                                    leaves=(0,),
                                    local=(), tail=((32, 1), (64, 2), (128, 2)),
                                    flags={"self_collide": True, "overlap": True}, refit_blocks=(0, 1024), reps=4,
                                    depth=depth, narrow=(0, 1))
    progress(f"self: tuned {tuned}")
    entry, expand = tuned["entry_level"], tuned["expand_levels"]
    stage(0)
    b.frame(entry, expand, self_collide=True, overlap=True)
    timer = api.Timer(ctx)
    k = max(1, args.self_frames)
    ctx.synchronize()
    if k == 1:
        b.set_timing(8)
    timer.start()
    for j in range(k):
        stage(j)
        b.frame(entry, expand, self_collide=True, asynchronous=True, overlap=True)
        if j == 0 and k > 1:
            b.set_timing(8)  # samples from the second frame, as in _time_batched
    ctx.synchronize()
    timer.stop()
    ms = timer.ms() / k
    tm = b.timing()
    b.set_timing(0)
    st = b.stats()
    pairs = b.pairs()
    tri = int(sum(f.shape[0] for _, f in base))
    # full-size cross-check of the last frame: every sheet as its own plain
    # tree and scene, self mode, eagerly, against the batched frame's pairs
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
                        f"refit + self-BVTT query (adjacent pairs excluded) per frame, frames overlapped"),
           "triangles": tri, "frames": k, "fps": 1000.0 / ms, "ms_per_frame": ms,
           "mtri_per_s": tri / (ms * 1e-3) / 1e6, "pairs_last_frame": int(len(pairs)),
           "adjacent_skipped": st["adjacent"], "candidates": st["candidates"], "bvtt_tests": st["bvtt_tests"],
           "async_errors": st["async_errors"], "retries_last_frame": st["retries"],
           "bvtt_levels_last_frame": b.level_counts().tolist(), "autotune": tuned, "host_gen_s": t_gen}
    # the walk's roofline: each tree's nodes read once by the broad phase,
    # positions and faces once by the narrow phase, 16 B per hit written
    det_ms = tm["detect_ms"] / max(1, tm["detects"])
    det_bytes = sum(24 * node_count(int(f.shape[0])) + 12 * int(x.shape[0]) + 12 * int(f.shape[0])
                    for x, f in base) + 16 * len(pairs)
    res["kernels"] = {"refit_stage1_ms": tm["refit_ms"] / max(1, tm["refits"]),
                      "detect": {"avg_ms": det_ms, "launches": tm["detects"], "alg_bytes": det_bytes,
                                 "alg_bytes_unit": "24 N per tree (broad) + 12 V + 12 F per tree (narrow) + "
                                                   "16 B per hit",
                                 "gbps": det_bytes / (det_ms * 1e-3) / 1e9 if det_ms else None,
                                 "frac": det_bytes / (det_ms * 1e-3) / HBM_PEAK if det_ms else None}}
    if plain_pairs is not None:
        res["pair_diffs_batched_vs_plain_full_size"] = _pair_diff(pairs, plain_pairs)
    if not args.no_verify:
        import oracle

        # against the oracle at a reduced size (the full sheet above is
        # checked against the plain form)
        x, f = M.cloth(256, 0.2, 0.05, M.CFG5_SEED)
        c = api.Batch(ctx)
        c.add_mesh(0, x, f)
        c.build()
        c.detect(2, 2, self_collide=True)
        want = oracle.detect([oracle.Tree(x, f, order="morton")], self_collide=True)
        res["oracle_check"] = {"cloth_nv": 256, "pairs": int(len(want)), "pair_diffs": _pair_diff(c.pairs(), want)}
        c.close()
    b.close()
    for buf in bufs:
        buf.free()
    ctx.set_tuning(**tb)
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
    out = {"metric": METRIC, "value": None, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup}
    W = None
    if "batched" in sections:
        base = ctx.get_tuning()
        t0 = time.time()
        W = run_batched(args, ctx, out)
        out["section_wall_s"] = time.time() - t0
        ctx.set_tuning(**base)
    if "instanced" in sections:
        from oibvh_amd import api as A
        from oibvh_amd import meshgen as M

        if W is None:
            W = BatchedWorkload(args, A, M)
        t0 = time.time()
        out["instanced"] = run_instanced(args, ctx, W)
        out["instanced"]["section_wall_s"] = time.time() - t0
    if W is not None:
        W.xbuf.free()
    for name, key, fn in [("deforming", "deforming_pair", run_deforming), ("precision", "precision", run_precision),
                          ("self", "self_collision", run_self)]:
        if name in sections:
            base = ctx.get_tuning()
            t0 = time.time()
            res = fn(args, ctx)
            res["section_wall_s"] = time.time() - t0
            out[key] = res
            ctx.set_tuning(**base)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
