"""Autotuner for the knobs that change speed but never results.

For a live scene it times short runs of refit + detect over a small grid of
settings -- the reduce subtree size (leaves per thread), the BVTT entry
level, the levels descended per BVTT step and the workgroups per launch --
with HIP events on the context's stream, and keeps the fastest. The caller
provides `frame(i)`, which refits the scene's trees for trial i (it may move
vertices); the tuner adds the detect.
"""
from __future__ import annotations

import itertools

from . import api


def _time(ctx: api.Context, scene: api.Scene, frame, entry: int, expand: int, reps: int) -> float:
    timer = api.Timer(ctx)
    frame(0)
    scene.detectCollision(entry, expand)
    scene.getIntTriPairCount()  # sizes buffers, checks errors
    ctx.synchronize()
    timer.start()
    for i in range(reps):
        frame(i + 1)
        scene.detectCollision(entry, expand, asynchronous=True)
    timer.stop()
    ms = timer.ms() / reps
    scene.getIntTriPairCount()
    return ms


def _time_refit(ctx: api.Context, frame, reps: int) -> float:
    timer = api.Timer(ctx)
    frame(0)
    ctx.synchronize()
    timer.start()
    for i in range(reps):
        frame(i + 1)
    timer.stop()
    return timer.ms() / reps


def tune(ctx: api.Context, scene: api.Scene, frame, *, entries=(4, 6, 8), expands=(1, 2, 3),
         blocks=(0, 2048), leaves=(16, 2, 4, 8), local=((64, 2), (128, 3), (256, 2), (256, 3)),
         tail=((32, 0), (64, 0), (64, 2), (128, 2)), narrow=(0, 1), reps: int = 5) -> dict:
    """Return {'entry_level', 'expand_levels', 'traverse_blocks',
    'reduce_leaves_per_thread', 'traverse_local', 'traverse_local_levels',
    'traverse_tail', 'traverse_tail_steps', 'narrow_separate', 'ms'} of the
    fastest setting and apply the tuning knobs to ctx. `local` lists the
    (initial pairs per workgroup, steps) settings of the head local walk
    tried at the best step shape, `tail` the (front pairs per workgroup pass,
    steps) settings of the tail walk tried after the best head setting, with
    leaf pairs tested inside the walk or in one k_narrow launch (`narrow`)."""
    best = None
    # reduce variant first, timed on the refit alone (the detect does not
    # depend on it), then the traversal shape
    base = ctx.get_tuning()
    lpt_best, lpt_ms = 0, None
    for lpt in leaves:
        ctx.set_tuning(reduce_leaves_per_thread=lpt)
        ms = min(_time_refit(ctx, frame, 4 * reps) for _ in range(2))
        if lpt_ms is None or ms < lpt_ms:
            lpt_best, lpt_ms = lpt, ms
    ctx.set_tuning(reduce_leaves_per_thread=lpt_best)
    for entry, expand, nb in itertools.product(entries, expands, blocks):
        ctx.set_tuning(traverse_blocks=nb, traverse_expand=expand, traverse_local=0, traverse_local_levels=0,
                       traverse_tail=0, traverse_tail_steps=0)
        ms = _time(ctx, scene, frame, entry, expand, reps)
        if best is None or ms < best["ms"]:
            best = {"entry_level": entry, "expand_levels": expand, "traverse_blocks": nb,
                    "reduce_leaves_per_thread": lpt_best, "traverse_local": 0, "traverse_local_levels": 0,
                    "traverse_tail": 0, "traverse_tail_steps": 0, "ms": ms}
    # then the first steps as one workgroup-local launch, at the best shape
    ctx.set_tuning(traverse_blocks=best["traverse_blocks"], traverse_expand=best["expand_levels"])
    for chunk, levels in local:
        ctx.set_tuning(traverse_local=chunk, traverse_local_levels=levels)
        ms = _time(ctx, scene, frame, best["entry_level"], best["expand_levels"], reps)
        if ms < best["ms"]:
            best = {**best, "traverse_local": chunk, "traverse_local_levels": levels, "ms": ms}
    # then the last steps as a tail walk, after the best head setting, and
    # where the leaf pairs are tested
    ctx.set_tuning(traverse_local=best["traverse_local"], traverse_local_levels=best["traverse_local_levels"])
    best["narrow_separate"] = 0
    for ns, (chunk, steps) in itertools.product(narrow, ((0, 0),) + tuple(tail)):
        if ns == 0 and chunk == 0:
            continue  # timed above
        ctx.set_tuning(narrow_separate=ns, traverse_tail=chunk, traverse_tail_steps=steps)
        ms = _time(ctx, scene, frame, best["entry_level"], best["expand_levels"], reps)
        if ms < best["ms"]:
            best = {**best, "narrow_separate": ns, "traverse_tail": chunk, "traverse_tail_steps": steps, "ms": ms}
    ctx.set_tuning(**{**{k: base[k] for k in base},
                      "reduce_leaves_per_thread": best["reduce_leaves_per_thread"],
                      "traverse_blocks": best["traverse_blocks"], "traverse_expand": best["expand_levels"],
                      "traverse_local": best["traverse_local"],
                      "traverse_local_levels": best["traverse_local_levels"],
                      "traverse_tail": best["traverse_tail"], "traverse_tail_steps": best["traverse_tail_steps"],
                      "narrow_separate": best["narrow_separate"]})
    return best


def deepen_last_step(depth: int, entry: int, expand: int) -> int:
    """traverse_schedule word for steps of `expand` levels from `entry` with a
    last step of expand + 1 that ends at `depth` (one front fewer than steps
    of `expand` only), or 0 when no such split exists."""
    if expand >= 3:
        return 0
    rest = depth - entry - (expand + 1)
    if rest <= 0 or rest % expand or rest // expand > 16:
        return 0
    word = 0
    for k in range(rest // expand):
        word |= expand << (2 * k)
    return word


def tune_batch(ctx: api.Context, batch: api.Batch, before_frame, *, entries=(0, 2, 4), expands=(1, 2, 3),
               blocks=(0, 4096, 8192), leaves=(0, 4, 8), local=((64, 2), (256, 3)), tail=(), refit_blocks=(0,),
               flags=None, reps: int = 3, depth: int = 0, narrow=(0,)) -> dict:
    """Pick the reduce subtree size (leaves per thread), the BVTT entry level,
    the levels per step and the workgroups per launch for a batch's captured
    frame (refit + detect). before_frame(i) stages frame i's inputs (model
    matrices or vertex pointers) in stream order. `tail` (none by default)
    lists tail-walk settings to try: a batch's late fronts are large, and a
    tail walk that outgrows LDS on a later frame costs that frame a re-run.
    `refit_blocks` lists caps on an overlapped frame's refit workgroups (with
    flags {"overlap": True}; none changes results). With `depth` (the deepest
    tree's depth) a last step one level deeper than the others is tried as
    well (deepen_last_step). `narrow` lists narrow_separate settings to try
    (FP32 queries: leaf pairs tested inside the steps, or in one k_narrow
    launch)."""
    flags = flags or {}
    base = ctx.get_tuning()
    timer = api.Timer(ctx)

    def run(entry, expand):
        before_frame(0)
        batch.frame(entry, expand, **flags)  # captures, sizes buffers
        ctx.synchronize()
        timer.start()
        for i in range(reps):
            before_frame(i + 1)
            batch.frame(entry, expand, asynchronous=True, **flags)
        # an overlapped frame's detect runs on one of the library's side
        # streams: wait for it before the stop event
        ctx.synchronize()
        timer.stop()
        ms = timer.ms() / reps
        batch.pair_count()
        return ms

    lpt_best, lpt_ms = 0, None
    for lpt in leaves:
        ctx.set_tuning(reduce_leaves_per_thread=lpt)
        ms = run(entries[0], expands[0])
        if lpt_ms is None or ms < lpt_ms:
            lpt_best, lpt_ms = lpt, ms
    ctx.set_tuning(reduce_leaves_per_thread=lpt_best)
    best = None
    for entry, expand, nb in itertools.product(entries, expands, blocks):
        ctx.set_tuning(traverse_blocks=nb, traverse_expand=expand, traverse_local=0, traverse_local_levels=0,
                       traverse_tail=0, traverse_tail_steps=0)
        ms = run(entry, expand)
        if best is None or ms < best["ms"]:
            best = {"entry_level": entry, "expand_levels": expand, "traverse_blocks": nb,
                    "reduce_leaves_per_thread": lpt_best, "traverse_local": 0, "traverse_local_levels": 0,
                    "traverse_tail": 0, "traverse_tail_steps": 0, "ms": ms}
    ctx.set_tuning(traverse_blocks=best["traverse_blocks"], traverse_expand=best["expand_levels"])
    best["traverse_schedule"] = 0
    word = deepen_last_step(depth, best["entry_level"], best["expand_levels"]) if depth else 0
    if word:
        e = best["expand_levels"] + 1
        ctx.set_tuning(traverse_expand=e, traverse_schedule=word)
        ms = run(best["entry_level"], e)
        if ms < best["ms"]:
            best = {**best, "expand_levels": e, "traverse_schedule": word, "ms": ms}
        ctx.set_tuning(traverse_expand=best["expand_levels"], traverse_schedule=best["traverse_schedule"])
    for chunk, levels in local:
        ctx.set_tuning(traverse_local=chunk, traverse_local_levels=levels)
        ms = run(best["entry_level"], best["expand_levels"])
        if ms < best["ms"]:
            best = {**best, "traverse_local": chunk, "traverse_local_levels": levels, "ms": ms}
    ctx.set_tuning(traverse_local=best["traverse_local"], traverse_local_levels=best["traverse_local_levels"])
    for chunk, steps in tail:
        ctx.set_tuning(traverse_tail=chunk, traverse_tail_steps=steps)
        ms = run(best["entry_level"], best["expand_levels"])
        if ms < best["ms"]:
            best = {**best, "traverse_tail": chunk, "traverse_tail_steps": steps, "ms": ms}
    ctx.set_tuning(traverse_tail=best["traverse_tail"], traverse_tail_steps=best["traverse_tail_steps"])
    best["narrow_separate"] = base.get("narrow_separate", 0)
    for ns in narrow:
        if ns == best["narrow_separate"]:
            continue
        ctx.set_tuning(narrow_separate=ns)
        ms = run(best["entry_level"], best["expand_levels"])
        if ms < best["ms"]:
            best = {**best, "narrow_separate": ns, "ms": ms}
    ctx.set_tuning(narrow_separate=best["narrow_separate"])
    best["refit_blocks"] = base.get("refit_blocks", 0)
    # a cap applies to the wave-subtree refit (16 leaves per thread) only:
    # each capped candidate runs with it
    for rb in refit_blocks:
        lpt = 16 if rb else best["reduce_leaves_per_thread"]
        ctx.set_tuning(refit_blocks=rb, reduce_leaves_per_thread=lpt)
        ms = run(best["entry_level"], best["expand_levels"])
        if ms < best["ms"]:
            best = {**best, "refit_blocks": rb, "reduce_leaves_per_thread": lpt, "ms": ms}
    ctx.set_tuning(refit_blocks=best["refit_blocks"], reduce_leaves_per_thread=best["reduce_leaves_per_thread"])
    # the reduce's subtree size again, now with the detect's shape it runs
    # beside (it was first picked with the shape of entries[0], expands[0])
    if len(leaves) > 1 and not best["refit_blocks"]:
        ms0 = run(best["entry_level"], best["expand_levels"])
        best = {**best, "ms": min(best["ms"], ms0)}
        for lpt in leaves:
            if lpt == best["reduce_leaves_per_thread"]:
                continue
            ctx.set_tuning(reduce_leaves_per_thread=lpt)
            ms = run(best["entry_level"], best["expand_levels"])
            if ms < best["ms"]:
                best = {**best, "reduce_leaves_per_thread": lpt, "ms": ms}
        ctx.set_tuning(reduce_leaves_per_thread=best["reduce_leaves_per_thread"])
    ctx.set_tuning(**{**base, "reduce_leaves_per_thread": best["reduce_leaves_per_thread"],
                      "traverse_blocks": best["traverse_blocks"], "traverse_expand": best["expand_levels"],
                      "traverse_schedule": best["traverse_schedule"],
                      "traverse_local": best["traverse_local"],
                      "traverse_local_levels": best["traverse_local_levels"],
                      "traverse_tail": best["traverse_tail"], "traverse_tail_steps": best["traverse_tail_steps"],
                      "narrow_separate": best["narrow_separate"], "refit_blocks": best["refit_blocks"]})
    return best


def tune_overlapped(ctx: api.Context, scene: api.Scene, run, candidates, *, frames: int = 40,
                    reps: int = 2) -> dict:
    """Pick the detect's shape by the rate of overlapped frames (a detect
    that is fastest alone can lose beside the refit of the next frame).
    run(i0, n, entry, expand) queues n overlapped frames from frame index
    i0; `candidates` are dicts of 'entry_level', 'expand_levels' and
    oibvh_tuning fields (the entry level and levels per step go to
    traverse_expand). Sets the best on ctx and returns it with its 'ms'."""
    timer = api.Timer(ctx)
    base = ctx.get_tuning()
    best = None
    for cand in candidates:
        knobs = {k: v for k, v in cand.items() if k not in ("entry_level", "expand_levels", "ms")}
        e, x = cand["entry_level"], cand["expand_levels"]
        ctx.set_tuning(**{**base, **knobs, "traverse_expand": x})
        run(0, 4, e, x)
        scene.getIntTriPairCount()
        ms = None
        for r in range(reps):
            ctx.synchronize()
            timer.start()
            run(4 + r * frames, frames, e, x)
            ctx.synchronize()
            timer.stop()
            t = timer.ms() / frames
            ms = t if ms is None else min(ms, t)
        scene.getIntTriPairCount()
        if best is None or ms < best["ms"]:
            best = {**cand, "ms": ms}
    knobs = {k: v for k, v in best.items() if k not in ("entry_level", "expand_levels", "ms")}
    ctx.set_tuning(**{**base, **knobs, "traverse_expand": best["expand_levels"]})
    return best


def tune_refit_blocks(ctx: api.Context, scene: api.Scene, run, *, blocks=(0, 384, 448, 512, 640),
                      frames: int = 40, reps: int = 2) -> dict:
    """Pick refit_blocks (the cap on the scene refit's k_reduce_w workgroups)
    by the rate of overlapped frames: run(i0, n) queues n overlapped frames
    starting at frame index i0. A capped refit leaves CU slots to the detect
    running beside it; the candidates are 1.5, 1.75, 2 and 2.5 workgroups per
    CU of an MI355X (256 CUs) and no cap. Sets the best on ctx and returns
    {'refit_blocks', 'ms', 'tried'}."""
    timer = api.Timer(ctx)
    tried = {}
    for nb in blocks:
        ctx.set_tuning(refit_blocks=nb)
        run(0, 4)
        scene.getIntTriPairCount()
        best = None
        for r in range(reps):
            ctx.synchronize()
            timer.start()
            run(4 + r * frames, frames)
            ctx.synchronize()
            timer.stop()
            ms = timer.ms() / frames
            best = ms if best is None else min(best, ms)
        scene.getIntTriPairCount()
        tried[nb] = best
    nb = min(tried, key=tried.get)
    ctx.set_tuning(refit_blocks=nb)
    return {"refit_blocks": nb, "ms": tried[nb], "tried": tried}

