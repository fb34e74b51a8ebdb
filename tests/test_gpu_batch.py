"""GPU parity of the batched form (oibvh_batch_*): every mesh's tree equals
the oracle's tree over the same transformed vertices, and every scene's pair
set equals the oracle's detect over that scene's meshes (integer outputs:
exact). Graph replays are checked replay by replay against eager runs."""
import sys

import numpy as np
import pytest

from oibvh_amd import meshgen as M

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    from oibvh_amd import api as A

    return A


def small_rigid(n_scenes=6, seed=0x5EED, meshes=(2, 4), faces=(20, 6000)):
    return M.RigidScenes(n_scenes=n_scenes, seed=seed, meshes=meshes, faces=faces, radius=0.3)


def make_batch(api, ctx, rs, instances=True):
    b = api.Batch(ctx)
    first = {}
    for scene, lib in rs.instances:
        if instances and lib in first:
            b.add_instance(scene, first[lib])
        else:
            xyz, f = rs.library[lib]
            idx = b.add_mesh(scene, xyz, f)
            first.setdefault(lib, idx)
    return b


def oracle_scene_pairs(orc, rs, mats, scene, fp64=False, self_collide=False, perms=None):
    """Oracle detect over one scene's meshes (batch order); rows mapped to
    batch mesh indices."""
    ids = [i for i, (s, _) in enumerate(rs.instances) if s == scene]
    trees = []
    for i in ids:
        xyz, f = rs.library[rs.instances[i][1]]
        y = orc.xform(mats[i], xyz)
        trees.append(orc.Tree(y, f, order="morton" if perms is None else perms[i]))
    if not trees:
        return np.zeros((0, 4), np.uint32)
    p = orc.detect(trees, self_collide=self_collide, fp64=fp64)
    if len(p) == 0:
        return np.zeros((0, 4), np.uint32)
    p = p.copy()
    p[:, 0] = np.asarray(ids, np.uint32)[p[:, 0]]
    p[:, 1] = np.asarray(ids, np.uint32)[p[:, 1]]
    return p


def all_pairs(orc, rs, mats, **kw):
    rows = [oracle_scene_pairs(orc, rs, mats, s, **kw) for s in range(rs.n_scenes)]
    rows = [r for r in rows if len(r)]
    if not rows:
        return np.zeros((0, 4), np.uint32)
    p = np.concatenate(rows)
    return p[np.lexsort((p[:, 3], p[:, 2], p[:, 1], p[:, 0]))]


def test_batch_trees_and_pairs(api, ctx, orc):
    rs = small_rigid()
    say(f"batch: {len(rs.instances)} meshes in {rs.n_scenes} scenes, {rs.triangles} triangles")
    b = make_batch(api, ctx, rs)
    mats = rs.transforms(0)
    b.set_transforms(mats)
    b.build()
    for i, (s, lib) in enumerate(rs.instances):
        xyz, f = rs.library[lib]
        o = orc.Tree(orc.xform(mats[i], xyz), f, order="morton")
        assert np.array_equal(b.perm(i), o.order), f"mesh {i}: leaf permutation differs"
        assert np.array_equal(bits(b.nodes(i)), bits(o.nodes())), f"mesh {i}: nodes differ"
    for entry, expand in [(0, 1), (2, 2), (3, 3)]:
        b.detect(entry, expand)
        want = all_pairs(orc, rs, mats)
        got = b.pairs()
        assert len(want) > 0
        assert np.array_equal(got, want), f"entry {entry} expand {expand}"
    pv = b.pair_vertices()
    pairs = b.pairs()
    assert pv.shape == (len(pairs), 6, 3)
    for i, (ma, mb, ta, tb) in enumerate(pairs[:: max(1, len(pairs) // 200)]):
        k = i * max(1, len(pairs) // 200)
        xa, fa = rs.library[rs.instances[ma][1]]
        xb, fb = rs.library[rs.instances[mb][1]]
        ya, yb = orc.xform(mats[ma], xa), orc.xform(mats[mb], xb)
        assert np.array_equal(bits(pv[k, :3]), bits(ya[fa[ta]])) and np.array_equal(bits(pv[k, 3:]), bits(yb[fb[tb]]))
    # refit after motion
    for frame in (3, 9):
        mats = rs.transforms(frame)
        b.set_transforms(mats)
        b.refit()
        b.detect(2, 2)
        perms = [b.perm(i) for i in range(b.mesh_count)]
        want = all_pairs(orc, rs, mats, perms=perms)
        assert np.array_equal(b.pairs(), want), f"frame {frame}"
        for i in (0, b.mesh_count - 1):
            xyz, f = rs.library[rs.instances[i][1]]
            o = orc.Tree(orc.xform(mats[i], xyz), f, order=perms[i])
            assert np.array_equal(bits(b.nodes(i)), bits(o.nodes()))
    b.close()


def test_batch_without_instances_equals_with(api, ctx, orc):
    rs = small_rigid(n_scenes=4, seed=0x1234)
    mats = rs.transforms(2)
    out = []
    for inst in (True, False):
        b = make_batch(api, ctx, rs, instances=inst)
        b.set_transforms(mats)
        b.build()
        b.detect(1, 2)
        out.append(b.pairs())
        b.close()
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], all_pairs(orc, rs, mats))


def test_batch_many_meshes_two_mesh_passes(api, ctx, orc):
    """> 256 meshes: the mesh-id sort takes two 8-bit passes."""
    rs = M.RigidScenes(n_scenes=90, seed=0x77, meshes=(2, 5), faces=(12, 400), radius=0.35)
    assert len(rs.instances) > 256
    say(f"batch: {len(rs.instances)} meshes")
    b = make_batch(api, ctx, rs)
    mats = rs.transforms(1)
    b.set_transforms(mats)
    b.build()
    for i in range(0, len(rs.instances), 37):
        xyz, f = rs.library[rs.instances[i][1]]
        o = orc.Tree(orc.xform(mats[i], xyz), f, order="morton")
        assert np.array_equal(b.perm(i), o.order)
        assert np.array_equal(bits(b.nodes(i)), bits(o.nodes()))
    b.detect(0, 2)
    assert np.array_equal(b.pairs(), all_pairs(orc, rs, mats))
    b.close()


def test_batch_morton63(api, ctx, orc):
    """63-bit keys in a batch: every tree equals the oracle's over the 64-bit
    stable order; the pair sets equal the 30-bit batch's."""
    rs = small_rigid(n_scenes=5, seed=0x6363)
    mats = rs.transforms(4)
    out = []
    for bits_ in (63, 30):
        say(f"batch {bits_}-bit Morton: {len(rs.instances)} meshes")
        b = make_batch(api, ctx, rs)
        b.set_morton_bits(bits_)
        b.set_transforms(mats)
        b.build()
        order = "morton63" if bits_ == 63 else "morton"
        for i, (s, lib) in enumerate(rs.instances):
            xyz, f = rs.library[lib]
            o = orc.Tree(orc.xform(mats[i], xyz), f, order=order)
            assert np.array_equal(b.perm(i), o.order), f"mesh {i}: leaf permutation differs"
            assert np.array_equal(bits(b.nodes(i)), bits(o.nodes())), f"mesh {i}: nodes differ"
        b.detect(1, 3)
        out.append(b.pairs())
        b.close()
    assert len(out[0]) > 0
    assert np.array_equal(out[0], out[1])
    assert np.array_equal(out[0], all_pairs(orc, rs, mats))


@pytest.mark.parametrize("fp64,levels,tail", [(False, 0, 0), (True, 0, 0), (False, 2, 0), (False, 1, 16),
                                              (True, 1, 16)])
def test_batch_local_walk(api, ctx, orc, fp64, levels, tail):
    """The workgroup-local walk (traverse_local = n) over a batch front, FP32
    fused and FP64-fallback, alone or followed by a tail walk (traverse_tail
    = n) for the steps after it: the same pairs as the oracle, eager and
    from a captured frame."""
    rs = small_rigid(n_scenes=6, seed=0x10CA)
    mats = rs.transforms(2)
    want = all_pairs(orc, rs, mats, fp64=fp64)
    assert len(want) > 0
    b = make_batch(api, ctx, rs)
    b.set_transforms(mats)
    b.build()
    ctx.set_tuning(traverse_local=32, traverse_local_levels=levels, traverse_tail=tail)
    try:
        say(f"batch local walk fp64={fp64} levels={levels} tail={tail}")
        b.detect(1, 2, fp64_fallback=fp64)
        assert np.array_equal(b.pairs(), want)
        b.frame(1, 2, fp64_fallback=fp64)
        assert np.array_equal(b.pairs(), want)
        assert b.stats()["retries"] == 0
    finally:
        ctx.set_tuning(traverse_local=0, traverse_local_levels=0, traverse_tail=0)
    b.close()


def test_batch_tiny_and_mixed_meshes(api, ctx, orc):
    """Meshes of 1, 2, 3 and 5 triangles next to a few thousand, scenes of one
    mesh and an empty scene index in between: every tree equals the oracle's
    and the pairs equal the oracle's detect, scene by scene (self mode too)."""
    r = M.Rng(0x7171)

    def blob(F, c, s):
        V = F + 2
        x = (c + s * r.uniform(V * 3, -1.0, 1.0).reshape(V, 3)).astype(np.float32)
        f = r.integers(F * 3, 0, V - 1).reshape(F, 3).astype(np.uint32)
        return x, f

    scenes = [
        [blob(1, 0.0, 0.5), blob(2, 0.1, 0.5), blob(3000, 0.0, 0.6)],
        [blob(5, 0.0, 0.4)],                     # one mesh: no pairs unless self mode
        [],                                      # scene 2 holds no mesh
        [blob(3, 0.2, 0.3), blob(1, 0.2, 0.3), blob(7, 0.25, 0.3), blob(2, 0.9, 0.1)],
    ]
    b = api.Batch(ctx)
    ids = []
    for si, ms in enumerate(scenes):
        for x, f in ms:
            ids.append((si, b.add_mesh(si, x, f), x, f))
    b.build()
    for si, idx, x, f in ids:
        o = orc.Tree(x, f, order="morton")
        assert np.array_equal(b.perm(idx), o.order), idx
        assert np.array_equal(bits(b.nodes(idx)), bits(o.nodes())), idx
    for self_c in (False, True):
        say(f"batch tiny meshes self={self_c}")
        want = []
        for si, ms in enumerate(scenes):
            sel = [(idx, x, f) for s_, idx, x, f in ids if s_ == si]
            if not sel:
                continue
            p = orc.detect([orc.Tree(x, f) for _, x, f in sel], self_collide=self_c)
            if len(p):
                p = p.copy()
                gid = np.asarray([idx for idx, _, _ in sel], np.uint32)
                p[:, 0] = gid[p[:, 0]]
                p[:, 1] = gid[p[:, 1]]
                want.append(p)
        want = np.concatenate(want) if want else np.zeros((0, 4), np.uint32)
        want = want[np.lexsort((want[:, 3], want[:, 2], want[:, 1], want[:, 0]))]
        for entry, expand in [(0, 1), (2, 3)]:
            b.detect(entry, expand, self_collide=self_c)
            assert np.array_equal(b.pairs(), want), (self_c, entry, expand)
    b.close()


def test_batch_frame_many_apex_meshes(api, ctx, orc):
    """Regression: a captured frame whose refit has more than eight apex work
    items (meshes taller than the stage-1 subtrees) carries the detect's
    counter reset in an extra apex workgroup; every mesh must still be climbed
    and no work item read past the list. FP32 and FP64-fallback frames equal
    the oracle."""
    rs = M.RigidScenes(n_scenes=6, seed=0xA9E7, meshes=(2, 4), faces=(2000, 6000), radius=0.3)
    b = make_batch(api, ctx, rs)
    mats = rs.transforms(1)
    b.set_transforms(mats)
    b.build()
    tall = sum(1 for _, lib in rs.instances if rs.library[lib][1].shape[0] > 1024)
    assert tall >= 9, tall
    for fp64 in (False, True):
        say(f"batch frame, {tall} apex meshes, fp64={fp64}")
        for frame in (2, 3):
            mats = rs.transforms(frame)
            b.set_transforms(mats)
            b.frame(1, 2, fp64_fallback=fp64)
            perms = [b.perm(i) for i in range(b.mesh_count)]
            assert np.array_equal(b.pairs(), all_pairs(orc, rs, mats, perms=perms, fp64=fp64)), (fp64, frame)
            for i in range(b.mesh_count):
                xyz, f = rs.library[rs.instances[i][1]]
                o = orc.Tree(orc.xform(mats[i], xyz), f, order=perms[i])
                assert np.array_equal(bits(b.nodes(i)), bits(o.nodes())), (fp64, frame, i)
    b.close()


@pytest.mark.parametrize("schedule", [0, 0b1010])
def test_batch_self_collision(api, ctx, orc, schedule):
    """Cloth sheets in self mode, as one batch, against the oracle; also with
    steps of 2 levels and a deeper uniform 3 after them (traverse_schedule),
    eager and from a captured overlapped frame."""
    say(f"batch: cloth self collision, schedule {schedule:#x}")
    cl = [M.cloth(24 + 6 * i, 0.2, 0.25, seed=40 + i) for i in range(3)]
    b = api.Batch(ctx)
    for s, (x, f) in enumerate(cl):
        b.add_mesh(s, x, f)
    b.build()
    expand = 3 if schedule else 2
    ctx.set_tuning(traverse_expand=expand if schedule else 0, traverse_schedule=schedule)
    try:
        b.detect(2, expand, self_collide=True)
        got = b.pairs()
        b.frame(2, expand, self_collide=True, overlap=True)
        b.frame(2, expand, self_collide=True, overlap=True)
        assert np.array_equal(b.pairs(), got)
    finally:
        ctx.set_tuning(traverse_expand=0, traverse_schedule=0)
    rows = []
    for s, (x, f) in enumerate(cl):
        p = orc.detect([orc.Tree(x, f, order="morton")], self_collide=True)
        if len(p):
            p = p.copy()
            p[:, 0] = s
            p[:, 1] = s
            rows.append(p)
    want = np.concatenate(rows)
    want = want[np.lexsort((want[:, 3], want[:, 2], want[:, 1], want[:, 0]))]
    assert len(want) > 0
    assert np.array_equal(got, want)
    assert b.stats()["adjacent"] > 0
    b.close()


def test_batch_precision_fp64_fallback(api, ctx, orc):
    """Coplanar / edge-touching scenes: FP32 filter + FP64 re-check equals the
    all-FP64 oracle exactly; the plain FP32 path equals the FP32 oracle."""
    scenes = M.precision_scenes(n_scenes=8, n_plane=24, n_box=10)
    b = api.Batch(ctx)
    for s, ms in enumerate(scenes):
        for x, f in ms:
            b.add_mesh(s, x, f)
    b.build()

    def want(fp64):
        rows = []
        for s, ms in enumerate(scenes):
            p = orc.detect([orc.Tree(x, f, order="morton") for x, f in ms], fp64=fp64)
            if len(p):
                p = p.copy()
                p[:, 0] = 2 * s + p[:, 0]
                p[:, 1] = 2 * s + p[:, 1]
                rows.append(p)
        w = np.concatenate(rows)
        return w[np.lexsort((w[:, 3], w[:, 2], w[:, 1], w[:, 0]))]

    b.detect(2, 2, fp64_fallback=True)
    w64 = want(True)
    assert np.array_equal(b.pairs(), w64)
    assert b.stats()["fp64_rechecks"] > 0
    b.detect(2, 2)
    w32 = want(False)
    assert np.array_equal(b.pairs(), w32)
    say(f"precision: {len(w64)} FP64 pairs, {len(w32)} FP32 pairs")
    b.close()


def test_batch_frame_graph_replays(api, ctx, orc):
    """refit + detect replayed from a captured graph: new transforms between
    replays (in stream order), unrelated eager work and synchronisations in
    between, every replay checked against an eager refit + detect."""
    rs = small_rigid(n_scenes=5, seed=0xAB)
    b = make_batch(api, ctx, rs)
    ref = make_batch(api, ctx, rs)
    b.set_transforms(rs.transforms(0))
    ref.set_transforms(rs.transforms(0))
    b.build()
    ref.build()
    side = api.Scene(ctx)
    (x0, f0), (x1, f1) = M.interpenetrating_icospheres(subdiv=2, offset=1.0)
    t0, t1 = api.OibvhTree(ctx, x0, f0), api.OibvhTree(ctx, x1, f1)
    t0.build()
    t1.build()
    side.addOibvhTree(t0)
    side.addOibvhTree(t1)
    side_want = orc.detect([orc.Tree(x0, f0, order="morton"), orc.Tree(x1, f1, order="morton")])
    for frame in range(1, 7):
        mats = rs.transforms(frame)
        b.set_transforms(mats)
        b.frame(2, 2)
        got = b.pairs()
        ref.set_transforms(mats)
        ref.refit()
        ref.detect(2, 2)
        assert np.array_equal(got, ref.pairs()), f"replay {frame}"
        for i in (0, b.mesh_count // 2):
            assert np.array_equal(bits(b.nodes(i)), bits(ref.nodes(i)))
        side.detectCollision(1, 1)
        assert np.array_equal(side.getIntTriPairs(), side_want)
        ctx.synchronize()
    perms = [b.perm(i) for i in range(b.mesh_count)]
    assert np.array_equal(got, all_pairs(orc, rs, mats, perms=perms))
    b.close()
    ref.close()


def test_batch_overflow_retry(api, orc):
    say("batch: tiny detect buffers")
    rs = small_rigid(n_scenes=4, seed=0x99)
    ctx2 = api.Context(0)
    ctx2.set_tuning(front_capacity=256, hits_capacity=16, candidate_capacity=64)
    b = make_batch(api, ctx2, rs)
    mats = rs.transforms(0)
    b.set_transforms(mats)
    b.build()
    b.frame(1, 1)
    got = b.pairs()
    assert b.stats()["retries"] > 0
    assert np.array_equal(got, all_pairs(orc, rs, mats))
    b.frame(1, 1)  # captured again over the grown buffers
    assert np.array_equal(b.pairs(), got)
    b.close()
    ctx2.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_batch_overlapped_frames(api, ctx, orc, fp64):
    """OIBVH_FRAME_OVERLAP: frame f's refit runs beside frame f - 1's detect,
    each half a captured graph per slot; matrices staged from HBM after each
    frame. Every checked frame equals an eager refit + detect of the same
    matrices on a second batch: synchronous overlapped frames, a run of
    asynchronous ones read at the end, non-overlapped calls in between (with
    inputs staged for an overlapped frame that never ran), and overlapped
    frames again."""
    say(f"batch: overlapped frames (fp64={fp64})")
    rs = small_rigid(n_scenes=7, seed=0xA7)
    b = make_batch(api, ctx, rs)
    ref = make_batch(api, ctx, rs)
    nf = 12
    mats = np.stack([rs.transforms(f) for f in range(nf)])
    buf = api.DeviceBuffer(mats.nbytes)
    buf.upload(np.ascontiguousarray(np.swapaxes(mats, -1, -2)))
    stride = mats.shape[1] * 64
    b.set_transforms(mats[0])
    ref.set_transforms(mats[0])
    b.build()
    ref.build()

    def want(f):
        ref.set_transforms(mats[f])
        ref.refit()
        ref.detect(2, 2, fp64_fallback=fp64)
        return ref.pairs()

    def check(f, msg):
        got = b.pairs()
        assert np.array_equal(got, want(f)), msg
        for i in (0, b.mesh_count - 1):
            assert np.array_equal(bits(b.nodes(i)), bits(ref.nodes(i))), msg

    # synchronous overlapped frames
    for f in range(0, 4):
        b.set_transforms_device(buf.at(f * stride))
        b.frame(2, 2, fp64_fallback=fp64, overlap=True)
        check(f, f"sync overlapped frame {f}")
    # asynchronous run, read at the end
    for f in range(4, 9):
        b.set_transforms_device(buf.at(f * stride))
        b.frame(2, 2, fp64_fallback=fp64, overlap=True, asynchronous=True)
    st = b.stats()
    assert st["async_errors"] == 0 and st["retries"] == 0
    check(8, "async overlapped run")
    # inputs staged for an overlapped frame, then a plain refit + detect
    b.set_transforms_device(buf.at(9 * stride))
    b.refit()
    b.detect(2, 2, fp64_fallback=fp64)
    check(9, "refit + detect after staging")
    b.frame(2, 2, fp64_fallback=fp64)  # non-overlapped graph frame, same matrices
    check(9, "plain frame")
    for f in (10, 11):
        b.set_transforms_device(buf.at(f * stride))
        b.frame(2, 2, fp64_fallback=fp64, overlap=True, asynchronous=f == 10)
    check(11, "overlapped again")
    perms = [b.perm(i) for i in range(b.mesh_count)]
    assert np.array_equal(b.pairs(), all_pairs(orc, rs, mats[11], perms=perms, fp64=fp64))
    b.close()
    ref.close()
    buf.free()


def test_batch_overlapped_positions(api, ctx, orc):
    """Overlapped frames of deforming meshes: each frame points every mesh
    at its own vertex memory (set_positions_device, borrowed), self mode;
    frames that stage only some meshes keep the others' last positions. Each
    frame equals the oracle."""
    say("batch: overlapped frames, borrowed positions")
    nv = 40
    sheets = [M.cloth(nv, 0.2, 0.25, 0x77 + i) for i in range(3)]
    frames = 5
    pos = [[M.cloth_positions(nv, 0.2, 0.25, 0x77 + i, phase=0.2 * f) for f in range(frames)] for i in range(3)]
    bufs = [[api.DeviceBuffer.from_array(p) for p in row] for row in pos]
    b = api.Batch(ctx)
    for i, (x, f) in enumerate(sheets):
        b.add_mesh(i, x, f)
    b.build()
    cur = [0, 0, 0]
    for f in range(frames):
        movers = range(3) if f % 2 == 0 else (1,)  # odd frames move sheet 1 only
        for i in movers:
            b.set_positions_device(i, bufs[i][f].at(0))
            cur[i] = f
        b.frame(2, 2, self_collide=True, overlap=True)
        got = b.pairs()
        rows = []
        for i, (x, fc) in enumerate(sheets):
            p = orc.detect([orc.Tree(pos[i][cur[i]], fc, order=b.perm(i))], self_collide=True)
            if len(p):
                p = p.copy()
                p[:, :2] = i
                rows.append(p)
        want = np.concatenate(rows)
        want = want[np.lexsort((want[:, 3], want[:, 2], want[:, 1], want[:, 0]))]
        assert len(want) > 0
        assert np.array_equal(got, want), f"frame {f}"
    b.close()
    for row in bufs:
        for x in row:
            x.free()


def test_batch_overlapped_overflow(api, orc):
    """Overlapped frames whose detect buffers overflow: a synchronous frame
    re-runs its detect over the grown buffers; an asynchronous run reports
    the overflow of frames it could not re-run as async_errors, and later
    frames are complete again."""
    say("batch: overlapped frames, tiny detect buffers")
    rs = small_rigid(n_scenes=4, seed=0x9A)
    ctx2 = api.Context(0)
    ctx2.set_tuning(front_capacity=256, hits_capacity=16, candidate_capacity=64)
    b = make_batch(api, ctx2, rs)
    b.set_transforms(rs.transforms(0))
    b.build()
    b.set_transforms(rs.transforms(1))
    b.frame(1, 1, overlap=True)
    assert b.stats()["retries"] > 0
    perms = [b.perm(i) for i in range(b.mesh_count)]
    assert np.array_equal(b.pairs(), all_pairs(orc, rs, rs.transforms(1), perms=perms))
    for f in (2, 3, 4):
        b.set_transforms(rs.transforms(f))
        b.frame(1, 1, overlap=True)
        assert np.array_equal(b.pairs(), all_pairs(orc, rs, rs.transforms(f), perms=perms)), f"frame {f}"
    b.close()
    ctx2.close()


def test_batch_timing(api, ctx):
    """oibvh_batch_set_timing: sampled refits run eagerly with an event pair
    on their stage-1 launch, sampled detects with a pair around them; both in
    overlapped and plain frames."""
    rs = small_rigid(n_scenes=5, seed=0xB1)
    b = make_batch(api, ctx, rs)
    b.set_transforms(rs.transforms(0))
    b.build()
    b.set_timing(2)
    for f in range(6):
        b.set_transforms(rs.transforms(f))
        b.frame(1, 2, overlap=f >= 3, asynchronous=True)
    b.refit()
    t = b.timing()
    assert t["refits"] == 4 and t["detects"] == 3
    assert t["refit_ms"] > 0 and t["detect_ms"] > 0
    b.set_timing(0)
    b.close()


def _scene_rows(p, lo, hi):
    if not len(p):
        return p
    a, b = np.searchsorted(p[:, 0], lo), np.searchsorted(p[:, 0], hi)
    return p[a:b]


def test_precision_full_size_vs_oracles(api, ctx, orc):
    """The precision workload at full size: 64 scenes of two ~250k-triangle
    planes or boxes, coplanar or edge-touching. On every scene where the FP32
    and the FP32 + FP64-fallback pair sets differ, and on one scene of each
    kind, the fallback's pairs equal the FP64 oracle's and the FP32 pairs the
    FP32 oracle's (the reference CPU path), at full size."""
    say("precision scenes at full size")
    scenes = M.precision_scenes()
    b = api.Batch(ctx)
    for si, ms in enumerate(scenes):
        for x, f in ms:
            b.add_mesh(si, x, f)
    b.build()
    b.detect(4, 2, fp64_fallback=True)
    p64 = b.pairs()
    assert b.stats()["fp64_rechecks"] > 0
    b.detect(4, 2)
    p32 = b.pairs()
    changed = [si for si in range(len(scenes))
               if not np.array_equal(_scene_rows(p32, 2 * si, 2 * si + 2), _scene_rows(p64, 2 * si, 2 * si + 2))]
    assert changed, "the workload should hold scenes where FP32 and FP64 disagree"
    for si in sorted(set(changed) | {0, 1, 2, 3}):
        say(f"  scene {si} ({'FP32 != FP64' if si in changed else 'kind ' + str(si % 4)})")
        ot = [orc.Tree(x, f, order=b.perm(2 * si + j)) for j, (x, f) in enumerate(scenes[si])]
        for fp64, got in ((True, p64), (False, p32)):
            want = orc.detect(ot, fp64=fp64)
            if len(want):
                want = want.copy()
                want[:, :2] += 2 * si
            assert np.array_equal(_scene_rows(got, 2 * si, 2 * si + 2), want), (si, fp64)
    b.close()
