"""GPU parity: the HIP path through the C ABI against the oracle.

Bar: bit-exact tree AABBs, identical leaf permutation and sorted faces, and
identical sorted colliding-pair sets (integer outputs, exact)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oibvh_amd import meshgen as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def api():
    from oibvh_amd import api as A

    return A


def gpu_tree(api, ctx, xyz, f):
    t = api.OibvhTree(ctx, xyz, f)
    t.build()
    return t


def check_tree(orc, t, xyz, f, order=None):
    o = orc.Tree(xyz, f, order="morton" if order is None else order)
    assert np.array_equal(t.perm(), o.order), "leaf permutation differs"
    assert np.array_equal(t.sorted_faces(), f[o.order]), "sorted faces differ"
    a, b = t.nodes(), o.nodes()
    assert a.shape == b.shape
    bad = np.nonzero((bits(a) != bits(b)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} nodes differ, first at real index {bad[:5]}"
    return o


def soup(F, seed):
    r = M.Rng(seed)
    V = F + 2
    xyz = r.uniform(V * 3, -1.0, 1.0).reshape(V, 3).astype(np.float32)
    f = r.integers(F * 3, 0, V - 1).reshape(F, 3).astype(np.uint32)
    return xyz, f


# --------------------------------------------------------------------- build
@pytest.mark.parametrize("F", [1, 2, 3, 4, 5, 7, 8, 9, 100, 255, 256, 257, 1000, 2047, 2048, 2049, 4096, 4097,
                               65537, 300001])
def test_build_soup_bit_exact(api, ctx, orc, F):
    say(f"build soup F={F}")
    xyz, f = soup(F, 1000 + F)
    check_tree(orc, gpu_tree(api, ctx, xyz, f), xyz, f)


@pytest.mark.parametrize("lpt", [1, 2, 4, 8, 16])
def test_build_subtree_sizes(api, ctx, orc, lpt):
    say(f"build leaves_per_thread={lpt}")
    ctx.set_tuning(reduce_leaves_per_thread=lpt)
    try:
        for xyz, f in [M.icosphere(5, 1.0, 0.02, 3), M.uv_sphere(300, 151), soup(12345, 7)]:
            check_tree(orc, gpu_tree(api, ctx, xyz, f), xyz, f)
    finally:
        ctx.set_tuning(reduce_leaves_per_thread=0)


def test_wave_reduce_all_heights(api, ctx, orc):
    """The wave-subtree reduce (tuning 16) across tree heights 0..13: wave
    subtrees shorter than 8 levels, one or two merged levels, partial waves,
    build, refit and a scene refit of mixed sizes."""
    ctx.set_tuning(reduce_leaves_per_thread=16)
    try:
        trees = []
        for F in [1, 2, 3, 5, 100, 255, 256, 257, 511, 513, 1023, 1024, 1025, 2049, 4097, 8193]:
            say(f"wave reduce F={F}")
            xyz, f = soup(F, 3000 + F)
            t = gpu_tree(api, ctx, xyz, f)
            check_tree(orc, t, xyz, f)
            x2 = (xyz * np.float32(0.75) + np.float32(0.125)).astype(np.float32)
            t.refit(x2)
            assert np.array_equal(bits(t.nodes()), bits(orc.Tree(x2, f, order=t.perm()).nodes()))
            trees.append((t, x2, f))
        sc = api.Scene(ctx)
        for t, _, _ in trees[-4:]:
            sc.addOibvhTree(t)
        for t, x2, f in trees[-4:]:
            t.set_positions(x2 * np.float32(1.5))
        sc.refit()
        for t, x2, f in trees[-4:]:
            x3 = (x2 * np.float32(1.5)).astype(np.float32)
            assert np.array_equal(bits(t.nodes()), bits(orc.Tree(x3, f, order=t.perm()).nodes()))
    finally:
        ctx.set_tuning(reduce_leaves_per_thread=0)


def test_build_flat_axis_morton_nan_policy(api, ctx, orc):
    """A plane has a zero-extent axis: offset/length is 0/0 = NaN there, which
    both sides quantise to 0."""
    say("build flat plane")
    xyz, f = M.plane(50)
    check_tree(orc, gpu_tree(api, ctx, xyz, f), xyz, f)


def test_build_cfg1_full_size(api, ctx, orc):
    say("build two 81,920-face icospheres")
    for xyz, f in M.interpenetrating_icospheres(subdiv=6):
        check_tree(orc, gpu_tree(api, ctx, xyz, f), xyz, f)


def test_build_is_deterministic(api, ctx):
    say("build determinism")
    xyz, f = M.icosphere(6, 1.0, 0.02, 1)
    t = gpu_tree(api, ctx, xyz, f)
    a = t.nodes().copy()
    p = t.perm().copy()
    for _ in range(3):
        t.build()
        assert np.array_equal(bits(t.nodes()), bits(a)) and np.array_equal(t.perm(), p)


def gpu_tree63(api, ctx, xyz, f):
    t = api.OibvhTree(ctx, xyz, f)
    t.set_morton_bits(63)
    t.build()
    return t


def cell_cluster(F, seed):
    """Two unreferenced corner vertices fix the Morton box at [-1, 1]^3; every
    centroid lies in [0.5, 0.5 + 2^-12]^3, one 30-bit cell (all 30-bit keys
    equal) but about 2^8 cells per axis at 21 bits."""
    r = M.Rng(seed)
    V = F + 2
    xyz = (0.5 + r.uniform(V * 3, 0.0, 1.0 / 4096)).reshape(V, 3).astype(np.float32)
    xyz[0] = -1.0
    xyz[1] = 1.0
    f = r.integers(F * 3, 2, V - 1).reshape(F, 3).astype(np.uint32)
    return xyz, f


@pytest.mark.parametrize("F", [1, 2, 5, 257, 4097, 65537, 300001])
def test_build_morton63_soup_bit_exact(api, ctx, orc, F):
    say(f"build 63-bit soup F={F}")
    xyz, f = soup(F, 2000 + F)
    check_tree(orc, gpu_tree63(api, ctx, xyz, f), xyz, f, order="morton63")


def test_build_morton63_splits_a_30bit_cell(api, ctx, orc):
    say("build 63-bit cell cluster")
    xyz, f = cell_cluster(20000, 5)
    t30 = gpu_tree(api, ctx, xyz, f)
    assert np.array_equal(t30.perm(), np.arange(len(f), dtype=np.uint32))  # one key: stable = face order
    t63 = gpu_tree63(api, ctx, xyz, f)
    check_tree(orc, t63, xyz, f, order="morton63")
    assert not np.array_equal(t63.perm(), t30.perm())
    # back to 30 bits on the same tree
    t63.set_morton_bits(30)
    t63.build()
    check_tree(orc, t63, xyz, f)


def test_morton63_flat_and_detect(api, ctx, orc):
    """63-bit trees: NaN policy on a flat axis, refit, and a detect whose pair
    set equals the 30-bit one."""
    say("63-bit plane + icospheres detect")
    xyz, f = M.plane(40)
    check_tree(orc, gpu_tree63(api, ctx, xyz, f), xyz, f, order="morton63")
    meshes = M.interpenetrating_icospheres(subdiv=5)
    trees = [gpu_tree63(api, ctx, x, fa) for x, fa in meshes]
    for t, (x, fa) in zip(trees, meshes):
        check_tree(orc, t, x, fa, order="morton63")
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    sc.detectCollision(2, 3)
    want = orc.detect([orc.Tree(x, fa) for x, fa in meshes])
    assert len(want) > 0
    assert np.array_equal(sc.getIntTriPairs(), want)
    x1 = (meshes[1][0] * np.float32(1.01)).astype(np.float32)
    trees[1].refit(x1)
    assert np.array_equal(bits(trees[1].nodes()), bits(orc.Tree(x1, meshes[1][1], order=trees[1].perm()).nodes()))


def expected_pair_vertices(pairs, meshes):
    """Scene::convertToVertexArray (scene.cu:69-93) from the oracle side:
    meshes[m] = (transformed xyz, faces)."""
    out = np.empty((len(pairs), 6, 3), np.float32)
    for i, (a, b, ta, tb) in enumerate(pairs):
        xa, fa = meshes[a]
        xb, fb = meshes[b]
        out[i, :3] = xa[fa[ta]]
        out[i, 3:] = xb[fb[tb]]
    return out


def test_pair_vertices_export(api, ctx, orc):
    """The vertex export gathers each pair's triangles through the model
    matrices, to host or device memory, bit for bit."""
    say("pair vertex export")
    (x0, f0), (x1, f1) = M.interpenetrating_icospheres(subdiv=4, offset=0.0)
    t0, t1 = api.OibvhTree(ctx, x0, f0), api.OibvhTree(ctx, x1, f1)
    A0, A1 = _motion(5, 0), _motion(6, 0)
    t0.set_transform(A0)
    t1.set_transform(A1)
    t0.build()
    t1.build()
    sc = api.Scene(ctx)
    sc.addOibvhTree(t0)
    sc.addOibvhTree(t1)
    sc.detectCollision(1, 2)
    pairs = sc.getIntTriPairs()
    assert len(pairs) > 0
    want = expected_pair_vertices(pairs, [(orc.xform(A0, x0), f0), (orc.xform(A1, x1), f1)])
    got = sc.pair_vertices()
    assert got.shape == want.shape and np.array_equal(bits(got), bits(want))
    buf = api.DeviceBuffer(want.nbytes)
    assert sc.pair_vertices_device(buf.at(0), want.size) == len(pairs)
    assert np.array_equal(bits(buf.download(np.float32, want.size)), bits(want.reshape(-1)))
    with pytest.raises(api.L.OibvhError):
        sc.pair_vertices_device(buf.at(0), want.size - 1)
    buf.free()
    # identity transform, self mode: the export reads the current positions
    t0.set_transform(None)
    t0.refit()
    sc2 = api.Scene(ctx)
    sc2.addOibvhTree(t0)
    sc2.detectCollision(0, 3, self_collide=True)
    p2 = sc2.getIntTriPairs()
    assert np.array_equal(bits(sc2.pair_vertices()), bits(expected_pair_vertices(p2, [(x0, f0)])))


# --------------------------------------------------------------------- refit
def test_refit_after_deformation(api, ctx, orc):
    say("refit deforming uv sphere")
    meshes, waves = M.deforming_uv_pair(200, 101)
    xyz, f = meshes[0]
    t = gpu_tree(api, ctx, xyz, f)
    perm = t.perm()
    for frame in (1, 7, 33):
        x = waves[0].frame(frame)
        t.refit(x)
        o = orc.Tree(x, f, order=perm)
        assert np.array_equal(bits(t.nodes()), bits(o.nodes()))


def test_deforming_pair_full_size_vs_oracle(api, ctx, orc):
    """The deforming workload at full size (two UV spheres of 1000 x 501,
    1,000,000 triangles each): both trees built by one scene build, refit to
    a later frame and rebuilt at a third, every tree's leaf order and node
    bits and the pair set of each state against the oracle (the reference's
    CPU path restated)."""
    say("deforming pair, full size")
    meshes, waves = M.deforming_uv_pair()
    trees = [api.OibvhTree(ctx, waves[m].frame(0), meshes[m][1]) for m in range(2)]
    for t in trees:
        t.build()
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    for step, frame in (("build", 3), ("refit", 37), ("rebuild", 50)):
        xs = [waves[m].frame(frame) for m in range(2)]
        for t, x in zip(trees, xs):
            t.set_positions(x)
        if step == "refit":
            sc.refit()
        else:
            sc.build()
        ot = []
        for t, x, (_, f) in zip(trees, xs, meshes):
            o = orc.Tree(x, f, order="morton") if step != "refit" else orc.Tree(x, f, order=t.perm())
            assert np.array_equal(t.perm(), o.order), step
            assert np.array_equal(bits(t.nodes()), bits(o.nodes())), step
            ot.append(o)
        sc.detectCollision(8, 2)
        want = orc.detect(ot)
        assert len(want) > 1000
        assert np.array_equal(sc.getIntTriPairs(), want), step
    sc.close()


def test_cloth_self_collision_vs_oracle(api, ctx, orc):
    """A folded cloth sheet of 512 x 512 vertices (522,242 triangles), the
    self-collision workload's shape at a quarter of its side: the self-mode
    pair set against the oracle, after the build and after a refit to the
    next phase of the height field."""
    say("cloth self-collision 512 x 512")
    x, f = M.cloth(512, amplitude=0.2, wavelength=0.05, seed=11)
    t = gpu_tree(api, ctx, x, f)
    sc = api.Scene(ctx)
    sc.addOibvhTree(t)
    for phase in (0.0, 0.02):
        xp = M.cloth_positions(512, amplitude=0.2, wavelength=0.05, seed=11, phase=phase)
        if phase:
            t.refit(xp)
        sc.detectCollision(4, 2, self_collide=True)
        want = orc.detect([orc.Tree(xp, f, order=t.perm())], self_collide=True)
        assert len(want) > 1000
        assert np.array_equal(sc.getIntTriPairs(), want), phase
    sc.close()


def test_refit_from_device_pointer(api, ctx, orc):
    say("refit from borrowed device positions")
    meshes, waves = M.deforming_uv_pair(120, 61)
    xyz, f = meshes[1]
    t = gpu_tree(api, ctx, xyz, f)
    perm = t.perm()
    x = waves[1].frame(5)
    buf = api.DeviceBuffer.from_array(x)
    t.refit_device(buf.at(0), borrow=True)
    assert np.array_equal(bits(t.nodes()), bits(orc.Tree(x, f, order=perm).nodes()))
    assert np.array_equal(bits(t.positions()), bits(x))
    t.refit(xyz)  # back to owned memory before the buffer goes away
    ctx.synchronize()
    buf.free()


def _motion(seed, k):
    r = M.Rng(seed * 97 + k)
    A = np.eye(4, dtype=np.float32)
    A[:3, :3] = M.random_rotation(r)
    A[:3, 3] = r.uniform(3, -0.5, 0.5)
    return A


def test_transform_build_refit_detect(api, ctx, orc):
    """Mesh::transform applied on the fly: build, refit and detect see
    mat4 @ (p, 1) in glm's order, bit for bit what the oracle computes from
    the transformed positions."""
    say("model matrices: build, refit, detect")
    (x0, f0), (x1, f1) = M.interpenetrating_icospheres(subdiv=4, offset=0.0)
    t0, t1 = api.OibvhTree(ctx, x0, f0), api.OibvhTree(ctx, x1, f1)
    A0, A1 = _motion(1, 0), _motion(2, 0)
    t0.set_transform(A0)
    t1.set_transform(A1)
    t0.build()
    t1.build()
    y0, y1 = orc.xform(A0, x0), orc.xform(A1, x1)
    check_tree(orc, t0, y0, f0)
    check_tree(orc, t1, y1, f1)
    sc = api.Scene(ctx)
    sc.addOibvhTree(t0)
    sc.addOibvhTree(t1)
    for k in range(1, 4):
        A0, A1 = _motion(1, k), _motion(2, k)
        t0.set_transform(A0)
        t1.set_transform(A1)
        t0.refit()
        t1.refit()
        y0, y1 = orc.xform(A0, x0), orc.xform(A1, x1)
        o0 = orc.Tree(y0, f0, order=t0.perm())
        o1 = orc.Tree(y1, f1, order=t1.perm())
        assert np.array_equal(bits(t0.nodes()), bits(o0.nodes()))
        assert np.array_equal(bits(t1.nodes()), bits(o1.nodes()))
        sc.detectCollision(2, 2)
        want = orc.detect([o0, o1])
        assert len(want) > 0
        assert np.array_equal(sc.getIntTriPairs(), want)
    t0.set_transform(None)
    t0.refit()
    assert np.array_equal(bits(t0.nodes()), bits(orc.Tree(x0, f0, order=t0.perm()).nodes()))


@pytest.mark.parametrize("kind", ["host", "device", "borrow"])
@pytest.mark.parametrize("n", [2, 6])
def test_scene_set_transforms(api, ctx, orc, kind, n):
    """oibvh_scene_set_transforms gives every tree its matrix in one call
    (host staging, one device copy, or matrices read in place from HBM):
    refit nodes and pairs equal the oracle's over the transformed positions,
    frame after frame, with scene frames (scenes above four trees take the
    device mesh table) and a plain refit + detect."""
    say(f"scene set_transforms ({kind}, {n} trees)")
    shapes = [M.icosphere(3, 1.0, 0.02, 20 + i) for i in range(n)]
    trees = [api.OibvhTree(ctx, x, f) for x, f in shapes]
    for t in trees:
        t.build()
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    frames = 4
    mats = np.stack([np.stack([_motion(30 + i, k) for i in range(n)]) for k in range(frames)])
    buf = api.DeviceBuffer(mats.nbytes)
    buf.upload(np.ascontiguousarray(np.swapaxes(mats, -1, -2)))  # column-major
    for k in range(frames):
        if kind == "host":
            sc.set_transforms(mats[k])
        else:
            sc.set_transforms(device=buf.at(k * n * 64), borrow=kind == "borrow")
        if k % 2:
            sc.frame(None, 1, 2)
        else:
            sc.refit()
            sc.detectCollision(1, 2)
        ot = [orc.Tree(orc.xform(mats[k][i], shapes[i][0]), shapes[i][1], order=trees[i].perm()) for i in range(n)]
        for t, o in zip(trees, ot):
            assert np.array_equal(bits(t.nodes()), bits(o.nodes()))
        want = orc.detect(ot)
        assert np.array_equal(sc.getIntTriPairs(), want)
    # a per-tree matrix afterwards replaces the scene's
    trees[0].set_transform(None)
    trees[0].refit()
    assert np.array_equal(bits(trees[0].nodes()), bits(orc.Tree(*shapes[0], order=trees[0].perm()).nodes()))
    sc.close()
    buf.free()


@pytest.mark.parametrize("n", [1, 3, 6])
def test_scene_refit_equals_tree_refits(api, ctx, orc, n):
    """oibvh_scene_refit (up to four trees per launch) gives every tree the
    nodes its own refit gives: mixed sizes (one without an apex stage), one
    tree with a model matrix."""
    say(f"scene refit over {n} trees")
    shapes = [M.icosphere(4, 1.0, 0.02, 5), soup(3000, 11), M.uv_sphere(200, 101), soup(700, 12),
              M.box(20, 0.5, 0.02, 3), M.icosphere(3, 1.0, 0.02, 6)][:n]
    sc = api.Scene(ctx)
    trees = []
    for k, (x, f) in enumerate(shapes):
        t = api.OibvhTree(ctx, x, f)
        t.build()
        sc.addOibvhTree(t)
        trees.append(t)
    if n > 1:
        trees[1].set_transform(_motion(7, 1))
    moved = []
    for k, (x, f) in enumerate(shapes):
        y = (x * np.float32(1.01) + np.float32(0.003 * k)).astype(np.float32)
        trees[k].set_positions(y)
        moved.append(y)
    sc.refit()
    for k, (x, f) in enumerate(shapes):
        y = moved[k]
        if n > 1 and k == 1:
            y = orc.xform(_motion(7, 1), y)
        o = orc.Tree(y, f, order=trees[k].perm())
        assert np.array_equal(bits(trees[k].nodes()), bits(o.nodes())), f"tree {k}"


@pytest.mark.parametrize("blocks", [8, 16, 40, 1 << 20])
def test_scene_refit_capped_workgroups(api, orc, blocks):
    """refit_blocks caps the scene refit's k_reduce_w launch; each workgroup
    then loops over subtrees of 1024 leaves (k_reduce_wl). Nodes stay those
    of the oracle's refit: two trees sharing a launch, sizes with a partial
    last subtree, one with a model matrix, in overlapped frames too."""
    say(f"scene refit, at most {blocks} workgroups")
    c = api.Context(0)
    c.set_tuning(reduce_leaves_per_thread=16, refit_blocks=blocks)
    shapes = [M.uv_sphere(200, 101), M.icosphere(5, 1.0, 0.02, 5), soup(3001, 13)]
    sc = api.Scene(c)
    trees = []
    for x, f in shapes:
        t = api.OibvhTree(c, x, f)
        t.build()
        sc.addOibvhTree(t)
        trees.append(t)
    trees[2].set_transform(_motion(3, 2))
    for rep in range(3):
        moved = [(x * np.float32(1.0 + 0.01 * rep) + np.float32(0.002 * k)).astype(np.float32)
                 for k, (x, f) in enumerate(shapes)]
        for t, y in zip(trees, moved):
            t.set_positions(y)
        if rep < 2:
            sc.refit()
        else:
            sc.frame(None, 3, 1, overlap=True)
            sc.getIntTriPairCount()
        for k, (x, f) in enumerate(shapes):
            y = orc.xform(_motion(3, 2), moved[k]) if k == 2 else moved[k]
            o = orc.Tree(y, f, order=trees[k].perm())
            assert np.array_equal(bits(trees[k].nodes()), bits(o.nodes())), f"rep {rep} tree {k}"


def test_scene_frame_and_timing(api, ctx, orc):
    """oibvh_scene_frame (borrow, refit, detect in one call) equals the
    separate calls, and the scene's event timing counts every refit and
    detect it bracketed."""
    say("scene frame + timing")
    meshes, waves = M.deforming_uv_pair(160, 81)
    trees = [gpu_tree(api, ctx, x, f) for x, f in meshes]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    frames = [[waves[m].frame(k) for m in range(2)] for k in (3, 8)]
    bufs = [[api.DeviceBuffer.from_array(x) for x in fr] for fr in frames]
    sc.set_timing(1)
    for k, fr in enumerate(frames):
        sc.frame([b.at(0) for b in bufs[k]], 2, 2)
        want = orc.detect([orc.Tree(fr[m], meshes[m][1], order=trees[m].perm()) for m in range(2)])
        assert np.array_equal(sc.getIntTriPairs(), want), k
        for m in range(2):
            o = orc.Tree(fr[m], meshes[m][1], order=trees[m].perm())
            assert np.array_equal(bits(trees[m].nodes()), bits(o.nodes()))
    sc.frame(None, 2, 2)  # keeps the last positions
    tm = sc.timing()
    assert tm["refits"] == 3 and tm["detects"] == 3 and tm["refit_ms"] > 0 and tm["detect_ms"] > 0
    assert sc.timing()["refits"] == 0
    sc.set_timing(2)  # every second one
    for _ in range(5):
        sc.frame(None, 2, 2)
    tm = sc.timing()
    assert tm["refits"] == 3 and tm["detects"] == 3
    sc.set_timing(0)
    for t, (x, f) in zip(trees, meshes):
        t.refit(x)  # owned memory again before the buffers go away
    ctx.synchronize()
    for bb in bufs:
        for b in bb:
            b.free()


@pytest.mark.parametrize("self_collide,fp64", [(False, False), (True, False), (False, True)])
def test_scene_frame_overlap(api, ctx, orc, self_collide, fp64):
    """Overlapped frames (OIBVH_FRAME_OVERLAP): frames queued back to back,
    each refit into the node buffer the previous frame's detect is not
    reading. The last queued frame's pairs and every tree's nodes equal the
    oracle's; so do the pairs of frames read one by one, a frame without the
    flag after them, the event timing, and calls that write the trees
    (refit from host memory, rebuild) between overlapped frames."""
    say(f"scene frame overlap self={self_collide} fp64={fp64}")
    meshes, waves = M.deforming_uv_pair(160, 81)
    trees = [gpu_tree(api, ctx, x, f) for x, f in meshes]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    ks = (3, 8, 13, 21, 34, 55, 89)
    frames = [[waves[m].frame(k) for m in range(2)] for k in ks]
    bufs = [[api.DeviceBuffer.from_array(x) for x in fr] for fr in frames]

    def want(k):
        fr = frames[k]
        return orc.detect([orc.Tree(fr[m], meshes[m][1], order=trees[m].perm()) for m in range(2)],
                          self_collide=self_collide, fp64=fp64)

    def check_nodes(k):
        for m in range(2):
            o = orc.Tree(frames[k][m], meshes[m][1], order=trees[m].perm())
            assert np.array_equal(bits(trees[m].nodes()), bits(o.nodes())), (k, m)

    sc.set_timing(1)
    for k in range(len(ks)):  # all queued, then the last one read
        sc.frame([b.at(0) for b in bufs[k]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, asynchronous=True,
                 overlap=True)
    assert np.array_equal(sc.getIntTriPairs(), want(len(ks) - 1))
    check_nodes(len(ks) - 1)
    tm = sc.timing()
    assert tm["refits"] == len(ks) and tm["detects"] == len(ks) and tm["detect_ms"] > 0
    sc.set_timing(0)
    for k in (2, 0, 5):  # read one by one (synchronous frames)
        sc.frame([b.at(0) for b in bufs[k]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, overlap=True)
        assert np.array_equal(sc.getIntTriPairs(), want(k)), k
        check_nodes(k)
    for k in (1, 4):  # queued, then a frame without the flag
        sc.frame([b.at(0) for b in bufs[k]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, asynchronous=True,
                 overlap=True)
    sc.frame([b.at(0) for b in bufs[6]], 3, 2, self_collide=self_collide, fp64_fallback=fp64)
    assert np.array_equal(sc.getIntTriPairs(), want(6))
    check_nodes(6)
    # a host refit and a rebuild between overlapped frames
    sc.frame([b.at(0) for b in bufs[3]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, asynchronous=True,
                 overlap=True)
    for t, x in zip(trees, frames[2]):
        t.refit(x)
    check_nodes(2)
    sc.frame([b.at(0) for b in bufs[5]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, asynchronous=True,
                 overlap=True)
    for t in trees:
        t.build()
    check_nodes(5)
    for k in (4, 1):
        sc.frame([b.at(0) for b in bufs[k]], 3, 2, self_collide=self_collide, fp64_fallback=fp64, asynchronous=True,
                 overlap=True)
    assert np.array_equal(sc.getIntTriPairs(), want(1))
    check_nodes(1)
    for t, (x, f) in zip(trees, meshes):
        t.refit(x)  # owned memory again before the buffers go away
    ctx.synchronize()
    for bb in bufs:
        for b in bb:
            b.free()


@pytest.mark.parametrize("overlap,cap", [(False, "hits"), (True, "hits"), (True, "front")])
def test_async_frames_report_earlier_errors(api, ctx, orc, overlap, cap):
    """Frames queued without a host wait whose hit lists (or BVTT fronts)
    overflow: reading the last one re-runs it with a larger buffer (the
    oracle's pairs) and reports the overflow of the unchecked frames before
    it in async_errors, with overlapped frames (a counter set and working
    buffers per frame in flight, detects on alternating streams) as without.
    Errors are reported once: clean asynchronous frames after that check
    report none, so every set's buffers have grown to the re-run's size."""
    say(f"async frame errors overlap={overlap} cap={cap}")
    meshes = M.interpenetrating_icospheres(subdiv=4, seed=27, offset=1.2)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    assert len(want) > 16
    ctx2 = api.Context(0)
    if cap == "hits":
        ctx2.set_tuning(hits_capacity=16)
        bit = 8  # hit-list overflow
    else:
        ctx2.set_tuning(front_capacity=64)
        bit = 1  # BVTT front overflow
    try:
        trees = [gpu_tree(api, ctx2, x, f) for x, f in meshes]
        sc = api.Scene(ctx2)
        for t in trees:
            sc.addOibvhTree(t)
        for _ in range(3):
            sc.frame(None, 2, 2, asynchronous=True, overlap=overlap)
        assert np.array_equal(sc.getIntTriPairs(), want)
        st = sc.stats()
        assert st["retries"] >= 1
        assert st["async_errors"] & bit  # the overflow of the earlier frames
        for _ in range(5):  # the grown buffer holds these frames' hits
            sc.frame(None, 2, 2, asynchronous=True, overlap=overlap)
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["async_errors"] == 0 and sc.stats()["retries"] == 0
        for _ in range(3):
            sc.frame(None, 2, 2, asynchronous=True, overlap=overlap)
        sc.frame(None, 2, 2, overlap=overlap)  # a checked frame starts afresh
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["async_errors"] == 0 and sc.stats()["retries"] == 0
        sc.close()
    finally:
        ctx2.close()


def test_scene_edge_cases(api, ctx, orc):
    """An empty scene, a one-tree scene without and with self mode, and the
    frame call on them: no pairs where the reference has none."""
    say("scene edge cases")
    sc = api.Scene(ctx)
    sc.detectCollision(0, 1)
    assert sc.getIntTriPairCount() == 0 and sc.getIntTriPairs().shape == (0, 4)
    sc.frame(None, 0, 1)
    assert sc.getIntTriPairCount() == 0
    x, f = M.cloth(24, amplitude=0.2, wavelength=0.2, seed=9)
    t = gpu_tree(api, ctx, x, f)
    sc.addOibvhTree(t)
    sc.frame(None, 1, 2)
    assert sc.getIntTriPairCount() == 0
    sc.frame(None, 1, 2, self_collide=True)
    assert np.array_equal(sc.getIntTriPairs(), orc.detect([orc.Tree(x, f)], self_collide=True))
    sc.close()


def test_create_like_then_refit(api, ctx, orc):
    """main.cpp:127-151: tree 2 copy-constructed from tree 1 over a
    translated mesh with the same faces, then refit."""
    say("copy-constructed tree")
    xyz, f = M.icosphere(4, 1.0, 0.02, 2)
    t1 = gpu_tree(api, ctx, xyz, f)
    x2 = (xyz + np.float32([1.0, 0.0, 0.0])).astype(np.float32)
    t2 = api.OibvhTree.like(t1, x2)
    t2.refit()
    o = orc.Tree(x2, f, order=t1.perm())
    assert np.array_equal(bits(t2.nodes()), bits(o.nodes()))
    assert t2.getPrimCount() == f.shape[0] and t2.getDepth() == t1.getDepth()


# --------------------------------------------------------------------- detect
def gpu_pairs(api, ctx, meshes, entry=0, expand=1, self_collide=False, fp64=False):
    trees = [gpu_tree(api, ctx, x, f) for x, f in meshes]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    sc.detectCollision(entry, expand, self_collide=self_collide, fp64_fallback=fp64)
    return sc.getIntTriPairs(), sc.stats(), trees


def test_detect_golden_scenes(api, ctx):
    g = np.load(os.path.join(GOLD, "scenes.npz"))
    for name in g["names"]:
        say(f"golden scene {name}")
        n = int(g[f"{name}_n"][0])
        self_c = bool(g[f"{name}_self"][0])
        meshes = [(g[f"{name}_xyz{k}"], g[f"{name}_faces{k}"]) for k in range(n)]
        p, _, trees = gpu_pairs(api, ctx, meshes, 0, 1, self_collide=self_c)
        for k, t in enumerate(trees):
            assert np.array_equal(t.perm(), g[f"{name}_perm{k}"])
            assert np.array_equal(bits(t.nodes()), bits(g[f"{name}_nodes{k}"]))
        assert np.array_equal(p, g[f"{name}_pairs"]), name
        p64, _, _ = gpu_pairs(api, ctx, meshes, 0, 1, self_collide=self_c, fp64=True)
        assert np.array_equal(p64, g[f"{name}_pairs64"]), name


def test_detect_reference_sat_vectors(api, ctx, orc):
    """The 8,000 SAT cases whose outcome the reference's own triangleIntersect
    fixed (tests/golden/sat_ref.npz), as two meshes: every overlapping face
    pair is tested, and the diagonal pairs (k, k) must match the reference."""
    say("reference SAT vectors on the GPU")
    g = np.load(os.path.join(GOLD, "sat_ref.npz"))
    P, Q, ref, f64 = g["P"].reshape(-1, 3), g["Q"].reshape(-1, 3), g["ref"], g["fp64"]
    n = ref.size
    fa = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    meshes = [(P, fa), (Q, fa)]
    p, _, _ = gpu_pairs(api, ctx, meshes, 2, 2)
    want = orc.detect([orc.Tree(P, fa), orc.Tree(Q, fa)])
    assert np.array_equal(p, want)
    diag = np.zeros(n, np.uint8)
    d = p[p[:, 2] == p[:, 3]]
    diag[d[:, 2]] = 1
    # a diagonal pair whose leaf boxes do not overlap never reaches the SAT;
    # every such pair is also a SAT miss in the reference
    assert np.all(diag <= ref) and np.array_equal(diag[ref == 1], ref[ref == 1])
    p64, st, _ = gpu_pairs(api, ctx, meshes, 2, 2, fp64=True)
    want64 = orc.detect([orc.Tree(P, fa), orc.Tree(Q, fa)], fp64=True)
    assert np.array_equal(p64, want64)
    d64 = np.zeros(n, np.uint8)
    d64[p64[p64[:, 2] == p64[:, 3]][:, 2]] = 1
    assert np.array_equal(d64[f64 == 1], f64[f64 == 1])
    assert st["fp64_rechecks"] > 0


def test_detect_cfg1_full_size(api, ctx, orc):
    say("detect cfg1 (2 x 81,920 faces)")
    meshes = M.interpenetrating_icospheres(subdiv=6)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    for entry, expand in [(4, 3), (0, 1), (2, 2)]:
        p, st, _ = gpu_pairs(api, ctx, meshes, entry, expand)
        assert np.array_equal(p, want), (entry, expand)
        assert st["pairs"] == len(want)


@pytest.mark.parametrize("entry", [0, 1, 3, 6, 30])
@pytest.mark.parametrize("expand", [1, 2, 3, 5])
@pytest.mark.parametrize("mode", [(0, 0), (64, 0), (7, 0), (64, 2), (16, 1),
                                  (0, 0, 32, 0), (64, 1, 16, 0), (0, 0, 64, 2), (7, 2, 5, 1)])
def test_detect_entry_expand_invariance(api, ctx, orc, entry, expand, mode):
    """mode: (traverse_local, traverse_local_levels[, traverse_tail,
    traverse_tail_steps]): steps only; the first steps (all of them, 2 or 1)
    as a head local walk of n initial pairs per workgroup; the last steps
    (all, or the last 2 or 1) as a tail walk of n front pairs per workgroup
    pass, alone or after a head walk."""
    say(f"detect entry={entry} expand={expand} local/tail={mode}")
    meshes = M.interpenetrating_icospheres(subdiv=4, seed=21, offset=1.3)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    tail = mode[2:] or (0, 0)
    ctx.set_tuning(traverse_local=mode[0], traverse_local_levels=mode[1], traverse_tail=tail[0],
                   traverse_tail_steps=tail[1])
    try:
        p, st, _ = gpu_pairs(api, ctx, meshes, entry, expand)
    finally:
        ctx.set_tuning(traverse_local=0, traverse_local_levels=0, traverse_tail=0, traverse_tail_steps=0)
    assert np.array_equal(p, want)
    assert st["retries"] == 0


@pytest.mark.parametrize("local,tail", [(0, 0), (64, 0), (64, 32)])
def test_detect_narrow_separate(api, ctx, orc, local, tail):
    """FP32 leaf pairs collected and tested by one k_narrow launch
    (narrow_separate = 1): the same pairs, pair sets and self mode, with
    steps, a head walk, or a head and a tail walk."""
    say(f"detect narrow_separate local={local} tail={tail}")
    meshes = M.interpenetrating_icospheres(subdiv=4, seed=23, offset=1.2)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    x, f = M.cloth(64, amplitude=0.2, wavelength=0.2, seed=5)
    want_self = orc.detect([orc.Tree(x, f)], self_collide=True)
    ctx.set_tuning(narrow_separate=1, traverse_local=local, traverse_local_levels=2, traverse_tail=tail)
    try:
        for entry, expand in [(0, 1), (3, 2), (2, 3)]:
            p, st, _ = gpu_pairs(api, ctx, meshes, entry, expand)
            assert np.array_equal(p, want), (entry, expand)
            assert st["candidates"] >= len(want)
        p, _, _ = gpu_pairs(api, ctx, [(x, f)], 2, 2, self_collide=True)
        assert np.array_equal(p, want_self)
    finally:
        ctx.set_tuning(narrow_separate=0, traverse_local=0, traverse_local_levels=0, traverse_tail=0)


def test_detect_local_walk_overflow_falls_back(api, ctx, orc):
    """Two identical spheres walked from the roots (one workgroup) three
    levels at a time: every node overlaps its twin and neighbours, so the
    workgroup's LDS stack overflows; the detect is re-run with steps (same
    pairs) and the next detects use steps."""
    say("local walk overflow")
    a = M.icosphere(5, 1.0, 0.0, 0)
    want = orc.detect([orc.Tree(*a), orc.Tree(*a)])
    ctx2 = api.Context(0)
    ctx2.set_tuning(traverse_local=64)
    try:
        trees = [gpu_tree(api, ctx2, *a), gpu_tree(api, ctx2, *a)]
        sc = api.Scene(ctx2)
        for t in trees:
            sc.addOibvhTree(t)
        sc.detectCollision(0, 3)
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["retries"] >= 1
        sc.detectCollision(0, 3)
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["retries"] == 0 and len(sc.level_counts()) > 1  # steps now
        sc.close()
    finally:
        ctx2.close()


def test_detect_tail_walk_overflow_falls_back(api, ctx, orc):
    """The same twin spheres walked from the roots by the tail walk alone
    (64 pairs per workgroup, every step): its LDS front overflows, the
    detect is re-run without it (same pairs) and the next detects use
    steps."""
    say("tail walk overflow")
    a = M.icosphere(5, 1.0, 0.0, 0)
    want = orc.detect([orc.Tree(*a), orc.Tree(*a)])
    ctx2 = api.Context(0)
    ctx2.set_tuning(traverse_tail=64)
    try:
        trees = [gpu_tree(api, ctx2, *a), gpu_tree(api, ctx2, *a)]
        sc = api.Scene(ctx2)
        for t in trees:
            sc.addOibvhTree(t)
        sc.detectCollision(0, 3)
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["retries"] >= 1
        sc.detectCollision(0, 3)
        assert np.array_equal(sc.getIntTriPairs(), want)
        assert sc.stats()["retries"] == 0
        lc = sc.level_counts()
        assert len(lc) > 1 and all(v > 0 for v in lc)  # every step wrote its front
        sc.close()
    finally:
        ctx2.close()


@pytest.mark.parametrize("expand,schedule", [(3, 0b1010), (3, 0b01), (2, 0b1111), (3, 0xAAAA), (1, 0b111011)])
def test_detect_step_schedule(api, ctx, orc, expand, schedule):
    """traverse_schedule: the first steps descend the listed levels (2 bits
    per step), later steps the uniform setting; pairs as the oracle's, for
    two meshes and for a cloth in self mode, including a schedule longer
    than the walk."""
    say(f"detect schedule {schedule:#x} then {expand}")
    meshes = M.interpenetrating_icospheres(subdiv=4, seed=31, offset=1.2)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    x, f = M.cloth(64, amplitude=0.2, wavelength=0.2, seed=9)
    want_self = orc.detect([orc.Tree(x, f)], self_collide=True)
    ctx.set_tuning(traverse_expand=expand, traverse_schedule=schedule)
    try:
        for entry in (0, 3):
            p, _, _ = gpu_pairs(api, ctx, meshes, entry, expand)
            assert np.array_equal(p, want), entry
        p, _, _ = gpu_pairs(api, ctx, [(x, f)], 2, expand, self_collide=True)
        assert np.array_equal(p, want_self)
    finally:
        ctx.set_tuning(traverse_expand=0, traverse_schedule=0)


@pytest.mark.parametrize("fp64", [False, True])
def test_scene_frame_overlap_many_trees(api, ctx, orc, fp64):
    """Overlapped frames on a scene of six trees (three deforming sphere
    pairs apart from each other), whose mesh table lives in device memory,
    one per detect counter set: frames queued back to back and read at the
    end, and frames read one by one, equal the oracle's pairs and nodes."""
    say(f"scene frame overlap, six trees, fp64={fp64}")
    meshes, waves, shift = [], [], []
    for j in range(3):
        ms, ws = M.deforming_uv_pair(96, 49, seed=700 + 10 * j)
        meshes += ms
        waves += ws
        shift += [np.array([0.0, 3.0 * j, 0.0], np.float32)] * 2
    ks = (2, 7, 19, 40)
    frames = [[(waves[m].frame(k) + shift[m]).astype(np.float32) for m in range(6)] for k in ks]
    bufs = [[api.DeviceBuffer.from_array(x) for x in fr] for fr in frames]
    trees = [gpu_tree(api, ctx, frames[0][m], meshes[m][1]) for m in range(6)]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)

    def want(k):
        return orc.detect([orc.Tree(frames[k][m], meshes[m][1], order=trees[m].perm()) for m in range(6)],
                          fp64=fp64)

    for k in range(len(ks)):  # all queued, then the last one read
        sc.frame([b.at(0) for b in bufs[k]], 2, 2, fp64_fallback=fp64, asynchronous=True, overlap=True)
    w = want(len(ks) - 1)
    assert len(w) > 0 and len(np.unique(w[:, 0])) == 3
    assert np.array_equal(sc.getIntTriPairs(), w)
    for m in range(6):
        o = orc.Tree(frames[-1][m], meshes[m][1], order=trees[m].perm())
        assert np.array_equal(bits(trees[m].nodes()), bits(o.nodes())), m
    for k in (1, 0, 2):  # read one by one
        sc.frame([b.at(0) for b in bufs[k]], 2, 2, fp64_fallback=fp64, overlap=True)
        assert np.array_equal(sc.getIntTriPairs(), want(k)), k
    st = sc.stats()
    assert st["retries"] == 0 and st["async_errors"] == 0
    sc.close()


def test_scene_frame_overlap_two_scenes(api, ctx, orc):
    """Two scenes on one context with their overlapped frames interleaved:
    the detects share the context's side streams and its ring of events
    (ten detects wrap the ring of four), each scene waits only for its own
    earlier detects, and both scenes' pairs and nodes equal the oracle's,
    for frames queued back to back and for frames read one by one."""
    say("scene frame overlap, two scenes on one context")
    sets = []
    for j in range(2):
        meshes, waves = M.deforming_uv_pair(112, 57, seed=900 + 10 * j)
        ks = (1, 6, 11, 17, 29)
        frames = [[waves[m].frame(k) for m in range(2)] for k in ks]
        bufs = [[api.DeviceBuffer.from_array(x) for x in fr] for fr in frames]
        trees = [gpu_tree(api, ctx, frames[0][m], meshes[m][1]) for m in range(2)]
        sc = api.Scene(ctx)
        for t in trees:
            sc.addOibvhTree(t)
        sets.append((meshes, frames, bufs, trees, sc))

    def want(j, k):
        meshes, frames, _, trees, _ = sets[j]
        return orc.detect([orc.Tree(frames[k][m], meshes[m][1], order=trees[m].perm()) for m in range(2)])

    def check_nodes(j, k):
        meshes, frames, _, trees, _ = sets[j]
        for m in range(2):
            o = orc.Tree(frames[k][m], meshes[m][1], order=trees[m].perm())
            assert np.array_equal(bits(trees[m].nodes()), bits(o.nodes())), (j, k, m)

    n = len(sets[0][1])
    for k in range(n):  # interleaved, all queued
        for j in range(2):
            sets[j][4].frame([b.at(0) for b in sets[j][2][k]], 3, 2, asynchronous=True, overlap=True)
    for j in range(2):
        assert np.array_equal(sets[j][4].getIntTriPairs(), want(j, n - 1)), j
        check_nodes(j, n - 1)
    for k in (3, 0, 2):  # interleaved, scene 1 queued, scene 0 read each time
        sets[1][4].frame([b.at(0) for b in sets[1][2][k]], 3, 2, asynchronous=True, overlap=True)
        sets[0][4].frame([b.at(0) for b in sets[0][2][k]], 3, 2, overlap=True)
        assert np.array_equal(sets[0][4].getIntTriPairs(), want(0, k)), k
    assert np.array_equal(sets[1][4].getIntTriPairs(), want(1, 2))
    check_nodes(1, 2)
    for j in range(2):
        st = sets[j][4].stats()
        assert st["retries"] == 0 and st["async_errors"] == 0
    for meshes, frames, bufs, trees, sc in sets:
        for t, x in zip(trees, frames[0]):
            t.refit(x)  # owned memory again before the buffers go away
        sc.close()
    ctx.synchronize()
    for meshes, frames, bufs, trees, sc in sets:
        for bb in bufs:
            for b in bb:
                b.free()


def test_detect_many_meshes(api, ctx, orc):
    say("detect 6 meshes")
    r = M.Rng(99)
    meshes = []
    for k in range(6):
        c = r.uniform(3, 0.0, 1.2)
        if k % 2:
            meshes.append(M.icosphere(3, 0.5, 0.02, 50 + k, center=tuple(c)))
        else:
            x, f = M.box(8, 0.4, noise=0.01, seed=60 + k)
            meshes.append(((x + c.astype(np.float32)).astype(np.float32), f))
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    assert len(want) > 0
    p, _, trees = gpu_pairs(api, ctx, meshes, 1, 2)
    assert np.array_equal(p, want)
    # frames with OIBVH_FRAME_OVERLAP on a scene of more than four trees (its
    # mesh table in device memory), with the same pairs
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    for _ in range(3):
        sc.frame(None, 1, 2, asynchronous=True, overlap=True)
    assert np.array_equal(sc.getIntTriPairs(), want)
    sc.frame(None, 1, 2, fp64_fallback=True, overlap=True)
    assert np.array_equal(sc.getIntTriPairs(), orc.detect([orc.Tree(x, f) for x, f in meshes], fp64=True))


def test_detect_separated_and_identical(api, ctx, orc):
    say("detect separated / identical meshes")
    a = M.icosphere(3, 1.0, 0.0, 0)
    b = M.icosphere(3, 1.0, 0.0, 0, center=(5.0, 0.0, 0.0))
    p, _, _ = gpu_pairs(api, ctx, [a, b])
    assert p.shape == (0, 4)
    p, _, _ = gpu_pairs(api, ctx, [a, a])
    assert np.array_equal(p, orc.detect([orc.Tree(*a), orc.Tree(*a)]))
    assert len(p) > a[1].shape[0]  # every face hits itself and its neighbours


def test_detect_self_collision(api, ctx, orc):
    say("self collision cloth")
    x, f = M.cloth(96, amplitude=0.2, wavelength=0.2, seed=4)
    want = orc.detect([orc.Tree(x, f)], self_collide=True)
    assert len(want) > 0
    p, st, _ = gpu_pairs(api, ctx, [(x, f)], 0, 1, self_collide=True)
    assert np.array_equal(p, want)
    assert st["adjacent"] > 0
    for entry, expand in [(3, 3), (2, 2)]:
        p, _, _ = gpu_pairs(api, ctx, [(x, f)], entry, expand, self_collide=True)
        assert np.array_equal(p, want)


def test_detect_overflow_retry(api, ctx, orc):
    say("detect with tiny queue and hit buffers")
    meshes = M.interpenetrating_icospheres(subdiv=5, seed=31, offset=1.2)
    want = orc.detect([orc.Tree(x, f) for x, f in meshes])
    ctx2 = api.Context(0)
    ctx2.set_tuning(front_capacity=256, hits_capacity=16, candidate_capacity=64)
    p, st, _ = gpu_pairs(api, ctx2, meshes, 0, 3)
    assert np.array_equal(p, want)
    assert st["retries"] > 0
    ctx2.close()


def test_detect_repeated_frames(api, ctx, orc):
    """refit + detect over several frames of the deforming pair, async detect."""
    say("deforming pair frames")
    meshes, waves = M.deforming_uv_pair(160, 81)
    trees = [gpu_tree(api, ctx, x, f) for x, f in meshes]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    perms = [t.perm() for t in trees]
    for frame in (0, 3, 9):
        xs = [w.frame(frame) for w in waves]
        for t, x in zip(trees, xs):
            t.refit(x)
        sc.detectCollision(4, 3, asynchronous=True)
        got = sc.getIntTriPairs()
        want = orc.detect([orc.Tree(x, m[1], order=p) for x, m, p in zip(xs, meshes, perms)])
        assert np.array_equal(got, want), frame


def test_fp64_fallback_coplanar_stress(api, ctx, orc):
    say("precision stress: near-coplanar planes")
    x, f = M.plane(40)
    r = M.Rng(5)
    x2 = x.copy()
    x2[:, 2] = r.uniform(x.shape[0], -1e-6, 1e-6).astype(np.float32)
    x2[:, :2] += np.float32(0.013)
    meshes = [(x, f), (x2.astype(np.float32), f)]
    want64 = orc.detect([orc.Tree(*m) for m in meshes], fp64=True)
    p64, st, _ = gpu_pairs(api, ctx, meshes, 0, 1, fp64=True)
    assert np.array_equal(p64, want64)
    assert st["fp64_rechecks"] > 0
    want32 = orc.detect([orc.Tree(*m) for m in meshes])
    p32, _, _ = gpu_pairs(api, ctx, meshes, 0, 1)
    assert np.array_equal(p32, want32)


def test_cpp_facade_sequence(ctx, orc, tmp_path):
    say("C++ facade program")
    exe = os.path.join(ROOT, "tests", "cpp", "facade_test")
    assert os.path.exists(exe), "facade_test not built"
    xyz, f = M.icosphere(4, 1.0, 0.02, 8)
    x2 = (xyz + np.float32([1.0, 0.0, 0.0])).astype(np.float32)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([xyz.shape[0], f.shape[0]], np.uint32).tofile(fh)
        xyz.tofile(fh)
        x2.tofile(fh)
        f.tofile(fh)
    mx, mf = M.icosphere(2, 1.0, 0.05, 9)
    obj = tmp_path / "model.obj"
    obj.write_text("".join(f"v {a:.9g} {b:.9g} {c:.9g}\n" for a, b, c in mx.astype(np.float64))
                   + "".join(f"f {a + 1}/1/1 {b + 1}/1/1 {c + 1}/1/1\n" for a, b, c in mf))
    r = subprocess.run([exe, str(inp), str(out), str(obj)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint32)
    N = int(raw[0])
    o = 1
    n1 = raw[o:o + 6 * N].view(np.float32).reshape(N, 6)
    o += 6 * N
    n2 = raw[o:o + 6 * N].view(np.float32).reshape(N, 6)
    o += 6 * N
    F = f.shape[0]
    perm = raw[o:o + F]
    o += F
    H = int(raw[o])
    o += 1
    pairs = raw[o:o + 4 * H].reshape(H, 4)
    o += 4 * H
    depth, prims = int(raw[o]), int(raw[o + 1])
    o += 2
    verts = raw[o:o + 18 * H].view(np.float32).reshape(H, 6, 3)
    o += 18 * H
    mv, mfc = int(raw[o]), int(raw[o + 1])
    o += 2
    assert (mv, mfc) == (mx.shape[0], mf.shape[0])
    assert np.array_equal(raw[o:o + 3 * mv], mx.view(np.uint32).reshape(-1))
    o += 3 * mv
    assert np.array_equal(raw[o:o + 3 * mfc], mf.reshape(-1))
    t1 = orc.Tree(xyz, f, order="morton")
    t2 = orc.Tree(x2, f, order=t1.order)
    assert np.array_equal(perm, t1.order)
    assert np.array_equal(bits(n1), bits(t1.nodes())) and np.array_equal(bits(n2), bits(t2.nodes()))
    assert np.array_equal(pairs, orc.detect([t1, t2]))
    assert prims == F and depth == int(np.log2(N))
    assert np.array_equal(bits(verts), bits(expected_pair_vertices(pairs, [(xyz, f), (x2, f)])))


def test_smoke_entry():
    say("smoke()")
    sys.path.insert(0, ROOT)
    import __graft_entry__

    __graft_entry__.smoke()


def test_pairs_sorted_on_device(api, ctx, orc):
    """The pair list sorted on the device (oibvh_scene_pairs_device) equals
    the oracle's sorted pairs; face ids above 2^16 (three key bytes) and a
    scene of more than 256 trees (two mesh-id bytes, many trees with no
    pair), both sorted by one workgroup (at most 8192 pairs); a cloth's
    18,350 self-collision pairs (the multi-workgroup passes); empty
    results."""
    say("pairs sorted on the device")
    (x0, f0), (x1, f1) = M.interpenetrating_icospheres(subdiv=6, offset=1.0)  # 81,920 faces each
    t0, t1 = api.OibvhTree(ctx, x0, f0), api.OibvhTree(ctx, x1, f1)
    t0.build()
    t1.build()
    sc = api.Scene(ctx)
    sc.addOibvhTree(t0)
    sc.addOibvhTree(t1)
    sc.detectCollision(4, 3)
    want = orc.detect([orc.Tree(x0, f0, order=t0.perm()), orc.Tree(x1, f1, order=t1.perm())])
    assert len(want) > 1000 and want[:, 2].max() > 65536
    p, n = sc.pairs_device()
    assert n == len(want)
    assert np.array_equal(api.download_u32(p, 4 * n).reshape(-1, 4), want)
    assert np.array_equal(sc.getIntTriPairs(), want)
    # 300 small trees in a row, neighbours touching
    trees, shapes = [], []
    for i in range(300):
        x, f = M.icosphere(1, 0.6, 0.02, 900 + i, center=(0.9 * i, 0.0, 0.0))
        t = api.OibvhTree(ctx, x, f)
        t.build()
        trees.append(t)
        shapes.append((x, f))
    big = api.Scene(ctx)
    for t in trees:
        big.addOibvhTree(t)
    big.detectCollision(0, 2)
    want = orc.detect([orc.Tree(x, f, order=t.perm()) for (x, f), t in zip(shapes, trees)])
    assert len(want) > 0 and want[:, 1].max() > 256
    p, n = big.pairs_device()
    assert np.array_equal(api.download_u32(p, 4 * n).reshape(-1, 4), want)
    # more pairs than one workgroup sorts
    x, f = M.cloth(256, amplitude=0.2, wavelength=0.05, seed=3)
    tc = api.OibvhTree(ctx, x, f)
    tc.build()
    cl = api.Scene(ctx)
    cl.addOibvhTree(tc)
    cl.detectCollision(2, 2, self_collide=True)
    want = orc.detect([orc.Tree(x, f, order=tc.perm())], self_collide=True)
    assert len(want) > 8192
    p, n = cl.pairs_device()
    assert np.array_equal(api.download_u32(p, 4 * n).reshape(-1, 4), want)
    # nothing collides
    far = api.Scene(ctx)
    far.addOibvhTree(trees[0])
    far.addOibvhTree(trees[10])
    far.detectCollision(0, 2)
    assert far.pairs_device()[1] == 0 and len(far.getIntTriPairs()) == 0


def test_scene_build_equals_tree_builds(api, ctx, orc):
    """oibvh_scene_build (the trees sharing each launch of the build chain)
    gives every tree what oibvh_tree_build gives it: leaf order, sorted faces
    and node bits; also right after overlapped frames (the side stream busy),
    followed by a detect and more frames."""
    say("scene build, trees sharing the launches")
    meshes = [M.icosphere(5, 1.0, 0.02, 60 + i, center=(0.8 * i, 0.0, 0.0)) for i in range(3)]
    trees = [api.OibvhTree(ctx, x, f) for x, f in meshes]
    for t in trees:
        t.build()
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    for k in range(3):
        sc.frame(None, 2, 2, overlap=True, asynchronous=True)
    moved = [(x * np.float32(1.0 + 0.01 * i)).astype(np.float32) for i, (x, _) in enumerate(meshes)]
    for t, x in zip(trees, moved):
        t.set_positions(x)
    sc.build()
    for t, x, (_, f) in zip(trees, moved, meshes):
        check_tree(orc, t, x, f)
    sc.detectCollision(2, 2)
    ot = [orc.Tree(x, f, order=t.perm()) for t, x, (_, f) in zip(trees, moved, meshes)]
    want = orc.detect(ot)
    assert len(want) > 0
    assert np.array_equal(sc.getIntTriPairs(), want)
    sc.frame(None, 2, 2, overlap=True)
    assert np.array_equal(sc.getIntTriPairs(), want)
    one = api.Scene(ctx)
    one.addOibvhTree(trees[0])
    one.build()  # a group of one tree
    check_tree(orc, trees[0], moved[0], meshes[0][1])


def test_scene_build_groups_mixed(api, ctx, orc):
    """Seven trees of very different sizes (one face to 70,001 faces, one to
    18 sort tiles), 30- and 63-bit Morton keys and model matrices on some:
    the scene build groups consecutive trees of one Morton width, at most
    four per launch, and every tree ends bit-identical to its own
    oibvh_tree_build (itself pinned to the oracle by the build tests), three
    times in a row (plain launches, then the captured graph, then its
    replay); the 30-bit trees without a matrix also against the oracle."""
    say("scene build, mixed groups")
    sizes = [1, 3, 4097, 70001, 20000, 2, 300]
    bits63 = [False, False, True, True, False, False, False]
    xforms = [None, _motion(2, 0), None, _motion(3, 1), None, None, _motion(4, 2)]
    meshes = [soup(F, 500 + i) for i, F in enumerate(sizes)]
    trees = []
    for (x, f), b63, A in zip(meshes, bits63, xforms):
        t = api.OibvhTree(ctx, x, f)
        if b63:
            t.set_morton_bits(63)
        if A is not None:
            t.set_transform(A)
        t.build()
        trees.append(t)
    want = [(t.perm().copy(), t.sorted_faces().copy(), bits(t.nodes()).copy()) for t in trees]
    sc = api.Scene(ctx)
    for t in trees:
        sc.addOibvhTree(t)
    for rep in range(3):
        for t, (x, _) in zip(trees, meshes):
            # positions mirrored through the origin: another leaf order and
            # other boxes over the outputs of the first builds
            t.set_positions(np.ascontiguousarray(-x[::-1]).astype(np.float32))
            t.build()
        for t, (x, _) in zip(trees, meshes):
            t.set_positions(x)
        sc.build()
        for i, (t, (p, sf, nb)) in enumerate(zip(trees, want)):
            assert np.array_equal(t.perm(), p), f"tree {i}: leaf permutation differs"
            assert np.array_equal(t.sorted_faces(), sf), f"tree {i}: sorted faces differ"
            assert np.array_equal(bits(t.nodes()), nb), f"tree {i}: node bits differ"
    for i in (0, 4, 5):
        check_tree(orc, trees[i], meshes[i][0], meshes[i][1])
