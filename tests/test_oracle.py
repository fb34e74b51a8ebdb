"""CPU tests that pin the oracle: closed forms, brute force, special cases,
the reference's own code (oracle/_ref, where built) and the golden fixtures
written from it (tests/golden/)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oibvh_amd import meshgen as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


# ----------------------------------------------------------------- layout
def _count_simplebvh_nodes(F):
    """node count of SimpleBVH::recursiveBuild by direct recursion (Python)"""
    depth = (F - 1).bit_length() if F > 1 else 0

    def rec(lo, hi, d):
        if lo == hi:
            return 1 + (rec(lo, hi, d + 1) if d < depth else 0)
        cnt = hi - lo + 1
        mid = lo + (1 << (cnt - 1).bit_length()) // 2 - 1
        if cnt <= (1 << (depth - d - 1)):
            return 1 + rec(lo, hi, d + 1)
        return 1 + rec(lo, mid, d + 1) + rec(mid + 1, hi, d + 1)

    return rec(0, F - 1, 0)


@pytest.mark.parametrize("F", list(range(1, 70)) + [127, 128, 129, 255, 256, 257, 1000, 1023, 1025])
def test_tree_size_closed_form(orc, F):
    assert orc.tree_size(F) == _count_simplebvh_nodes(F)
    # sum of per-level real counts
    L = (F - 1).bit_length() if F > 1 else 0
    vl = (1 << L) - F
    assert sum(orc.level_real_count(l, L, vl) for l in range(L + 1)) == orc.tree_size(F)
    # level counts halve with ceiling (left-aligned real prefix)
    for l in range(L):
        assert orc.level_real_count(l, L, vl) == -(-orc.level_real_count(l + 1, L, vl) // 2)


@pytest.mark.parametrize("F", [1, 2, 3, 5, 6, 7, 9, 33, 100, 513, 4097])
def test_bfs_topology_is_oibvh_layout(orc, F):
    """SimpleBVH's breadth-first order == real order of the implicit layout:
    node p of level l has children 2p, 2p+1 (right real iff < count)."""
    xyz = np.zeros((3 * F, 3), np.float32)
    xyz[:, 0] = np.arange(3 * F, dtype=np.float32)
    faces = np.arange(3 * F, dtype=np.uint32).reshape(F, 3)
    t = orc.Tree(xyz, faces)
    tri, nch = t.topology()
    L = (F - 1).bit_length() if F > 1 else 0
    vl = (1 << L) - F
    r = 0
    for l in range(L + 1):
        cnt = orc.level_real_count(l, L, vl)
        nxt = orc.level_real_count(l + 1, L, vl) if l < L else 0
        for p in range(cnt):
            if l == L:
                assert nch[r] == 0 and tri[r] == p
            else:
                want = 1 + (2 * p + 1 < nxt)
                assert nch[r] == want and tri[r] == -1
            assert orc.real_to_implicit(r, L, vl) == (1 << l) - 1 + p
            assert orc.implicit_to_real((1 << l) - 1 + p, L, vl) == r
            r += 1
    assert r == t.node_count


# ----------------------------------------------------------------- morton
def test_quantise_known_answers(orc):
    assert orc.quantise10(0.0) == 0
    assert orc.quantise10(-0.0) == 0
    assert orc.quantise10(0.5) == 512
    assert orc.quantise10(1.0) == 1023
    assert orc.quantise10(0.99999) == 1023
    assert orc.quantise10(-3.0) == 0
    assert orc.quantise10(float("nan")) == 0
    assert orc.quantise10(float("inf")) == 1023
    assert orc.quantise10(float("-inf")) == 0
    assert orc.quantise10(1.0 / 1024) == 1
    assert orc.quantise10(np.nextafter(np.float32(1.0 / 1024), np.float32(0))) == 0


def test_morton_bit_interleave(orc):
    def spread(x):
        return sum(((x >> b) & 1) << (3 * b) for b in range(10))

    rng = M.Rng(7)
    for _ in range(300):
        q = rng.integers(3, 0, 1023)
        f = [(int(v) + 0.5) / 1024.0 for v in q]
        want = (spread(int(q[0])) << 2) | (spread(int(q[1])) << 1) | spread(int(q[2]))
        assert orc.morton30(*f) == want


def soup_small(F, seed):
    r = M.Rng(seed)
    V = F + 2
    xyz = r.uniform(V * 3, -1.0, 1.0).reshape(V, 3).astype(np.float32)
    # repeated vertices make equal keys (ties)
    xyz[V // 2:] = xyz[: V - V // 2]
    f = r.integers(F * 3, 0, V - 1).reshape(F, 3).astype(np.uint32)
    return xyz, f


def _interleave(qx, qy, qz, bits):
    k = 0
    for b in range(bits):
        k |= ((qx >> b) & 1) << (3 * b + 2) | ((qy >> b) & 1) << (3 * b + 1) | ((qz >> b) & 1) << (3 * b)
    return k


def test_morton63_bit_interleave(orc):
    """63-bit codes against a bit-by-bit interleave of the 21-bit
    quantisations, and the quantisation's closed form."""
    assert orc.quantise21(0.0) == 0 and orc.quantise21(-1.0) == 0
    assert orc.quantise21(1.0) == (1 << 21) - 1 and orc.quantise21(float("inf")) == (1 << 21) - 1
    assert orc.quantise21(float("nan")) == 0
    assert orc.quantise21(0.5) == 1 << 20
    r = M.Rng(0x63)
    for x, y, z in r.uniform(3 * 300, -0.1, 1.1).reshape(300, 3):
        qx, qy, qz = (orc.quantise21(float(np.float32(v))) for v in (x, y, z))
        assert orc.morton63(float(np.float32(x)), float(np.float32(y)), float(np.float32(z))) == \
            _interleave(qx, qy, qz, 21)
    assert orc.morton63(1.0, 1.0, 1.0) == (1 << 63) - 1


def test_morton63_stable_order(orc):
    """The 64-bit order is sorted by key and keeps equal keys in face order."""
    xyz, f = soup_small(500, 17)
    k63 = orc.leaf_keys63(f, xyz)
    o = orc.stable_order64(k63)
    assert np.all(np.diff(k63[o].astype(np.float64)) >= 0)
    ties = np.nonzero(k63[o][1:] == k63[o][:-1])[0]
    assert np.all(o[ties] < o[ties + 1])  # stable


def test_leaf_keys_and_stable_order(orc):
    xyz, f = M.icosphere(3, 1.0, 0.02, 5)
    box = orc.mesh_box(xyz)
    keys, leaves = orc.leaf_keys(f, xyz, box, with_leaves=True)
    # leaf boxes = compare-select min/max of the 3 vertices
    tri = xyz[f]
    assert np.array_equal(bits(leaves[:, :3]), bits(tri.min(axis=1)))
    assert np.array_equal(bits(leaves[:, 3:]), bits(tri.max(axis=1)))
    # keys recomputed with numpy float32 arithmetic, same op order
    c = (leaves[:, :3] + leaves[:, 3:]) * np.float32(0.5)
    p = (c - box[:3]) / (box[3:] - box[:3])
    q = np.clip(p * np.float32(1024.0), np.float32(0.0), np.float32(1023.0)).astype(np.uint32)

    def spread(x):
        x = x.astype(np.uint64) & 0x3FF
        out = np.zeros_like(x)
        for b in range(10):
            out |= ((x >> b) & 1) << (3 * b)
        return out.astype(np.uint32)

    want = (spread(q[:, 0]) << 2) | (spread(q[:, 1]) << 1) | spread(q[:, 2])
    assert np.array_equal(keys, want)
    assert np.array_equal(orc.stable_order(keys), np.argsort(keys, kind="stable").astype(np.uint32))


def test_mesh_box_matches_numpy(orc):
    xyz, _ = M.uv_sphere(40, 21)
    b = orc.mesh_box(xyz)
    assert np.array_equal(b[:3], xyz.min(axis=0)) and np.array_equal(b[3:], xyz.max(axis=0))


# ----------------------------------------------------------------- tree AABBs
@pytest.mark.parametrize("order", [None, "morton"])
def test_tree_parent_is_merge_of_children(orc, order):
    xyz, f = M.icosphere(3, 1.0, 0.02, 9)
    t = orc.Tree(xyz, f, order=order)
    nodes = t.nodes()
    F = f.shape[0]
    L = (F - 1).bit_length()
    vl = (1 << L) - F
    starts = np.cumsum([0] + [orc.level_real_count(l, L, vl) for l in range(L + 1)])
    perm = np.arange(F) if order is None else t.order
    tri = xyz[f[perm]]
    leaf = nodes[starts[L]:]
    assert np.array_equal(bits(leaf[:, :3]), bits(tri.min(axis=1)))
    assert np.array_equal(bits(leaf[:, 3:]), bits(tri.max(axis=1)))
    for l in range(L):
        par = nodes[starts[l]:starts[l + 1]]
        ch = nodes[starts[l + 1]:starts[l + 2]]
        n = ch.shape[0]
        left = ch[0::2]
        right = np.vstack([ch[1::2], left[n // 2:]]) if n % 2 else ch[1::2]
        mn = np.where(right[:, :3] < left[:, :3], right[:, :3], left[:, :3])
        mx = np.where(left[:, 3:] < right[:, 3:], right[:, 3:], left[:, 3:])
        assert np.array_equal(bits(par[:, :3]), bits(mn)) and np.array_equal(bits(par[:, 3:]), bits(mx))


def test_refit_equals_rebuild(orc):
    meshes, waves = M.deforming_uv_pair(60, 31)
    xyz, f = meshes[0]
    t = orc.Tree(xyz, f, order="morton")
    x2 = waves[0].frame(17)
    t.refit(x2)
    t2 = orc.Tree(x2, f, order=t.order)
    assert np.array_equal(bits(t.nodes()), bits(t2.nodes()))


# ----------------------------------------------------------------- SAT
def test_sat_special_cases(orc):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    assert orc.tri_intersect(P, P)
    assert not orc.tri_intersect(P, P + np.float32(5))
    # shared vertex only: touching counts (strict separation test)
    Q = np.array([[1, 0, 0], [2, 0, 1], [2, 1, -1]], np.float32)
    assert orc.tri_intersect(P, Q)
    # coplanar, overlapping
    Q = np.array([[0.2, 0.2, 0], [2, 0.2, 0], [0.2, 2, 0]], np.float32)
    assert orc.tri_intersect(P, Q)
    # coplanar, disjoint across the hypotenuse
    Q = np.array([[0.8, 0.8, 0], [2, 0.8, 0], [0.8, 2, 0]], np.float32)
    assert not orc.tri_intersect(P, Q)
    # piercing
    Q = np.array([[0.2, 0.2, -1], [0.2, 0.2, 1], [3, 3, 0.5]], np.float32)
    assert orc.tri_intersect(P, Q)
    # degenerate (all vertices equal) inside P: zero axes never separate
    Q = np.array([[0.1, 0.1, 0]] * 3, np.float32)
    assert orc.tri_intersect(P, Q)


def test_sat_matches_reference_golden(orc):
    g = np.load(os.path.join(GOLD, "sat_ref.npz"))
    P, Q, ref, f64 = g["P"], g["Q"], g["ref"], g["fp64"]
    got = np.array([orc.tri_intersect(P[i], Q[i]) for i in range(P.shape[0])], np.uint8)
    assert np.array_equal(got, ref)
    got64 = np.array([orc.tri_intersect(P[i], Q[i], fp64=True) for i in range(P.shape[0])], np.uint8)
    assert np.array_equal(got64, f64)
    assert (ref != f64).sum() > 0  # the fixture exercises FP32/FP64 disagreements


def test_sat_matches_reference_live(orc):
    ref = orc.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted)")
    r = M.Rng(0x11)
    n = 4000
    P = r.uniform(n * 9, -1, 1).reshape(n, 9).astype(np.float32)
    Q = r.uniform(n * 9, -1, 1).reshape(n, 9).astype(np.float32)
    Q[: n // 2, 6:9] = P[: n // 2, 0:3]  # shared vertices
    import ctypes as C

    want = np.zeros(n, np.uint8)
    ref.ref_tri_intersect_many(P.ctypes.data_as(C.POINTER(C.c_float)), Q.ctypes.data_as(C.POINTER(C.c_float)), n,
                               want.ctypes.data_as(C.POINTER(C.c_uint8)))
    got = np.array([orc.tri_intersect(P[i], Q[i]) for i in range(n)], np.uint8)
    assert np.array_equal(got, want)


def test_aabb_matches_reference_golden(orc):
    g = np.load(os.path.join(GOLD, "aabb_ref.npz"))
    # leaf boxes through the oracle's tree builder on 1-face meshes
    V = g["V"]
    for i in range(0, V.shape[0], 7):
        t = orc.Tree(V[i].reshape(3, 3), np.array([[0, 1, 2]], np.uint32))
        assert np.array_equal(bits(t.nodes()[0]), bits(g["leaf"][i]))
    # mesh box fold incl. signed-zero ties
    off = 0
    for k, n in enumerate(g["cloud_len"]):
        c = g["cloud"][off:off + n]
        off += n
        assert np.array_equal(bits(orc.mesh_box(c)), bits(g["cloud_box"][k]))
    # merge: a 2-face mesh whose leaves are the A and B boxes; its root is
    # init(A) merge(B)
    A, B = g["A"], g["B"]
    checked = 0
    fp = C.POINTER(C.c_float)
    for i in range(A.shape[0]):
        a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
        assert orc.lib().orc_box_overlap(a.ctypes.data_as(fp), b.ctypes.data_as(fp)) == g["overlap"][i]
        if np.any(a[:3] > a[3:]) or np.any(b[:3] > b[3:]):
            continue
        xyz = np.array([a[:3], a[3:], a[:3], b[:3], b[3:], b[:3]], np.float32)
        nodes = orc.Tree(xyz, np.array([[0, 1, 2], [3, 4, 5]], np.uint32)).nodes()
        # a degenerate triangle reproduces its box only without +-0 ties
        if not (np.array_equal(bits(nodes[1]), bits(a)) and np.array_equal(bits(nodes[2]), bits(b))):
            continue
        assert np.array_equal(bits(nodes[0]), bits(g["merged"][i]))
        checked += 1
    assert checked > 1000


# ----------------------------------------------------------------- transform
def test_xform_matches_reference_golden(orc):
    """glm's own mat4 * vec4 (via the reference's vendored glm) on seeded
    matrices and points, bit for bit."""
    g = np.load(os.path.join(GOLD, "xform_ref.npz"))
    for k in range(g["M"].shape[0]):
        assert np.array_equal(bits(orc.xform(g["M"][k], g["P"])), bits(g["out"][k]))


def test_xform_matches_reference_live(orc):
    ref = orc.load_ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted)")
    r = M.Rng(0x7F1)
    fp = C.POINTER(C.c_float)
    for k in range(16):
        A = np.eye(4, dtype=np.float32)
        A[:3, :] = r.uniform(12, -3.0, 3.0).reshape(3, 4)
        P = r.uniform(3 * 500, -50.0, 50.0).reshape(500, 3).astype(np.float32)
        want = np.empty_like(P)
        cm = np.ascontiguousarray(A.T).reshape(16)
        ref.ref_xform(cm.ctypes.data_as(fp), P.ctypes.data_as(fp), 500, want.ctypes.data_as(fp))
        assert np.array_equal(bits(orc.xform(A, P)), bits(want))


def test_xform_closed_forms(orc):
    """Translation and axis permutation are exact; the grouping
    (m0 x + m1 y) + (m2 z + m3) is pinned by a case where the other
    summation orders round differently, and the column-major layout by an
    asymmetric matrix."""
    P = np.array([[1.5, -2.0, 0.25], [3.0, 4.0, -5.0]], np.float32)
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = [10.0, -20.0, 0.5]
    assert np.array_equal(orc.xform(T, P), P + T[:3, 3])
    S = np.zeros((4, 4), np.float32)
    S[0, 1] = S[1, 2] = S[2, 0] = S[3, 3] = 1.0  # (x, y, z) -> (y, z, x)
    assert np.array_equal(orc.xform(S, P), P[:, [1, 2, 0]])
    # row (1, 1, 1, e) on p = (1, e, e), e = 2^-24: glm's grouping
    # (1 + e) + (e + e) = 1 + 2^-23, a left-to-right sum rounds to 1
    e = np.float32(2.0 ** -24)
    A = np.zeros((4, 4), np.float32)
    A[0, :] = [1.0, 1.0, 1.0, 1.0]
    A[3, 3] = 1.0
    p = np.array([[1.0, e, e]], np.float32)
    A[0, 3] = e
    got = orc.xform(A, p)[0, 0]
    glm = (np.float32(1.0) + e) + (e + e)            # = 1 + 2^-23
    ltr = ((np.float32(1.0) + e) + e) + e            # = 1
    assert glm != ltr and got == glm


# ----------------------------------------------------------------- traversal
def _scene(orc, meshes, order):
    return [orc.Tree(x, f, order=order) for x, f in meshes]


@pytest.mark.parametrize("subdiv,offset", [(2, 1.5), (3, 1.7), (3, 2.5), (2, 0.0)])
def test_detect_equals_brute_force(orc, subdiv, offset):
    meshes = M.interpenetrating_icospheres(subdiv=subdiv, seed=subdiv * 7, offset=offset)
    t0 = _scene(orc, meshes, None)
    t1 = _scene(orc, meshes, "morton")
    bf = orc.brute_force(t0)
    assert np.array_equal(orc.detect(t0), bf)
    assert np.array_equal(orc.detect(t1), bf)
    assert np.array_equal(orc.detect(t1, broad_only=True), orc.brute_force(t0, broad_only=True))
    assert np.array_equal(orc.detect(t1, fp64=True), orc.brute_force(t0, fp64=True))


def test_detect_three_meshes(orc):
    meshes = [M.icosphere(2, 1.0, 0.02, 1), M.icosphere(2, 1.0, 0.02, 2, center=(1.2, 0, 0)),
              M.icosphere(2, 1.0, 0.02, 3, center=(0.6, 1.0, 0))]
    t = _scene(orc, meshes, "morton")
    p = orc.detect(t)
    assert np.array_equal(p, orc.brute_force(_scene(orc, meshes, None)))
    assert set(map(tuple, p[:, :2].tolist())) == {(0, 1), (0, 2), (1, 2)}


def test_self_collision_equals_brute_force(orc):
    x, f = M.cloth(24, amplitude=0.2, wavelength=0.3, seed=2)
    t = orc.Tree(x, f, order="morton")
    p = orc.detect([t], self_collide=True)
    assert len(p) > 0
    assert np.array_equal(p, orc.brute_force([orc.Tree(x, f)], self_collide=True))
    # no pair shares a vertex; ids ordered
    fa, fb = f[p[:, 2]], f[p[:, 3]]
    assert not np.any((fa[:, :, None] == fb[:, None, :]).any(axis=(1, 2)))
    assert np.all(p[:, 2] < p[:, 3])


def test_golden_scenes_reproduce(orc):
    g = np.load(os.path.join(GOLD, "scenes.npz"))
    for name in g["names"]:
        n = int(g[f"{name}_n"][0])
        self_c = bool(g[f"{name}_self"][0])
        trees = []
        for k in range(n):
            t = orc.Tree(g[f"{name}_xyz{k}"], g[f"{name}_faces{k}"], order="morton")
            assert np.array_equal(t.order, g[f"{name}_perm{k}"])
            assert np.array_equal(bits(t.nodes()), bits(g[f"{name}_nodes{k}"]))
            trees.append(t)
        assert np.array_equal(orc.detect(trees, self_collide=self_c), g[f"{name}_pairs"])
        assert np.array_equal(orc.detect(trees, self_collide=self_c, fp64=True), g[f"{name}_pairs64"])


def test_generators_shapes():
    x, f = M.icosphere(6)
    assert x.shape == (40962, 3) and f.shape == (81920, 3)
    x, f = M.uv_sphere(1000, 501)
    assert x.shape == (500002, 3) and f.shape == (1000000, 3)
    x, f = M.plane(354)
    assert f.shape[0] == 250632
    assert M.uniform01(1, 3).tolist() == M.uniform01(1, 3).tolist()
    v = M.icosphere_unit(2)[0]
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    assert math.isclose(float(M.uniform01(123, 1)[0] * (1 << 24)) % 1.0, 0.0)
