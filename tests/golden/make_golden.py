"""Write the golden fixtures in tests/golden/.

Run in the container where the reference tree is mounted (oracle/_ref must be
built: `make -C oracle ref`). The fixtures hold inputs and expected outputs
only:

  sat_ref.npz    triangle pairs and the reference's own triangleIntersect
                 result (src/utils/utils.cpp:97-169, via oracle/_ref), plus
                 the FP64 evaluation of the same expression (oracle)
  aabb_ref.npz   aabb_box_t::merge / overlap and the SimpleBVH leaf box
                 (utils.h:33-44, simpleBVH.cpp:136-139) from the reference
                 header, and Mesh::setupAABB's fold (mesh.cpp:91-98)
  scenes.npz     small seeded scenes with the oracle's trees (Morton order)
                 and sorted colliding pairs, inputs included
  xform_ref.npz  model matrices, points and glm's mat4 * vec4(p, 1) as the
                 reference's transform kernel applies it (transform.cu:35-41)

usage: python tests/golden/make_golden.py [sat aabb scenes xform]
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oibvh_amd import meshgen as M  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def sat_cases(n_rand: int = 3000, seed: int = 0x5A7):
    """Random, near-coplanar, shared-vertex, shared-edge and degenerate pairs."""
    r = M.Rng(seed)
    P, Q = [], []
    # random overlapping-ish pairs in a unit box
    p = r.uniform(n_rand * 9, -1.0, 1.0).reshape(n_rand, 9)
    q = r.uniform(n_rand * 9, -1.0, 1.0).reshape(n_rand, 9)
    P.append(p)
    Q.append(q)
    # near-coplanar: Q in the plane z ~ 0 of P with tiny offsets
    m = 2000
    p = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    p[:, :, 2] = 0.0
    q = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    q[:, :, 2] = r.uniform(m * 3, -1e-6, 1e-6).reshape(m, 3)
    P.append(p.reshape(m, 9))
    Q.append(q.reshape(m, 9))
    # touching: Q shares one vertex of P, or an edge, with random far parts
    m = 1000
    p = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    q = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    q[: m // 2, 0] = p[: m // 2, 1]
    q[m // 2:, 0] = p[m // 2:, 0]
    q[m // 2:, 1] = p[m // 2:, 2]
    P.append(p.reshape(m, 9))
    Q.append(q.reshape(m, 9))
    # degenerate (zero-area / repeated vertices) and identical triangles
    m = 500
    p = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    q = r.uniform(m * 9, -1.0, 1.0).reshape(m, 3, 3)
    p[: m // 3, 2] = p[: m // 3, 1]
    q[m // 3: 2 * m // 3, 1] = q[m // 3: 2 * m // 3, 0]
    q[2 * m // 3:] = p[2 * m // 3:]
    P.append(p.reshape(m, 9))
    Q.append(q.reshape(m, 9))
    # edge-on-edge crossings with tiny separations (sign tests near zero)
    m = 1500
    a = r.uniform(m * 3, -1.0, 1.0).reshape(m, 3)
    b = r.uniform(m * 3, -1.0, 1.0).reshape(m, 3)
    eps = (10.0 ** r.uniform(m, -8.0, -3.0)) * np.where(r.u01(m) < 0.5, -1.0, 1.0)
    p = np.stack([a, b, a + r.uniform(m * 3, -1, 1).reshape(m, 3)], 1)
    mid = 0.5 * (a + b)
    nrm = np.cross(b - a, p[:, 2] - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    q0 = mid + nrm * eps[:, None]
    q = np.stack([q0, q0 + r.uniform(m * 3, -1, 1).reshape(m, 3), q0 + r.uniform(m * 3, -1, 1).reshape(m, 3)], 1)
    P.append(p.reshape(m, 9))
    Q.append(q.reshape(m, 9))
    return (np.concatenate(P).astype(np.float32), np.concatenate(Q).astype(np.float32))


def make_sat(ref):
    P, Q = sat_cases()
    n = P.shape[0]
    out = np.zeros(n, np.uint8)
    ref.ref_tri_intersect_many(_fp(P), _fp(Q), n, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    f64 = np.array([oracle.tri_intersect(P[i], Q[i], fp64=True) for i in range(n)], np.uint8)
    np.savez_compressed(os.path.join(OUT, "sat_ref.npz"), P=P, Q=Q, ref=out, fp64=f64)
    print(f"sat_ref.npz: {n} pairs, {int(out.sum())} hits, {int((out != f64).sum())} fp32/fp64 differences")


def make_aabb(ref):
    r = M.Rng(0xAABB)
    n = 2000
    # values with many ties and signed zeros
    pool = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0, 1e-30, -1e-30, 3.4e38, -3.4e38], np.float32)
    idx = r.integers(n * 9, 0, pool.size - 1).reshape(n, 9)
    V = pool[idx].astype(np.float32)
    V[n // 2:] = r.uniform((n - n // 2) * 9, -2, 2).reshape(-1, 9).astype(np.float32)
    leaf = np.zeros((n, 6), np.float32)
    for i in range(n):
        v = np.ascontiguousarray(V[i])
        o = np.zeros(6, np.float32)
        ref.ref_leaf_box(_fp(v[0:3].copy()), _fp(v[3:6].copy()), _fp(v[6:9].copy()), _fp(o))
        leaf[i] = o
    A = pool[r.integers(n * 6, 0, pool.size - 1).reshape(n, 6)].astype(np.float32)
    B = pool[r.integers(n * 6, 0, pool.size - 1).reshape(n, 6)].astype(np.float32)
    A[n // 2:] = r.uniform((n - n // 2) * 6, -1, 1).reshape(-1, 6).astype(np.float32)
    B[n // 2:] = r.uniform((n - n // 2) * 6, -1, 1).reshape(-1, 6).astype(np.float32)
    # well-formed boxes (min <= max by value; ties and signed zeros kept)
    for X in (A, B):
        lo = np.where(X[:, 3:] < X[:, :3], X[:, 3:], X[:, :3])
        hi = np.where(X[:, 3:] < X[:, :3], X[:, :3], X[:, 3:])
        X[:, :3], X[:, 3:] = lo, hi
    merged = np.zeros((n, 6), np.float32)
    ov = np.zeros(n, np.uint8)
    for i in range(n):
        a, b = np.ascontiguousarray(A[i]), np.ascontiguousarray(B[i])
        o = np.zeros(6, np.float32)
        ref.ref_merge(_fp(a), _fp(b), _fp(o))
        merged[i] = o
        ov[i] = ref.ref_overlap(_fp(a), _fp(b))
    # mesh box fold over point clouds with +-0 ties
    clouds = []
    boxes = []
    for k in range(20):
        m = 50 + 7 * k
        c = pool[r.integers(m * 3, 0, 5).reshape(m, 3)].astype(np.float32)
        o = np.zeros(6, np.float32)
        ref.ref_mesh_box(_fp(c), m, _fp(o))
        clouds.append(c)
        boxes.append(o)
    lens = np.array([c.shape[0] for c in clouds], np.int64)
    np.savez_compressed(os.path.join(OUT, "aabb_ref.npz"), V=V, leaf=leaf, A=A, B=B, merged=merged, overlap=ov,
                        cloud=np.concatenate(clouds), cloud_len=lens, cloud_box=np.stack(boxes))
    print(f"aabb_ref.npz: {n} leaf boxes, {n} merges, {len(clouds)} clouds")


def make_scenes():
    """Small scenes: two noisy icospheres (subdiv 2 and 3), a box pair, a
    folded cloth for self collision."""
    arrays = {}
    cases = [
        ("ico2", M.interpenetrating_icospheres(subdiv=2, seed=11, offset=1.5), False),
        ("ico3", M.interpenetrating_icospheres(subdiv=3, seed=12, offset=1.7), False),
        ("box", [M.box(6, noise=0.01, seed=3), tuple(x if i else (x + np.float32(0.7)).astype(np.float32)
                                                      for i, x in enumerate(M.box(5, noise=0.01, seed=4)))], False),
        ("cloth", [M.cloth(40, amplitude=0.2, wavelength=0.3, seed=5)], True),
    ]
    for name, meshes, self_c in cases:
        trees = [oracle.Tree(x, f, order="morton") for x, f in meshes]
        pairs = oracle.detect(trees, self_collide=self_c)
        pairs64 = oracle.detect(trees, self_collide=self_c, fp64=True)
        for k, (x, f) in enumerate(meshes):
            arrays[f"{name}_xyz{k}"] = x
            arrays[f"{name}_faces{k}"] = f
            arrays[f"{name}_nodes{k}"] = trees[k].nodes()
            arrays[f"{name}_perm{k}"] = trees[k].order
        arrays[f"{name}_pairs"] = pairs
        arrays[f"{name}_pairs64"] = pairs64
        arrays[f"{name}_self"] = np.array([self_c])
        arrays[f"{name}_n"] = np.array([len(meshes)])
        print(f"scenes.npz: {name}: {[m[1].shape[0] for m in meshes]} faces, {len(pairs)} pairs")
    arrays["names"] = np.array([c[0] for c in cases])
    np.savez_compressed(os.path.join(OUT, "scenes.npz"), **arrays)


def make_xform(ref, n_mat: int = 64, n_pts: int = 256, seed: int = 0x7F0):
    """Rigid motions with scale, shear and translation; points of mixed scale."""
    r = M.Rng(seed)
    mats = []
    for k in range(n_mat):
        R = M.random_rotation(r)
        S = np.diag(r.uniform(3, 0.25, 4.0)) if k % 3 == 0 else np.eye(3)
        A = np.eye(4)
        A[:3, :3] = R @ S
        if k % 5 == 1:
            A[0, 1] += r.uniform(1, -0.5, 0.5)[0]  # shear
        A[:3, 3] = r.uniform(3, -10.0, 10.0)
        mats.append(A.astype(np.float32))
    pts = r.uniform(n_pts * 3, -1.0, 1.0).reshape(n_pts, 3)
    pts[: n_pts // 4] *= 1e3
    pts[n_pts // 4: n_pts // 2] *= 1e-3
    pts = pts.astype(np.float32)
    out = np.empty((n_mat, n_pts, 3), np.float32)
    colmajor = np.stack([np.ascontiguousarray(m.T).reshape(16) for m in mats])
    for k in range(n_mat):
        ref.ref_xform(_fp(np.ascontiguousarray(colmajor[k])), _fp(pts), n_pts, _fp(out[k]))
    np.savez_compressed(os.path.join(OUT, "xform_ref.npz"), M=np.stack(mats), P=pts, out=out)
    print(f"xform_ref.npz: {n_mat} matrices x {n_pts} points")


if __name__ == "__main__":
    oracle.build()
    ref = oracle.load_ref()
    if ref is None:
        sys.exit("oracle/_ref/libref.so is not built (the reference tree must be mounted)")
    which = set(sys.argv[1:]) or {"sat", "aabb", "scenes", "xform"}
    if "sat" in which:
        make_sat(ref)
    if "aabb" in which:
        make_aabb(ref)
    if "scenes" in which:
        make_scenes()
    if "xform" in which:
        make_xform(ref)
