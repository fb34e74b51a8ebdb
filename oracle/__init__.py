"""ctypes front end of the CPU oracle (oracle/oibvh_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the product path (oibvh_amd) never does.

``load_ref()`` returns oracle/_ref/libref.so -- the reference's own
src/utils/utils.cpp compiled from the reference tree -- or None where that
build is absent (a machine with the repository but not the reference tree
runs the checks without it).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libref.so")

SELF = 1
FP64 = 2
BROAD = 4

_lib = None
_ref = None


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (and _ref/libref.so when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference/src")
    if ref:
        subprocess.run(["make", "-s", "-C", _HERE, "-B", "ref"], check=True)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build(ref=False)
        L = C.CDLL(_LIB)
        u32, f32p, u32p = C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        vp = C.c_void_p
        for name, res, args in [
            ("orc_next_pow2", u32, [u32]),
            ("orc_ilog2", u32, [u32]),
            ("orc_tree_size", u32, [u32]),
            ("orc_level_real_count", u32, [u32, u32, u32]),
            ("orc_implicit_to_real", u32, [u32, u32, u32]),
            ("orc_real_to_implicit", u32, [u32, u32, u32]),
            ("orc_box_overlap", C.c_int, [f32p, f32p]),
            ("orc_mesh_box", None, [f32p, u32, f32p]),
            ("orc_xform", None, [f32p, f32p, u32, f32p]),
            ("orc_quantise10", u32, [C.c_float]),
            ("orc_morton30", u32, [C.c_float, C.c_float, C.c_float]),
            ("orc_quantise21", u32, [C.c_float]),
            ("orc_morton63", C.c_uint64, [C.c_float, C.c_float, C.c_float]),
            ("orc_leaf_keys63", None, [u32p, f32p, u32, f32p, C.POINTER(C.c_uint64)]),
            ("orc_stable_order64", None, [C.POINTER(C.c_uint64), u32, u32p]),
            ("orc_leaf_keys", None, [u32p, f32p, u32, f32p, u32p, f32p]),
            ("orc_stable_order", None, [u32p, u32, u32p]),
            ("orc_tree_build", vp, [u32p, u32, f32p, u32, u32p]),
            ("orc_tree_free", None, [vp]),
            ("orc_tree_node_count", u32, [vp]),
            ("orc_tree_depth", u32, [vp]),
            ("orc_tree_nodes", None, [vp, f32p]),
            ("orc_tree_topology", None, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
            ("orc_tree_refit", None, [vp, f32p]),
            ("orc_tri_intersect", C.c_int, [f32p, f32p]),
            ("orc_tri_intersect_f64", C.c_int, [f32p, f32p]),
            ("orc_detect", C.c_uint64, [C.POINTER(vp), u32, u32, C.POINTER(u32p),
                                        C.POINTER(C.c_uint64)]),
            ("orc_brute_force", C.c_uint64, [C.POINTER(vp), u32, u32, C.POINTER(u32p)]),
            ("orc_free", None, [vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def load_ref():
    """The reference's own SAT/AABB code (oracle/_ref/libref.so) or None."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        R = C.CDLL(_REF)
        f32p = C.POINTER(C.c_float)
        R.ref_tri_intersect.argtypes = [f32p, f32p]
        R.ref_tri_intersect.restype = C.c_int
        R.ref_tri_intersect_many.argtypes = [f32p, f32p, C.c_uint64, C.POINTER(C.c_uint8)]
        R.ref_leaf_box.argtypes = [f32p, f32p, f32p, f32p]
        R.ref_merge.argtypes = [f32p, f32p, f32p]
        R.ref_overlap.argtypes = [f32p, f32p]
        R.ref_overlap.restype = C.c_int
        R.ref_mesh_box.argtypes = [f32p, C.c_uint32, f32p]
        R.ref_xform.argtypes = [f32p, f32p, C.c_uint32, f32p]
        _ref = R
    return _ref


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- scalars
def next_pow2(x: int) -> int:
    return lib().orc_next_pow2(x)


def tree_size(F: int) -> int:
    return lib().orc_tree_size(F)


def level_real_count(l: int, L: int, vl: int) -> int:
    return lib().orc_level_real_count(l, L, vl)


def implicit_to_real(i: int, L: int, vl: int) -> int:
    return lib().orc_implicit_to_real(i, L, vl)


def real_to_implicit(r: int, L: int, vl: int) -> int:
    return lib().orc_real_to_implicit(r, L, vl)


def quantise21(p: float) -> int:
    return lib().orc_quantise21(p)


def morton63(x: float, y: float, z: float) -> int:
    return lib().orc_morton63(x, y, z)


def leaf_keys63(faces, xyz, box=None) -> np.ndarray:
    """63-bit Morton keys of the leaf-box centroids (no reference counterpart)."""
    f = _u32(faces).reshape(-1, 3)
    x = _f32(xyz).reshape(-1, 3)
    b = _f32(mesh_box(x) if box is None else box)
    out = np.empty(f.shape[0], np.uint64)
    lib().orc_leaf_keys63(_p(f, C.c_uint32), _p(x, C.c_float), f.shape[0], _p(b, C.c_float),
                          _p(out, C.c_uint64))
    return out


def stable_order64(keys) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64)
    out = np.empty(k.size, np.uint32)
    lib().orc_stable_order64(_p(k, C.c_uint64), k.size, _p(out, C.c_uint32))
    return out


def xform(mat4, xyz) -> np.ndarray:
    """Mesh::transform's vertex map: mat4 @ (x, y, z, 1) in glm's evaluation
    order, FP32 without contraction. mat4 is a 4x4 matrix (row, column)."""
    m = _f32(np.asarray(mat4, np.float32).reshape(4, 4).T)  # column-major
    x = _f32(xyz).reshape(-1, 3)
    out = np.empty_like(x)
    lib().orc_xform(m.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)), x.shape[0],
                    out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def quantise10(p: float) -> int:
    return lib().orc_quantise10(p)


def morton30(x: float, y: float, z: float) -> int:
    return lib().orc_morton30(x, y, z)


# ---------------------------------------------------------------- arrays
def mesh_box(xyz) -> np.ndarray:
    xyz = _f32(xyz)
    out = np.empty(6, np.float32)
    lib().orc_mesh_box(_p(xyz, C.c_float), xyz.shape[0], _p(out, C.c_float))
    return out


def leaf_keys(faces, xyz, box=None, with_leaves: bool = False):
    faces, xyz = _u32(faces), _f32(xyz)
    box = mesh_box(xyz) if box is None else _f32(box)
    F = faces.shape[0]
    keys = np.empty(F, np.uint32)
    leaves = np.empty((F, 6), np.float32) if with_leaves else None
    lib().orc_leaf_keys(_p(faces, C.c_uint32), _p(xyz, C.c_float), F, _p(box, C.c_float),
                        _p(keys, C.c_uint32), _p(leaves, C.c_float) if with_leaves else None)
    return (keys, leaves) if with_leaves else keys


def stable_order(keys) -> np.ndarray:
    keys = _u32(keys)
    perm = np.empty(keys.size, np.uint32)
    lib().orc_stable_order(_p(keys, C.c_uint32), keys.size, _p(perm, C.c_uint32))
    return perm


def tri_intersect(P, Q, fp64: bool = False) -> bool:
    P, Q = _f32(P).ravel(), _f32(Q).ravel()
    f = lib().orc_tri_intersect_f64 if fp64 else lib().orc_tri_intersect
    return bool(f(_p(P, C.c_float), _p(Q, C.c_float)))


class Tree:
    """SimpleBVH-shaped tree (simpleBVH.cpp:122-160) over a face order.

    order=None builds over the mesh's own face order, exactly like src/cpu;
    order="morton" builds over the Morton-stable order the oibvh GPU build
    uses (oibvhTree.cu:287-297), which makes the node AABBs comparable bit for
    bit with the oibvh tree."""

    def __init__(self, xyz, faces, order=None, box=None):
        self.xyz = _f32(xyz)
        self.faces = _u32(faces)
        F = self.faces.shape[0]
        if isinstance(order, str) and order == "morton":
            order = stable_order(leaf_keys(self.faces, self.xyz, box))
        elif isinstance(order, str) and order == "morton63":
            order = stable_order64(leaf_keys63(self.faces, self.xyz, box))
        self.order = None if order is None else _u32(order)
        h = lib().orc_tree_build(_p(self.faces, C.c_uint32), F, _p(self.xyz, C.c_float),
                                 self.xyz.shape[0],
                                 None if self.order is None else _p(self.order, C.c_uint32))
        if not h:
            raise ValueError("empty mesh")
        self.h = C.c_void_p(h)
        self.F = F

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tree_free(self.h)
            self.h = None

    @property
    def node_count(self) -> int:
        return lib().orc_tree_node_count(self.h)

    @property
    def depth(self) -> int:
        return lib().orc_tree_depth(self.h)

    def nodes(self) -> np.ndarray:
        out = np.empty((self.node_count, 6), np.float32)
        lib().orc_tree_nodes(self.h, _p(out, C.c_float))
        return out

    def topology(self):
        n = self.node_count
        tri = np.empty(n, np.int32)
        nch = np.empty(n, np.int32)
        lib().orc_tree_topology(self.h, _p(tri, C.c_int32), _p(nch, C.c_int32))
        return tri, nch

    def refit(self, xyz) -> None:
        self.xyz = _f32(xyz)
        lib().orc_tree_refit(self.h, _p(self.xyz, C.c_float))


def _pairs(fn, trees, flags, *extra):
    arr = (C.c_void_p * len(trees))(*[t.h.value for t in trees])
    out = C.POINTER(C.c_uint32)()
    n = fn(arr, len(trees), flags, C.byref(out), *extra)
    if n:
        res = np.ctypeslib.as_array(out, shape=(n, 4)).copy()
    else:
        res = np.zeros((0, 4), np.uint32)
    lib().orc_free(C.cast(out, C.c_void_p))
    return sort_pairs(res)


def sort_pairs(p: np.ndarray) -> np.ndarray:
    """Lexicographic (meshA, meshB, triA, triB) order."""
    p = np.asarray(p, dtype=np.uint32).reshape(-1, 4)
    if p.shape[0] == 0:
        return p
    idx = np.lexsort((p[:, 3], p[:, 2], p[:, 1], p[:, 0]))
    return p[idx]


def detect(trees, self_collide: bool = False, fp64: bool = False, broad_only: bool = False,
           stats: bool = False):
    """SimpleCollide::detect (simpleCollide.cpp:47-161) -> sorted (n, 4) pairs."""
    flags = (SELF if self_collide else 0) | (FP64 if fp64 else 0) | (BROAD if broad_only else 0)
    st = (C.c_uint64 * 2)()
    res = _pairs(lib().orc_detect, trees, flags, st)
    return (res, (int(st[0]), int(st[1]))) if stats else res


def brute_force(trees, self_collide: bool = False, fp64: bool = False, broad_only: bool = False):
    flags = (SELF if self_collide else 0) | (FP64 if fp64 else 0) | (BROAD if broad_only else 0)
    return _pairs(lib().orc_brute_force, trees, flags)
