"""Python mirror of the reference's OibvhTree / Scene interface over the C ABI.

Names and argument meanings follow include/cuda/oibvhTree.cuh:41-174 and
include/cuda/scene.cuh:26-182; errors raise OibvhError instead of asserting.
Mesh data are numpy arrays: xyz float32 (V, 3), faces uint32 (F, 3).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L


def _check(ctx_h, st: int, what: str) -> None:
    if st != L.OK:
        msg = L.load().oibvh_last_error(ctx_h) if ctx_h else b""
        raise L.OibvhError(st, f"{what}: {msg.decode() if msg else ''}")


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class Context:
    """Device + stream that every tree and scene of this context uses."""

    def __init__(self, device: int = 0):
        self._lib = L.load()
        h = C.c_void_p()
        st = self._lib.oibvh_ctx_create(device, None, C.byref(h))
        if st != L.OK:
            raise L.OibvhError(st, f"oibvh_ctx_create(device={device})")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self._lib.oibvh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def stream(self) -> int:
        return self._lib.oibvh_ctx_stream(self.h) or 0

    def synchronize(self) -> None:
        _check(self.h, self._lib.oibvh_ctx_synchronize(self.h), "synchronize")

    def set_tuning(self, **kw) -> None:
        """Set tuning knobs (oibvh_tuning field names); unspecified ones keep
        their current values."""
        t = self.get_tuning()
        for k, v in kw.items():
            if k not in t:
                raise KeyError(k)
            t[k] = v
        s = L.Tuning(**t)
        _check(self.h, self._lib.oibvh_ctx_set_tuning(self.h, C.byref(s)), "set_tuning")

    def get_tuning(self) -> dict:
        s = L.Tuning()
        _check(self.h, self._lib.oibvh_ctx_get_tuning(self.h, C.byref(s)), "get_tuning")
        return {f: getattr(s, f) for f, _ in L.Tuning._fields_}


class DeviceBuffer:
    """Plain hipMalloc'd buffer (harness plumbing: device-resident inputs)."""

    def __init__(self, nbytes: int):
        H = L.hip()
        p = C.c_void_p()
        L.hip_check(H.hipMalloc(C.byref(p), max(1, nbytes)), "hipMalloc")
        self.ptr = p
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        L.hip_check(L.hip().hipMemcpy(b.ptr, _ptr(a), a.nbytes, 1), "hipMemcpy H2D")
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        L.hip_check(L.hip().hipMemcpy(C.c_void_p(self.ptr.value + offset), _ptr(a), a.nbytes, 1), "hipMemcpy H2D")

    def at(self, offset: int) -> int:
        return self.ptr.value + offset

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype)
        if count:
            L.hip_check(L.hip().hipMemcpy(_ptr(out), C.c_void_p(self.ptr.value + offset), out.nbytes, 2),
                        "hipMemcpy D2H")
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None) and self.ptr.value:
            L.hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        self.free()


class OibvhTree:
    """OibvhTree (oibvhTree.cuh:41-174)."""

    def __init__(self, ctx: Context, xyz=None, faces=None, *, _handle=None, _V=0, _F=0):
        self.ctx = ctx
        self._lib = ctx._lib
        if _handle is not None:
            self.h, self.V, self.F = _handle, _V, _F
            return
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        _check(ctx.h, self._lib.oibvh_tree_create(ctx.h, _ptr(xyz), xyz.shape[0], _ptr(faces), faces.shape[0],
                                                  L.PTR_HOST, C.byref(h)), "oibvh_tree_create")
        self.h, self.V, self.F = h, xyz.shape[0], faces.shape[0]

    @classmethod
    def like(cls, other: "OibvhTree", xyz) -> "OibvhTree":
        """OibvhTree(shared_ptr<OibvhTree> other, shared_ptr<Mesh>) (oibvhTree.cu:17-33)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        h = C.c_void_p()
        _check(other.ctx.h, other._lib.oibvh_tree_create_like(other.h, _ptr(xyz), xyz.shape[0], L.PTR_HOST,
                                                              C.byref(h)), "oibvh_tree_create_like")
        return cls(other.ctx, _handle=h, _V=other.V, _F=other.F)

    def close(self) -> None:
        if getattr(self, "h", None):
            self._lib.oibvh_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def build(self) -> None:
        _check(self.ctx.h, self._lib.oibvh_tree_build(self.h), "build")

    def refit(self, xyz=None) -> None:
        """refit() with the current positions, or with new host positions."""
        if xyz is None:
            _check(self.ctx.h, self._lib.oibvh_tree_refit(self.h, None, L.PTR_HOST), "refit")
        else:
            xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
            if xyz.shape[0] != self.V:
                raise ValueError("vertex count changed")
            _check(self.ctx.h, self._lib.oibvh_tree_refit(self.h, _ptr(xyz), L.PTR_HOST), "refit")

    def set_positions_device(self, dptr: int, borrow: bool = True) -> None:
        kind = L.PTR_DEVICE_BORROW if borrow else L.PTR_DEVICE
        _check(self.ctx.h, self._lib.oibvh_tree_set_positions(self.h, C.c_void_p(dptr), kind), "set_positions")

    def set_positions(self, xyz) -> None:
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        if xyz.shape[0] != self.V:
            raise ValueError("vertex count changed")
        _check(self.ctx.h, self._lib.oibvh_tree_set_positions(self.h, _ptr(xyz), L.PTR_HOST), "set_positions")

    def set_transform(self, mat4) -> None:
        """Mesh::transform: every later build, refit and detect sees each
        vertex as mat4 @ (x, y, z, 1) (glm evaluation order). mat4 is a 4x4
        matrix in the usual row/column sense (stored column-major like
        glm::mat4 when passed down); None restores the identity."""
        if mat4 is None:
            _check(self.ctx.h, self._lib.oibvh_tree_set_transform(self.h, None), "set_transform")
            return
        m = np.ascontiguousarray(np.asarray(mat4, np.float32).reshape(4, 4).T)  # column-major
        _check(self.ctx.h, self._lib.oibvh_tree_set_transform(self.h, m.ctypes.data_as(L._f32p)), "set_transform")

    def set_morton_bits(self, bits: int) -> None:
        """Key width of later builds: 30 (morton3D, oibvh.cu:6-22) or 63."""
        _check(self.ctx.h, self._lib.oibvh_tree_set_morton_bits(self.h, bits), "set_morton_bits")

    def refit_device(self, dptr: int, borrow: bool = True) -> None:
        kind = L.PTR_DEVICE_BORROW if borrow else L.PTR_DEVICE
        _check(self.ctx.h, self._lib.oibvh_tree_refit(self.h, C.c_void_p(dptr), kind), "refit")

    def _u32(self, fn) -> int:
        v = C.c_uint32()
        _check(self.ctx.h, fn(self.h, C.byref(v)), fn.__name__)
        return v.value

    def getDepth(self) -> int:
        return self._u32(self._lib.oibvh_tree_get_depth)

    def getPrimCount(self) -> int:
        return self._u32(self._lib.oibvh_tree_get_prim_count)

    @property
    def node_count(self) -> int:
        return self._u32(self._lib.oibvh_tree_get_node_count)

    def nodes(self) -> np.ndarray:
        n = self.node_count
        out = np.empty((n, 6), np.float32)
        got = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_tree_get_nodes(self.h, out.ctypes.data_as(L._f32p), n, C.byref(got)),
               "get_nodes")
        return out

    def sorted_faces(self) -> np.ndarray:
        out = np.empty((self.F, 3), np.uint32)
        _check(self.ctx.h, self._lib.oibvh_tree_get_sorted_faces(self.h, out.ctypes.data_as(L._u32p), self.F),
               "get_sorted_faces")
        return out

    def perm(self) -> np.ndarray:
        out = np.empty(self.F, np.uint32)
        _check(self.ctx.h, self._lib.oibvh_tree_get_perm(self.h, out.ctypes.data_as(L._u32p), self.F), "get_perm")
        return out

    def positions(self) -> np.ndarray:
        out = np.empty((self.V, 3), np.float32)
        _check(self.ctx.h, self._lib.oibvh_tree_get_positions(self.h, out.ctypes.data_as(L._f32p), 3 * self.V),
               "get_positions")
        return out

    def positions_device(self) -> int:
        p = C.c_void_p()
        _check(self.ctx.h, self._lib.oibvh_tree_positions_device(self.h, C.byref(p)), "positions_device")
        return p.value




def load_obj(path: str, ctx: "Context | None" = None):
    """Model(path) (model.cpp:40-186) reduced to the collision path: (xyz
    (V, 3) float32, faces (F, 3) uint32) of a Wavefront OBJ file, read by the
    library's native reader (oibvh_obj_read). Host only."""
    lib = L.load()
    h = ctx.h if ctx is not None else None
    V, F = C.c_uint32(), C.c_uint32()
    bpath = os.fsencode(path)
    st = lib.oibvh_obj_read(h, bpath, None, 0, None, 0, C.byref(V), C.byref(F))
    if st not in (L.OK, L.ERR_CAPACITY):
        _check(h, st, f"load_obj({path})")
    xyz = np.empty((V.value, 3), np.float32)
    faces = np.empty((F.value, 3), np.uint32)
    _check(h, lib.oibvh_obj_read(h, bpath, xyz.ctypes.data_as(L._f32p), xyz.size, faces.ctypes.data_as(L._u32p),
                                 faces.size, C.byref(V), C.byref(F)), f"load_obj({path})")
    return xyz, faces

def _pair_vertices(obj, fn, n: int) -> np.ndarray:
    out = np.empty((max(n, 1), 6, 3), np.float32)
    got = C.c_uint32()
    _check(obj.ctx.h, fn(obj.h, out.ctypes.data, out.size, L.PTR_HOST, C.byref(got)), "pair_vertices")
    return out[: got.value]


def download_u32(dptr: int, count: int) -> np.ndarray:
    """count uint32 words from device memory (harness plumbing)."""
    out = np.empty(count, np.uint32)
    if count:
        L.hip_check(L.hip().hipMemcpy(_ptr(out), C.c_void_p(dptr), out.nbytes, 2), "hipMemcpy D2H")
    return out


class Scene:
    """Scene (scene.cuh:26-182)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _check(ctx.h, self._lib.oibvh_scene_create(ctx.h, C.byref(h)), "oibvh_scene_create")
        self.h = h
        self.trees: list[OibvhTree] = []

    def close(self) -> None:
        if getattr(self, "h", None):
            self._lib.oibvh_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def addOibvhTree(self, tree: OibvhTree) -> None:
        _check(self.ctx.h, self._lib.oibvh_scene_add_tree(self.h, tree.h), "addOibvhTree")
        self.trees.append(tree)

    def detectCollision(self, entry_level: int = 0, expand_levels: int = 1, *, self_collide: bool = False,
                        fp64_fallback: bool = False, asynchronous: bool = False) -> None:
        flags = ((L.DETECT_SELF if self_collide else 0) | (L.DETECT_FP64_FALLBACK if fp64_fallback else 0)
                 | (L.DETECT_ASYNC if asynchronous else 0))
        _check(self.ctx.h, self._lib.oibvh_scene_detect(self.h, entry_level, expand_levels, flags), "detectCollision")

    def refit(self) -> None:
        """Refit every tree of the scene in one launch per stage."""
        _check(self.ctx.h, self._lib.oibvh_scene_refit(self.h), "scene refit")

    def build(self) -> None:
        """Rebuild every tree of the scene (oibvh_scene_build): up to four trees
        share each launch of the build chain."""
        _check(self.ctx.h, self._lib.oibvh_scene_build(self.h), "scene build")

    def set_transforms(self, mats=None, *, device: int | None = None, borrow: bool = False) -> None:
        """Model matrices of every tree (addOibvhTree order): `mats` is an
        (n, 4, 4) array in the usual row/column sense (host memory), or
        `device` the address of n column-major matrices in device memory,
        copied (borrow=False) or read in place (borrow=True)."""
        if device is not None:
            kind = L.PTR_DEVICE_BORROW if borrow else L.PTR_DEVICE
            _check(self.ctx.h, self._lib.oibvh_scene_set_transforms(self.h, C.c_void_p(device), kind),
                   "set_transforms")
            return
        m = np.ascontiguousarray(np.swapaxes(np.asarray(mats, np.float32).reshape(-1, 4, 4), -1, -2))
        _check(self.ctx.h, self._lib.oibvh_scene_set_transforms(self.h, _ptr(m), L.PTR_HOST), "set_transforms")

    def getIntTriPairCount(self) -> int:
        v = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_scene_get_pair_count(self.h, C.byref(v)), "getIntTriPairCount")
        return v.value

    def getIntTriPairs(self) -> np.ndarray:
        """(n, 4) uint32 rows (meshA, meshB, origTriA, origTriB), sorted."""
        n = self.getIntTriPairCount()
        out = np.empty((max(n, 1), 4), np.uint32)
        got = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_scene_get_pairs(self.h, out.ctypes.data_as(L._u32p), max(n, 1),
                                                           C.byref(got)), "get_pairs")
        return out[: got.value]

    def frame(self, positions=None, entry_level: int = 0, expand_levels: int = 1, *, self_collide: bool = False,
              fp64_fallback: bool = False, asynchronous: bool = False, overlap: bool = False) -> None:
        """One frame: tree i borrows device vertex memory positions[i] (an int
        device address, or None to keep its positions), every tree is refit,
        then detectCollision (oibvh_scene_frame). overlap: this frame's refit
        may run while the previous frame's detect does (OIBVH_FRAME_OVERLAP;
        keep each frame's position memory unchanged until its results are
        read)."""
        arr = positions if isinstance(positions, C.Array) else None
        if positions is not None and arr is None:
            arr = position_array(positions)
        flags = _flags(self_collide, fp64_fallback, asynchronous) | (L.FRAME_OVERLAP if overlap else 0)
        _check(self.ctx.h, self._lib.oibvh_scene_frame(self.h, arr, entry_level, expand_levels, flags), "frame")

    def set_timing(self, every: int = 1) -> None:
        """Time every `every`-th refit and detect with HIP events (0 stops)."""
        _check(self.ctx.h, self._lib.oibvh_scene_set_timing(self.h, int(every)), "set_timing")

    def timing(self) -> dict:
        """Summed event durations since the last call: refit_ms (stage-1
        reduce launches), detect_ms, and their counts."""
        t = L.SceneTiming()
        _check(self.ctx.h, self._lib.oibvh_scene_get_timing(self.h, C.byref(t)), "get_timing")
        return {f: getattr(t, f) for f, _ in L.SceneTiming._fields_}

    def pair_vertices(self) -> np.ndarray:
        """Scene::convertToVertexArray: (n, 6, 3) float32, triangle A's three
        vertices then triangle B's, in getIntTriPairs order."""
        return _pair_vertices(self, self._lib.oibvh_scene_get_pair_vertices, self.getIntTriPairCount())

    def pairs_device(self) -> tuple[int, int]:
        """(device address, n) of the pairs in getIntTriPairs order, sorted on
        the device (4 uint32 each; valid until the next detect)."""
        p, n = C.c_void_p(), C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_scene_pairs_device(self.h, C.byref(p), C.byref(n)), "pairs_device")
        return p.value or 0, n.value

    def pair_vertices_device(self, dptr: int, cap_floats: int) -> int:
        """The same into device memory (18 floats per pair); returns n."""
        got = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_scene_get_pair_vertices(self.h, dptr, cap_floats, L.PTR_DEVICE,
                                                                   C.byref(got)), "pair_vertices")
        return got.value

    def level_counts(self) -> np.ndarray:
        """BVTT front length per level of the last detect."""
        out = np.zeros(64, np.uint32)
        n = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_scene_get_level_counts(self.h, out.ctypes.data_as(L._u32p), out.size,
                                                                  C.byref(n)), "get_level_counts")
        return out[: n.value].copy()

    def stats(self) -> dict:
        s = L.DetectStats()
        _check(self.ctx.h, self._lib.oibvh_scene_get_stats(self.h, C.byref(s)), "get_stats")
        return {f: getattr(s, f) for f, _ in L.DetectStats._fields_}


def position_array(positions):
    """The ctypes array of device addresses Scene.frame takes (int or None
    per tree); building it once per frame buffer saves the conversion on
    every call."""
    return (C.c_void_p * len(positions))(*[C.c_void_p(p) if p else None for p in positions])


def _flags(self_collide: bool, fp64_fallback: bool, asynchronous: bool) -> int:
    return ((L.DETECT_SELF if self_collide else 0) | (L.DETECT_FP64_FALLBACK if fp64_fallback else 0)
            | (L.DETECT_ASYNC if asynchronous else 0))


class Batch:
    """Many independent scenes built, refit and queried together
    (oibvh_batch_*). Meshes of one scene collide with each other, never with
    other scenes' meshes."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _check(ctx.h, self._lib.oibvh_batch_create(ctx.h, C.byref(h)), "oibvh_batch_create")
        self.h = h
        self.scene_of: list[int] = []
        self.faces_of: list[int] = []

    def close(self) -> None:
        if getattr(self, "h", None):
            self._lib.oibvh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def add_mesh(self, scene: int, xyz, faces) -> int:
        x = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, np.uint32).reshape(-1, 3)
        idx = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_add_mesh(self.h, scene, x.ctypes.data, x.shape[0], f.ctypes.data,
                                                          f.shape[0], C.byref(idx)), "add_mesh")
        self.scene_of.append(scene)
        self.faces_of.append(f.shape[0])
        return idx.value

    def add_instance(self, scene: int, source: int) -> int:
        idx = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_add_instance(self.h, scene, source, C.byref(idx)), "add_instance")
        self.scene_of.append(scene)
        self.faces_of.append(self.faces_of[source])
        return idx.value

    @property
    def mesh_count(self) -> int:
        return len(self.scene_of)

    def set_transforms(self, mats) -> None:
        """mats: (n, 4, 4) matrices (row, column); stored column-major."""
        m = np.ascontiguousarray(np.asarray(mats, np.float32).reshape(-1, 4, 4).transpose(0, 2, 1))
        _check(self.ctx.h, self._lib.oibvh_batch_set_transforms(self.h, m.ctypes.data, L.PTR_HOST), "set_transforms")

    def set_morton_bits(self, bits: int) -> None:
        """Key width (30 or 63) of the next build of every mesh."""
        _check(self.ctx.h, self._lib.oibvh_batch_set_morton_bits(self.h, bits), "set_morton_bits")

    def set_transforms_device(self, dptr: int) -> None:
        """n x 16 column-major floats already in device memory."""
        _check(self.ctx.h, self._lib.oibvh_batch_set_transforms(self.h, dptr, L.PTR_DEVICE), "set_transforms")

    def positions_device(self, mesh: int) -> int:
        p = C.c_void_p()
        _check(self.ctx.h, self._lib.oibvh_batch_positions_device(self.h, mesh, C.byref(p)), "positions_device")
        return p.value

    def set_positions(self, mesh: int, xyz) -> None:
        x = np.ascontiguousarray(xyz, np.float32)
        _check(self.ctx.h, self._lib.oibvh_batch_set_positions(self.h, mesh, x.ctypes.data, L.PTR_HOST),
               "batch set_positions")

    def set_positions_device(self, mesh: int, dptr: int, borrow: bool = True) -> None:
        kind = L.PTR_DEVICE_BORROW if borrow else L.PTR_DEVICE
        _check(self.ctx.h, self._lib.oibvh_batch_set_positions(self.h, mesh, dptr, kind), "batch set_positions")

    def build(self) -> None:
        _check(self.ctx.h, self._lib.oibvh_batch_build(self.h), "batch build")

    def refit(self) -> None:
        _check(self.ctx.h, self._lib.oibvh_batch_refit(self.h), "batch refit")

    def detect(self, entry_level: int = 0, expand_levels: int = 1, *, self_collide: bool = False,
               fp64_fallback: bool = False, asynchronous: bool = False) -> None:
        _check(self.ctx.h, self._lib.oibvh_batch_detect(self.h, entry_level, expand_levels,
                                                        _flags(self_collide, fp64_fallback, asynchronous)),
               "batch detect")

    def frame(self, entry_level: int = 0, expand_levels: int = 1, *, self_collide: bool = False,
              fp64_fallback: bool = False, asynchronous: bool = False, overlap: bool = False) -> None:
        """refit + detect replayed from a captured hipGraph. overlap: this
        frame's refit may run while the previous frame's detect does
        (OIBVH_FRAME_OVERLAP)."""
        flags = _flags(self_collide, fp64_fallback, asynchronous) | (L.FRAME_OVERLAP if overlap else 0)
        _check(self.ctx.h, self._lib.oibvh_batch_frame(self.h, entry_level, expand_levels, flags), "batch frame")

    def set_timing(self, every: int = 1) -> None:
        """Time every `every`-th refit (its stage-1 launch) and detect with
        HIP events (0 stops)."""
        _check(self.ctx.h, self._lib.oibvh_batch_set_timing(self.h, int(every)), "batch set_timing")

    def timing(self) -> dict:
        t = L.SceneTiming()
        _check(self.ctx.h, self._lib.oibvh_batch_get_timing(self.h, C.byref(t)), "batch get_timing")
        return {f: getattr(t, f) for f, _ in L.SceneTiming._fields_}

    def pair_count(self) -> int:
        v = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_get_pair_count(self.h, C.byref(v)), "batch pair count")
        return v.value

    def pairs(self) -> np.ndarray:
        """(n, 4) uint32 rows (meshA, meshB, origTriA, origTriB), batch mesh
        indices, sorted."""
        n = self.pair_count()
        out = np.empty((max(n, 1), 4), np.uint32)
        got = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_get_pairs(self.h, out.ctypes.data_as(L._u32p), max(n, 1),
                                                           C.byref(got)), "batch pairs")
        return out[: got.value]

    def pairs_device(self) -> tuple[int, int]:
        """(device address, n) of pairs() sorted on the device."""
        p, n = C.c_void_p(), C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_pairs_device(self.h, C.byref(p), C.byref(n)), "pairs_device")
        return p.value or 0, n.value

    def pair_vertices(self) -> np.ndarray:
        """(n, 6, 3) float32 triangle vertices of every pair, pairs() order."""
        return _pair_vertices(self, self._lib.oibvh_batch_get_pair_vertices, self.pair_count())

    def stats(self) -> dict:
        s = L.DetectStats()
        _check(self.ctx.h, self._lib.oibvh_batch_get_stats(self.h, C.byref(s)), "batch stats")
        return {f: getattr(s, f) for f, _ in L.DetectStats._fields_}

    def level_counts(self) -> np.ndarray:
        out = np.zeros(256, np.uint32)
        n = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_get_level_counts(self.h, out.ctypes.data_as(L._u32p), out.size,
                                                                  C.byref(n)), "batch level counts")
        return out[: n.value].copy()

    def nodes(self, mesh: int) -> np.ndarray:
        F = self.faces_of[mesh]
        N = 2 * F + bin((1 << (F - 1).bit_length()) - F).count("1") - 1
        out = np.empty((N, 6), np.float32)
        n = C.c_uint32()
        _check(self.ctx.h, self._lib.oibvh_batch_get_nodes(self.h, mesh, out.ctypes.data_as(L._f32p), N,
                                                           C.byref(n)), "batch nodes")
        return out[: n.value]

    def perm(self, mesh: int) -> np.ndarray:
        F = self.faces_of[mesh]
        out = np.empty(F, np.uint32)
        _check(self.ctx.h, self._lib.oibvh_batch_get_perm(self.h, mesh, out.ctypes.data_as(L._u32p), F),
               "batch perm")
        return out


class Timer:
    """HIP events recorded on the library's stream."""

    def __init__(self, ctx: Context):
        H = L.hip()
        self.H = H
        self.stream = C.c_void_p(ctx.stream)
        self.a, self.b = C.c_void_p(), C.c_void_p()
        L.hip_check(H.hipEventCreate(C.byref(self.a)), "hipEventCreate")
        L.hip_check(H.hipEventCreate(C.byref(self.b)), "hipEventCreate")

    def start(self) -> None:
        L.hip_check(self.H.hipEventRecord(self.a, self.stream), "hipEventRecord")

    def stop(self) -> None:
        L.hip_check(self.H.hipEventRecord(self.b, self.stream), "hipEventRecord")

    def ms(self) -> float:
        L.hip_check(self.H.hipEventSynchronize(self.b), "hipEventSynchronize")
        v = C.c_float()
        L.hip_check(self.H.hipEventElapsedTime(C.byref(v), self.a, self.b), "hipEventElapsedTime")
        return v.value

    def __del__(self):
        try:
            self.H.hipEventDestroy(self.a)
            self.H.hipEventDestroy(self.b)
        except Exception:
            pass
