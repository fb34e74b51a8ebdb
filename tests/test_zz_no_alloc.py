"""oibvh_build_noalloc (the build over caller-owned device buffers with its
scratch in the library's fixed device array) against oibvh_tree_build, the
entry point whose launches it shares: nodes, leaf permutation and sorted
faces bit for bit, on calls in a row whose sizes shrink (the array keeps the
previous call's contents), through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from oibvh_amd import _lib as L
from oibvh_amd import api
from oibvh_amd import meshgen as M

pytestmark = pytest.mark.gpu


def _node_count(F):
    np2 = 1 << max(0, (F - 1).bit_length())
    return 2 * F + bin(np2 - F).count("1") - 1


def _tree(ctx, xyz, faces):
    t = api.OibvhTree(ctx, xyz, faces)
    t.build()
    out = (t.nodes(), t.perm(), t.sorted_faces())
    t.close()
    return out


def _noalloc(ctx, xyz, faces):
    lib = L.load()
    xyz = np.ascontiguousarray(xyz, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    V, F = xyz.shape[0], faces.shape[0]
    N = _node_count(F)
    bx, bf = api.DeviceBuffer.from_array(xyz), api.DeviceBuffer.from_array(faces)
    bn, bp, bs = api.DeviceBuffer(24 * N), api.DeviceBuffer(4 * F), api.DeviceBuffer(12 * F)
    # outputs start as all-ones bits, so an unwritten word cannot pass
    bn.upload(np.full(6 * N, 0xFFFFFFFF, np.uint32))
    bp.upload(np.full(F, 0xFFFFFFFF, np.uint32))
    bs.upload(np.full(3 * F, 0xFFFFFFFF, np.uint32))
    st = lib.oibvh_build_noalloc(ctx.h, bx.ptr, V, bf.ptr, F, bn.ptr, bp.ptr, bs.ptr)
    assert st == L.OK, lib.oibvh_last_error(ctx.h)
    ctx.synchronize()
    out = (bn.download(np.float32, 6 * N).reshape(N, 6), bp.download(np.uint32, F),
           bs.download(np.uint32, 3 * F).reshape(F, 3))
    for b in (bx, bf, bn, bp, bs):
        b.free()
    return out


def _same(a, b):
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(np.ascontiguousarray(x).view(np.uint32), np.ascontiguousarray(y).view(np.uint32))


def test_noalloc_build_equals_tree_build(ctx):
    """Two calls in a row, the second smaller, then a third at about the
    maximum size (the wave-subtree reduce path), each equal to
    oibvh_tree_build on the same mesh."""
    big = M.icosphere(6, noise=0.02, seed=0x0B1B0001)  # 81,920 faces
    small = M.icosphere(3, noise=0.02, seed=0x0B1B0002)  # 1,280 faces
    for xyz, faces in (big, small):
        print(f"noalloc build: F={faces.shape[0]}", flush=True)
        _same(_noalloc(ctx, xyz, faces), _tree(ctx, xyz, faces))
    xyz, faces = M.uv_sphere(1000, 501)  # 1,000,000 faces
    print(f"noalloc build: F={faces.shape[0]}", flush=True)
    _same(_noalloc(ctx, xyz, faces), _tree(ctx, xyz, faces))
    # and small again after the large one
    _same(_noalloc(ctx, *small), _tree(ctx, *small))


def test_noalloc_build_odd_sizes(ctx):
    """Face counts around powers of two, the sort's 8192-key tiles and the
    maximum (the scratch's sizing)."""
    for F in (1, 2, 3, 255, 4095, 4096, 4097, 8191, 8192, 8193, 70001, L.NOALLOC_MAX_FACES):
        r = M.Rng(0x0B1B0100 + F)
        xyz = r.uniform(3 * (F + 2), -1, 1).reshape(-1, 3).astype(np.float32)
        faces = (np.arange(3 * F, dtype=np.uint32).reshape(F, 3) % np.uint32(xyz.shape[0]))
        print(f"noalloc build: F={F}", flush=True)
        _same(_noalloc(ctx, xyz, faces), _tree(ctx, xyz, faces))


def test_noalloc_rejects_too_many_faces(ctx):
    """Sizes beyond the maximum are refused before any device access."""
    lib = L.load()
    F = L.NOALLOC_MAX_FACES + 1
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    st = lib.oibvh_build_noalloc(ctx.h, dummy, 3, dummy, F, dummy, dummy, dummy)
    assert st == L.ERR_CAPACITY
    assert b"exceed the maximum" in lib.oibvh_last_error(ctx.h)
    st = lib.oibvh_build_noalloc(ctx.h, None, 3, dummy, 1, dummy, dummy, dummy)
    assert st == L.ERR_INVALID
