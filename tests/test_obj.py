"""The native OBJ reader (oibvh_obj_read), host only: no device work, so these
run without a GPU. It stands in for the reference's assimp Model
(model.cpp:40-186) on the collision path: positions and triangles."""
import numpy as np
import pytest

from oibvh_amd import meshgen as M


@pytest.fixture(scope="module")
def api():
    from oibvh_amd import _lib as L
    from oibvh_amd import api as A

    try:
        L.load()
    except OSError as e:  # pragma: no cover
        pytest.skip(f"library not built: {e}")
    return A


def write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_obj_corner_forms_fans_and_skips(api, tmp_path):
    path = write(tmp_path, "\n".join([
        "# a comment", "mtllib x.mtl", "o thing", "v 0 0 0", "v 1 0 0 1.0", "v 1 1 0", "v 0 1 0",
        "vt 0 0", "vn 0 0 1", "g grp", "usemtl m", "s off",
        "f 1/1/1 2/1/1 3/1/1 4/1/1",      # quad -> (1,2,3), (1,3,4)
        "  v 0.5 0.5 2.5\t",
        "f -1//1 1//1 2//1",               # negative index = last position
        "f 3/1 4/1 5/1\r",
        "l 1 2", "p 3", "",
    ]))
    xyz, f = api.load_obj(path)
    assert xyz.dtype == np.float32 and f.dtype == np.uint32
    np.testing.assert_array_equal(xyz, np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 2.5]],
                                                np.float32))
    np.testing.assert_array_equal(f, np.array([[0, 1, 2], [0, 2, 3], [4, 0, 1], [2, 3, 4]], np.uint32))


def test_obj_pentagon_fan(api, tmp_path):
    lines = [f"v {k} {k * k} 0" for k in range(5)] + ["f 1 2 3 4 5"]
    xyz, f = api.load_obj(write(tmp_path, "\n".join(lines) + "\n"))
    np.testing.assert_array_equal(f, np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4]], np.uint32))


def test_obj_round_trip_bit_exact(api, tmp_path):
    """%.9g prints every float32 exactly: a generated mesh reads back bit for
    bit."""
    xyz, f = M.icosphere(3, 1.0, 0.02, 7)
    text = "".join(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in xyz.astype(np.float64))
    text += "".join(f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in f)
    rx, rf = api.load_obj(write(tmp_path, text))
    assert np.array_equal(rx.view(np.uint32), xyz.view(np.uint32))
    assert np.array_equal(rf, f)


def test_obj_long_line(api, tmp_path):
    n = 3000  # one polygon line far longer than any read buffer
    lines = [f"v {k}.0 {k % 7}.5 -{k}.25" for k in range(n)]
    lines.append("f " + " ".join(str(k + 1) for k in range(n)))
    xyz, f = api.load_obj(write(tmp_path, "\n".join(lines)))
    assert xyz.shape == (n, 3) and f.shape == (n - 2, 3)
    assert np.array_equal(f[:, 0], np.zeros(n - 2, np.uint32))
    assert np.array_equal(f[:, 1], np.arange(1, n - 1, dtype=np.uint32))


@pytest.mark.parametrize("body", [
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",     # index beyond the positions read
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",     # OBJ indices start at 1
    "v 0 0 0\nv 1 0 0\nf 1 2\n",                # two corners
    "v 0 0\n",                                  # short position
    "f 1 2 3\nv 0 0 0\nv 1 0 0\nv 0 1 0\n",     # face before its positions
    "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -4\n",    # negative index past the start
])
def test_obj_errors(api, tmp_path, body):
    from oibvh_amd import _lib as L

    with pytest.raises(L.OibvhError) as e:
        api.load_obj(write(tmp_path, body))
    assert e.value.status == L.ERR_INVALID


def test_obj_missing_file(api, tmp_path):
    from oibvh_amd import _lib as L

    with pytest.raises(L.OibvhError):
        api.load_obj(str(tmp_path / "absent.obj"))


def test_obj_capacity_protocol(api, tmp_path):
    import ctypes as C

    from oibvh_amd import _lib as L

    path = write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3 4\n")
    lib = L.load()
    V, F = C.c_uint32(), C.c_uint32()
    assert lib.oibvh_obj_read(None, path.encode(), None, 0, None, 0, C.byref(V), C.byref(F)) == L.ERR_CAPACITY
    assert (V.value, F.value) == (4, 2)
    xyz = np.empty(11, np.float32)
    idx = np.empty(6, np.uint32)
    st = lib.oibvh_obj_read(None, path.encode(), xyz.ctypes.data_as(L._f32p), xyz.size,
                            idx.ctypes.data_as(L._u32p), idx.size, C.byref(V), C.byref(F))
    assert st == L.ERR_CAPACITY
