"""The library's host/device index math (oibvh_amd/csrc/oibvh_layout.h),
compiled from its own header with the host C compiler and called through
ctypes: the XCD remap of workgroup indices (oi_xcd_remap) -- a launch of n
work items plus one extra workgroup (the refit's apex launch carrying the
detect's counter reset) must still hand every item to exactly one
workgroup -- and the implicit-layout functions against a brute-force
enumeration of the layout."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "oibvh_amd", "csrc")

SHIM = r"""
#include "oibvh_layout.h"
uint32_t shim_xcd_remap(uint32_t b, uint32_t n) { return oi_xcd_remap(b, n); }
uint32_t shim_tree_size(uint32_t f) { return oi_tree_size(f); }
uint32_t shim_lvl_cnt(uint32_t l, uint32_t L, uint32_t vl) { return oi_lvl_cnt(l, L, vl); }
uint32_t shim_lvl_start(uint32_t l, uint32_t L, uint32_t vl) { return oi_lvl_start(l, L, vl); }
uint32_t shim_lvl_last_implicit(uint32_t l, uint32_t L, uint32_t vl) { return oi_lvl_last_implicit(l, L, vl); }
"""


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    d = tmp_path_factory.mktemp("xcd")
    src, lib = d / "shim.c", d / "libshim.so"
    src.write_text(SHIM)
    subprocess.run([cc, "-O1", "-std=c99", "-shared", "-fPIC", f"-I{CSRC}", str(src), "-o", str(lib)], check=True)
    L = C.CDLL(str(lib))
    for name, n in [("shim_xcd_remap", 2), ("shim_tree_size", 1), ("shim_lvl_cnt", 3), ("shim_lvl_start", 3),
                    ("shim_lvl_last_implicit", 3)]:
        fn = getattr(L, name)
        fn.restype = C.c_uint32
        fn.argtypes = [C.c_uint32] * n
    return L


@pytest.fixture(scope="module")
def remap(shim):
    return shim.shim_xcd_remap


@pytest.mark.parametrize("F", list(range(1, 130)) + [1000, 4097, 81920, 1000000])
def test_layout_math_brute_force(shim, F):
    """Build the complete tree of height L over F real leaves level by level
    (a node is real iff its leftmost descendant leaf index is < F), number
    real nodes breadth first, and compare counts, level starts, the right-most
    real implicit index and the node count with the header's closed forms."""
    L = (F - 1).bit_length() if F > 1 else 0
    vl = (1 << L) - F
    real = 0
    for l in range(L + 1):
        cnt = sum(1 for p in range(1 << l) if (p << (L - l)) < F)
        assert shim.shim_lvl_cnt(l, L, vl) == cnt
        assert shim.shim_lvl_start(l, L, vl) == real
        assert shim.shim_lvl_last_implicit(l, L, vl) == (1 << l) - 1 + cnt - 1
        real += cnt
    assert shim.shim_tree_size(F) == real


NS = list(range(1, 70)) + [100, 255, 256, 257, 1000, 2048, 4633]


@pytest.mark.parametrize("n_items", NS)
def test_xcd_remap_with_extra_workgroup(remap, n_items):
    """k_reduce_apex with a counter reset: N = n + 1 workgroups, the one whose
    remapped index is N - 1 resets, the others take item remap(b, N). They
    cover the items 0 .. n - 1 exactly once (the remap is a bijection of
    0 .. N - 1)."""
    N = n_items + 1
    got = [remap(b, N) for b in range(N)]
    assert sorted(got) == list(range(N))
    real = sorted(w for w in got if w != N - 1)
    assert real == list(range(n_items))


@pytest.mark.parametrize("n_items", NS)
def test_xcd_remap_contiguous_per_xcd(remap, n_items):
    """Workgroup b runs on XCD b % 8; each XCD's workgroups take one
    contiguous range of items, in launch order."""
    per = {}
    for b in range(n_items):
        per.setdefault(b & 7, []).append(remap(b, n_items))
    start = 0
    for x in range(8):
        ws = per.get(x, [])
        assert ws == list(range(start, start + len(ws)))
        start += len(ws)
    assert start == n_items


@pytest.mark.parametrize("n_items,grid", [(100, 8), (4633, 512), (1954, 448), (2048, 640), (17, 16)])
def test_xcd_remap_looped_launch(remap, n_items, grid):
    """k_reduce_wl: `grid` workgroups (a multiple of 8) loop over virtual
    workgroups v = b, b + grid, ... < n and take item remap(v, n); every item
    is taken once and v keeps workgroup b's XCD."""
    seen = []
    for b in range(grid):
        for v in range(b, n_items, grid):
            assert v % 8 == b % 8
            seen.append(remap(v, n_items))
    assert sorted(seen) == list(range(n_items))
