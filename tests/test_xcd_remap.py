"""The XCD remap of workgroup indices (oi_xcd_remap, oibvh_common.h),
restated here: a launch of n work items plus one extra workgroup must still
hand every item to exactly one workgroup."""


def test_xcd_remap_with_extra_workgroup():
    """A launch of n work items plus one extra workgroup (the refit's apex
    launch carrying the detect's counter reset): with the extra workgroup
    chosen as the one whose remapped index is the last, the others cover the
    items 0 .. n - 1 exactly once, for every n (the remap is a bijection of
    0 .. N - 1 that sends each XCD's workgroups to one contiguous range)."""
    def remap(b, n):
        q, r = n >> 3, n & 7
        x, k = b & 7, b >> 3
        return x * q + min(x, r) + k

    for n_items in list(range(1, 70)) + [100, 255, 256, 257, 4633]:
        N = n_items + 1
        got = [remap(b, N) for b in range(N)]
        assert sorted(got) == list(range(N))
        real = sorted(w for w in got if w != N - 1)
        assert real == list(range(n_items))
