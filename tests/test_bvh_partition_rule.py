"""The partition rule the GPU BVH builder uses (craystal_amd/csrc/bvh_device.hip)
against libstdc++'s bidirectional two-pointer std::partition, the one
BVH::build calls (Core/BVH.cu:177-181): swapping the r-th element that fails
the predicate counted from the left with the r-th element that passes it
counted from the right, for every r below the number of failing elements
left of the split, gives the same permutation and split point.  Both are
written out here in Python on random arrays (CPU only)."""
import numpy as np


def _two_pointer(a, pred):
    a = list(a)
    first, last = 0, len(a)
    while True:
        while True:
            if first == last:
                return a, first
            if pred(a[first]):
                first += 1
            else:
                break
        last -= 1
        while True:
            if first == last:
                return a, first
            if not pred(a[last]):
                last -= 1
            else:
                break
        a[first], a[last] = a[last], a[first]
        first += 1


def _rank_pairs(a, pred):
    a = list(a)
    p = [pred(x) for x in a]
    lc = sum(p)
    left, right = {}, {}
    before = 0
    for i, ok in enumerate(p):
        if i < lc and not ok:
            left[i - before] = i            # failing elements before it
        elif i >= lc and ok:
            right[lc - 1 - before] = i      # passing elements after it
        before += ok
    assert sorted(left) == sorted(right) == list(range(len(left)))
    for r in range(len(left)):
        x, y = left[r], right[r]
        a[x], a[y] = a[y], a[x]
    return a, lc


def test_rank_pairs_equal_two_pointer_partition():
    rng = np.random.default_rng(0)
    for _ in range(20000):
        n = int(rng.integers(0, 48))
        a = [int(v) for v in rng.integers(0, 64, n)]
        th = int(rng.integers(0, 65))
        assert _two_pointer(a, lambda x: x < th) == _rank_pairs(a, lambda x: x < th), (a, th)
