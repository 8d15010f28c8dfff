"""The two closed forms the device pickpeaks (hicops_amd/csrc/k_pick.hip)
uses in place of keyval.h's sequential Hoare partition, checked on the CPU
against the oracle's quicksort (orc_keyval_sort) on tie-heavy arrays:

* whole-range form (coop_partition): with L the ascending positions of keys
  >= pivot and R the descending positions of keys <= pivot, the scans swap
  L[k] with R[k] exactly while L[k] <= R[k]; the final scan positions follow
  from L[K-1], R[K-1], L[K], R[K];
* per-key form (window_sort): a key >= pivot is swapped exactly when the keys
  <= pivot from it on are at least as many as the keys >= pivot up to it (its
  partner is the a-th key <= pivot from the right); a key <= pivot
  symmetrically.

Either form, applied to every range of the recursion (in any order, the
ranges being disjoint) and skipping ranges wholly below a cut, must leave the
same suffix as the reference sort."""
import numpy as np
import pytest

from oracle import binding as O


def partition_whole(key, val, l, r):
    p = key[(l + r) >> 1]
    xs = np.arange(l, r + 1)
    seg = key[l:r + 1]
    ge, le = seg >= p, seg <= p
    a = np.cumsum(ge)                    # keys >= pivot in [l, x]
    b = np.cumsum(le[::-1])[::-1]        # keys <= pivot in [x, r]
    K = int(np.max(np.minimum(a, b)))
    L, R = xs[ge], xs[le][::-1]
    k0, v0 = key.copy(), val.copy()
    for t in range(K):
        key[L[t]], key[R[t]] = k0[R[t]], k0[L[t]]
        val[L[t]], val[R[t]] = v0[R[t]], v0[L[t]]
    i0, j0 = L[K - 1] + 1, R[K - 1] - 1
    if i0 > j0:
        return i0, j0
    i = L[K] if K < len(L) and L[K] <= j0 else R[K - 1]
    j = R[K] if K < len(R) and R[K] >= i0 else L[K - 1]
    return i, j


def partition_per_key(key, val, l, r):
    p = key[(l + r) >> 1]
    seg = key[l:r + 1]
    ge, le = seg >= p, seg <= p
    a = np.cumsum(ge)
    b = np.cumsum(le[::-1])[::-1]
    gpos, lpos = np.nonzero(ge)[0], np.nonzero(le)[0]
    src = np.arange(len(seg))
    for x in range(len(seg)):
        if ge[x] and b[x] >= a[x]:
            src[x] = lpos[len(lpos) - a[x]]      # the a-th key <= pivot from the right
        elif le[x] and a[x] >= b[x]:
            src[x] = gpos[b[x] - 1]              # the b-th key >= pivot from the left
    K = int(np.sum(ge & (b >= a)))
    key[l:r + 1], val[l:r + 1] = seg[src], val[l:r + 1][src]
    lk1, rk1 = gpos[K - 1], lpos[len(lpos) - K]
    i0, j0 = lk1 + 1, rk1 - 1
    if i0 > j0:
        return l + i0, l + j0
    lk = gpos[K] if K < len(gpos) else len(seg)
    rk = lpos[len(lpos) - 1 - K] if K < len(lpos) else -1
    i = lk if lk <= j0 else rk1
    j = rk if rk >= i0 else lk1
    return l + i, l + j


def sort_by(part, key, val, cut):
    todo = [(0, len(key) - 1)] if len(key) >= 2 else []
    while todo:
        l, r = todo.pop()
        i, j = part(key, val, l, r)
        if l < j and j >= cut:
            todo.append((l, j))
        if i < r:
            todo.append((i, r))


@pytest.mark.parametrize("part", [partition_whole, partition_per_key])
def test_closed_forms_match_the_reference_sort(part):
    rng = np.random.default_rng(23)
    for trial in range(400):
        n = int(rng.integers(0, 200))
        hi = int(rng.choice([1, 2, 3, 5, 40, 1 << 31]))
        key = rng.integers(0, hi, n).astype(np.int64)
        val = np.arange(n, dtype=np.int64)
        cut = max(0, n - int(rng.integers(1, 150)))
        k_ref, v_ref = O.keyval_sort(key, val)
        k, v = key.copy(), val.copy()
        sort_by(part, k, v, cut)
        assert np.array_equal(k[cut:], k_ref[cut:].astype(np.int64)), (trial, n, hi)
        assert np.array_equal(v[cut:], v_ref[cut:].astype(np.int64)), (trial, n, hi)
