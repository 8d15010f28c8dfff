"""CPU check of the closed form the device mass sort is built on.

The device reproduces libstdc++'s std::sort (the sort LBE_Initialize applies
to pepEntries, lbe_internal.cpp:184-187) without running its scans: each
unguarded Hoare partition is computed from two position lists (G: keys not
below the pivot, ascending; S: keys not above it, descending), the pairs
G[k], S[k] swapped for k < K = max_x min(#G before x, #S from x on), the cut
at min(G[K], S[K-1]) (G[0] when K = 0); ranges of at most 16 keys end stably
sorted, ranges whose depth is used up are heap-sorted.  This file restates
that model in Python and checks it against the library's std::sort on
tie-heavy, sorted, adversarial and random arrays.  The heap fallback is
reached with McIlroy's adversary.
"""
import numpy as np
import pytest

from oracle import binding as O


def _partition(k, v, f, l, stats):
    # std::__move_median_to_first(first, first + 1, mid, last - 1)
    a, b, c = f + 1, f + (l - f) // 2, l - 1
    if k[a] < k[b]:
        x = b if k[b] < k[c] else (c if k[a] < k[c] else a)
    elif k[a] < k[c]:
        x = a
    elif k[b] < k[c]:
        x = c
    else:
        x = b
    k[f], k[x] = k[x], k[f]
    v[f], v[x] = v[x], v[f]
    P = k[f]
    seg = k[f + 1:l]
    G = [f + 1 + i for i in np.nonzero(~(seg < P))[0]]
    S = [f + 1 + i for i in np.nonzero(~(P < seg))[0]][::-1]
    # K = max over x in [f + 1, l] of min(#G in [f + 1, x), #S in [x, l))
    gpre = np.concatenate([[0], np.cumsum(~(seg < P))])
    spre = np.concatenate([[0], np.cumsum(~(P < seg))])
    K = int(np.max(np.minimum(gpre, spre[-1] - spre)))
    for i in range(K):
        g, s = G[i], S[i]
        k[g], k[s] = k[s], k[g]
        v[g], v[s] = v[s], v[g]
    gK = G[K] if K < len(G) else 1 << 62
    cut = G[0] if K == 0 else min(gK, S[K - 1])
    stats["partitions"] += 1
    return cut


def _heap_sort(k, v, f, l):
    # std::__partial_sort(first, last, last): __make_heap then __sort_heap
    K_ = k[f:l].copy()
    V_ = v[f:l].copy()
    n = l - f

    def adjust(hole, length, kv, vv):
        top = hole
        second = hole
        while second < (length - 1) // 2:
            second = 2 * (second + 1)
            if K_[second] < K_[second - 1]:
                second -= 1
            K_[hole], V_[hole] = K_[second], V_[second]
            hole = second
        if length % 2 == 0 and second == (length - 2) // 2:
            second = 2 * (second + 1)
            K_[hole], V_[hole] = K_[second - 1], V_[second - 1]
            hole = second - 1
        parent = (hole - 1) // 2
        while hole > top and K_[parent] < kv:
            K_[hole], V_[hole] = K_[parent], V_[parent]
            hole = parent
            parent = (hole - 1) // 2
        K_[hole], V_[hole] = kv, vv

    if n >= 2:
        parent = (n - 2) // 2
        while True:
            adjust(parent, n, K_[parent], V_[parent])
            if parent == 0:
                break
            parent -= 1
        last = n
        while last > 1:
            last -= 1
            kv, vv = K_[last], V_[last]
            K_[last], V_[last] = K_[0], V_[0]
            adjust(0, last, kv, vv)
    k[f:l], v[f:l] = K_, V_


def model_sort(keys, vals):
    k = np.array(keys, np.float32)
    v = np.array(vals, np.uint32)
    n = len(k)
    stats = {"partitions": 0, "heaps": 0}
    if n < 2:
        return k, v, stats
    level = [(0, n, 2 * (int(n).bit_length() - 1))]
    while level:
        nxt = []
        for f, l, d in level:
            if l - f <= 16:
                order = np.argsort(k[f:l], kind="stable")
                k[f:l], v[f:l] = k[f:l][order], v[f:l][order]
            elif d == 0:
                _heap_sort(k, v, f, l)
                stats["heaps"] += 1
            else:
                cut = _partition(k, v, f, l, stats)
                nxt += [(f, cut, d - 1), (cut, l, d - 1)]
        level = nxt
    return k, v, stats


def _check(keys):
    vals = np.arange(len(keys), dtype=np.uint32)
    k_ref, v_ref = O.std_sort_mass(keys, vals)
    k, v, st = model_sort(keys, vals)
    assert np.array_equal(k, k_ref)
    assert np.array_equal(v, v_ref), "tie order differs from std::sort"
    return st


@pytest.mark.parametrize("n", [1, 2, 3, 16, 17, 18, 33, 100, 1000, 5000])
@pytest.mark.parametrize("distinct", [1, 2, 5, 50, 0])
def test_model_matches_std_sort_random(n, distinct):
    rng = np.random.default_rng(n * 31 + distinct)
    if distinct:
        keys = rng.integers(0, distinct, n).astype(np.float32)
    else:
        keys = rng.random(n, dtype=np.float32) * 3000
    _check(keys)


@pytest.mark.parametrize("shape", ["sorted", "reverse", "organ", "sawtooth", "masses"])
def test_model_matches_std_sort_shapes(shape):
    n = 3000
    i = np.arange(n)
    keys = {"sorted": i, "reverse": n - i, "organ": np.minimum(i, n - i),
            "sawtooth": i % 37,
            # peptide-like masses: few distinct values shared by variants
            "masses": np.repeat(np.random.default_rng(1).random(n // 6) * 4000 + 500, 6)[:n]}[shape]
    _check(np.asarray(keys, np.float32))


@pytest.mark.parametrize("n", [600, 3000])
def test_model_heap_fallback_on_adversary(n):
    keys = O.antiqsort(n)
    st = _check(keys)
    assert st["heaps"] > 0, "the adversary should exhaust std::sort's depth limit"
