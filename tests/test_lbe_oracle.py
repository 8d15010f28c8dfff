"""Pins of the oracle's peptide database stage (LBE + MODS) on the CPU.

  * count() (mods.cpp:290-315) against brute force: every subset of the
    modifiable residues within the per-type caps and the per-form limit;
  * MODS_ModList's emission order and glibc qsort with cmpvarEntries
    (mods.cpp:333-384, 105-126, 491) on hand-traced sequences;
  * cmpvarEntries under qsort is a stable sort by ascending span (last
    modified residue minus first), the rule the device applies;
  * the rank share of LBE_CreatePartitions / LBE_GeneratePeps /
    MODS_GenerateMods (lbe_internal.cpp:308-352, 209-232; mods.cpp:448-496)
    on a hand-worked three-rank split;
  * masses: the unmodified mass and the first mod only (utils.cpp:290-362);
  * the mass sort is std::sort (lbe_internal.cpp:184-187).
"""
import itertools

import numpy as np
import pytest

from hicops_amd import workload as W
from oracle import binding as O


def _params(mods=(1599,)):
    p = W.make_params(W.scaled("small"))
    p.nmods = len(mods)
    for i, m in enumerate(mods):
        p.mod_mass[i] = m
    return p


def _brute_forms(seq: str, limit: int, conds: list[tuple[str, int]]) -> int:
    typ = {}
    for t, (res, _) in enumerate(conds):
        for ch in res:
            typ[ch] = t
    sites = [i for i, ch in enumerate(seq) if ch in typ]
    n = 0
    for r in range(0, min(limit, len(sites)) + 1):
        for comb in itertools.combinations(sites, r):
            per = [0] * len(conds)
            for i in comb:
                per[typ[seq[i]]] += 1
            if all(per[t] <= conds[t][1] for t in range(len(conds))):
                n += 1
    return n


@pytest.mark.parametrize("seed", range(6))
def test_count_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    alphabet = "ACDEFGHIKLMNPQRSTVWY"
    for _ in range(40):
        L = int(rng.integers(4, 16))
        seq = "".join(rng.choice(list(alphabet + "MMMSTY"), L))
        limit = int(rng.integers(0, 5))
        conds = [("M", int(rng.integers(0, 4))), ("STY", int(rng.integers(0, 3)))][: int(rng.integers(1, 3))]
        text = f"{limit} " + " ".join(f"{r} {c}" for r, c in conds)
        expect = _brute_forms(seq, limit, conds) if limit > 0 else 1
        got = O.lbe_count(seq, text)
        if limit == 0:
            # MODS_ModCounter skips counting when the limit is 0 (mods.cpp:403)
            got = 1
        assert got == expect, (seq, text)


def _one(seq, mc, mods=(1599,), sort=False):
    p = _params(mods)
    ents, cnt = O.lbe_generate(p, np.frombuffer(seq.encode(), np.uint8), len(seq), mc, sort=sort)
    return p, ents, cnt


def test_modlist_order_and_qsort_one_type():
    # MAMK, M-ox at most 2: MODS_ModList emits {0}, {0,2}, {2}; qsort by site
    # distance (larger s1 + 60 - s3 first, ties kept) gives {0}, {2}, {0,2}
    p, e, cnt = _one("MAMK", "2 M 2")
    assert cnt == (1, 3, 1, 3)
    assert list(e["sites"]) == [0, 0b1, 0b100, 0b101]
    assert list(e["mod_num"]) == [0, 0x1, 0x1, 0x11]
    assert list(e["seq_id"]) == [0, 0, 0, 0]
    base = O.pep_mass("MAMK")
    assert e["mass"][0] == np.float32(base)
    for k in range(1, 4):
        assert e["mass"][k] == np.float32(O.mod_mass("MAMK", int(e["mod_num"][k]), p))


def test_modlist_order_and_qsort_two_types():
    # MSTK with "2 M 1 ST 1": types M -> 1, S/T -> 2 (nibble values).  Emission
    # {0} 0x1, {0,1} 0x21, {0,2} 0x21, {1} 0x2, {2} 0x2; spans 0, 1, 2, 0, 0;
    # after qsort: {0}, {1}, {2}, {0,1}, {0,2}
    p, e, cnt = _one("MSTK", "2 M 1 ST 1", mods=(1599, 7997))
    assert cnt == (1, 5, 1, 5)
    assert list(e["sites"][1:]) == [0b1, 0b10, 0b100, 0b11, 0b101]
    assert list(e["mod_num"][1:]) == [0x1, 0x2, 0x2, 0x21, 0x21]
    # first mod only: {1} and {2} carry type 2's mass, {0,1} type 1's
    assert e["mass"][2] == np.float32(O.pep_mass("MSTK") + np.float32(7997) / np.float32(100))
    assert e["mass"][4] == np.float32(O.pep_mass("MSTK") + np.float32(1599) / np.float32(100))


def test_modlist_caps_and_limit():
    # MMMK, M at most 2 per form, limit 3: the 3-site form is not generated
    p, e, cnt = _one("MMMK", "3 M 2")
    assert cnt[1] == 3 + 3
    assert 0b111 not in set(int(x) for x in e["sites"])
    # emission {0},{0,1},{0,2},{1},{1,2},{2}; spans 0,1,2,0,1,0 ->
    # {0},{1},{2},{0,1},{1,2},{0,2}
    assert list(e["sites"][1:]) == [0b1, 0b10, 0b100, 0b11, 0b110, 0b101]
    # no mods
    p, e, cnt = _one("MMMK", "0 M 0")
    assert cnt == (1, 0, 1, 0) and len(e) == 1


def test_qsort_is_stable_by_span():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 33, 200, 1000):
        sites = np.zeros(n, np.uint64)
        for i in range(n):
            bits = rng.choice(40, int(rng.integers(1, 5)), replace=False)
            sites[i] = np.uint64(sum(1 << int(b) for b in bits))
        vals = np.arange(n, dtype=np.uint32)
        s_ref, v_ref = O.qsort_sites(sites, vals)
        span = np.array([int(s).bit_length() - 1 - (int(s) & -int(s)).bit_length() + 1 for s in sites])
        order = np.argsort(span, kind="stable")
        assert np.array_equal(v_ref, vals[order])


def test_rank_share_three_ranks_cyclic():
    # 5 sequences, 3 ranks; rank 1 holds unmodified rows 1 and 4
    # (DSLIM_GenerateIndex: key * nodes + myid) and every third variant
    seqs = ["MAMKR", "GGGGK", "AMAAK", "MMGGK", "PEPMK"]
    p = _params()
    buf = np.frombuffer("".join(seqs).encode(), np.uint8)
    full, cnt_full = O.lbe_generate(p, buf, 5, "2 M 2", sort=False)
    # variants per sequence: 3, 0, 1, 3, 1 -> 8 in all
    assert cnt_full == (5, 8, 5, 8)
    e1, cnt1 = O.lbe_generate(p, buf, 5, "2 M 2", nodes=3, myid=1, sort=False)
    assert cnt1 == (5, 8, 2, 3)   # 5 // 3 + (5 % 3 > 1), 8 // 3 + (8 % 3 > 1)
    assert list(e1["seq_id"][:2]) == [1, 4]
    # global variant indices 1, 4, 7: sequence 0's second form, sequence 3's
    # first, sequence 4's only one
    # each sequence's local forms are qsorted among themselves: sequence 0
    # keeps one (global 1 in emission order = {0,2}); 3 keeps emission 0 = {0}
    # of its forms; 4 keeps its only form
    got = [(int(e["seq_id"]), int(e["sites"])) for e in e1[2:]]
    assert got == [(0, 0b101), (3, 0b1), (4, 0b1000)]


def test_mass_sort_is_std_sort():
    pr = W.scaled("small", npep=400)
    db = W.Database(pr)
    p = W.make_params(pr)
    for L in list(db.lengths)[:6]:
        seqs = db.seqs(L)
        unsorted, _ = O.lbe_generate(p, seqs, L, None, sort=False)
        srt, _ = O.lbe_generate(p, seqs, L, None, sort=True)
        k, v = O.std_sort_mass(unsorted["mass"], np.arange(len(unsorted), dtype=np.uint32))
        assert np.array_equal(srt, unsorted[v])


def test_filter_bounds_and_case():
    # GGGGGGG: 18.015 + 7 * 57.02146 = 417.17 Da; PEPTIDE: 799.37 Da
    # (utils.cpp:41-98 residue masses); lower case is upper-cased first
    lines = np.frombuffer(b"GGGGGGG" + b"peptide", np.uint8)
    assert bytes(O.lbe_filter(lines, 7, 500, 5000)) == b"PEPTIDE"
    assert bytes(O.lbe_filter(lines, 7, 400, 5000)) == b"GGGGGGGPEPTIDE"
    assert bytes(O.lbe_filter(lines, 7, 400, 799)) == b"GGGGGGG"
    assert bytes(O.lbe_filter(lines, 7, 800, 5000)) == b""
