"""GPU parity of the peptide database stage (LBE + MODS) against the oracle.

The oracle runs the reference's own sorts (glibc qsort with cmpvarEntries,
libstdc++ std::sort with CmpPepEntries); the device reproduces their orders
without calling them.  Checked bit for bit: kept sequences, counts (pepCount,
modCount, the rank share), every pepEntry field (mass bits, seqID, sites,
modNum) in order, and the index built from the device-resident entries.
"""
import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine
from oracle import binding as O

pytestmark = pytest.mark.gpu

FIELDS = ("mass", "seq_id", "sites", "mod_num")


def _same_entries(a, b):
    assert len(a) == len(b)
    for f in FIELDS:
        x, y = a[f], b[f]
        if f == "mass":
            x, y = x.view(np.uint32), y.view(np.uint32)
        bad = np.nonzero(x != y)[0]
        assert len(bad) == 0, f"field {f} differs at {bad[:5]} of {len(a)}"


@pytest.fixture(scope="module")
def eng():
    p = W.scaled("rank_share", npep=4000, nspec=16)
    e = Engine(W.make_params(p))
    yield e
    e.close()


@pytest.mark.parametrize("n,distinct", [(2, 0), (17, 0), (100, 3), (4096, 7), (4097, 7), (5000, 0),
                                        (70000, 40), (300000, 0), (300000, 1000), (1 << 20, 5)])
def test_sort_mass_matches_std_sort(eng, n, distinct):
    print(f"[sort] n={n} distinct={distinct}", flush=True)
    rng = np.random.default_rng(n + distinct)
    keys = (rng.integers(0, distinct, n) if distinct else rng.random(n) * 4000).astype(np.float32)
    vals = np.arange(n, dtype=np.uint32)
    k_ref, v_ref = O.std_sort_mass(keys, vals)
    k, v = eng.sort_mass(keys, vals)
    assert np.array_equal(k, k_ref)
    assert np.array_equal(v, v_ref), "tie order differs from std::sort"


@pytest.mark.parametrize("shape", ["sorted", "reverse", "organ", "sawtooth", "equal"])
def test_sort_mass_shapes(eng, shape):
    n = 200000
    i = np.arange(n)
    keys = {"sorted": i, "reverse": n - i, "organ": np.minimum(i, n - i), "sawtooth": i % 1001,
            "equal": np.zeros(n)}[shape].astype(np.float32)
    print(f"[sort] shape={shape}", flush=True)
    vals = np.arange(n, dtype=np.uint32)
    k_ref, v_ref = O.std_sort_mass(keys, vals)
    k, v = eng.sort_mass(keys, vals)
    assert np.array_equal(v, v_ref)


@pytest.mark.parametrize("n", [3000, 12000])
def test_sort_mass_heap_fallback(eng, n):
    # McIlroy's adversary exhausts std::sort's depth limit: at 3000 keys the
    # heap sort runs in LDS, at 12000 over HBM
    print(f"[sort] adversary n={n}", flush=True)
    keys = O.antiqsort(n)
    vals = np.arange(n, dtype=np.uint32)
    k_ref, v_ref = O.std_sort_mass(keys, vals)
    k, v = eng.sort_mass(keys, vals)
    assert np.array_equal(v, v_ref)


def _lbe_both(eng, db, p, lengths, nodes=1, myid=0, sort=True, mc=None):
    prm = eng.params
    mc = W.modconditions(p) if mc is None else mc
    dbs = [eng.count_peps(L, db.seqs(L), mc, nodes=nodes, myid=myid, sort=sort) for L in lengths]
    eng.initialize(dbs)
    for L, d in zip(lengths, dbs):
        seqs_o = O.lbe_filter(db.seqs(L), L, prm.min_mass, prm.max_mass)
        ents_o, cnt_o = O.lbe_generate(prm, seqs_o, L, mc, nodes=nodes, myid=myid, sort=sort)
        seqs_g, ents_g = d.export()
        assert np.array_equal(seqs_g, seqs_o)
        assert (d.pep_count, d.mod_count, d.lclpep, d.lclmod) == cnt_o
        _same_entries(ents_g, ents_o)
    return dbs


@pytest.mark.parametrize("nodes,myid,sort", [(1, 0, True), (3, 1, True), (3, 2, False), (1, 0, False)])
def test_pepdb_matches_oracle_mods(eng, nodes, myid, sort):
    p = W.scaled("rank_share", npep=4000)
    db = W.Database(p)
    print(f"[lbe] rank_share 4000 nodes={nodes} myid={myid} sort={sort}", flush=True)
    dbs = _lbe_both(eng, db, p, db.lengths, nodes, myid, sort)
    for d in dbs:
        d.free()


def test_pepdb_two_mod_types_and_lower_case(eng):
    p = W.scaled("rank_share", npep=3000)
    db = W.Database(p)
    prm = W.make_params(p)
    prm.nmods = 2
    prm.mod_mass[1] = 7997   # phospho, (uint)(79.9663 * 100)
    e2 = Engine(prm)
    mc = "3 M 2 STY 2"
    print("[lbe] two mod types", flush=True)
    for L in db.lengths[:8]:
        lines = db.seqs(L).copy()
        sel = lines[::7]
        lines[::7] = np.where((sel >= 65) & (sel <= 90), sel + 32, sel)
        d = e2.count_peps(L, lines, mc)
        e2.initialize([d])
        seqs_o = O.lbe_filter(lines, L, prm.min_mass, prm.max_mass)
        ents_o, cnt_o = O.lbe_generate(prm, seqs_o, L, mc)
        seqs_g, ents_g = d.export()
        assert np.array_equal(seqs_g, seqs_o)
        assert (d.pep_count, d.mod_count) == cnt_o[:2]
        _same_entries(ents_g, ents_o)
        d.free()
    e2.close()


def test_index_from_pepdb_equals_host_build(eng):
    p = W.scaled("rank_share", npep=3000)
    db = W.Database(p)
    mc = W.modconditions(p)
    print("[lbe] index from device entries", flush=True)
    for L in db.lengths[:5]:
        d = eng.count_peps(L, db.seqs(L), mc)
        eng.initialize([d])
        seqs, ents = d.export()
        ix_dev = eng.construct_pepdb(d)
        ix_host = eng.construct(L, seqs, ents)
        a, b = ix_dev.export(), ix_host.export()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        iA_o, bA_o = O.index_build(eng.params, L, seqs, ents)
        assert np.array_equal(a[0], iA_o) and np.array_equal(a[1], bA_o)
        ix_dev.free()
        ix_host.free()
        d.free()


def test_pepdb_errors(eng):
    L = 10
    lines = np.frombuffer(b"PEPTIDEKMK" * 3, np.uint8)
    for mc, code in (("9 M 2", N.ERR_INVLD_MOD), ("2 M 1 M 1", N.ERR_INVLD_MOD), ("2 m1 1", N.ERR_INVLD_MOD),
                     ("x M 1", N.ERR_INVLD_MOD), ("2 M -1", N.ERR_INVLD_MOD)):
        with pytest.raises(N.HcxError) as ei:
            eng.count_peps(L, lines, mc)
        assert ei.value.code == code, mc
    with pytest.raises(N.HcxError) as ei:
        eng.count_peps(L, lines, "1 M 1", nodes=2, myid=0, policy=N.POLICY_CHUNK)
    assert ei.value.code == N.ERR_INVLD_PARAM
    with pytest.raises(N.HcxError) as ei:
        eng.count_peps(L, lines, "1 M 1", nodes=2, myid=2)
    assert ei.value.code == N.ERR_INVLD_NODE_RANK
    bad = lines.copy()
    bad[3] = ord("1")
    with pytest.raises(N.HcxError) as ei:
        eng.count_peps(L, bad, "1 M 1")
    assert ei.value.code == N.ERR_INVLD_PARAM
    # an unsorted database cannot back the query's window search
    d = eng.count_peps(L, lines, "1 M 1", sort=False)
    eng.initialize([d])
    with pytest.raises(N.HcxError) as ei:
        eng.construct_pepdb(d)
    assert ei.value.code == N.ERR_INVLD_SORT
    d.free()
    # empty input
    d = eng.count_peps(L, np.zeros(0, np.uint8), "1 M 1")
    eng.initialize([d])
    s, e = d.export()
    assert len(s) == 0 and len(e) == 0
    d.free()


def _full_size(name, min_forms):
    from concurrent.futures import ThreadPoolExecutor
    import time
    p = W.PRESETS[name]
    t0 = time.perf_counter()
    db = W.Database(p)
    prm = W.make_params(p)
    mc = W.modconditions(p)
    print(f"[lbe] {name}: {db.nforms} generator forms over {len(db.lengths)} lengths "
          f"({time.perf_counter() - t0:.1f}s)", flush=True)
    eng = Engine(prm)
    dbs = [eng.count_peps(L, db.seqs(L), mc) for L in db.lengths]
    eng.initialize(dbs)
    dev_ms = sum(d.timing()[1] for d in dbs) + dbs[0].timing()[2]
    total = sum(d.lcltot for d in dbs)
    print(f"[lbe] device: {total} forms in {dev_ms:.1f} ms", flush=True)
    assert total == db.nforms and total > min_forms

    def orc(L):
        s = O.lbe_filter(db.seqs(L), L, prm.min_mass, prm.max_mass)
        return O.lbe_generate(prm, s, L, mc)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(16) as ex:
        ref = list(ex.map(orc, db.lengths))
    print(f"[lbe] oracle ({time.perf_counter() - t0:.1f}s)", flush=True)
    ties = 0
    for L, d, (ents_o, cnt_o) in zip(db.lengths, dbs, ref):
        seqs_g, ents_g = d.export()
        assert np.array_equal(seqs_g, db.seqs(L))
        assert (d.pep_count, d.mod_count, d.lclpep, d.lclmod) == cnt_o, L
        _same_entries(ents_g, ents_o)
        m = ents_o["mass"]
        ties += int(np.count_nonzero(m[1:] == m[:-1]))
        d.free()
    print(f"[lbe] {name}: every length's pepEntries bit-exact ({ties} adjacent equal masses)", flush=True)
    assert ties > 0
    eng.close()


def test_rank_share_pepdb_full_size():
    # the one-rank share: 2M tryptic peptides (7-40) + Met-oxidation <= 2
    _full_size("rank_share", 5_000_000)


@pytest.mark.timeout(600)
def test_large40m_pepdb_full_size():
    # 12.7M tryptic peptides (7-40) + Met-oxidation <= 3: ~40M forms
    _full_size("large40m", 38_000_000)
