"""Full-size configurations against the oracle: the one-rank share (2M
tryptic peptides with Met-oxidation <= 2, ~5.6M forms, charge-1 fragments;
50,000 spectra, open search +-500 Da), the small CPU-runnable case and the
benchmark workload itself: the GPU index and both GPU query forms, bit-exact
counts, candidates' top hits, histograms and work counters; and, on the
benchmark's spectra before normalisation, the device pickpeaks."""
import sys
import time

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine
from oracle import binding as O
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_rank_share_full_size_bitexact():
    p = W.PRESETS["rank_share"]
    t0 = time.perf_counter()
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_rank_share_full_size_bitexact: {db.nforms} forms, {q.nspec} spectra "
        f"({time.perf_counter() - t0:.1f}s)")
    assert db.nforms > 5_000_000 and q.nspec == 50_000

    t0 = time.perf_counter()
    odb = O.OracleDB(prm, db)
    log(f"  oracle index {time.perf_counter() - t0:.1f}s")
    eng = Engine(prm)
    idx = eng.construct_db(db)
    for L, ix in zip(db.lengths, idx):
        iA, bA = ix.export()
        assert np.array_equal(iA, odb.iA[L]) and np.array_equal(bA, odb.bA[L]), L
    log("  index: iA and bA bit-exact for every length")

    t0 = time.perf_counter()
    res_o, hist_o, st_o = odb.search(q)
    log(f"  oracle search {time.perf_counter() - t0:.1f}s")
    b = eng.upload(q)
    for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
        out = eng.query_spectrum(b, kernel=kernel)
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"rank_share-full-{kernel}")
        log(f"  kernel {kernel}: {n} spectra with e-values; counts, top hits, histograms, work match")
        assert n > 40_000
    eng.close()


def test_small_config_full_size():
    # 20,000 tryptic peptides (7-30), charge 1-2 fragments, 1,000 spectra, +-10 Da
    p = W.PRESETS["small"]
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_small_config_full_size: {db.nforms} forms, {q.nspec} spectra")
    odb = O.OracleDB(prm, db)
    res_o, hist_o, st_o = odb.search(q)
    eng = Engine(prm)
    idx = eng.construct_db(db)
    for L, ix in zip(db.lengths, idx):
        iA, bA = ix.export()
        assert np.array_equal(iA, odb.iA[L]) and np.array_equal(bA, odb.bA[L]), L
    b = eng.upload(q)
    for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
        out = eng.query_spectrum(b, kernel=kernel)
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"small-full-{kernel}")
        log(f"  kernel {kernel}: {n} spectra with e-values")
        assert n > 300   # narrow windows: about 430 of the 1,000 spectra have >= min_cpsm candidates
    eng.close()


def test_open10m_benchmark_workload_full_size():
    # the benchmark's workload itself: 10M peptides (890M ions), all 200,000
    # spectra, +-500 Da
    p = W.PRESETS["open10m"]
    db = W.Database(p)
    raw, q = W.make_raw_spectra(db, p)
    prm = W.make_params(p)
    log(f"test_open10m_benchmark_workload_full_size: {db.nforms} forms, {q.nspec} spectra")
    t0 = time.perf_counter()
    odb = O.OracleDB(prm, db)
    log(f"  oracle index {time.perf_counter() - t0:.1f}s")
    eng = Engine(prm)
    idx = eng.construct_db(db)
    for L, ix in zip(db.lengths, idx):
        iA, bA = ix.export()
        assert np.array_equal(iA, odb.iA[L]) and np.array_equal(bA, odb.bA[L]), L
    t0 = time.perf_counter()
    res_o, hist_o, st_o = odb.search(q)
    log(f"  oracle search {time.perf_counter() - t0:.1f}s")
    b = eng.upload(q)
    for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
        out = eng.query_spectrum(b, kernel=kernel)
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"open10m-{kernel}")
        log(f"  kernel {kernel}: {n} spectra with e-values")
        assert n > 190_000
    # the Queries producer on the same 200,000 spectra before normalisation
    t0 = time.perf_counter()
    ref = O.pickpeaks(raw, p.peaks_max, p.base_int, p.min_int)
    log(f"  oracle pickpeaks {time.perf_counter() - t0:.1f}s, {len(raw.mz)} raw -> {len(ref.moz)} kept peaks")
    got = eng.pickpeaks(raw, p.peaks_max, p.base_int, p.min_int).export()
    for f in ("idx", "moz", "intensity", "precurse", "charges", "rtimes"):
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f
    log("  pickpeaks: bit-exact")
    eng.close()


@pytest.mark.timeout(900)
def test_large40m_full_size_index_properties_and_sample():
    # the large workload: 12.7M tryptic peptides (7-40) with Met-oxidation
    # (<= 3 per peptide), ~40M forms, ~4.2G fragment ions, charge 1-2
    # fragments; 1,000,000 spectra at 0.01 Da fragment bins (dF = 1), open
    # search +-500 Da.  Every length's iA/bA is checked by its defining
    # properties (a bin-sorted permutation of the ion ids, ascending inside
    # each bin, bA the bin offsets, sampled peptides' ions in their own
    # bins); then 5,000 random spectra are compared field by field with the
    # oracle searching the exported index, and the same spectra inside the
    # full 1M-spectrum search equal that.
    p = W.PRESETS["large40m"]
    t0 = time.perf_counter()
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_large40m: {db.nforms} forms, {q.nspec} spectra ({time.perf_counter() - t0:.1f}s)")
    assert db.nforms > 38_000_000 and q.nspec == 1_000_000
    eng = Engine(prm)
    idx = eng.construct_db(db)
    NB = eng.NB
    nions = sum(ix.nions for ix in idx)
    log(f"  GPU index: {nions} ions, {sum(ix.build_ms for ix in idx):.1f} ms")
    assert nions > 4_000_000_000
    rng = np.random.default_rng(5)
    built = {}
    for L, ix in zip(db.lengths, idx):
        ents = db.entries(L)
        seqs = db.seqs(L)
        n = len(ents)
        speclen = 2 * prm.maxz * (L - 1)
        total = n * speclen
        iA, bA = ix.export()
        assert len(iA) == total and bA[0] == 0 and int(bA[NB]) == total, L
        assert np.all(bA[1:] >= bA[:-1]), L
        inv = np.full(total, 0xFFFFFFFF, np.uint32)
        inv[iA] = np.arange(total, dtype=np.uint32)
        assert not np.any(inv == 0xFFFFFFFF), L          # every ion id exactly once
        desc = np.nonzero(iA[1:] <= iA[:-1])[0] + 1
        assert np.all(np.isin(desc, bA)), L               # a descent only where a bin starts
        for k in rng.choice(n, min(n, 64), replace=False):
            e = ents[k]
            sid = int(e["seq_id"])
            seq = bytes(seqs[sid * L:(sid + 1) * L]).decode()
            m, bins = O.gen_spectrum(seq, prm, int(e["sites"]), int(e["mod_num"]))
            if m < prm.min_mass or m > prm.max_mass:
                bins[:] = 0
            else:
                bins = np.minimum(bins, NB - 1)
            pos = inv[int(k) * speclen + np.arange(speclen)].astype(np.int64)
            assert np.all((bA[bins] <= pos) & (pos < bA[bins + 1])), (L, int(k))
        del inv, desc
        built[L] = (iA, bA)
        log(f"  length {L}: {n} forms, {total} ions: properties hold")

    sel = np.sort(rng.choice(q.nspec, 5000, replace=False))
    sample = q.subset(sel)
    odb = O.OracleDB.from_built(prm, db, built)
    t0 = time.perf_counter()
    res_o, hist_o, st_o = odb.search(sample)
    log(f"  oracle search of 5,000 spectra {time.perf_counter() - t0:.1f}s")
    del odb, built
    bs = eng.upload(sample)
    out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED)
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, "large40m-sample")
    log(f"  sample: {n} spectra with e-values match the oracle")
    assert n > 4000
    bs.free()
    b = eng.upload(q)
    full = eng.query_spectrum(b, kernel=N.KERNEL_BATCHED, want_hist=False)
    assert np.array_equal(full.results[sel], out.results)
    assert np.array_equal(full.stats[sel], out.stats)
    log(f"  full 1M-spectrum search: the sample's rows equal the sample search")
    eng.close()
