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
