"""The Queries producer on the GPU: MSQuery::pickpeaks (msquery.cpp:509-569)
over raw spectra, bit-exact against the oracle -- including the reference
quicksort's order of equal intensities (keyval.h:48-132) -- and searches of
the device-produced batch equal to searches of the oracle's Queries."""
import sys

import numpy as np
import pytest

from hicops_amd import ingest
from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine
from oracle import binding as O
from tests.test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_queries(a, b, tag):
    for f in ("idx", "moz", "intensity", "precurse", "charges", "rtimes"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape, (tag, f, x.shape, y.shape)
        if x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (tag, f)
        else:
            assert np.array_equal(x, y), (tag, f)


def random_raw(rng, n, tie_hi):
    """Raw spectra of 0..3000 peaks (some above the 1024 staged in LDS),
    intensities drawn from [0, tie_hi) so equal intensities are common."""
    sizes = rng.choice([0, 1, 2, 3, 17, 64, 65, 150, 400, 1024, 1025, 3000], n,
                       p=[.03, .03, .03, .03, .1, .1, .1, .2, .2, .08, .06, .04])
    idx = np.zeros(n + 1, np.uint32)
    idx[1:] = np.cumsum(sizes)
    k = int(idx[-1])
    mz = np.sort(rng.integers(10000, 200000, k)).astype(np.uint32)
    it = rng.integers(0, tie_hi, k).astype(np.uint32)
    prec = rng.uniform(-10, 5000, n)
    z = rng.integers(-1, 6, n).astype(np.int32)
    rt = rng.uniform(-100, 5000, n)
    return W.RawSpectra(mz, it, idx, prec, z, rt)


@pytest.fixture(scope="module")
def eng():
    prm = W.make_params(W.PRESETS["small"])
    e = Engine(prm)
    yield e
    e.close()


@pytest.mark.parametrize("qalen,tie_hi,base_int,min_int", [
    (100, 4, 100000, 1000),          # heavy ties, QALEN default
    (7, 50, 100000, 0),              # short QALEN: ties straddle the cut
    (500, 1 << 20, 100000000, 0),    # base_int x1000 (hicops.cpp:606-610), QALEN 500
    (1, 3, 1000, 500),               # the top peak only
    (4096, 10, 100000, 30000),       # the largest QALEN
])
def test_pickpeaks_random_bitexact(eng, qalen, tie_hi, base_int, min_int):
    rng = np.random.default_rng(qalen * 7 + tie_hi)
    raw = random_raw(rng, 600, tie_hi)
    log(f"test_pickpeaks_random_bitexact: qalen {qalen} ties<{tie_hi} base {base_int} min {min_int}, "
        f"{raw.nspec} spectra, {len(raw.mz)} raw peaks")
    ref = O.pickpeaks(raw, qalen, base_int, min_int)
    b = eng.pickpeaks(raw, qalen, base_int, min_int)
    got = b.export()
    same_queries(got, ref, f"qalen{qalen}")
    b.free()


def test_pickpeaks_raw_batch_reusable_and_empty(eng):
    rng = np.random.default_rng(99)
    raw = random_raw(rng, 300, 5)
    log("test_pickpeaks_raw_batch_reusable_and_empty")
    rb = eng.upload_raw(raw)
    ref = O.pickpeaks(raw, 50, 100000, 100)
    for rep in range(3):                       # the device raw copy is never modified
        b, kms, tms = eng.pickpeaks_timed(rb, 50, 100000, 100)
        same_queries(b.export(), ref, f"rep{rep}")
        assert 0 < kms <= tms
        b.free()
    rb.free()
    empty = W.RawSpectra(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(1, np.uint32),
                         np.zeros(0), np.zeros(0, np.int32), np.zeros(0))
    b = eng.pickpeaks(empty, 100, 100000, 0)
    q = b.export()
    assert q.nspec == 0 and len(q.moz) == 0 and q.idx.tolist() == [0]
    b.free()
    # no retention times: zeros (msquery.cpp:517 with rtime 0)
    nort = W.RawSpectra(raw.mz, raw.intensity, raw.idx, raw.prec_mz, raw.z, None)
    b = eng.pickpeaks(nort, 50, 100000, 100)
    got = b.export()
    same_queries(got, O.pickpeaks(nort, 50, 100000, 100), "no-rtime")
    assert np.all(got.rtimes == 0)
    b.free()
    # an uploaded batch without retention times exports zeros as well
    q0 = got
    q0.rtimes = None
    up = eng.upload(q0)
    assert np.all(up.export().rtimes == 0)
    up.free()
    with pytest.raises(N.HcxError):
        eng.pickpeaks(raw, 0, 100000, 0)
    with pytest.raises(N.HcxError):
        eng.pickpeaks(raw, 4097, 100000, 0)


@pytest.mark.parametrize("preset", ["small", "rank_share"])
def test_pickpeaks_then_search_end_to_end(preset):
    # generated spectra before normalisation -> GPU pickpeaks -> GPU search,
    # against oracle pickpeaks -> oracle search
    p = W.PRESETS[preset] if preset == "small" else W.scaled("rank_share", npep=20000, nspec=2000)
    db = W.Database(p)
    raw, _ = W.make_raw_spectra(db, p)
    prm = W.make_params(p)
    qalen = p.peaks_max
    log(f"test_pickpeaks_then_search_end_to_end[{preset}]: {raw.nspec} spectra, {len(raw.mz)} raw peaks, "
        f"QALEN {qalen}")
    ref_q = O.pickpeaks(raw, qalen, p.base_int, p.min_int)
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.pickpeaks(raw, qalen, p.base_int, p.min_int)
    same_queries(b.export(), ref_q, preset)
    odb = O.OracleDB(prm, db)
    res_o, hist_o, st_o = odb.search(ref_q)
    for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
        out = eng.query_spectrum(b, kernel=kernel)
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"pick-{preset}-{kernel}")
        log(f"  kernel {kernel}: {n} spectra with e-values match")
        assert n > 100
    # the device-produced batch searches exactly like the same Queries uploaded
    up = eng.upload(ref_q)
    a = eng.query_spectrum(up, kernel=N.KERNEL_BATCHED)
    c = eng.query_spectrum(b, kernel=N.KERNEL_BATCHED)
    assert np.array_equal(a.results, c.results) and np.array_equal(a.hist, c.hist)
    eng.close()


def test_ms2_file_to_device_batch(eng, tmp_path):
    # MS2 text -> host reader (MSQuery::readspectrum rules) -> device pickpeaks,
    # against the oracle's pickpeaks of the same raw spectra
    from tests.test_ms2_ingest import write_ms2
    p = W.scaled("small", npep=3000, nspec=400)
    db = W.Database(p)
    raw0, _ = W.make_raw_spectra(db, p)
    f = tmp_path / "q.ms2"
    write_ms2(f, raw0, 100)
    raw = ingest.read_ms2(f, 100)
    log(f"test_ms2_file_to_device_batch: {raw.nspec} spectra, {len(raw.mz)} raw peaks")
    b = eng.pickpeaks(raw, 100, p.base_int, p.min_int)
    same_queries(b.export(), O.pickpeaks(raw, 100, p.base_int, p.min_int), "ms2")
    b.free()


def _facts_ref(moz, inten, idx, span):
    """Per spectrum, the most peaks and the largest intensity sum inside
    span + 1 consecutive bins: a sort and a sliding window (the definitions
    the exact scan's bound uses)."""
    n = len(idx) - 1
    mult = np.zeros(n, np.uint32)
    wsum = np.zeros(n, np.uint64)
    for s in range(n):
        a, b = int(idx[s]), int(idx[s + 1])
        o = np.argsort(moz[a:b], kind="stable")
        t = moz[a:b][o].astype(np.int64)
        w = inten[a:b][o].astype(np.int64)
        if len(t):
            e = np.searchsorted(t, t + span, side="right")
            c = np.concatenate([[0], np.cumsum(w)])
            i = np.arange(len(t))
            mult[s] = int((e - i).max())
            wsum[s] = int((c[e] - c[i]).max())
    return mult, np.minimum(wsum, 0xFFFFFFFF).astype(np.uint32)


def test_uploaded_batch_facts(eng):
    # spectra of 0..5000 peaks (above the 4096 staged in LDS too), bins drawn
    # from narrow ranges so that duplicates and dense windows are common
    rng = np.random.default_rng(11)
    sizes = rng.choice([0, 1, 2, 7, 64, 65, 300, 4096, 4097, 5000], 60)
    idx = np.zeros(len(sizes) + 1, np.uint32)
    idx[1:] = np.cumsum(sizes)
    k = int(idx[-1])
    moz = np.concatenate([rng.integers(1000, 1000 + int(rng.choice([5, 50, 5000])), n) for n in sizes]
                         ).astype(np.uint32) if k else np.zeros(0, np.uint32)
    # 1 % of the intensities near 2^31: dense windows saturate the sums
    inten = np.where(rng.random(k) < 0.01, rng.integers(1 << 30, 1 << 31, k),
                     rng.integers(0, 1 << 20, k)).astype(np.uint32)
    n = len(sizes)
    q = W.Queries(moz, inten, idx, np.full(n, 1500.0, np.float32), np.full(n, 2, np.int32),
                  np.zeros(n, np.float32), np.zeros(n, np.uint32), np.zeros(n, np.int32))
    b = eng.upload(q)
    f = b.facts()
    span = 2 * eng.params.dF
    ref, wref = _facts_ref(moz, inten, idx, span)
    log(f"[facts] {n} spectra, {k} peaks, span {span}: max_mult {f['max_mult']} (ref {ref.max()})")
    assert np.array_equal(f["mult"], ref)
    assert np.array_equal(f["wsum"], wref)
    assert f["max_mult"] == ref.max()
    assert f["max_int"] == (int(inten.max()) if k else 0)
    assert f["max_peaks"] == int(sizes.max())
    b.free()
    # the device pickpeaks path computes the same facts for its own batch
    raw = W.make_raw_spectra(W.Database(W.PRESETS["small"]), W.PRESETS["small"], nspec=300)[0]
    pb = eng.pickpeaks(raw, 100, 1000000, 10)
    pf = pb.facts()
    pq = pb.export()
    pm, pw = _facts_ref(pq.moz, pq.intensity, pq.idx, span)
    assert np.array_equal(pf["mult"], pm)
    assert np.array_equal(pf["wsum"], pw)
    assert pf["max_int"] == int(pq.intensity.max())
    assert pf["max_peaks"] == int(np.diff(pq.idx.astype(np.int64)).max())
    pb.free()
    # an empty batch: no facts, and a search of it returns nothing
    z = W.Queries(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(1, np.uint32), np.zeros(0, np.float32),
                  np.zeros(0, np.int32), np.zeros(0, np.float32), np.zeros(0, np.uint32), np.zeros(0, np.int32))
    e2 = Engine(eng.params)
    e2.construct_db(W.Database(W.scaled("small", npep=500)))
    zb = e2.upload(z)
    zf = zb.facts()
    assert (zf["max_peaks"], zf["max_int"], zf["max_mult"]) == (0, 0, 0) and len(zf["mult"]) == 0
    out = e2.query_spectrum(zb, kernel=N.KERNEL_BATCHED)
    assert len(out.results) == 0
    zb.free()
    e2.close()
