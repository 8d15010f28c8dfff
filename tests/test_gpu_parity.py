"""HIP path vs the CPU oracle on the same seeded inputs (bit-exact counts,
candidate sets, top hits, histograms; hyperscore <= 1e-5 rel; e-value <= 1e-3 rel)."""
import sys

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine
from oracle import binding as O

pytestmark = pytest.mark.gpu

CHECK_EXACT = ["cpsms", "psid", "idxoffset", "sharedions", "totalions", "bc", "yc", "maxhypscore",
               "minhypscore", "nexthypscore", "mu", "beta"]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class TwinDB:
    """A database whose every sequence appears twice: every top hit is tied."""

    def __init__(self, db):
        self.lengths = db.lengths
        self.tables = {}
        for L in db.lengths:
            seqs, ents = db.seqs(L), db.entries(L)
            ns = len(seqs) // L
            e2 = ents.copy()
            e2["seq_id"] += ns
            allents = np.concatenate([ents, e2])
            order = np.argsort(allents["mass"], kind="stable")
            self.tables[L] = (np.concatenate([seqs, seqs]), np.ascontiguousarray(allents[order]))

    def seqs(self, L):
        return self.tables[L][0]

    def entries(self, L):
        return self.tables[L][1]


def compare(res_g, hist_g, st_g, res_o, hist_o, st_o, label):
    valid = (res_o["flags"] & N.RES_VALID) != 0
    # every flag the oracle sets (the GPU adds HCX_RES_TIE_REPLAYED on its own)
    assert np.array_equal(res_g["flags"] & 0x3B, res_o["flags"] & 0x3B), label
    assert np.array_equal(res_g["cpsms"], res_o["cpsms"]), label
    for f in CHECK_EXACT:
        g, o = res_g[f][valid], res_o[f][valid]
        bad = np.nonzero(g != o)[0]
        assert len(bad) == 0, (label, f, bad[:5], g[bad[:5]], o[bad[:5]])
    hs_g, hs_o = res_g["hyperscore"][valid], res_o["hyperscore"][valid]
    assert np.all(np.abs(hs_g - hs_o) <= 1e-5 * np.abs(hs_o)), label
    e_g, e_o = res_g["e_x"][valid], res_o["e_x"][valid]
    assert np.all(np.abs(e_g - e_o) <= 1e-3 * np.abs(e_o)), label
    assert np.array_equal(hist_g, hist_o), label
    nh = int(((res_g["flags"] & N.RES_FIT_HOST) != 0).sum())
    if nh:
        log(f"  {label}: {nh} tail fits decided on the host")
    assert np.array_equal(st_g["n_ia"], st_o["n_ia"]) and np.array_equal(st_g["n_probe"], st_o["n_probe"]), label
    return int(valid.sum())


@pytest.fixture(scope="module")
def small_case():
    p = W.scaled("small", npep=4000, nspec=300, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    return p, db, q, prm


def test_index_parity_small(small_case):
    p, db, q, prm = small_case
    log("test_index_parity_small: building on GPU")
    eng = Engine(prm)
    for L in db.lengths:
        ix = eng.construct(L, db.seqs(L), db.entries(L))
        iA, bA = ix.export()
        oiA, obA = O.index_build(prm, L, db.seqs(L), db.entries(L))
        assert np.array_equal(bA, obA), L
        assert np.array_equal(iA, oiA), L
    eng.close()


@pytest.mark.parametrize("maxz", [1, 2, 3])
def test_index_parity_mods(maxz):
    p = W.scaled("rank_share", npep=5000, nspec=10, maxz=maxz, min_len=7, max_len=25)
    db = W.Database(p)
    prm = W.make_params(p)
    log(f"test_index_parity_mods maxz={maxz}")
    eng = Engine(prm)
    for L in db.lengths:
        ix = eng.construct(L, db.seqs(L), db.entries(L))
        iA, bA = ix.export()
        oiA, obA = O.index_build(prm, L, db.seqs(L), db.entries(L))
        assert np.array_equal(bA, obA), L
        assert np.array_equal(iA, oiA), L
    eng.close()


KERNELS = {"plain": N.KERNEL_PLAIN, "batched": N.KERNEL_BATCHED, "cursor": N.KERNEL_CURSOR}


SCANS = {"auto": N.SCAN_AUTO, "two-pass": N.SCAN_TWO_PASS}


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
@pytest.mark.parametrize("kernel", ["plain", "batched", "cursor"])
def test_query_parity_small(small_case, kernel, scan):
    p, db, q, prm = small_case
    if kernel == "cursor" and scan == "two-pass":
        pytest.skip("the cursor kernel has no scan modes")
    log(f"test_query_parity_small kernel={kernel} scan={scan}")
    odb = O.OracleDB(prm, db)
    res_o, hist_o, st_o = odb.search(q)
    eng = Engine(prm)
    eng.set_scan(SCANS[scan])
    eng.construct_db(db)
    b = eng.upload(q)
    out = eng.query_spectrum(b, kernel=KERNELS[kernel])
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"small-{kernel}-{scan}")
    assert n > 100
    two = (out.results["flags"] & N.RES_TWO_PASS) != 0
    if kernel != "cursor":
        # nearly every spectrum of this batch is within the packed scan's exact range
        assert two.all() if scan == "two-pass" else two.sum() < 0.05 * q.nspec, int(two.sum())
    # a second search on the same batch (work items rebuilt) gives the same answer
    out2 = eng.query_spectrum(b, kernel=KERNELS[kernel])
    assert np.array_equal(out2.results, out.results) and np.array_equal(out2.hist, out.hist)
    eng.close()


def crowd_peaks(q, every, copies):
    """Every `every`-th spectrum gets `copies` extra copies of its most intense
    peak (the last: peaks are stored in ascending intensity): more peaks and
    more intensity within 2 dF + 1 bins than the packed scan admits, so those
    spectra go to the two-pass kernel."""
    moz, inten, idx = [], [], [0]
    for s in range(q.nspec):
        m = q.moz[q.idx[s]:q.idx[s + 1]]
        i = q.intensity[q.idx[s]:q.idx[s + 1]]
        if s % every == 0 and len(m):
            m = np.concatenate([m, np.full(copies, m[-1], np.uint32)])
            i = np.concatenate([i, np.full(copies, i[-1], np.uint32)])
        moz.append(m)
        inten.append(i)
        idx.append(idx[-1] + len(m))
    return W.Queries(np.concatenate(moz).astype(np.uint32), np.concatenate(inten).astype(np.uint32),
                     np.asarray(idx, np.uint32), q.precurse.copy(), q.charges.copy(), q.rtimes.copy(),
                     q.src_len.copy(), q.src_psid.copy())


@pytest.mark.parametrize("kernel", ["plain", "batched"])
def test_packed_and_two_pass_mixed(small_case, kernel):
    # a batch split between the packed kernel and the two-pass kernel's list
    p, db, q0, prm = small_case
    q = crowd_peaks(q0, 3, 4)
    log(f"test_packed_and_two_pass_mixed kernel={kernel}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.construct_db(db)
    out = eng.query_spectrum(eng.upload(q), kernel=KERNELS[kernel])
    compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"mixed-{kernel}")
    two = (out.results["flags"] & N.RES_TWO_PASS) != 0
    crowded = (np.arange(q.nspec) % 3 == 0) & (np.diff(q0.idx.astype(np.int64)) > 0)
    assert two[crowded].all() and two[~crowded].sum() < 0.05 * q.nspec
    eng.close()


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
@pytest.mark.parametrize("grab", [1, 7, 256])
def test_batched_grab_sizes(small_case, scan, grab):
    # grabs of one item (a merge per item), of an odd size and of the largest
    p, db, q, prm = small_case
    log(f"test_batched_grab_sizes scan={scan} grab={grab}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm, chunk=1024)
    eng.set_scan(SCANS[scan])
    eng.set_grab(grab)
    eng.construct_db(db)
    out = eng.query_spectrum(eng.upload(q), kernel=N.KERNEL_BATCHED)
    compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"grab-{scan}-{grab}")
    eng.close()


@pytest.mark.parametrize("tile", [64, 256, 4096])
def test_cursor_kernel_tiles(small_case, tile):
    p, db, q, prm = small_case
    log(f"test_cursor_kernel_tiles tile={tile}")
    odb = O.OracleDB(prm, db)
    res_o, hist_o, st_o = odb.search(q)
    eng = Engine(prm)
    eng.construct_db(db)
    out = eng.query_spectrum(eng.upload(q), tile=tile, kernel=N.KERNEL_CURSOR)
    compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"tile{tile}")
    eng.close()


@pytest.fixture(scope="module")
def edge_case():
    p = W.scaled("rank_share", npep=20000, nspec=120, maxz=2)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    return db, q, prm, O.OracleDB(prm, db).search(q)


@pytest.mark.parametrize("kernel", ["plain", "batched"])
@pytest.mark.parametrize("chunk,slots", [(256, 64), (1024, 64), (8192, 256)])
@pytest.mark.parametrize("hitlist", [-1, 0, 64])
def test_chunked_edges_and_slot_rounds(edge_case, kernel, chunk, slots, hitlist):
    # small chunks: windows span many chunks and start/end inside chunks;
    # few slots: peptides reaching min_shp overflow the slot table (several rounds);
    # hit list off / 64 entries: the exact-count pass reads iA for all or most entries
    # (the two-pass kernel's own knobs: the packed scan is off)
    db, q, prm, (res_o, hist_o, st_o) = edge_case
    log(f"test_chunked_edges_and_slot_rounds kernel={kernel} chunk={chunk} slots={slots} hitlist={hitlist}")
    eng = Engine(prm, chunk=chunk, slots=slots, hitlist=hitlist)
    eng.set_scan(N.SCAN_TWO_PASS)
    eng.construct_db(db)
    out = eng.query_spectrum(eng.upload(q), kernel=KERNELS[kernel])
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"{kernel}-{chunk}-{slots}-{hitlist}")
    assert n > 50
    eng.close()


@pytest.mark.parametrize("kernel", ["plain", "batched"])
@pytest.mark.parametrize("chunk", [256, 1024, 8192])
def test_packed_chunk_edges(edge_case, kernel, chunk):
    # the packed scan with windows that span many chunks and start / end
    # inside chunks
    db, q, prm, (res_o, hist_o, st_o) = edge_case
    log(f"test_packed_chunk_edges kernel={kernel} chunk={chunk}")
    eng = Engine(prm, chunk=chunk)
    eng.construct_db(db)
    out = eng.query_spectrum(eng.upload(q), kernel=KERNELS[kernel])
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"packed-{kernel}-{chunk}")
    assert n > 50 and ((out.results["flags"] & N.RES_TWO_PASS) != 0).sum() < 0.05 * q.nspec
    eng.close()


@pytest.mark.parametrize("hitlist", [0, 256])
def test_open_search_groups_without_hit_list(hitlist):
    # 250-400 peaks x 5 bins (dF = 2): two run-table groups per item
    p = W.scaled("rank_share", npep=20000, nspec=100, maxz=2, peaks_min=250, peaks_max=400)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_open_search_groups_without_hit_list hitlist={hitlist}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm, hitlist=hitlist)
    eng.construct_db(db)
    b = eng.upload(q)
    for scan in ("auto", "two-pass"):
        eng.set_scan(SCANS[scan])
        for kernel in ("plain", "batched"):
            out = eng.query_spectrum(b, kernel=KERNELS[kernel])
            n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"groups-{hitlist}-{kernel}-{scan}")
            assert n > 50
    eng.close()


def test_candidate_sets_bitexact(small_case):
    p, db, q, prm = small_case
    log("test_candidate_sets_bitexact")
    odb = O.OracleDB(prm, db)
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.upload(q)
    g = eng.candidates(b)
    g = g[np.lexsort((g["psid"], g["idxoffset"], g["spec"]))]
    o = np.concatenate([odb.candidates(q, s) for s in range(q.nspec)])
    assert len(g) == len(o) and len(o) > 1000
    for f in ["spec", "idxoffset", "psid", "bc", "yc", "ibc", "iyc"]:
        assert np.array_equal(g[f], o[f]), f
    assert np.all(np.abs(g["hyperscore"] - o["hyperscore"]) <= 1e-5 * np.abs(o["hyperscore"]))
    eng.close()


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
@pytest.mark.parametrize("kernel", ["plain", "batched"])
def test_ties_replayed_in_reference_order(small_case, kernel, scan):
    p, db, q, prm = small_case
    log(f"test_ties_replayed_in_reference_order kernel={kernel} scan={scan}")
    twin = TwinDB(db)
    odb = O.OracleDB(prm, twin)
    res_o, hist_o, st_o = odb.search(q)
    eng = Engine(prm)
    eng.set_scan(SCANS[scan])
    eng.construct_db(twin)
    out = eng.query_spectrum(eng.upload(q), kernel=KERNELS[kernel])
    valid = (res_o["flags"] & N.RES_VALID) != 0
    assert np.all((out.results["flags"][valid] & N.RES_TIE_REPLAYED) != 0)
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, "twin")
    assert n > 100
    eng.close()


@pytest.mark.parametrize("maxz", [1, 2])
def test_query_parity_open_search_mods(maxz):
    p = W.scaled("rank_share", npep=30000, nspec=200, maxz=maxz)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_query_parity_open_search_mods maxz={maxz}: {db.nforms} forms")
    odb = O.OracleDB(prm, db)
    res_o, hist_o, st_o = odb.search(q)
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.upload(q)
    for kernel in ("plain", "batched"):
        out = eng.query_spectrum(b, kernel=KERNELS[kernel])
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"open{maxz}-{kernel}")
        assert n > 100
    eng.close()


def test_open_search_uneven_windows():
    # the benchmark's shape at 1/100 scale: 20-500 peaks, windows of 1e2-1e6 peptides
    p = W.scaled("open10m", npep=100000, nspec=300)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_open_search_uneven_windows: {db.nforms} peptides")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.upload(q)
    for kernel in ("plain", "batched"):
        out = eng.query_spectrum(b, kernel=KERNELS[kernel])
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"uneven-{kernel}")
        assert n > 200
    eng.close()


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
def test_bins_not_a_whole_number_of_offset_groups(scan):
    # NB = max_mass * scale = 499 900 bins: the (bin, chunk) offsets are stored
    # in groups of 8 bins, so the last group holds 4 bins and 4 padding words
    p = W.scaled("small", npep=4000, nspec=300, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    prm.max_mass = 4999
    assert (prm.max_mass * prm.scale) % 8 == 4
    log(f"test_bins_not_a_whole_number_of_offset_groups[{scan}]: NB {prm.max_mass * prm.scale}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.set_scan(SCANS[scan])
    eng.construct_db(db)
    b = eng.upload(q)
    for kernel in ("plain", "batched"):
        out = eng.query_spectrum(b, kernel=KERNELS[kernel])
        n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"nb-odd-{scan}-{kernel}")
        assert n > 100
    eng.close()


def test_tailfit_batch_hand_traced():
    prm = W.make_params(W.PRESETS["small"])
    eng = Engine(prm)
    H = np.zeros((4, 1002), np.uint32)
    H[0, 2:10] = [5, 10, 8, 6, 3, 1, 0, 1]
    H[0, 10] = 1
    H[1, 5], H[1, 8] = 4, 1
    H[2, 7], H[2, 8] = 3, 1
    H[3, 4], H[3, 6], H[3, 10] = 2, 1, 1
    cps = np.array([35, 5, 4, 4], np.uint32)
    hyp = np.array([10, 8, 8, 10], np.int32)
    out = eng.model_tail_fit(H, cps, hyp)
    for i in range(4):
        mu, beta, mn, nx, fl = O.tailfit(H[i].astype(np.float64), int(cps[i]), int(hyp[i]))
        assert (out["mu"][i], out["beta"][i], out["minhyp"][i], out["nexthyp"][i]) == (mu, beta, mn, nx), i
    eng.close()


def test_tailfit_batch_random_histograms():
    rng = np.random.default_rng(11)
    prm = W.make_params(W.PRESETS["small"])
    eng = Engine(prm)
    n = 2000
    H = np.zeros((n, 1002), np.uint32)
    hyp = rng.integers(20, 900, n).astype(np.int32)
    for i in range(n):
        k = int(rng.integers(3, 400))
        centre = rng.uniform(5, hyp[i] * 0.6)
        b = np.clip(rng.normal(centre, rng.uniform(3, 40), k), 0, hyp[i] - 1).astype(int)
        np.add.at(H[i], b, 1)
        H[i, hyp[i]] += 1
    cps = H.sum(axis=1).astype(np.uint32)
    out = eng.model_tail_fit(H, cps, hyp)
    mism = 0
    for i in range(n):
        mu, beta, mn, nx, fl = O.tailfit(H[i].astype(np.float64), int(cps[i]), int(hyp[i]))
        if (out["mu"][i], out["beta"][i], out["minhyp"][i], out["nexthyp"][i]) != (mu, beta, mn, nx):
            mism += 1
    assert mism == 0
    eng.close()


def random_histograms(rng, n, hbins=1002):
    H = np.zeros((n, hbins), np.uint32)
    hyp = rng.integers(20, hbins - 102, n).astype(np.int32)
    for i in range(n):
        k = int(rng.integers(3, 400))
        centre = rng.uniform(5, hyp[i] * 0.6)
        b = np.clip(rng.normal(centre, rng.uniform(3, 40), k), 0, hyp[i] - 1).astype(int)
        np.add.at(H[i], b, 1)
        H[i, hyp[i]] += 1
    return H, H.sum(axis=1).astype(np.uint32), hyp


@pytest.mark.parametrize("force", [0, 2])
def test_tailfit_host_refit_exact(force):
    # slope and intercept are FP64 values x 1e6 truncated to ints; fits whose
    # values sit within the two log10s' error bound of an integer (or of a
    # window threshold) are decided on the host: every int equals the
    # oracle's (host libm) exactly.  force=2: every other fit goes through
    # the host path on purpose.
    rng = np.random.default_rng(23 + force)
    prm = W.make_params(W.PRESETS["small"])
    eng = Engine(prm)
    eng.set_option(N.OPT_FORCE_HOST_FIT, force)
    n = 20000
    H, cps, hyp = random_histograms(rng, n)
    out = eng.model_tail_fit(H, cps, hyp)
    bad = []
    for i in range(n):
        mu, beta, mn, nx, fl = O.tailfit(H[i].astype(np.float64), int(cps[i]), int(hyp[i]))
        if (out["mu"][i], out["beta"][i], out["minhyp"][i], out["nexthyp"][i]) != (mu, beta, mn, nx):
            bad.append(i)
    host = (out["flags"] & N.RES_FIT_HOST) != 0
    log(f"test_tailfit_host_refit_exact force={force}: {int(host.sum())} of {n} fitted on the host")
    assert not bad, bad[:5]
    if force:
        assert np.array_equal(host[::2], np.ones(len(host[::2]), bool))
    eng.close()


@pytest.mark.parametrize("kernel", ["plain", "batched"])
def test_search_host_refit(small_case, kernel):
    # every third spectrum's tail fit recomputed on the host inside the
    # search (gathered from the device histogram, written back): results
    # equal the oracle's
    p, db, q, prm = small_case
    log(f"test_search_host_refit kernel={kernel}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.set_option(N.OPT_FORCE_HOST_FIT, 3)
    eng.construct_db(db)
    b = eng.upload(q)
    out = eng.query_spectrum(b, kernel=KERNELS[kernel])
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"host-refit-{kernel}")
    assert n > 100
    valid = (res_o["flags"] & N.RES_VALID) != 0
    host = (out.results["flags"] & N.RES_FIT_HOST) != 0
    third = np.arange(q.nspec) % 3 == 0
    assert np.all(host[valid & third])
    g = eng.query_spectrum(b, kernel=KERNELS[kernel], use_graph=1)
    assert np.array_equal(g.results, out.results)
    eng.close()


def test_graph_replays_match_eager(small_case):
    # the captured search sequence, replayed several times with eager searches
    # and synchronisations in between; every replay equals the eager result
    p, db, q, prm = small_case
    log("test_graph_replays_match_eager")
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.upload(q)
    for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
        ref = eng.query_spectrum(b, kernel=kernel)
        for i in range(3):
            g = eng.query_spectrum(b, kernel=kernel, use_graph=1)
            assert np.array_equal(g.results, ref.results), (kernel, i)
            assert np.array_equal(g.hist, ref.hist) and np.array_equal(g.stats, ref.stats), (kernel, i)
            e = eng.query_spectrum(b, kernel=kernel)
            assert np.array_equal(e.results, ref.results), (kernel, i)
        k, t = eng.query_device(b, kernel=kernel, use_graph=1)
        assert k > 0 and t >= k
    eng.close()


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
def test_autotune_keeps_parity(scan):
    p = W.scaled("open10m", npep=100000, nspec=300)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    log(f"test_autotune_keeps_parity scan={scan}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.set_scan(SCANS[scan])
    eng.construct_db(db)
    b = eng.upload(q)
    best = eng.autotune(b, N.KERNEL_BATCHED)
    log(f"autotune: {best}")
    assert best["kernel_ms"] > 0 and best["grab"] in (8, 16, 32, 64, 128, 256)
    assert best["slots"] in (512, 1024, 1536, 2048) and best["blocks_per_cu"] in (1, 2, 3)
    t = eng.tuning()
    assert (t["slots"], t["blocks_per_cu"], t["grab"]) == (best["slots"], best["blocks_per_cu"], best["grab"])
    out = eng.query_spectrum(b, kernel=N.KERNEL_BATCHED)
    compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"autotuned-{scan}")
    eng.close()


def test_index_build_tiny_illegal_and_empty():
    # bA read off the sorted bins: one form, forms whose precursor mass is
    # below min_mass (every ion in bin 0, dslim.cpp:317-322), a mix, and no forms
    prm = W.make_params(W.PRESETS["small"])
    log("test_index_build_tiny_illegal_and_empty")
    L = 7
    seqs = np.frombuffer(b"GGGGGGK" + b"PEPTIDK" + b"AAAAAAR" + b"WWWWWWR", np.uint8).copy()

    def entries(ids):
        e = np.zeros(len(ids), N.PEP_DTYPE)
        for k, sid in enumerate(ids):
            e[k]["seq_id"] = sid
            e[k]["mass"] = O.pep_mass(bytes(seqs[sid * L:(sid + 1) * L]).decode())
        order = np.argsort(e["mass"], kind="stable")
        return e[order]

    eng = Engine(prm)
    # GGGGGGK (488 Da) and AAAAAAR (~600 Da) straddle min_mass 500
    for ids in ([1], [0], [0, 0, 0], [0, 1, 2, 3, 0, 2], []):
        ents = entries(ids)
        ix = eng.construct(L, seqs, ents)
        iA, bA = ix.export()
        oiA, obA = O.index_build(prm, L, seqs, ents)
        assert np.array_equal(bA, obA), ids
        assert np.array_equal(iA, oiA), ids
        assert bA[0] == 0 and bA[-1] == len(ids) * 2 * prm.maxz * (L - 1)
    eng.close()


@pytest.mark.parametrize("scan", ["auto", "two-pass"])
@pytest.mark.parametrize("kernel", ["plain", "batched"])
def test_forced_redo_on_cursor_kernel(small_case, kernel, scan):
    # spectra listed for an exact re-run (what k_finalize does for spectra
    # whose 8-bit counters neared overflow) are re-scored by the cursor kernel
    # and their ties replayed there; forced here for every third spectrum,
    # on a database whose every top hit is tied
    import ctypes as C
    p, db, q, prm = small_case
    log(f"test_forced_redo_on_cursor_kernel kernel={kernel} scan={scan}")
    twin = TwinDB(db)
    res_o, hist_o, st_o = O.OracleDB(prm, twin).search(q)
    eng = Engine(prm)
    eng.set_scan(SCANS[scan])
    N.check(N.hcx().hcx_set_option(eng._h, N.OPT_FORCE_REDO, C.c_int64(3)))
    eng.construct_db(twin)
    out = eng.query_spectrum(eng.upload(q), kernel=KERNELS[kernel])
    n = compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"redo-{kernel}-{scan}")
    assert n > 100
    redone = (out.results["flags"] & N.RES_COUNT_OVF) != 0
    assert np.array_equal(redone, np.arange(q.nspec) % 3 == 0)
    valid = (res_o["flags"] & N.RES_VALID) != 0
    assert np.all((out.results["flags"][valid] & N.RES_TIE_REPLAYED) != 0)
    eng.close()


def test_log2_estimate_error_bound():
    # the packed scan settles most scores from log2 of the exact integer
    # sums (1 .. 2^24 + 1) by v_log_f32, assuming its error is below a bound:
    # measure it over the whole range
    import ctypes as C
    prm = W.make_params(W.PRESETS["small"])
    eng = Engine(prm)
    n = (1 << 24) + 1
    out = np.zeros(n, np.float32)
    bound = C.c_double()
    N.check(N.hcx().hcx_selftest_log2(eng._h, C.c_uint32(n), N.ptr(out), C.byref(bound)))
    err = np.abs(out.astype(np.float64) - np.log2(np.arange(1, n + 1, dtype=np.float64)))
    log(f"v_log_f32 over 1..2^24+1: max error {err.max():.3e}, assumed bound {bound.value:.1e}")
    assert err.max() < bound.value / 4
    eng.close()


class HalfDB:
    """The first half of every length's pepEntries (the same sequences)."""

    def __init__(self, db):
        self.lengths = db.lengths
        self.db = db

    def seqs(self, L):
        return self.db.seqs(L)

    def entries(self, L):
        e = self.db.entries(L)
        return e[:max(1, len(e) // 2)]


@pytest.mark.parametrize("use_graph", [0, 1])
def test_rebuilt_index_same_batch(small_case, use_graph):
    # search a batch, free the index, rebuild it with fewer peptides (the
    # allocator may hand back the same addresses), search the same uploaded
    # batch again: both searches equal the oracle on their own index
    p, db, q, prm = small_case
    log(f"test_rebuilt_index_same_batch use_graph={use_graph}")
    eng = Engine(prm)
    b = eng.upload(q)
    for d in (db, HalfDB(db), db):
        ix = eng.construct_db(d)
        res_o, hist_o, st_o = O.OracleDB(prm, d).search(q)
        for kernel in (N.KERNEL_BATCHED, N.KERNEL_PLAIN):
            out = eng.query_spectrum(b, indices=ix, kernel=kernel, use_graph=use_graph)
            compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"rebuilt-{type(d).__name__}-{kernel}")
        for x in ix:
            x.free()
        assert not eng.indices
    eng.close()


@pytest.mark.parametrize("kernel", ["plain", "batched"])
def test_packed_bound_at_its_limit(kernel):
    # Spectra made of one longest-length peptide's own ion bins, peaks more
    # than 2 dF apart, every intensity I: each ion takes one hit of I, so its
    # b series sums to half * I with half = speclen / 2 of the longest length.
    # At I = (2^24 - 1) // half the packed words sit one step below a carry
    # into the count field and the spectrum must still be packed and exact;
    # at I + 1 the bound sends it to the two-pass kernel.
    p = W.scaled("small", npep=3000, nspec=10, dM=100.0)
    db = W.Database(p)
    prm = W.make_params(p)
    L = max(db.lengths)
    half = prm.maxz * (L - 1)
    seq = db.sequence(L, 0)
    mass, bins = O.gen_spectrum(seq, prm)
    span = 2 * prm.dF
    kept = []
    for b in sorted(set(int(x) for x in bins)):
        if prm.dF < b < prm.max_mass * prm.scale - 1 - prm.dF and (not kept or b - kept[-1] > span):
            kept.append(b)
    I = ((1 << 24) - 1) // half
    k = len(kept)
    moz = np.array(kept * 2, np.uint32)
    inten = np.array([I] * k + [I + 1] * k, np.uint32)
    q = W.Queries(moz, inten, np.array([0, k, 2 * k], np.uint32), np.full(2, mass, np.float32),
                  np.full(2, 2, np.int32), np.zeros(2, np.float32), np.zeros(2, np.uint32), np.zeros(2, np.int32))
    log(f"test_packed_bound_at_its_limit kernel={kernel}: L={L} half={half} I={I} peaks={k}")
    res_o, hist_o, st_o = O.OracleDB(prm, db).search(q)
    eng = Engine(prm)
    eng.construct_db(db)
    b = eng.upload(q)
    f = b.facts()
    assert list(f["mult"]) == [1, 1] and list(f["wsum"]) == [I, I + 1]
    out = eng.query_spectrum(b, kernel=KERNELS[kernel])
    compare(out.results, out.hist, out.stats, res_o, hist_o, st_o, f"bound-{kernel}")
    two = (out.results["flags"] & N.RES_TWO_PASS) != 0
    assert list(two) == [False, True]
    # the peptide's b series in the packed spectrum: nearly every ion hit once,
    # its sum at most one count step (I) short of 2^24
    kept_a = np.array(kept, np.int64)
    hit_b = sum(1 for x in bins[: len(bins) // 2] if np.any(np.abs(kept_a - int(x)) <= prm.dF))
    log(f"  b ions hit {hit_b} of {half}: b sum {hit_b * I}, limit {(1 << 24) - 1}")
    assert hit_b >= half - 4 and hit_b * I <= (1 << 24) - 1
    eng.close()
