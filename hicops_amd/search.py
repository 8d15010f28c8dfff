"""Python face of the search core, mirroring the reference's entry points.

  Engine(params)                  gParams + DSLIM_InitializeScorecard (dslim.cpp:390-437)
  Engine.count_peps(L, lines, ...) LBE_CountPeps + MODS_ModCounter for one peptide length
                                  (lbe_internal.cpp:369-467, mods.cpp:399-432)
  Engine.initialize(dbs)          LBE_Initialize (lbe_internal.cpp:119-202) for several lengths at
                                  once: LBE_GeneratePeps, MODS_GenerateMods, std::sort by mass
  Engine.construct_pepdb(db)      DSLIM_Construct over a device-resident PepDb
  Engine.construct(L, seqs, ents) DSLIM_Construct(Index*) for one peptide length (dslim.cpp:46-138)
  Engine.pickpeaks(raw, ...)      MSQuery::pickpeaks (msquery.cpp:509-569) over a batch, the
                                  Queries producer (device-resident result)
  Engine.query_spectrum(queries)  DSLIM_QuerySpectrum(Queries*, Index*, nIdx) (dslim_query.cpp:515-842)
                                  including expeRT::ModelTailFit and the e-value block (:771-818)
  Engine.model_tail_fit(...)      expeRT::ModelTailFit(Results*) (expeRT.cpp:572-718) on histograms

All work runs in libhcx.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native as N


@dataclass
class SearchOutput:
    results: np.ndarray           # RESULT_DTYPE per spectrum
    hist: np.ndarray | None       # (nspec, hbins) u32 score histograms
    stats: np.ndarray | None      # STATS_DTYPE per spectrum


class Index:
    """One peptide length's DSLIM index, resident in HBM."""

    def __init__(self, engine: "Engine", peplen: int, seqs: np.ndarray, entries: np.ndarray):
        self.engine = engine
        self.peplen = peplen
        seqs = np.ascontiguousarray(seqs, dtype=np.uint8)
        entries = np.ascontiguousarray(entries, dtype=N.PEP_DTYPE)
        nseqs = len(seqs) // peplen
        h = C.c_void_p()
        N.check(N.hcx().hcx_index_build(engine._h, peplen, N.ptr(seqs), nseqs, N.ptr(entries),
                                        len(entries), C.byref(h)))
        self._h = h
        pl, sl, ne, ni = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        N.check(N.hcx().hcx_index_info(h, C.byref(pl), C.byref(sl), C.byref(ne), C.byref(ni)))
        self.speclen, self.n, self.nions = sl.value, ne.value, ni.value

    @property
    def build_ms(self) -> float:
        ms = C.c_float()
        N.check(N.hcx().hcx_index_build_ms(self._h, C.byref(ms)))
        return ms.value

    def export(self) -> tuple[np.ndarray, np.ndarray]:
        iA = np.zeros(self.nions, np.uint32)
        bA = np.zeros(self.engine.NB + 1, np.uint32)
        N.check(N.hcx().hcx_index_export(self._h, N.ptr(iA), N.ptr(bA)))
        return iA, bA

    def free(self) -> None:
        if getattr(self, "_h", None):
            N.hcx().hcx_index_free(self._h)
            self._h = None
            eng = getattr(self, "engine", None)
            if eng is not None and self in getattr(eng, "indices", []):
                eng.indices.remove(self)

    def __del__(self):
        self.free()


class PepDb:
    """One peptide length's database (the reference's per-length Index before
    DSLIM_Construct: sequences and pepEntries), resident in HBM."""

    def __init__(self, engine: "Engine", peplen: int, lines: np.ndarray, modconditions: str | None,
                 nodes: int = 1, myid: int = 0, policy: int = N.POLICY_CYCLIC, sort: bool = True):
        self.engine = engine
        self.peplen = peplen
        lines = np.ascontiguousarray(lines, dtype=np.uint8)
        cfg = N.LbeCfg()
        cfg.modconditions = modconditions.encode() if modconditions is not None else None
        cfg.nodes, cfg.myid, cfg.policy, cfg.sort = nodes, myid, policy, int(bool(sort))
        h = C.c_void_p()
        N.check(N.hcx().hcx_pepdb_create(engine._h, C.byref(cfg), C.c_uint32(peplen), N.ptr(lines),
                                         C.c_uint32(len(lines) // peplen), C.byref(h)))
        self._h = h
        cnt = np.zeros(6, np.uint32)
        N.check(N.hcx().hcx_pepdb_counts(h, N.ptr(cnt)))
        (self.pep_count, self.mod_count, self.total_count, self.lclpep, self.lclmod,
         self.lcltot) = (int(x) for x in cnt)

    def timing(self) -> tuple[float, float, float]:
        """(upload ms, filter/count ms, generate ms) on the device."""
        a, b, g = C.c_float(), C.c_float(), C.c_float()
        N.check(N.hcx().hcx_pepdb_ms(self._h, C.byref(a), C.byref(b), C.byref(g)))
        return a.value, b.value, g.value

    def export(self) -> tuple[np.ndarray, np.ndarray]:
        """(kept sequences as pep_count * peplen bytes, pepEntries)."""
        seqs = np.zeros(self.pep_count * self.peplen, np.uint8)
        ents = np.zeros(self.lcltot, N.PEP_DTYPE)
        N.check(N.hcx().hcx_pepdb_export(self._h, N.ptr(seqs), N.ptr(ents)))
        return seqs, ents

    def free(self) -> None:
        if getattr(self, "_h", None):
            N.hcx().hcx_pepdb_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


class Batch:
    """A Queries batch resident in HBM: uploaded from the host, or produced on
    the device by Engine.pickpeaks."""

    def __init__(self, engine: "Engine", queries, handle: C.c_void_p | None = None):
        self.engine = engine
        self.queries = queries
        if handle is None:
            h = C.c_void_p()
            v = queries.view()
            N.check(N.hcx().hcx_queries_upload(engine._h, C.byref(v), C.byref(h)))
            handle = h
        self._h = handle
        n = C.c_int32()
        N.check(N.hcx().hcx_queries_info(handle, C.byref(n), None))
        self.nspec = n.value

    def export(self):
        """Host copy of the batch as a workload.Queries (no source fields)."""
        from .workload import Queries
        n, k = C.c_int32(), C.c_uint64()
        N.check(N.hcx().hcx_queries_info(self._h, C.byref(n), C.byref(k)))
        n, k = n.value, k.value
        q = Queries(np.zeros(k, np.uint32), np.zeros(k, np.uint32), np.zeros(n + 1, np.uint32),
                    np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.float32),
                    np.zeros(n, np.uint32), np.zeros(n, np.int32))
        N.check(N.hcx().hcx_queries_export(self._h, N.ptr(q.moz), N.ptr(q.intensity), N.ptr(q.idx),
                                           N.ptr(q.precurse), N.ptr(q.charges), N.ptr(q.rtimes)))
        return q

    def facts(self) -> dict:
        """The batch facts the search sizes its exact scan by: largest
        spectrum, largest intensity, most peaks within 2*dF+1 bins (batch and
        per spectrum) and, per spectrum, the largest intensity sum within
        2*dF+1 bins."""
        mp, mi, mm = C.c_uint32(), C.c_uint32(), C.c_uint32()
        mult = np.zeros(self.nspec, np.uint32)
        wsum = np.zeros(self.nspec, np.uint32)
        N.check(N.hcx().hcx_queries_facts(self._h, C.byref(mp), C.byref(mi), C.byref(mm), N.ptr(mult), N.ptr(wsum)))
        return {"max_peaks": mp.value, "max_int": mi.value, "max_mult": mm.value, "mult": mult, "wsum": wsum}

    def free(self) -> None:
        if getattr(self, "_h", None):
            N.hcx().hcx_queries_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


class RawBatch:
    """Raw MS2 spectra (workload.RawSpectra) resident in HBM, before pickpeaks."""

    def __init__(self, engine: "Engine", raw):
        self.engine = engine
        h = C.c_void_p()
        v = raw.view()
        N.check(N.hcx().hcx_raw_upload(engine._h, C.byref(v), C.byref(h)))
        self._h = h
        self.nspec = raw.nspec

    def free(self) -> None:
        if getattr(self, "_h", None):
            N.hcx().hcx_raw_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


class Engine:
    def __init__(self, params: N.Params, device: int = 0, chunk: int = 0, slots: int = 0,
                 blocks_per_cu: int = 0, hitlist: int = -1):
        self.params = params
        h = C.c_void_p()
        N.check(N.hcx().hcx_create(C.byref(params), device, C.byref(h)))
        self._h = h
        self.NB = params.max_mass * params.scale
        self.hbins = params.hbins
        self.indices: list[Index] = []
        self.set_tuning(chunk, slots, blocks_per_cu)
        self.set_hitlist(hitlist)

    def set_tuning(self, chunk: int = 0, slots: int = 0, blocks_per_cu: int = 0) -> None:
        """chunk must be set before indices are built (they record it)."""
        N.check(N.hcx().hcx_set_tuning(self._h, C.c_uint32(chunk), C.c_uint32(slots), C.c_uint32(blocks_per_cu)))

    def set_hitlist(self, entries: int = -1) -> None:
        """LDS hit-list entries per work item: -1 as many as fit, 0 none."""
        N.check(N.hcx().hcx_set_option(self._h, N.OPT_HITLIST, C.c_int64(entries)))

    def set_option(self, option: int, value: int) -> None:
        N.check(N.hcx().hcx_set_option(self._h, option, C.c_int64(value)))

    def set_scan(self, mode: int) -> None:
        """N.SCAN_AUTO: the packed single-pass scan where it is exact, the
        two-pass scan elsewhere; N.SCAN_TWO_PASS: the two-pass scan throughout."""
        N.check(N.hcx().hcx_set_option(self._h, N.OPT_SCAN, C.c_int64(mode)))

    def set_grab(self, items: int) -> None:
        """Batched kernels: work items per queue grab, at most (1..256)."""
        N.check(N.hcx().hcx_set_option(self._h, N.OPT_GRAB, C.c_int64(items)))

    def set_group(self, probes: int) -> None:
        """(peak, bin) probes per run table of the chunked kernels: 0 automatic, else 64..2560."""
        N.check(N.hcx().hcx_set_option(self._h, N.OPT_GROUP, C.c_int64(probes)))

    def tuning(self) -> dict:
        ch, sl, bp = C.c_uint32(), C.c_uint32(), C.c_uint32()
        cu, lds = C.c_int32(), C.c_int32()
        N.check(N.hcx().hcx_get_tuning(self._h, C.byref(ch), C.byref(sl), C.byref(bp), C.byref(cu), C.byref(lds)))
        grab, scan = C.c_int64(), C.c_int64()
        N.check(N.hcx().hcx_get_option(self._h, N.OPT_GRAB, C.byref(grab)))
        N.check(N.hcx().hcx_get_option(self._h, N.OPT_SCAN, C.byref(scan)))
        return {"chunk": ch.value, "slots": sl.value, "blocks_per_cu": bp.value, "num_cu": cu.value,
                "max_lds": lds.value, "grab": grab.value, "scan": "auto" if scan.value == N.SCAN_AUTO else "two-pass"}

    # DSLIM_Construct, one length at a time (hicops.cpp:200-300 loop)
    def construct(self, peplen: int, seqs: np.ndarray, entries: np.ndarray) -> Index:
        ix = Index(self, peplen, seqs, entries)
        self.indices.append(ix)
        return ix

    def construct_db(self, db) -> list[Index]:
        return [self.construct(L, db.seqs(L), db.entries(L)) for L in db.lengths]

    # the peptide database stage (LBE + MODS) on the device
    def count_peps(self, peplen: int, lines: np.ndarray, modconditions: str | None, nodes: int = 1,
                   myid: int = 0, policy: int = N.POLICY_CYCLIC, sort: bool = True) -> PepDb:
        return PepDb(self, peplen, lines, modconditions, nodes, myid, policy, sort)

    def initialize(self, dbs: list) -> None:
        arr = (C.c_void_p * max(1, len(dbs)))(*[d._h for d in dbs])
        N.check(N.hcx().hcx_pepdb_generate(arr, C.c_uint32(len(dbs))))

    def construct_pepdb(self, db: PepDb) -> Index:
        ix = Index.__new__(Index)
        ix.engine = self
        ix.peplen = db.peplen
        h = C.c_void_p()
        N.check(N.hcx().hcx_index_build_pepdb(self._h, db._h, C.byref(h)))
        ix._h = h
        pl, sl, ne, ni = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        N.check(N.hcx().hcx_index_info(h, C.byref(pl), C.byref(sl), C.byref(ne), C.byref(ni)))
        ix.speclen, ix.n, ix.nions = sl.value, ne.value, ni.value
        self.indices.append(ix)
        return ix

    def sort_mass(self, keys: np.ndarray, vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """The device mass sort alone (libstdc++ std::sort order by key)."""
        k = np.ascontiguousarray(keys, np.float32).copy()
        v = np.ascontiguousarray(vals, np.uint32).copy()
        N.check(N.hcx().hcx_sort_mass(self._h, N.ptr(k), N.ptr(v), C.c_uint32(len(k))))
        return k, v

    def upload(self, queries) -> Batch:
        return Batch(self, queries)

    def upload_raw(self, raw) -> RawBatch:
        return RawBatch(self, raw)

    def pickpeaks(self, raw, qalen: int, base_int: int, min_int: int) -> Batch:
        """MSQuery::pickpeaks (msquery.cpp:509-569) over a batch of raw spectra
        (workload.RawSpectra or a RawBatch) on the device; the Queries batch
        stays in HBM for query_spectrum."""
        return self.pickpeaks_timed(raw, qalen, base_int, min_int)[0]

    def pickpeaks_timed(self, raw, qalen: int, base_int: int, min_int: int) -> tuple[Batch, float, float]:
        """pickpeaks returning (batch, normalisation kernel ms, whole pass ms)."""
        rb = raw if isinstance(raw, RawBatch) else RawBatch(self, raw)
        pp = N.PickParams()
        pp.qalen, pp.base_int, pp.min_int = qalen, base_int, min_int
        h = C.c_void_p()
        km, tm = C.c_float(), C.c_float()
        N.check(N.hcx().hcx_queries_pick(self._h, rb._h, C.byref(pp), C.byref(h), C.byref(km), C.byref(tm)))
        if rb is not raw:
            rb.free()
        return Batch(self, None, handle=h), km.value, tm.value

    def _idx_array(self, indices):
        indices = self.indices if indices is None else indices
        arr = (C.c_void_p * max(1, len(indices)))(*[ix._h for ix in indices])
        return arr, len(indices)

    @staticmethod
    def _opts(kernel=N.KERNEL_PLAIN, evalue=N.EVAL_TAILFIT, tile=0, use_graph=0, audit_every=0,
              audit_tol=1e-3) -> N.SearchOpts:
        o = N.SearchOpts()
        o.kernel, o.evalue, o.tile, o.use_graph = kernel, evalue, tile, use_graph
        o.audit_every, o.audit_tol = audit_every, audit_tol
        return o

    # DSLIM_QuerySpectrum + ModelTailFit + e-value
    def query_spectrum(self, batch: Batch, indices=None, want_hist=True, want_stats=True,
                       **opts) -> SearchOutput:
        arr, n = self._idx_array(indices)
        res = np.zeros(batch.nspec, N.RESULT_DTYPE)
        hist = np.zeros((batch.nspec, self.hbins), np.uint32) if want_hist else None
        stats = np.zeros(batch.nspec, N.STATS_DTYPE) if want_stats else None
        o = self._opts(**opts)
        N.check(N.hcx().hcx_search(self._h, arr, n, batch._h, C.byref(o), N.ptr(res),
                                   N.ptr(hist), N.ptr(stats)))
        return SearchOutput(res, hist, stats)

    def query_device(self, batch: Batch, indices=None, **opts) -> tuple[float, float]:
        """Run a search leaving results on the device; returns (query kernel ms, total ms)."""
        arr, n = self._idx_array(indices)
        o = self._opts(**opts)
        k, t = C.c_float(), C.c_float()
        N.check(N.hcx().hcx_search_device(self._h, arr, n, batch._h, C.byref(o), C.byref(k), C.byref(t)))
        return k.value, t.value

    def autotune(self, batch: Batch, kernel: int = N.KERNEL_BATCHED, indices=None) -> dict:
        """Time the scan for each candidate tuning on this batch; keep the fastest."""
        arr, n = self._idx_array(indices)
        t = N.Tuning()
        N.check(N.hcx().hcx_autotune(self._h, arr, n, batch._h, kernel, C.byref(t)))
        return {f: getattr(t, f) for f, _ in N.Tuning._fields_}

    # ---- learned tail-parameter surrogate ----
    def surrogate_init(self, seed: int = 1) -> None:
        N.check(N.hcx().hcx_surrogate_init(self._h, C.c_uint64(seed)))

    def surrogate_train(self, batch: Batch, steps: int = 2000, batch_size: int = 1024, lr: float = 1e-3,
                        seed: int = 7) -> float:
        """Train on `batch`, which must have been searched with the plain fit."""
        loss = C.c_float()
        N.check(N.hcx().hcx_surrogate_train(self._h, batch._h, C.c_uint32(steps), C.c_uint32(batch_size),
                                            C.c_float(lr), C.c_uint64(seed), C.byref(loss)))
        return loss.value

    def surrogate_params(self) -> tuple[np.ndarray, np.ndarray]:
        n = N.hcx().hcx_surrogate_param_count()
        p = np.zeros(n, np.float32)
        nrm = np.zeros(6, np.float32)
        N.check(N.hcx().hcx_surrogate_get(self._h, N.ptr(p), C.c_uint32(n), N.ptr(nrm)))
        return p, nrm

    def surrogate_set(self, params: np.ndarray, norm6: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        nrm = np.ascontiguousarray(norm6, np.float32)
        N.check(N.hcx().hcx_surrogate_set(self._h, N.ptr(p), C.c_uint32(len(p)), N.ptr(nrm)))

    def surrogate_gradient(self, X: np.ndarray, T: np.ndarray) -> tuple[np.ndarray, float]:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(N.hcx().hcx_surrogate_param_count(), np.float32)
        loss = C.c_float()
        N.check(N.hcx().hcx_surrogate_gradient(self._h, N.ptr(X), N.ptr(T), C.c_uint32(len(X)), N.ptr(g),
                                               C.byref(loss)))
        return g, loss.value

    def last_timing(self) -> dict:
        """Device ms of the last query_device: scan kernel, e-value stage, whole search."""
        q, e, t = C.c_float(), C.c_float(), C.c_float()
        N.check(N.hcx().hcx_last_timing(self._h, C.byref(q), C.byref(e), C.byref(t)))
        pk, tp = C.c_float(), C.c_float()
        N.check(N.hcx().hcx_last_scan_split(self._h, C.byref(pk), C.byref(tp)))
        return {"scan_ms": q.value, "evalue_ms": e.value, "total_ms": t.value,
                "packed_ms": pk.value, "two_pass_ms": tp.value}

    def last_audit(self) -> dict:
        a = N.Audit()
        N.check(N.hcx().hcx_surrogate_last_audit(self._h, C.byref(a)))
        return {"audited": a.audited, "disagreed": a.disagreed, "max_rel_err": a.max_rel_err,
                "fell_back": bool(a.fell_back), "located": a.located, "plain_fit": a.plain_fit}

    def selftest_mfma(self, A: np.ndarray, B: np.ndarray) -> np.ndarray:
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        D = np.zeros((16, 16), np.float32)
        N.check(N.hcx().hcx_selftest_mfma(self._h, N.ptr(A), N.ptr(B), N.ptr(D)))
        return D

    def fetch(self, batch: Batch, want_hist=True, want_stats=True) -> SearchOutput:
        res = np.zeros(batch.nspec, N.RESULT_DTYPE)
        hist = np.zeros((batch.nspec, self.hbins), np.uint32) if want_hist else None
        stats = np.zeros(batch.nspec, N.STATS_DTYPE) if want_stats else None
        N.check(N.hcx().hcx_search_fetch(self._h, batch._h, N.ptr(res), N.ptr(hist), N.ptr(stats)))
        return SearchOutput(res, hist, stats)

    def candidates(self, batch: Batch, indices=None, cap: int = 1 << 20) -> np.ndarray:
        arr, n = self._idx_array(indices)
        out = np.zeros(cap, N.CAND_DTYPE)
        cnt = C.c_uint64()
        N.check(N.hcx().hcx_search_candidates(self._h, arr, n, batch._h, N.ptr(out), cap, C.byref(cnt)))
        if cnt.value > cap:
            raise N.HcxError(N.ERR_INVLD_SIZE, f"candidate dump needs {cnt.value} > {cap}")
        return out[:cnt.value]

    # expeRT::ModelTailFit over a batch of histograms
    def model_tail_fit(self, hist: np.ndarray, cpsms: np.ndarray, maxhyp: np.ndarray):
        hist = np.ascontiguousarray(hist, np.uint32)
        n, hb = hist.shape
        cpsms = np.ascontiguousarray(cpsms, np.uint32)
        maxhyp = np.ascontiguousarray(maxhyp, np.int32)
        out = {k: np.zeros(n, np.int32) for k in ("mu", "beta", "minhyp", "nexthyp")}
        flags = np.zeros(n, np.uint32)
        N.check(N.hcx().hcx_tailfit_batch(self._h, N.ptr(hist), hb, n, N.ptr(cpsms), N.ptr(maxhyp),
                                          N.ptr(out["mu"]), N.ptr(out["beta"]), N.ptr(out["minhyp"]),
                                          N.ptr(out["nexthyp"]), N.ptr(flags)))
        out["flags"] = flags
        return out

    # partial results of index shards and their merge (the reference's
    # multi-rank wire format, dslim_query.cpp:733-763 / dslim_score.cpp:198-386)
    def part_export(self, batch: Batch, qid0: int = 0, slot: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """The last search's partial result: (packs, slices [nspec, slot] u16);
        slot 0: one histogram's bins."""
        slot = slot or self.hbins
        packs = np.zeros(batch.nspec, N.PART_DTYPE)
        slices = np.zeros((batch.nspec, slot), np.uint16)
        N.check(N.hcx().hcx_part_export(self._h, batch._h, C.c_int32(qid0), C.c_uint32(slot), N.ptr(packs),
                                        N.ptr(slices)))
        return packs, slices

    def combine(self, packs: list, slices: list) -> tuple[np.ndarray, np.ndarray]:
        """CombineResults over shards in order: (fResult rows, keys)."""
        g = len(packs)
        n = len(packs[0])
        slot = slices[0].shape[1]
        packs = [np.ascontiguousarray(x, N.PART_DTYPE) for x in packs]
        slices = [np.ascontiguousarray(x, np.uint16) for x in slices]
        P = C.c_void_p * g
        out = np.zeros(n, N.FRES_DTYPE)
        keys = np.zeros(n, np.int32)
        N.check(N.hcx().hcx_combine(self._h, C.c_uint32(g), C.c_uint32(n), C.c_uint32(slot),
                                    P(*[x.ctypes.data for x in packs]), P(*[x.ctypes.data for x in slices]),
                                    N.ptr(out), N.ptr(keys)))
        return out, keys

    def close(self) -> None:
        for ix in list(self.indices):
            ix.free()
        self.indices = []
        if getattr(self, "_h", None):
            N.hcx().hcx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
