"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the reported CPU baseline.
The product package (hicops_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import time
from pathlib import Path

import numpy as np

from hicops_amd import native as N

LIB = Path(__file__).resolve().parent / "liboracle.so"
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle not built ({LIB}); run `make -C oracle`")
        L = C.CDLL(str(LIB))
        L.orc_pep_mass.restype = C.c_float
        L.orc_mod_mass.restype = C.c_float
        L.orc_gen_spectrum.restype = C.c_float
        L.orc_gen_mod_spectrum.restype = C.c_float
        L.orc_aa_mass.restype = C.c_float
        L.orc_stat_mod.restype = C.c_float
        L.orc_aa_mass.argtypes = [C.c_char]
        L.orc_stat_mod.argtypes = [C.c_char]
        L.orc_tailfit_merged.restype = C.c_double
        L.orc_tailfit_merged.argtypes = [C.c_void_p, C.c_uint32, C.c_int32, C.c_int32, C.c_void_p]
        for f in ("orc_part_export", "orc_combine", "orc_index_build", "orc_index_build_all", "orc_window", "orc_search", "orc_candidates", "orc_count_work",
                  "orc_tailfit", "orc_heap_replay", "orc_heap_capacity", "orc_pickpeaks"):
            getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


class OrcIndex(C.Structure):
    _fields_ = [("peplen", C.c_uint32), ("n", C.c_uint32), ("mass", C.c_void_p),
                ("iA", C.c_void_p), ("bA", C.c_void_p)]


def lgfact(n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    lib().orc_lgfact(N.ptr(out), C.c_uint32(n))
    return out


def pep_mass(seq: str) -> float:
    b = seq.encode()
    return lib().orc_pep_mass(b, C.c_uint32(len(b)))


def mod_mass(seq: str, mod_num: int, params: N.Params) -> float:
    b = seq.encode()
    return lib().orc_mod_mass(b, C.c_uint32(len(b)), C.c_uint32(mod_num), C.byref(params))


def gen_spectrum(seq: str, params: N.Params, sites: int = 0, mod_num: int = 0) -> tuple[float, np.ndarray]:
    b = seq.encode()
    out = np.zeros(2 * params.maxz * (len(b) - 1), np.uint32)
    if mod_num == 0:
        m = lib().orc_gen_spectrum(b, C.c_uint32(len(b)), C.byref(params), N.ptr(out))
    else:
        e = np.zeros(1, N.PEP_DTYPE)
        e["sites"], e["mod_num"] = sites, mod_num
        m = lib().orc_gen_mod_spectrum(b, C.c_uint32(len(b)), N.ptr(e), C.byref(params), N.ptr(out))
    return m, out


def index_build(params: N.Params, L: int, seqs: np.ndarray, entries: np.ndarray):
    n = len(entries)
    speclen = 2 * params.maxz * (L - 1)
    iA = np.zeros(n * speclen, np.uint32)
    bA = np.zeros(params.max_mass * params.scale + 1, np.uint32)
    seqs = np.ascontiguousarray(seqs, np.uint8)
    entries = np.ascontiguousarray(entries, N.PEP_DTYPE)
    rc = lib().orc_index_build(C.byref(params), C.c_uint32(L), N.ptr(seqs), N.ptr(entries),
                               C.c_uint32(n), N.ptr(iA), N.ptr(bA))
    if rc != 0:
        raise RuntimeError(f"orc_index_build failed: {rc}")
    return iA, bA


def window(mass: np.ndarray, prec: float, dM: float):
    lo, hi = C.c_int32(), C.c_int32()
    mass = np.ascontiguousarray(mass, np.float32)
    ok = lib().orc_window(N.ptr(mass), C.c_uint32(len(mass)), C.c_float(prec), C.c_double(dM),
                          C.byref(lo), C.byref(hi))
    return bool(ok), lo.value, hi.value


def index_build_all(params: N.Params, db, threads: int = 0, out=None):
    """orc_index_build for every length of db at once (OpenMP over the
    lengths).  Returns ({L: (iA, bA)}, seconds inside the C call); `out`, a
    previous result for the same db, is filled again instead of allocating."""
    Ls = list(db.lengths)
    k = len(Ls)
    seqs = [np.ascontiguousarray(db.seqs(L), np.uint8) for L in Ls]
    ents = [np.ascontiguousarray(db.entries(L), N.PEP_DTYPE) for L in Ls]
    if out is None:
        out = {L: (np.zeros(len(e) * 2 * params.maxz * (L - 1), np.uint32),
                   np.zeros(params.max_mass * params.scale + 1, np.uint32)) for L, e in zip(Ls, ents)}
    iA = [out[L][0] for L in Ls]
    bA = [out[L][1] for L in Ls]
    P = C.c_void_p * max(1, k)
    lens = np.asarray(Ls, np.uint32)
    counts = np.asarray([len(e) for e in ents], np.uint32)
    t0 = time.perf_counter()
    rc = lib().orc_index_build_all(C.byref(params), C.c_uint32(k), N.ptr(lens),
                                   P(*[x.ctypes.data for x in seqs]), P(*[x.ctypes.data for x in ents]),
                                   N.ptr(counts), P(*[x.ctypes.data for x in iA]), P(*[x.ctypes.data for x in bA]),
                                   C.c_int(threads))
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"orc_index_build_all failed: {rc}")
    return out, dt


class OracleDB:
    """The oracle's view of a database: per-length iA/bA built on the CPU."""

    def __init__(self, params: N.Params, db, lengths=None):
        self.params = params
        self.lengths = list(db.lengths if lengths is None else lengths)
        self.keep = []
        self.idx = (OrcIndex * max(1, len(self.lengths)))()
        self.iA, self.bA = {}, {}
        for i, L in enumerate(self.lengths):
            ents = db.entries(L)
            iA, bA = index_build(params, L, db.seqs(L), ents)
            mass = np.ascontiguousarray(ents["mass"])
            self.iA[L], self.bA[L] = iA, bA
            self.keep += [iA, bA, mass]
            self.idx[i] = OrcIndex(L, len(ents), mass.ctypes.data, iA.ctypes.data, bA.ctypes.data)

    @classmethod
    def from_built(cls, params: N.Params, db, built):
        """An OracleDB over indices already built ({L: (iA, bA)}, as
        index_build_all returns them)."""
        o = cls.__new__(cls)
        o.params, o.lengths, o.keep = params, list(db.lengths), []
        o.idx = (OrcIndex * max(1, len(o.lengths)))()
        o.iA, o.bA = {}, {}
        for i, L in enumerate(o.lengths):
            iA, bA = built[L]
            mass = np.ascontiguousarray(db.entries(L)["mass"])
            o.iA[L], o.bA[L] = iA, bA
            o.keep += [iA, bA, mass]
            o.idx[i] = OrcIndex(L, len(mass), mass.ctypes.data, iA.ctypes.data, bA.ctypes.data)
        return o

    def search(self, queries, threads: int = 0):
        n = queries.nspec
        res = np.zeros(n, N.RESULT_DTYPE)
        hist = np.zeros((n, self.params.hbins), np.uint32)
        stats = np.zeros(n, N.STATS_DTYPE)
        v = queries.view()
        rc = lib().orc_search(C.byref(self.params), self.idx, len(self.lengths), C.byref(v), N.ptr(res),
                              N.ptr(hist), N.ptr(stats), threads)
        if rc != 0:
            raise RuntimeError(f"orc_search failed: {rc}")
        return res, hist, stats

    def candidates(self, queries, s: int, cap: int = 1 << 20) -> np.ndarray:
        out = np.zeros(cap, N.CAND_DTYPE)
        cnt = C.c_uint64()
        v = queries.view()
        rc = lib().orc_candidates(C.byref(self.params), self.idx, len(self.lengths), C.byref(v), s, N.ptr(out),
                                  C.c_uint64(cap), C.byref(cnt))
        if rc != 0:
            raise RuntimeError(f"orc_candidates failed: {rc}")
        return out[:min(cnt.value, cap)]

    def count_work(self, queries, threads: int = 0) -> np.ndarray:
        st = np.zeros(queries.nspec, N.STATS_DTYPE)
        v = queries.view()
        rc = lib().orc_count_work(C.byref(self.params), self.idx, len(self.lengths), C.byref(v), N.ptr(st), threads)
        if rc != 0:
            raise RuntimeError(f"orc_count_work failed: {rc}")
        return st


def tailfit(hist: np.ndarray, cpsms: int, maxhyp: int):
    h = np.ascontiguousarray(hist, np.float64)
    mu, beta, mn, nx = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    fl = C.c_uint32(0)
    lib().orc_tailfit(N.ptr(h), C.c_uint32(len(h)), C.c_uint32(cpsms), C.c_int32(maxhyp), C.byref(mu), C.byref(beta),
                      C.byref(mn), C.byref(nx), C.byref(fl))
    return mu.value, beta.value, mn.value, nx.value, fl.value


def part_export(res: np.ndarray, hist: np.ndarray, qid0: int, slot: int):
    """One shard's partRes + u16 slices of searched spectra (res, hist)."""
    res = np.ascontiguousarray(res, N.RESULT_DTYPE)
    hist = np.ascontiguousarray(hist, np.uint32)
    n, hb = hist.shape
    packs = np.zeros(n, N.PART_DTYPE)
    slices = np.zeros((n, slot), np.uint16)
    rc = lib().orc_part_export(N.ptr(res), N.ptr(hist), C.c_uint32(hb), C.c_uint32(n), C.c_int32(qid0),
                               C.c_uint32(slot), N.ptr(packs), N.ptr(slices))
    if rc != 0:
        raise RuntimeError(f"orc_part_export failed: {rc}")
    return packs, slices


def tailfit_merged(pdata: np.ndarray, pN: int, hyp: int):
    pdata = np.ascontiguousarray(pdata, np.float64)
    fl = C.c_uint32(0)
    e = lib().orc_tailfit_merged(pdata.ctypes.data, len(pdata), pN, hyp, C.byref(fl))
    return e, fl.value


def combine(params: N.Params, packs: list, slices: list, slot: int):
    """CombineResults over shards (lists of per-shard arrays, shard order)."""
    g = len(packs)
    n = len(packs[0])
    packs = [np.ascontiguousarray(x, N.PART_DTYPE) for x in packs]
    slices = [np.ascontiguousarray(x, np.uint16) for x in slices]
    P = C.c_void_p * max(1, g)
    out = np.zeros(n, N.FRES_DTYPE)
    keys = np.zeros(n, np.int32)
    rc = lib().orc_combine(C.byref(params), C.c_uint32(g), C.c_uint32(n), C.c_uint32(slot),
                           P(*[x.ctypes.data for x in packs]), P(*[x.ctypes.data for x in slices]),
                           N.ptr(out), N.ptr(keys))
    if rc != 0:
        raise RuntimeError(f"orc_combine failed: {rc}")
    return out, keys


def heap_replay(hs, cap: int):
    a = np.ascontiguousarray(hs, np.float32)
    am, sz = C.c_int32(), C.c_int32()
    rc = lib().orc_heap_replay(N.ptr(a), C.c_uint32(len(a)), C.c_int32(cap), C.byref(am), C.byref(sz))
    if rc != 0:
        raise RuntimeError("heap replay failed")
    return am.value, sz.value


def heap_capacity(t: int) -> int:
    return lib().orc_heap_capacity(C.c_uint32(t))


def keyval_sort(key: np.ndarray, val: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """keyval.h:48-132 on copies of (key, val)."""
    k = np.ascontiguousarray(key, np.uint32).copy()
    v = np.ascontiguousarray(val, np.uint32).copy()
    lib().orc_keyval_sort(N.ptr(k), N.ptr(v), C.c_uint32(len(k)))
    return k, v


def pickpeaks(raw, qalen: int, base_int: int, min_int: int):
    """msquery.cpp:509-569 over a batch of raw spectra (workload.RawSpectra);
    returns a workload.Queries (no source fields)."""
    from hicops_amd import workload as W
    n = raw.nspec
    cap = int(np.minimum(np.diff(raw.idx.astype(np.int64)), qalen).sum()) if n else 0
    q = W.Queries(np.zeros(cap, np.uint32), np.zeros(cap, np.uint32), np.zeros(n + 1, np.uint32),
                  np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.float32),
                  np.zeros(n, np.uint32), np.zeros(n, np.int32))
    rt = N.ptr(raw.rtime) if raw.rtime is not None else None
    rc = lib().orc_pickpeaks(N.ptr(raw.mz), N.ptr(raw.intensity), N.ptr(raw.idx), N.ptr(raw.prec_mz),
                             N.ptr(raw.z), rt, C.c_int32(n), C.c_uint32(qalen), C.c_uint32(base_int),
                             C.c_uint32(min_int), N.ptr(q.moz), N.ptr(q.intensity), N.ptr(q.idx),
                             N.ptr(q.precurse), N.ptr(q.charges), N.ptr(q.rtimes))
    if rc != 0:
        raise RuntimeError(f"orc_pickpeaks failed: {rc}")
    k = int(q.idx[n]) if n else 0
    q.moz, q.intensity = q.moz[:k].copy(), q.intensity[:k].copy()
    return q


# ---- peptide database stage (LBE + MODS) ----

def lbe_count(seq: str, modconditions: str) -> int:
    """mods.cpp:290-315 count(): forms of one sequence, unmodified included."""
    b = seq.encode()
    out = C.c_uint64()
    rc = lib().orc_lbe_count(b, C.c_uint32(len(b)), modconditions.encode(), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_lbe_count failed: {rc}")
    return out.value


def lbe_filter(lines: np.ndarray, L: int, min_mass: int, max_mass: int) -> np.ndarray:
    """LBE_CountPeps' upper-casing and mass filter: the kept sequences."""
    lines = np.ascontiguousarray(lines, np.uint8)
    n = len(lines) // L
    out = np.zeros(n * L, np.uint8)
    k = C.c_uint32()
    rc = lib().orc_lbe_filter(N.ptr(lines), C.c_uint32(n), C.c_uint32(L), C.c_uint32(min_mass),
                              C.c_uint32(max_mass), N.ptr(out), C.byref(k))
    if rc != 0:
        raise RuntimeError(f"orc_lbe_filter failed: {rc}")
    return out[:k.value * L].copy()


def lbe_generate(params: N.Params, seqs: np.ndarray, L: int, modconditions: str | None, nodes: int = 1,
                 myid: int = 0, policy: int = N.POLICY_CYCLIC, sort: bool = True):
    """LBE_Initialize of one length: (pepEntries, (pepCount, modCount, lclpep, lclmod))."""
    seqs = np.ascontiguousarray(seqs, np.uint8)
    n = len(seqs) // L
    mc = (modconditions if modconditions is not None else "0 M 0").encode()
    cnt = np.zeros(4, np.uint32)
    args = (C.byref(params), N.ptr(seqs), C.c_uint32(n), C.c_uint32(L), mc, C.c_uint32(nodes), C.c_uint32(myid),
            C.c_int(policy), C.c_int(int(bool(sort))))
    rc = lib().orc_lbe_generate(*args, None, N.ptr(cnt))
    if rc != 0:
        raise RuntimeError(f"orc_lbe_generate failed: {rc}")
    out = np.zeros(int(cnt[2]) + int(cnt[3]), N.PEP_DTYPE)
    rc = lib().orc_lbe_generate(*args, N.ptr(out), N.ptr(cnt))
    if rc != 0:
        raise RuntimeError(f"orc_lbe_generate failed: {rc}")
    return out, tuple(int(x) for x in cnt)


def std_sort_mass(keys: np.ndarray, vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """std::sort of (key, value) records by key, as LBE_Initialize sorts pepEntries."""
    k = np.ascontiguousarray(keys, np.float32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_std_sort_mass(N.ptr(k), N.ptr(v), C.c_uint32(len(k)))
    return k, v


def qsort_sites(sites: np.ndarray, vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """glibc qsort with cmpvarEntries on (sites, value) records."""
    s = np.ascontiguousarray(sites, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().orc_qsort_sites(N.ptr(s), N.ptr(v), C.c_uint32(len(s)))
    return s, v


def antiqsort(n: int) -> np.ndarray:
    """Keys on which std::sort falls back to its heap sort (McIlroy's adversary)."""
    k = np.zeros(n, np.float32)
    lib().orc_antiqsort(C.c_uint32(n), N.ptr(k))
    return k
