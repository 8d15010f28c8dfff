"""Synthetic databases and query batches (seeded, reproducible).

Each preset is one of the benchmark configurations, written out with its sizes:

  small          20,000 tryptic peptides (7-30), charge 1-2 fragments, 1,000 spectra,
                 +-10 Da, 0.02 Da fragments, top 100 peaks
  rank_share     2M peptides (7-40) + M-oxidation (<= 2 mods), charge-1 fragments,
                 50,000 spectra (charge 2-4, 50% retention, top 150), +-500 Da
  open10m        10M peptides (7-40), charge 1-2 fragments, 200,000 spectra with
                 20-500 peaks (log-uniform) and windows of 1e2-1e6 candidates, +-500 Da
  large40m       12.7M peptides (7-40) + Met-oxidation (<= 3 mods): ~40M forms, ~4.2G
                 ions; 1,000,000 spectra as open10m, 0.01 Da fragments, +-500 Da

``scaled(name, npep=..., nspec=...)`` keeps a preset's shape at other sizes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import native as N

AA = "ACDEFGHIKLMNPQRSTVWY"


@dataclass
class Preset:
    name: str
    npep: int
    min_len: int
    max_len: int
    maxz: int
    nspec: int
    dM: float
    dF: int
    min_charge: int
    max_charge: int
    keep_prob: float
    n_noise: int
    peaks_min: int
    peaks_max: int
    window_mode: int = 0
    wmin: float = 1e2
    wmax: float = 1e6
    mods: list = field(default_factory=list)   # [(residues, Da, max per peptide)]
    max_mods: int = 0
    max_hyperscore: int = 100
    max_shdpeaks: int = 80
    seed: int = 0x48C095
    base_int: int = 100000
    min_int: int = 1000


PRESETS = {
    "small": Preset("small", 20000, 7, 30, 2, 1000, 10.0, 2, 2, 3, 0.6, 60, 100, 100),
    "rank_share": Preset("rank_share", 2_000_000, 7, 40, 1, 50000, 500.0, 2, 2, 4, 0.5, 100, 150, 150,
                         mods=[("M", 15.9949, 2)], max_mods=2, seed=0x48C095 ^ 1,
                         max_hyperscore=200, max_shdpeaks=160),
    "open10m": Preset("open10m", 10_000_000, 7, 40, 2, 200000, 500.0, 2, 2, 4, 0.5, 100, 20, 500,
                      window_mode=1, wmin=1e2, wmax=1e6, seed=0x48C095 ^ 2,
                      max_hyperscore=200, max_shdpeaks=160),
    # 12.7M base peptides with Met-oxidation (<= 3 per peptide): ~40M forms,
    # ~4.2G fragment ions; 1M spectra at 0.01 Da fragment bins (dF = 1)
    "large40m": Preset("large40m", 12_700_000, 7, 40, 2, 1_000_000, 500.0, 1, 2, 4, 0.5, 100, 20, 500,
                       window_mode=1, wmin=1e2, wmax=4e6, mods=[("M", 15.9949, 3)], max_mods=3,
                       seed=0x48C095 ^ 4, max_hyperscore=200, max_shdpeaks=160),
}


def scaled(name: str, **kw) -> Preset:
    return replace(PRESETS[name], **kw)


def make_params(p: Preset, topmatches: int = 10, expect_max: float = 1e300) -> N.Params:
    prm = N.Params()
    prm.maxz = p.maxz
    prm.scale = 100
    prm.min_mass = 500
    prm.max_mass = 5000
    prm.dF = p.dF
    prm.min_shp = 4
    prm.min_cpsm = 4
    prm.topmatches = topmatches
    prm.max_hyperscore = p.max_hyperscore
    prm.max_shdpeaks = p.max_shdpeaks
    prm.nmods = len(p.mods)
    for i, (_, da, _) in enumerate(p.mods):
        prm.mod_mass[i] = int(da * 100)          # (uint)(Da*scale), hicops.cpp:661
    prm.dM = p.dM
    prm.expect_max = expect_max
    return prm


def modconditions(p: Preset) -> str:
    """gParams.modconditions for the preset's variable mods, as hicops.cpp:636-666
    builds it: the per-form limit, then residues and per-form cap per mod."""
    if not p.mods or p.max_mods == 0:
        return "0 M 0"
    return " ".join([str(p.max_mods)] + [f"{res} {mx}" for res, _, mx in p.mods])


class Database:
    """Per-length sequence tables and mass-sorted pepEntries."""

    def __init__(self, p: Preset):
        cfg = N.DbGenCfg()
        cfg.seed = p.seed
        cfg.npep = p.npep
        cfg.min_len = p.min_len
        cfg.max_len = p.max_len
        cfg.min_mass = 500
        cfg.max_mass = 5000
        cfg.scale = 100
        cfg.nmods = len(p.mods)
        cfg.max_mods = p.max_mods
        for i, (res, da, mx) in enumerate(p.mods):
            cfg.mod_residues[i].value = res.encode()[:4]
            cfg.mod_max[i] = mx
            cfg.mod_mass[i] = int(da * 100)
        h = C.c_void_p()
        lib = N.host()
        N.check(lib.hcx_db_create(C.byref(cfg), C.byref(h)), lib)
        self._h = h
        self.preset = p
        n = C.c_uint32()
        N.check(lib.hcx_db_lengths(h, None, C.byref(n)), lib)
        lens = np.zeros(n.value, np.uint32)
        N.check(lib.hcx_db_lengths(h, N.ptr(lens), C.byref(n)), lib)
        self.lengths = [int(x) for x in lens]
        self.tables = {}
        for L in self.lengths:
            ns, ne = C.c_uint32(), C.c_uint32()
            N.check(lib.hcx_db_length_info(h, L, C.byref(ns), C.byref(ne)), lib)
            seqs = np.zeros(ns.value * L, np.uint8)
            ents = np.zeros(ne.value, N.PEP_DTYPE)
            N.check(lib.hcx_db_length_copy(h, L, N.ptr(seqs), N.ptr(ents)), lib)
            self.tables[L] = (seqs, ents)

    def seqs(self, L: int) -> np.ndarray:
        return self.tables[L][0]

    def entries(self, L: int) -> np.ndarray:
        return self.tables[L][1]

    def sequence(self, L: int, psid: int) -> str:
        seqs, ents = self.tables[L]
        sid = int(ents[psid]["seq_id"])
        return bytes(seqs[sid * L:(sid + 1) * L]).decode()

    @property
    def nforms(self) -> int:
        return sum(len(e) for _, e in self.tables.values())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.host().hcx_db_free(h)
            self._h = None


@dataclass
class Queries:
    moz: np.ndarray
    intensity: np.ndarray
    idx: np.ndarray
    precurse: np.ndarray
    charges: np.ndarray
    rtimes: np.ndarray
    src_len: np.ndarray
    src_psid: np.ndarray

    @property
    def nspec(self) -> int:
        return len(self.precurse)

    def view(self) -> N.QueriesView:
        v = N.QueriesView()
        v.moz, v.intensity, v.idx = N.ptr(self.moz), N.ptr(self.intensity), N.ptr(self.idx)
        v.precurse, v.charges, v.rtimes = N.ptr(self.precurse), N.ptr(self.charges), N.ptr(self.rtimes)
        v.num_specs = self.nspec
        v.file_num = 0
        return v

    def subset(self, sel) -> "Queries":
        sel = np.asarray(sel)
        lens = self.idx[sel + 1] - self.idx[sel]
        idx = np.zeros(len(sel) + 1, np.uint32)
        idx[1:] = np.cumsum(lens)
        parts = [np.arange(self.idx[s], self.idx[s + 1]) for s in sel]
        take = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        return Queries(np.ascontiguousarray(self.moz[take]), np.ascontiguousarray(self.intensity[take]),
                       idx, self.precurse[sel].copy(), self.charges[sel].copy(), self.rtimes[sel].copy(),
                       self.src_len[sel].copy(), self.src_psid[sel].copy())


def _spec_cfg(p: Preset, nspec: int | None, seed: int | None) -> N.SpecGenCfg:
    cfg = N.SpecGenCfg()
    cfg.seed = p.seed ^ 0x5EC7 if seed is None else seed
    cfg.nspec = p.nspec if nspec is None else nspec
    cfg.min_charge, cfg.max_charge = p.min_charge, p.max_charge
    cfg.frag_maxz = p.maxz
    cfg.n_noise = p.n_noise
    cfg.peaks_min, cfg.peaks_max = p.peaks_min, p.peaks_max
    cfg.scale = 100
    cfg.base_int, cfg.min_int = p.base_int, p.min_int
    cfg.window_mode = p.window_mode
    cfg.keep_prob = p.keep_prob
    cfg.noise_lo, cfg.noise_hi = 100.0, 2000.0
    cfg.jitter_sd = 0.01
    cfg.sig_mu, cfg.sig_sigma = 1.0, 1.0
    cfg.noise_mu, cfg.noise_sigma = 0.0, 1.0
    cfg.dM = p.dM
    cfg.wmin, cfg.wmax = p.wmin, p.wmax
    return cfg


def _copy_queries(lib, h) -> Queries:
    ns, npk = C.c_uint32(), C.c_uint64()
    N.check(lib.hcx_spec_info(h, C.byref(ns), C.byref(npk)), lib)
    n, k = ns.value, npk.value
    q = Queries(np.zeros(k, np.uint32), np.zeros(k, np.uint32), np.zeros(n + 1, np.uint32),
                np.zeros(n, np.float32), np.zeros(n, np.int32), np.zeros(n, np.float32),
                np.zeros(n, np.uint32), np.zeros(n, np.int32))
    N.check(lib.hcx_spec_copy(h, N.ptr(q.moz), N.ptr(q.intensity), N.ptr(q.idx), N.ptr(q.precurse),
                              N.ptr(q.charges), N.ptr(q.rtimes), N.ptr(q.src_len), N.ptr(q.src_psid)), lib)
    return q


def make_queries(db: Database, p: Preset, nspec: int | None = None, seed: int | None = None) -> Queries:
    cfg = _spec_cfg(p, nspec, seed)
    lib = N.host()
    h = C.c_void_p()
    N.check(lib.hcx_spec_create(db._h, C.byref(cfg), C.byref(h)), lib)
    try:
        return _copy_queries(lib, h)
    finally:
        lib.hcx_spec_free(h)


@dataclass
class RawSpectra:
    """Raw MS2 peaks as MSQuery::readspectrum stores them (msquery.cpp:300-412):
    mz = (uint)(m/z * scale), intensity = (uint)(I * 1000), in file (m/z) order,
    with each spectrum's Z-line precursor and charge and its retention time."""
    mz: np.ndarray
    intensity: np.ndarray
    idx: np.ndarray
    prec_mz: np.ndarray
    z: np.ndarray
    rtime: np.ndarray

    @property
    def nspec(self) -> int:
        return len(self.prec_mz)

    def view(self) -> N.RawSpectraView:
        v = N.RawSpectraView()
        v.mz, v.intensity, v.idx = N.ptr(self.mz), N.ptr(self.intensity), N.ptr(self.idx)
        v.prec_mz, v.z, v.rtime = N.ptr(self.prec_mz), N.ptr(self.z), N.ptr(self.rtime)
        v.num_specs = self.nspec
        v.file_num = 0
        return v


def make_raw_spectra(db: Database, p: Preset, nspec: int | None = None,
                     seed: int | None = None) -> tuple[RawSpectra, Queries]:
    """The spectra of make_queries (same seed, same spectra) before pickpeaks,
    and the generator's own normalised batch (stable tie order)."""
    cfg = _spec_cfg(p, nspec, seed)
    lib = N.host()
    h = C.c_void_p()
    N.check(lib.hcx_spec_create_raw(db._h, C.byref(cfg), C.byref(h)), lib)
    try:
        q = _copy_queries(lib, h)
        npk = C.c_uint64()
        N.check(lib.hcx_spec_raw_info(h, C.byref(npk)), lib)
        n, k = q.nspec, npk.value
        r = RawSpectra(np.zeros(k, np.uint32), np.zeros(k, np.uint32), np.zeros(n + 1, np.uint32),
                       np.zeros(n, np.float64), np.zeros(n, np.int32), np.zeros(n, np.float64))
        N.check(lib.hcx_spec_raw_copy(h, N.ptr(r.mz), N.ptr(r.intensity), N.ptr(r.idx), N.ptr(r.prec_mz),
                                      N.ptr(r.z), N.ptr(r.rtime)), lib)
    finally:
        lib.hcx_spec_free(h)
    return r, q


def write_driver_case(d, db: Database, q: Queries, prm: N.Params, kernel: int, **extra) -> None:
    """The input directory of the C++ driver of the reference-shaped entry
    points (tests/cpp/dslim_driver.cpp): params.txt (gParams fields, limits,
    kernel and any extra keys), lengths.txt, each length's sequences and
    pepEntries, and the CSR spectra."""
    from pathlib import Path
    d = Path(d)
    kv = {"maxz": prm.maxz, "scale": prm.scale, "min_mass": prm.min_mass, "max_mass": prm.max_mass,
          "dF": prm.dF, "min_shp": prm.min_shp, "min_cpsm": prm.min_cpsm, "topmatches": prm.topmatches,
          "dM": prm.dM, "expect_max": 1e300, "nmods": prm.nmods, "max_hyperscore": prm.max_hyperscore,
          "max_shdpeaks": prm.max_shdpeaks, "kernel": kernel}
    for i in range(prm.nmods):
        kv[f"mod{i}"] = prm.mod_mass[i]
    kv.update(extra)
    (d / "params.txt").write_text("".join(f"{k} {v!r}\n" for k, v in kv.items()))
    (d / "lengths.txt").write_text("".join(f"{L}\n" for L in db.lengths))
    for L in db.lengths:
        np.ascontiguousarray(db.seqs(L)).tofile(d / f"seqs_{L}.bin")
        np.ascontiguousarray(db.entries(L)).tofile(d / f"ents_{L}.bin")
    q.moz.astype(np.uint32).tofile(d / "moz.bin")
    q.intensity.astype(np.uint32).tofile(d / "inten.bin")
    q.idx.astype(np.uint32).tofile(d / "idx.bin")
    q.precurse.astype(np.float32).tofile(d / "prec.bin")
    q.charges.astype(np.int32).tofile(d / "chg.bin")
