"""Writes tests/golden/mini_case.npz: a small search case frozen as data.

Inputs: a mini database (lengths 7-12, fragment charge 1-2, Met-ox variable
mods) and 24 spectra, as the reference would hold them (per-length sequences
and mass-sorted pepEntry records; CSR peaks).  Expected outputs:
  * per spectrum, every candidate with bc+yc >= min_shp and hs > 0, as
    (length index, peptide row, bc, yc, ibc, iyc, hs) -- by brute force:
    each candidate's fragment bins compared with every peak (no index);
  * per length, FNV-1a 64 of the ion array iA and the bucket array bA of the
    oracle's index build (dslim.cpp:46-482), so a platform whose libm or
    sort order differs is caught;
  * per spectrum, the top hit the reference heap returns (minheap.h:86-277),
    cpsms and the e-value of the tail fit.

Run from the repository root: python tests/golden/make_golden.py
"""
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from hicops_amd import workload as W  # noqa: E402
from oracle import binding as O  # noqa: E402


def fnv1a64(a: np.ndarray) -> int:
    h = 1469598103934665603
    for b in np.ascontiguousarray(a).view(np.uint8).tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def candidates(prm, lengths, seqs, ents, moz, inten, prec):
    NB = prm.max_mass * prm.scale
    lgf = [math.lgamma(n + 1) / math.log(10) for n in range(prm.max_shdpeaks + 1)]
    moz = moz.astype(np.int64)
    inten = inten.astype(np.int64)
    keep = (moz > prm.dF) & (moz < NB - 1 - prm.dF)
    moz, inten = moz[keep], inten[keep]
    p1 = np.float32(float(prec) - prm.dM)
    p2 = np.float32(float(prec) + prm.dM)
    out = []
    for ixx, L in enumerate(lengths):
        e_all = ents[ixx]
        for psid in np.nonzero((e_all["mass"] >= p1) & (e_all["mass"] <= p2))[0]:
            e = e_all[psid]
            sid = int(e["seq_id"])
            seq = bytes(seqs[ixx][sid * L:(sid + 1) * L]).decode()
            m, bins = O.gen_spectrum(seq, prm, int(e["sites"]), int(e["mod_num"]))
            if not (prm.min_mass <= m <= prm.max_mass):
                bins = np.zeros_like(bins)
            bins = np.minimum(bins.astype(np.int64), NB - 1)
            half = len(bins) // 2
            hit = np.abs(bins[:, None] - moz[None, :]) <= prm.dF
            bc, yc = int(hit[:half].sum()), int(hit[half:].sum())
            ibc = int((hit[:half] * inten[None, :]).sum())
            iyc = int((hit[half:] * inten[None, :]).sum())
            if bc + yc < prm.min_shp:
                continue
            hs = lgf[bc] + lgf[yc] + math.log10(1 + ibc) + math.log10(1 + iyc) - 6
            if hs > 0:
                out.append((ixx, int(psid), bc, yc, ibc, iyc, hs))
    return out


def main() -> None:
    p = W.scaled("small", npep=400, min_len=7, max_len=12, maxz=2, nspec=24, dM=150.0,
                 mods=[("M", 15.9949, 2)], max_mods=2, seed=0x60D)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    lengths = list(db.lengths)
    seqs = [np.ascontiguousarray(db.seqs(L)) for L in lengths]
    ents = [np.ascontiguousarray(db.entries(L)) for L in lengths]
    arrays = {"lengths": np.array(lengths, np.int32), "moz": q.moz, "intensity": q.intensity, "idx": q.idx,
              "precurse": q.precurse, "charges": q.charges,
              "params": np.array([prm.maxz, prm.scale, prm.min_mass, prm.max_mass, prm.dF, prm.min_shp,
                                  prm.min_cpsm, prm.topmatches, prm.max_hyperscore, prm.max_shdpeaks,
                                  prm.nmods, prm.mod_mass[0]], np.int64),
              "dM": np.array([prm.dM])}
    for i, L in enumerate(lengths):
        arrays[f"seqs_{L}"] = seqs[i]
        arrays[f"ents_{L}"] = ents[i]
        iA, bA = O.index_build(prm, L, seqs[i], ents[i])
        arrays[f"hash_{L}"] = np.array([fnv1a64(iA), fnv1a64(bA), len(iA)], np.uint64)
    rows = []
    for s in range(q.nspec):
        mz = q.moz[q.idx[s]:q.idx[s + 1]]
        it = q.intensity[q.idx[s]:q.idx[s + 1]]
        for c in candidates(prm, lengths, seqs, ents, mz, it, q.precurse[s]):
            rows.append((s,) + c)
    cand = np.array([r[:7] for r in rows], np.int64).reshape(-1, 7)
    arrays["cand"] = cand
    arrays["cand_hs"] = np.array([r[7] for r in rows], np.float64)
    res, _, _ = O.OracleDB(prm, db).search(q)
    arrays["top"] = np.stack([res["cpsms"].astype(np.int64), res["psid"].astype(np.int64),
                              res["idxoffset"].astype(np.int64), res["flags"].astype(np.int64)], 1)
    arrays["top_hs"] = res["hyperscore"].astype(np.float64)
    arrays["top_e"] = res["e_x"].astype(np.float64)
    out = Path(__file__).resolve().parent / "mini_case.npz"
    np.savez_compressed(out, **arrays)
    print(f"wrote {out}: {len(lengths)} lengths, {q.nspec} spectra, {len(cand)} candidates")


if __name__ == "__main__":
    main()
