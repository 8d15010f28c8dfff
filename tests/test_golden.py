"""The frozen mini case (tests/golden/mini_case.npz, written by
tests/golden/make_golden.py): brute-force candidate lists, index hashes and
top hits, checked against the oracle here and against the GPU path (-m gpu)."""
from pathlib import Path

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from oracle import binding as O

GOLD = Path(__file__).resolve().parent / "golden" / "mini_case.npz"


class FixtureDB:
    def __init__(self, z):
        self.lengths = [int(L) for L in z["lengths"]]
        self._s = {L: z[f"seqs_{L}"] for L in self.lengths}
        self._e = {L: z[f"ents_{L}"] for L in self.lengths}

    def seqs(self, L):
        return self._s[L]

    def entries(self, L):
        return self._e[L]


def load():
    z = np.load(GOLD)
    db = FixtureDB(z)
    q = W.Queries(z["moz"], z["intensity"], z["idx"], z["precurse"], z["charges"],
                  np.zeros(len(z["precurse"]), np.float32), np.zeros(len(z["precurse"]), np.int32),
                  np.zeros(len(z["precurse"]), np.int32))
    v = [int(x) for x in z["params"]]
    prm = N.Params()
    (prm.maxz, prm.scale, prm.min_mass, prm.max_mass, prm.dF, prm.min_shp, prm.min_cpsm, prm.topmatches,
     prm.max_hyperscore, prm.max_shdpeaks, prm.nmods) = v[:11]
    prm.mod_mass[0] = v[11]
    prm.dM = float(z["dM"][0])
    prm.expect_max = 1e300
    return z, db, q, prm


def fnv1a64(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a).view(np.uint8).tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def cand_rows(cands, s):
    return [(s, int(c["idxoffset"]), int(c["psid"]), int(c["bc"]), int(c["yc"]), int(c["ibc"]), int(c["iyc"]))
            for c in cands]


def test_oracle_reproduces_golden_case():
    z, db, q, prm = load()
    for L in db.lengths:
        iA, bA = O.index_build(prm, L, db.seqs(L), db.entries(L))
        assert [fnv1a64(iA), fnv1a64(bA), len(iA)] == [int(x) for x in z[f"hash_{L}"]], L
    odb = O.OracleDB(prm, db)
    got = []
    for s in range(q.nspec):
        got += cand_rows(odb.candidates(q, s), s)
    assert got == [tuple(int(x) for x in r) for r in z["cand"]]
    res, _, _ = odb.search(q)
    top = np.stack([res["cpsms"], res["psid"], res["idxoffset"], res["flags"]], 1).astype(np.int64)
    assert np.array_equal(top, z["top"])
    assert np.array_equal(res["hyperscore"].astype(np.float64), z["top_hs"])


@pytest.mark.gpu
def test_gpu_reproduces_golden_case():
    from hicops_amd.search import Engine
    z, db, q, prm = load()
    eng = Engine(prm)
    idx = eng.construct_db(db)
    for L, ix in zip(db.lengths, idx):
        iA, bA = ix.export()
        assert [fnv1a64(iA), fnv1a64(bA), len(iA)] == [int(x) for x in z[f"hash_{L}"]], L
    b = eng.upload(q)
    g = eng.candidates(b)
    g = g[np.lexsort((g["psid"], g["idxoffset"], g["spec"]))]
    got = [(int(c["spec"]), int(c["idxoffset"]), int(c["psid"]), int(c["bc"]), int(c["yc"]), int(c["ibc"]),
            int(c["iyc"])) for c in g]
    assert got == [tuple(int(x) for x in r) for r in z["cand"]]
    assert np.all(np.abs(g["hyperscore"] - z["cand_hs"]) <= 1e-5 * np.abs(z["cand_hs"]))
    for kernel in (N.KERNEL_PLAIN, N.KERNEL_BATCHED, N.KERNEL_CURSOR):
        out = eng.query_spectrum(b, kernel=kernel)
        r = out.results
        valid = (z["top"][:, 3] & N.RES_VALID) != 0
        assert np.array_equal(r["cpsms"].astype(np.int64), z["top"][:, 0])
        assert np.array_equal(r["psid"][valid].astype(np.int64), z["top"][valid, 1])
        assert np.array_equal(r["idxoffset"][valid].astype(np.int64), z["top"][valid, 2])
        assert np.all(np.abs(r["hyperscore"][valid] - z["top_hs"][valid]) <= 1e-5 * np.abs(z["top_hs"][valid]))
        assert np.all(np.abs(r["e_x"][valid] - z["top_e"][valid]) <= 1e-3 * np.abs(z["top_e"][valid]))
    eng.close()
