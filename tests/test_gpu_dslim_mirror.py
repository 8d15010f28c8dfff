"""The reference-shaped C++ entry points (include/hcx_dslim.hpp) driven the
way the reference application drives its own (tests/cpp/dslim_driver.cpp):
DSLIM_Construct per length, DSLIM_InitializeScorecard, DSLIM_QuerySpectrum
per batch, every reported PSM through DFile_PrintScore, ion-index export,
expeRT::ModelTailFit.  Checked against the oracle on the same inputs."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from oracle import binding as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "cpp" / "dslim_driver"


def write_case(d: Path, p, db, q, prm, kernel: int) -> None:
    W.write_driver_case(d, db, q, prm, kernel)


@pytest.mark.parametrize("kernel", ["batched", "plain"])
def test_dslim_entry_points_match_oracle(tmp_path, kernel):
    assert DRIVER.exists(), "tests/cpp/dslim_driver is built by __graft_entry__.build()"
    p = W.scaled("small", npep=4000, nspec=300, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    kc = N.KERNEL_BATCHED if kernel == "batched" else N.KERNEL_PLAIN
    write_case(tmp_path, p, db, q, prm, kc)
    print(f"test_dslim_entry_points_match_oracle kernel={kernel}", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path), "3"], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0, r.stderr

    res_o, _, _ = O.OracleDB(prm, db).search(q)
    reported = np.nonzero((res_o["flags"] & N.RES_VALID) != 0)[0]
    rows = [ln.split() for ln in (tmp_path / "psms.txt").read_text().splitlines()]
    assert len(rows) == len(reported) and len(rows) > 100
    got = {int(x[0]): x for x in rows}
    assert sorted(got) == [int(s) for s in reported]      # running spectrum ids across 3 batches
    for s in reported:
        x = got[int(s)]
        o = res_o[s]
        assert int(x[1]) == o["psid"] and int(x[2]) == o["idxoffset"], s
        assert int(x[3]) == o["sharedions"] and int(x[4]) == o["totalions"], s
        hs = np.frombuffer(bytes.fromhex(x[5])[::-1], np.float32)[0]
        assert abs(hs - o["hyperscore"]) <= 1e-5 * abs(o["hyperscore"]), s
        assert abs(float(x[6]) - o["e_x"]) <= 1e-3 * abs(o["e_x"]), s
        assert int(x[7]) == o["cpsms"], s
        assert int(x[8]) == int(q.charges[s]) and int(x[9]) == 3, s

    # DSLIM_ExportIonIndex: the host copy of the first length's iA / bA
    L0 = db.lengths[0]
    oiA, obA = O.index_build(prm, L0, db.seqs(L0), db.entries(L0))
    raw = np.fromfile(tmp_path / "ion0.bin", np.uint32)
    assert np.array_equal(raw[:len(oiA)], oiA) and np.array_equal(raw[len(oiA):], obA)

    # expeRT::ModelTailFit(Results*) on a hand-traced histogram
    H = np.zeros(10 * prm.max_hyperscore + 2)
    H[2:11] = [5, 10, 8, 6, 3, 1, 0, 1, 1]
    mu, beta, mn, nx, _ = O.tailfit(H, 35, 10)
    assert [int(v) for v in (tmp_path / "tailfit.txt").read_text().split()] == [mu, beta, mn, nx]


TSV_DRIVER = ROOT / "tests" / "cpp" / "tsv_driver"
TSV_HEADER = ("file\tscan_num\tprec_mass\tcharge\tretention_time\tpeptide\tmatched_ions\ttotal_ions\t"
              "calc_pep_mass\tmass_diff\tmod_info\thyperscore\texpectscore\tnum_hits")


def test_tsv_writer_matches_reference_format(tmp_path):
    # DFile_InitFiles / DFile_PrintScore / DFile_DeinitFiles (dslim_fileout.cpp:45-152)
    assert TSV_DRIVER.exists(), "tests/cpp/tsv_driver is built by __graft_entry__.build()"
    p = W.scaled("small", npep=4000, nspec=300, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    q.rtimes[:] = np.linspace(0.0, 4321.123, q.nspec, dtype=np.float32)
    prm = W.make_params(p)
    write_case(tmp_path, p, db, q, prm, N.KERNEL_BATCHED)
    q.rtimes.astype(np.float32).tofile(tmp_path / "rt.bin")
    print("test_tsv_writer_matches_reference_format", flush=True)
    r = subprocess.run([str(TSV_DRIVER), str(tmp_path), "2"], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0, r.stderr
    files = sorted(tmp_path.glob("*.tsv"))
    assert len(files) == 2                                           # params.threads files
    assert [f.name[-6:] for f in files] == ["_0.tsv", "_1.tsv"]
    lines = []
    for f in files:
        t = f.read_text().split("\n")
        assert t[0] == TSV_HEADER and t[-1] == ""
        lines += t[1:-1]

    res_o, _, _ = O.OracleDB(prm, db).search(q)
    reported = np.nonzero((res_o["flags"] & N.RES_VALID) != 0)[0]
    assert len(lines) == len(reported) and len(lines) > 100
    for ln, s in zip(lines, reported):
        f = ln.split("\t")
        assert len(f) == 14, ln
        o = res_o[s]
        L = db.lengths[o["idxoffset"]]
        mass = np.float32(db.entries(L)["mass"][o["psid"]])
        pm = np.float32(q.precurse[s])
        # std::to_string: "%f" for float and double, "%d" / "%u" for integers
        exact = ["sample_3.ms2", str(s + 1), f"{float(pm):f}", str(int(q.charges[s])), f"{float(q.rtimes[s]):f}",
                 db.sequence(L, int(o["psid"])), str(int(o["sharedions"])), str(int(o["totalions"])),
                 f"{float(mass):f}", f"{float(np.float32(pm - mass)):f}", ""]
        assert f[:11] == exact, (s, f, exact)
        assert abs(float(f[11]) - o["hyperscore"]) <= 1e-5 * abs(o["hyperscore"]) + 5e-7, s
        assert abs(float(f[12]) - o["e_x"]) <= 1e-3 * abs(o["e_x"]) + 5e-7, s
        assert f[13] == str(int(o["cpsms"])), s

    # the reference's post-processing (tools/runtime/psm2tsv.py): every partial
    # TSV of the directory read with pandas (tab-separated, header row), the
    # frames concatenated in the directory's listing order, written as one
    # results.tsv with a fresh index column.  The partial files must go through
    # it unchanged: one row per reported PSM, the 14 columns by header name.
    pd = pytest.importorskip("pandas")
    frames = [pd.read_csv(f, sep="\t", index_col=None, header=0) for f in tmp_path.glob("*.tsv")]
    frame = pd.concat(frames, axis=0, ignore_index=True)
    assert list(frame.columns) == TSV_HEADER.split("\t")
    assert len(frame) == len(reported)
    frame.to_csv(tmp_path / "results.tsv", sep="\t")
    merged = pd.read_csv(tmp_path / "results.tsv", sep="\t", index_col=0, header=0)
    got = merged.sort_values("scan_num").reset_index(drop=True)
    assert list(got["scan_num"]) == [int(s) + 1 for s in reported]
    assert list(got["matched_ions"]) == [int(res_o[s]["sharedions"]) for s in reported]
    assert list(got["num_hits"]) == [int(res_o[s]["cpsms"]) for s in reported]
    assert list(got["peptide"]) == [db.sequence(db.lengths[res_o[s]["idxoffset"]], int(res_o[s]["psid"]))
                                    for s in reported]


class _LbeDb:
    """The oracle's database (sequences and pepEntries per length) in the
    shape OracleDB reads."""

    def __init__(self, lengths, tables):
        self.lengths = lengths
        self.tables = tables

    def seqs(self, L):
        return self.tables[L][0]

    def entries(self, L):
        return self.tables[L][1]


def test_lbe_entry_points_match_oracle(tmp_path):
    # MODS_Initialize, LBE_CountPeps (FASTA per length), LBE_CreatePartitions,
    # LBE_Initialize, LBE_Distribute, then DSLIM_Construct from the entries the
    # stage left on the device, DSLIM_QuerySpectrum, LBE_Deinitialize
    p = W.scaled("rank_share", npep=3000, nspec=200)
    db = W.Database(p)
    q = W.make_queries(db, p)
    prm = W.make_params(p)
    mc = W.modconditions(p)
    write_case(tmp_path, p, db, q, prm, N.KERNEL_BATCHED)
    with open(tmp_path / "params.txt", "a") as f:
        f.write("lbe 1\n")
    (tmp_path / "mods.txt").write_text(mc + "\n")
    rng = np.random.default_rng(3)
    for L in db.lengths:
        s = db.seqs(L).reshape(-1, L)
        with open(tmp_path / f"fasta_{L}.txt", "wb") as f:
            for i, row in enumerate(s):
                line = bytes(row)
                if rng.random() < 0.2:
                    line = line.lower()
                f.write(b">p%d\n" % i + line + (b"\r\n" if i % 3 == 0 else b"\n"))
    print("test_lbe_entry_points_match_oracle", flush=True)
    r = subprocess.run([str(DRIVER), str(tmp_path), "2"], capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0, r.stderr

    tables = {}
    for L in db.lengths:
        so = O.lbe_filter(db.seqs(L), L, prm.min_mass, prm.max_mass)
        eo, _ = O.lbe_generate(prm, so, L, mc)
        eg = np.fromfile(tmp_path / f"lbe_ents_{L}.bin", N.PEP_DTYPE)
        sg = np.fromfile(tmp_path / f"lbe_seqs_{L}.bin", np.uint8)
        assert np.array_equal(sg, so), L
        for f in ("seq_id", "sites", "mod_num"):
            assert np.array_equal(eg[f], eo[f]), (L, f)
        assert np.array_equal(eg["mass"].view(np.uint32), eo["mass"].view(np.uint32)), L
        tables[L] = (so, eo)
    res_o, _, _ = O.OracleDB(prm, _LbeDb(db.lengths, tables)).search(q)
    reported = np.nonzero((res_o["flags"] & N.RES_VALID) != 0)[0]
    rows = [ln.split() for ln in (tmp_path / "psms.txt").read_text().splitlines()]
    assert len(rows) == len(reported) and len(rows) > 50
    got = {int(x[0]): x for x in rows}
    for s in reported:
        x, o = got[int(s)], res_o[s]
        assert int(x[1]) == o["psid"] and int(x[2]) == o["idxoffset"] and int(x[3]) == o["sharedions"], s
        assert abs(float(x[6]) - o["e_x"]) <= 1e-3 * abs(o["e_x"]), s
