"""MS2 text ingest (include/hcx_ms2.h) against the reference's reading rules:
MSQuery::initialize's spectrum count (msquery.cpp:96-170) and
MSQuery::readspectrum (msquery.cpp:300-505), pinned by hand-written files
whose expected values follow from those rules, and by a round trip of
generated spectra through MS2 text."""
import numpy as np
import pytest

from hicops_amd import ingest
from hicops_amd import native as N
from hicops_amd import workload as W

QUIRKS = (
    "H\tCreationDate\ttoday\n"
    "H\tExtractor\tx\n"
    "S\t1\t1\t500.25\n"
    "I\tRTime\t12.5\n"
    "Z\t2\t1001.5\n"
    "100.25 2000.5\n"
    "200.5\t3000\n"          # peaks split on spaces only: one token, intensity missing
    "300.75\n"               # first spectrum: a missing token reads "0.01"
    "D\tsomething\n"
    "S\t2\t2\t600.5\n"
    "Z\t3\t1500.75\n"
    "I\tOther\t5\n"          # not RTime: ignored
    "50.5 100\n"
    "60.5\t7\n"              # later spectra: a missing token reads "0.0"
    "70.5\n"
    "\n"
    "S\t3\t3\t700.0\n"       # no Z / I lines: charge, precursor and RTime carry over
    "1e2 5e-1\n"
    "   80.5 9\n"            # leading spaces are skipped by strtok
    "\r\n"                   # split at '\n' only: a lone '\r' is a peak line
    "S\t4\t4\t800\n"
    "Z\t1\n"                 # precursor token missing: "0.01"
    "I\tRTime\n"             # value missing: "0.00"
    "90.25 1.5\n"
    "S\t5\t5\t900\n"         # empty last spectrum
)


def bins(tok, scale):
    return int(float(tok) * scale)


def test_reading_rules(tmp_path):
    f = tmp_path / "q.ms2"
    f.write_bytes(QUIRKS.encode())
    r = ingest.read_ms2(f, 100)
    assert r.nspec == 5
    want = [
        [(bins("100.25", 100), bins("2000.5", 1000)), (bins("200.5", 100), bins("0.01", 1000)),
         (bins("300.75", 100), bins("0.01", 1000))],
        [(bins("50.5", 100), bins("100", 1000)), (bins("60.5", 100), 0), (bins("70.5", 100), 0)],
        [(bins("1e2", 100), bins("5e-1", 1000)), (bins("80.5", 100), bins("9", 1000)), (0, 0)],
        [(bins("90.25", 100), bins("1.5", 1000))],
        [],
    ]
    assert r.idx.tolist() == list(np.cumsum([0] + [len(w) for w in want]))
    got = [list(zip(r.mz[r.idx[s]:r.idx[s + 1]].tolist(), r.intensity[r.idx[s]:r.idx[s + 1]].tolist()))
           for s in range(5)]
    assert got == want
    assert r.z.tolist() == [2, 3, 3, 1, 1]
    assert r.prec_mz.tolist() == [1001.5, 1500.75, 1500.75, 0.01, 0.01]
    assert r.rtime.tolist() == [12.5, 12.5, 12.5, 0.0, 0.0]


def test_peaks_before_the_first_scan_and_errors(tmp_path):
    f = tmp_path / "a.ms2"
    f.write_text("10.5 1\nS 1\nZ 2 500.5\n20.5 2\nS 2\n30.5 3\n")
    r = ingest.read_ms2(f, 100)
    # the first spectrum runs from the file start to the second S
    assert r.idx.tolist() == [0, 2, 3]
    assert r.mz.tolist() == [1050, 2050, 3050] and r.intensity.tolist() == [1000, 2000, 3000]
    assert r.z.tolist() == [2, 2] and r.prec_mz.tolist() == [500.5, 500.5]
    g = tmp_path / "b.ms2"
    g.write_text("H x\nS 1\nZ 2 100\n")
    with pytest.raises(N.HcxError) as e:
        ingest.read_ms2(g, 100)
    assert e.value.code == N.ERR_INVLD_SIZE
    with pytest.raises(N.HcxError) as e:
        ingest.read_ms2(tmp_path / "missing.ms2", 100)
    assert e.value.code == N.ERR_FILE_NOT_FOUND


def write_ms2(path, raw, scale):
    """Raw spectra as MS2 text (m/z and intensity with 4 decimals)."""
    lines, toks = [], []
    for s in range(raw.nspec):
        lines += [f"S\t{s + 1}\t{s + 1}\t{raw.prec_mz[s]:.4f}", f"I\tRTime\t{raw.rtime[s]:.4f}",
                  f"Z\t{raw.z[s]}\t{raw.prec_mz[s]:.4f}"]
        for k in range(raw.idx[s], raw.idx[s + 1]):
            m, i = f"{raw.mz[k] / scale + 0.004:.4f}", f"{raw.intensity[k] / 1000:.4f}"
            lines.append(f"{m} {i}")
            toks.append((m, i))
    path.write_text("\n".join(lines) + "\n")
    return toks


def test_round_trip_of_generated_spectra(tmp_path):
    p = W.scaled("small", npep=2000, nspec=300)
    db = W.Database(p)
    raw, _ = W.make_raw_spectra(db, p)
    f = tmp_path / "gen.ms2"
    toks = write_ms2(f, raw, 100)
    r = ingest.read_ms2(f, 100)
    assert r.nspec == raw.nspec and np.array_equal(r.idx, raw.idx)
    assert r.mz.tolist() == [bins(m, 100) for m, _ in toks]
    assert r.intensity.tolist() == [bins(i, 1000) for _, i in toks]
    assert np.array_equal(r.z, raw.z)
    assert r.prec_mz.tolist() == [float(f"{x:.4f}") for x in raw.prec_mz]
    assert r.rtime.tolist() == [float(f"{x:.4f}") for x in raw.rtime]
