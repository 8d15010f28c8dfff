"""The PSM TSV writer of the C++ mirror (dslim_fileout.cpp:45-152) on the CPU:
a small program fills an Index by hand and calls DFile_InitFiles,
DFile_PrintScore, DFile_PrintPartials and DFile_DeinitFiles; the files must
hold the reference's header and lines, each field as std::to_string prints it
("%f" for float and double, "%d" / "%u" for integers)."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "hicops_amd"

HEADER = ("file\tscan_num\tprec_mass\tcharge\tretention_time\tpeptide\tmatched_ions\ttotal_ions\t"
          "calc_pep_mass\tmass_diff\tmod_info\thyperscore\texpectscore\tnum_hits")

PROGRAM = r'''
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "hcx_dslim.hpp"

gParams params;
std::vector<std::string> queryfiles = {"a.ms2", "run_two.ms2"};

int main(int argc, char **argv) {
    params.workspace = argv[1];
    params.threads = 3;
    params.myid = 7;
    // two peptide lengths: 4 ("PEPK", "GGGR") and 6 ("ACDEFK")
    static char s4[] = "PEPKGGGR";
    static char s6[] = "ACDEFK";
    static pepEntry e4[2], e6[1];
    e4[0].Mass = 471.2336f; e4[0].seqID = 1;    // psid 0 -> "GGGR"
    e4[1].Mass = 1234.5f;   e4[1].seqID = 0;    // psid 1 -> "PEPK"
    e6[0].Mass = 711.2911f; e6[0].seqID = 0;
    Index ix[2];
    std::memset(ix, 0, sizeof ix);
    ix[0].pepIndex.seqs = s4; ix[0].pepIndex.peplen = 4; ix[0].pepEntries = e4;
    ix[1].pepIndex.seqs = s6; ix[1].pepIndex.peplen = 6; ix[1].pepEntries = e6;
    if (DFile_InitFiles() != SLM_SUCCESS) return 1;
    hCell h;
    std::memset(&h, 0, sizeof h);
    h.idxoffset = 0; h.psid = 0; h.sharedions = 7; h.totalions = 12; h.fileIndex = 1;
    h.pchg = 2; h.rtime = 12.3456789f; h.hyperscore = 23.45678f;
    if (DFile_PrintScore(ix, 41, 470.9876f, &h, 1.2345678e-7, 321)) return 2;
    h.idxoffset = 1; h.psid = 0; h.sharedions = 11; h.totalions = 20; h.fileIndex = 0;
    h.pchg = -1; h.rtime = 0.0f; h.hyperscore = 0.5f;
    if (DFile_PrintScore(ix, 0, 1500.25f, &h, 98765.4321, 4)) return 3;
    h.idxoffset = 0; h.psid = 1; h.fileIndex = 9;              // beyond queryfiles: empty name
    h.pchg = 3; h.rtime = 5000.5f; h.hyperscore = 100.125f;
    if (DFile_PrintScore(ix, 199999, 1234.5f, &h, 3.0e300, 1000000)) return 4;
    Results r;
    std::memset(&r, 0, sizeof r);
    r.cpsms = 17; r.mu = -123456; r.beta = 2345678; r.minhypscore = 3; r.nexthypscore = 41;
    if (DFile_PrintPartials(5, &r)) return 5;
    if (DFile_DeinitFiles() != SLM_SUCCESS) return 6;
    return 0;
}
'''


def f32(x):
    return float(np.float32(x))


def test_tsv_writer_format(tmp_path):
    lib = PKG / "libhcx_dslim.so"
    if not lib.exists():
        pytest.skip("libhcx_dslim.so is built by __graft_entry__.build()")
    src = tmp_path / "w.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "w"
    subprocess.run(["g++", "-std=c++17", "-I", str(ROOT / "include"), "-o", str(exe), str(src),
                    "-L", str(PKG), "-lhcx_dslim", "-lhcx", f"-Wl,-rpath,{PKG}"], check=True)
    ws = tmp_path / "ws"
    ws.mkdir()
    subprocess.run([str(exe), str(ws)], check=True, timeout=60)
    files = sorted(ws.glob("*.tsv"))
    # <%m.%d.%Y_%H.%M.%S_><myid>_<f>.tsv, one per params.threads
    assert len(files) == 3
    for f, name in enumerate(sorted(p.name for p in files)):
        assert re.fullmatch(r"\d\d\.\d\d\.\d{4}_\d\d\.\d\d\.\d\d_7_%d\.tsv" % f, name), name
    texts = [p.read_text() for p in files]
    for t in texts:
        assert t.split("\n")[0] == HEADER
    assert texts[1] == HEADER + "\n" and texts[2] == HEADER + "\n"
    lines = texts[0].split("\n")[1:-1]
    m0, m1, m6 = f32(471.2336), f32(1234.5), f32(711.2911)
    want = [
        ["run_two.ms2", "42", "%f" % f32(470.9876), "2", "%f" % f32(12.3456789), "GGGR", "7", "12",
         "%f" % m0, "%f" % f32(np.float32(470.9876) - np.float32(471.2336)), "", "%f" % f32(23.45678),
         "%f" % 1.2345678e-7, "321"],
        ["a.ms2", "1", "%f" % f32(1500.25), "-1", "%f" % 0.0, "ACDEFK", "11", "20",
         "%f" % m6, "%f" % f32(np.float32(1500.25) - np.float32(711.2911)), "", "%f" % 0.5,
         "%f" % 98765.4321, "4"],
        ["", "200000", "%f" % m1, "3", "%f" % f32(5000.5), "PEPK", "11", "20",
         "%f" % m1, "%f" % 0.0, "", "%f" % 100.125, "%f" % 3.0e300, "1000000"],
    ]
    assert [ln.split("\t") for ln in lines[:3]] == want
    # DFile_PrintPartials: specid+1, cpsms, mu, beta, minhypscore, nexthypscore
    assert lines[3] == "6\t17\t-123456\t2345678\t3\t41"
    assert len(lines) == 4
