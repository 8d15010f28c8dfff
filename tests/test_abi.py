"""The C ABI: every declared entry point is exported, struct layouts match the
Python views, and host-side argument checks behave (no GPU needed)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from hicops_amd import native as N

ROOT = Path(__file__).resolve().parent.parent


def declared(header: Path) -> list[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hcx_[a-z0-9_]+)\s*\(", text)))


def test_hcx_exports_every_declared_symbol():
    names = declared(ROOT / "include" / "hcx.h")
    assert len(names) >= 15
    lib = N.hcx()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_datagen_exports_every_declared_symbol():
    names = declared(ROOT / "include" / "hcx_datagen.h") + declared(ROOT / "include" / "hcx_ms2.h")
    lib = N.host()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts(tmp_path):
    src = tmp_path / "lay.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "hcx.h"
#include "hcx_datagen.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(hcx_params), sizeof(hcx_pep_entry),
         sizeof(hcx_result), sizeof(hcx_stats), sizeof(hcx_cand), sizeof(hcx_queries_view),
         sizeof(hcx_search_opts), sizeof(hcx_dbgen_cfg), sizeof(hcx_specgen_cfg));
  printf("%zu %zu %zu %zu\n", offsetof(hcx_result, e_x), offsetof(hcx_result, cpsms),
         offsetof(hcx_cand, ibc), offsetof(hcx_params, dM));
  return 0;
}''')
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    sizes = [int(x) for x in out[0].split()]
    assert sizes == [C.sizeof(N.Params), N.PEP_DTYPE.itemsize, N.RESULT_DTYPE.itemsize,
                     N.STATS_DTYPE.itemsize, N.CAND_DTYPE.itemsize, C.sizeof(N.QueriesView),
                     C.sizeof(N.SearchOpts), C.sizeof(N.DbGenCfg), C.sizeof(N.SpecGenCfg)]
    offs = [int(x) for x in out[1].split()]
    assert offs == [N.RESULT_DTYPE.fields["e_x"][1], N.RESULT_DTYPE.fields["cpsms"][1],
                    N.CAND_DTYPE.fields["ibc"][1], N.Params.dM.offset]
    # pepEntry of the reference is 24 bytes (slm_dsts.h:146-282)
    assert sizes[1] == 24


def test_divider_matches_integer_division():
    rng = np.random.default_rng(0)
    nums = np.concatenate([rng.integers(0, 2**32, 20000, dtype=np.uint64),
                           np.array([0, 1, 2**32 - 1, 2**32 - 2, 2**31, 2**31 - 1], np.uint64)]).astype(np.uint32)
    out = np.zeros_like(nums)
    for d in [1, 2, 3, 12, 24, 32, 36, 48, 52, 64, 100, 116, 156, 234, 354, 1000, 65535]:
        extra = np.array([d * k + r for k in (1, 7, 1234, (2**32 - 1) // d) for r in (0, d - 1)
                          if d * k + r < 2**32], np.uint32)
        arr = np.concatenate([nums, extra])
        out = np.zeros_like(arr)
        N.check(N.hcx().hcx_selftest_divider(C.c_uint32(d), N.ptr(arr), C.c_uint32(len(arr)), N.ptr(out)))
        assert np.array_equal(out, arr // np.uint32(d)), d


def test_heap_capacity_matches_oracle():
    from oracle import binding as O
    for t in range(0, 70):
        assert N.hcx().hcx_heap_capacity(C.c_uint32(t)) == O.heap_capacity(t)


def test_create_rejects_bad_params():
    from hicops_amd import workload as W
    p = W.make_params(W.PRESETS["small"])
    bad = N.Params.from_buffer_copy(p)
    bad.topmatches = 1          # zero-capacity heap (reference UB, Q14)
    h = C.c_void_p()
    rc = N.hcx().hcx_create(C.byref(bad), 0, C.byref(h))
    assert rc == N.ERR_INVLD_PARAM
    bad = N.Params.from_buffer_copy(p)
    bad.max_mass = 0
    assert N.hcx().hcx_create(C.byref(bad), 0, C.byref(h)) == N.ERR_INVLD_PARAM
    assert N.hcx().hcx_create(None, 0, C.byref(h)) == N.ERR_INVLD_PTR


def test_null_pointer_errors():
    lib = N.hcx()
    assert lib.hcx_index_info(None, None, None, None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_search(None, None, 0, None, None, None, None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_queries_upload(None, None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_index_free(None) == 0


def test_datagen_deterministic_and_sorted():
    from hicops_amd import workload as W
    p = W.scaled("small", npep=2000, nspec=50)
    a, b = W.Database(p), W.Database(p)
    assert a.lengths == b.lengths
    for L in a.lengths:
        assert np.array_equal(a.entries(L), b.entries(L))
        m = a.entries(L)["mass"]
        assert np.all(np.diff(m) >= 0)
        seqs = a.seqs(L).reshape(-1, L)
        assert np.all(np.isin(seqs[:, -1], [ord("K"), ord("R")]))
    qa, qb = W.make_queries(a, p), W.make_queries(b, p)
    assert np.array_equal(qa.moz, qb.moz) and np.array_equal(qa.intensity, qb.intensity)
    # pickpeaks: ascending intensities, top peak = base_int, all >= min_int, <= QALEN peaks
    for s in range(qa.nspec):
        it = qa.intensity[qa.idx[s]:qa.idx[s + 1]]
        assert len(it) <= p.peaks_max
        assert np.all(np.diff(it.astype(np.int64)) >= 0)
        assert it[-1] == p.base_int and it.min() >= p.min_int


def test_datagen_mod_variants():
    from hicops_amd import workload as W
    p = W.scaled("rank_share", npep=3000, nspec=10)
    db = W.Database(p)
    nmod = 0
    for L in db.lengths:
        ents = db.entries(L)
        seqs = db.seqs(L).reshape(-1, L)
        for e in ents[ents["mod_num"] != 0]:
            sites = [i for i in range(L) if (int(e["sites"]) >> i) & 1]
            assert 1 <= len(sites) <= 2
            assert all(seqs[e["seq_id"], i] == ord("M") for i in sites)
            assert int(e["mod_num"]) == (0x1 if len(sites) == 1 else 0x11)
            nmod += 1
        # every M subset of size <= 2 appears once per sequence
        for sid in range(len(seqs)):
            m = int((seqs[sid, :-1] == ord("M")).sum())
            cnt = int(((ents["seq_id"] == sid) & (ents["mod_num"] != 0)).sum())
            assert cnt == m + m * (m - 1) // 2
    assert nmod > 100


# The reference's own C++ entry points, by their mangled names, so that an
# application built against source/core/include/dslim.h and expeRT.h links to
# libhcx_dslim.so (dslim.h:49,53,149,164,184; expeRT.h:143; dslim_fileout.h:29;
# the peptide database stage of lbe.h and mods.h).
REFERENCE_SYMBOLS = [
    "_Z15DSLIM_ConstructP6_Index",
    "_Z25DSLIM_InitializeScorecardP6_Indexj",
    "_Z19DSLIM_QuerySpectrumP8_queriesP6_Indexj",
    "_Z24DSLIM_DeallocateIonIndexP6_Index",
    "_Z18DSLIM_DeallocateSCv",
    "_ZN6expeRT12ModelTailFitEP8_Results",
    "_Z16DFile_PrintScoreP6_IndexjfP10_heapEntrydj",
    "_Z15DFile_InitFilesv",
    "_Z17DFile_DeinitFilesv",
    "_Z19DFile_PrintPartialsjP8_Results",
    # lbe.h:45,74,90,92; mods.h:63; lbe_internal.cpp:241
    "_Z14LBE_InitializeP6_Index",
    "_Z14LBE_DistributeP6_Index",
    "_Z13LBE_CountPepsPcP6_Indexj",
    "_Z20LBE_CreatePartitionsP6_Index",
    "_Z16LBE_DeinitializeP6_Index",
    "_Z15MODS_Initializev",
]


def test_mirror_exports_reference_entry_points():
    # like the reference core, the library leaves `gParams params` to the
    # application, so it is inspected rather than loaded
    so = Path(__file__).resolve().parent.parent / "hicops_amd" / "libhcx_dslim.so"
    out = subprocess.run(["nm", "-D", str(so)], capture_output=True, text=True, check=True).stdout
    kinds = {}
    for ln in out.splitlines():
        parts = ln.split()
        if len(parts) >= 2:
            kinds[parts[-1]] = parts[-2]
    for name in REFERENCE_SYMBOLS:
        assert kinds.get(name) in ("T", "W"), name
    assert kinds.get("params") == "U"


def test_pickpeaks_struct_layouts_and_null_checks(tmp_path):
    src = tmp_path / "lay2.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "hcx.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(hcx_raw_spectra_view), sizeof(hcx_pick_params),
         offsetof(hcx_raw_spectra_view, num_specs), offsetof(hcx_pick_params, min_int));
  return 0;
}''')
    exe = tmp_path / "lay2"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert out == [C.sizeof(N.RawSpectraView), C.sizeof(N.PickParams),
                   N.RawSpectraView.num_specs.offset, N.PickParams.min_int.offset]
    lib = N.hcx()
    h = C.c_void_p()
    # argument checks that return before any device call
    assert lib.hcx_raw_upload(None, None, C.byref(h)) == N.ERR_INVLD_PTR
    assert lib.hcx_queries_pick(None, None, None, C.byref(h), None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_queries_info(None, None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_queries_export(None, None, None, None, None, None, None) == N.ERR_INVLD_PTR
    assert lib.hcx_queries_facts(None, None, None, None, None, None) == N.ERR_INVLD_PTR


def test_status_codes_are_the_reference_values(tmp_path):
    # source/core/include/slmerr.h:24-43: SLM_ERROR_BASE (-3000) minus 1 ... 13
    ref = {"INVLD_MEMORY": -3001, "INVLD_PARAM": -3002, "INVLD_INDEX": -3003, "INVLD_SIZE": -3004,
           "SLM_ALREADY_INIT": -3005, "FILE_NOT_FOUND": -3006, "BAD_MEM_ALLOC": -3007, "INVLD_PTR": -3008,
           "INVLD_SORT": -3009, "INVLD_MOD": -3010, "INVLD_NODE_RANK": -3011, "NOT_ENOUGH_DATA": -3012,
           "FILE_ERROR": -3013}
    src = tmp_path / "codes.cpp"
    src.write_text("#include <cstdio>\n#include \"hcx_dslim.hpp\"\nint main() {\n" + "".join(
        f'  std::printf("%d %d\\n", HCX_ERR_{k}, ERR_{k});\n' for k in ref) + "  return SLM_SUCCESS;\n}\n")
    exe = tmp_path / "codes"
    subprocess.run(["g++", "-std=c++17", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for (k, v), line in zip(ref.items(), out):
        assert [int(x) for x in line.split()] == [v, v], k
    for k in ("INVLD_MEMORY", "INVLD_PARAM", "INVLD_SIZE", "FILE_NOT_FOUND", "BAD_MEM_ALLOC", "INVLD_PTR",
              "NOT_ENOUGH_DATA"):
        assert getattr(N, "ERR_" + k) == ref[k], k
