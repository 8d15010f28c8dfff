"""ctypes bindings of libhcx.so (include/hcx.h) and libhcx_host.so (include/hcx_datagen.h,
include/hcx_ms2.h).

The search path runs only through libhcx.so.  If the library is missing or a
call fails, an exception is raised: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIBHCX = PKG / "libhcx.so"
LIBHOST = PKG / "libhcx_host.so"

HCX_MAX_MOD_TYPES = 15

# status codes (slmerr.h values)
SUCCESS = 0
# source/core/include/slmerr.h:24-43
ERR_INVLD_MEMORY = -3001
ERR_INVLD_PARAM = -3002
ERR_INVLD_SIZE = -3004
ERR_FILE_NOT_FOUND = -3006
ERR_BAD_MEM_ALLOC = -3007
ERR_INVLD_PTR = -3008
ERR_NOT_ENOUGH_DATA = -3012
ERR_DEVICE = -3100

RES_VALID = 0x1
RES_REPORTED = 0x2
RES_TIE_REPLAYED = 0x4
RES_HIST_CLAMP = 0x8
RES_LGF_CLAMP = 0x10
RES_FIT_UB = 0x20
RES_SURROGATE = 0x40
RES_AUDITED = 0x80
RES_COUNT_OVF = 0x100
RES_TWO_PASS = 0x200
RES_FIT_HOST = 0x400

KERNEL_PLAIN = 0
KERNEL_BATCHED = 1
KERNEL_CURSOR = 2
OPT_HITLIST = 1
OPT_GROUP = 3
OPT_SCAN = 4
OPT_GRAB = 5
OPT_FORCE_REDO = 6
OPT_FORCE_HOST_FIT = 7
SCAN_AUTO = 0
SCAN_TWO_PASS = 1
EVAL_TAILFIT = 0
EVAL_SURROGATE = 1
POLICY_CHUNK = 0
POLICY_CYCLIC = 1
ERR_INVLD_SORT = -3009
ERR_INVLD_MOD = -3010
ERR_INVLD_NODE_RANK = -3011


class HcxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hcx error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("maxz", C.c_uint32), ("scale", C.c_uint32), ("min_mass", C.c_uint32),
        ("max_mass", C.c_uint32), ("dF", C.c_uint32), ("min_shp", C.c_uint32),
        ("min_cpsm", C.c_uint32), ("topmatches", C.c_uint32), ("max_hyperscore", C.c_uint32),
        ("max_shdpeaks", C.c_uint32), ("nmods", C.c_uint32),
        ("mod_mass", C.c_uint32 * HCX_MAX_MOD_TYPES),
        ("dM", C.c_double), ("expect_max", C.c_double),
    ]

    @property
    def hbins(self) -> int:
        return 10 * self.max_hyperscore + 2


class QueriesView(C.Structure):
    _fields_ = [
        ("moz", C.c_void_p), ("intensity", C.c_void_p), ("idx", C.c_void_p),
        ("precurse", C.c_void_p), ("charges", C.c_void_p), ("rtimes", C.c_void_p),
        ("num_specs", C.c_int32), ("file_num", C.c_int32),
    ]


class RawSpectraView(C.Structure):
    _fields_ = [
        ("mz", C.c_void_p), ("intensity", C.c_void_p), ("idx", C.c_void_p),
        ("prec_mz", C.c_void_p), ("z", C.c_void_p), ("rtime", C.c_void_p),
        ("num_specs", C.c_int32), ("file_num", C.c_int32),
    ]


class PickParams(C.Structure):
    _fields_ = [("qalen", C.c_uint32), ("base_int", C.c_uint32), ("min_int", C.c_uint32),
                ("pad_", C.c_uint32)]


class SearchOpts(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("evalue", C.c_int32), ("tile", C.c_uint32),
        ("audit_every", C.c_uint32), ("audit_tol", C.c_double), ("use_graph", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("chunk", C.c_uint32), ("slots", C.c_uint32), ("blocks_per_cu", C.c_uint32),
        ("hitlist", C.c_int32), ("grab", C.c_int32), ("kernel_ms", C.c_float),
    ]


class Audit(C.Structure):
    _fields_ = [
        ("audited", C.c_uint32), ("disagreed", C.c_uint32), ("max_rel_err", C.c_double),
        ("fell_back", C.c_int32), ("located", C.c_uint32), ("plain_fit", C.c_uint32), ("pad_", C.c_uint32),
    ]


class LbeCfg(C.Structure):
    _fields_ = [("modconditions", C.c_char_p), ("nodes", C.c_uint32), ("myid", C.c_uint32),
                ("policy", C.c_int32), ("sort", C.c_int32)]


class DbGenCfg(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("npep", C.c_uint32), ("min_len", C.c_uint32),
        ("max_len", C.c_uint32), ("min_mass", C.c_uint32), ("max_mass", C.c_uint32),
        ("scale", C.c_uint32), ("nmods", C.c_uint32), ("max_mods", C.c_uint32),
        ("mod_residues", (C.c_char * 4) * HCX_MAX_MOD_TYPES),
        ("mod_max", C.c_uint32 * HCX_MAX_MOD_TYPES),
        ("mod_mass", C.c_uint32 * HCX_MAX_MOD_TYPES),
    ]


class SpecGenCfg(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("nspec", C.c_uint32), ("min_charge", C.c_uint32),
        ("max_charge", C.c_uint32), ("frag_maxz", C.c_uint32), ("n_noise", C.c_uint32),
        ("peaks_min", C.c_uint32), ("peaks_max", C.c_uint32), ("scale", C.c_uint32),
        ("base_int", C.c_uint32), ("min_int", C.c_uint32), ("window_mode", C.c_uint32),
        ("keep_prob", C.c_double), ("noise_lo", C.c_double), ("noise_hi", C.c_double),
        ("jitter_sd", C.c_double), ("sig_mu", C.c_double), ("sig_sigma", C.c_double),
        ("noise_mu", C.c_double), ("noise_sigma", C.c_double), ("dM", C.c_double),
        ("wmin", C.c_double), ("wmax", C.c_double),
    ]


PEP_DTYPE = np.dtype([("mass", "<f4"), ("seq_id", "<u4"), ("sites", "<u8"),
                      ("mod_num", "<u4"), ("pad_", "<u4")])
RESULT_DTYPE = np.dtype([
    ("hyperscore", "<f4"), ("psid", "<i4"), ("idxoffset", "<u2"), ("sharedions", "<u2"),
    ("totalions", "<u2"), ("flags", "<u2"), ("cpsms", "<u4"), ("maxhypscore", "<i4"),
    ("mu", "<i4"), ("beta", "<i4"), ("minhypscore", "<i4"), ("nexthypscore", "<i4"),
    ("bc", "<u4"), ("yc", "<u4"), ("e_x", "<f8")])
STATS_DTYPE = np.dtype([("n_ia", "<u8"), ("n_probe", "<u8")])
CAND_DTYPE = np.dtype([
    ("spec", "<u4"), ("idxoffset", "<u2"), ("bc", "<u2"), ("psid", "<i4"), ("yc", "<u2"),
    ("pad_", "<u2"), ("ibc", "<u8"), ("iyc", "<u8"), ("hyperscore", "<f4"), ("pad2_", "<u4")])

PART_DTYPE = np.dtype([("min", "<u2"), ("max2", "<u2"), ("max", "<f4"), ("N", "<i4"), ("qID", "<i4")])
FRES_DTYPE = np.dtype([("eValue", "<i4"), ("specID", "<i4"), ("npsms", "<i4")])
assert PEP_DTYPE.itemsize == 24 and RESULT_DTYPE.itemsize == 56
assert PART_DTYPE.itemsize == 16 and FRES_DTYPE.itemsize == 12
assert CAND_DTYPE.itemsize == 40 and C.sizeof(Params) == 120


def ptr(a: np.ndarray | None) -> C.c_void_p:
    if a is None:
        return C.c_void_p(0)
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return C.c_void_p(a.ctypes.data)


_hcx = None
_host = None


def _load(path: Path, what: str) -> C.CDLL:
    if not path.exists():
        raise RuntimeError(
            f"{what} is not built ({path}); run `python -m hicops_amd.build` "
            "(the search path has no fallback implementation)")
    return C.CDLL(str(path))


def hcx() -> C.CDLL:
    """The HIP library.  Raises if it is not built."""
    global _hcx
    if _hcx is None:
        lib = _load(LIBHCX, "libhcx.so")
        lib.hcx_last_error.restype = C.c_char_p
        for name in ["hcx_create", "hcx_destroy", "hcx_index_build", "hcx_index_info",
                     "hcx_index_export", "hcx_index_build_ms", "hcx_index_free",
                     "hcx_queries_upload", "hcx_queries_free", "hcx_search",
                     "hcx_search_device", "hcx_search_fetch", "hcx_search_candidates",
                     "hcx_tailfit_batch", "hcx_lgfact_table", "hcx_heap_capacity",
                     "hcx_selftest_divider", "hcx_set_tuning", "hcx_get_tuning", "hcx_set_option", "hcx_get_option", "hcx_autotune",
                     "hcx_surrogate_param_count", "hcx_surrogate_init", "hcx_surrogate_set",
                     "hcx_surrogate_get", "hcx_surrogate_train", "hcx_surrogate_gradient",
                     "hcx_surrogate_last_audit", "hcx_selftest_mfma", "hcx_last_timing", "hcx_last_scan_split", "hcx_selftest_log2", "hcx_search_noalloc",
                     "hcx_raw_upload", "hcx_raw_free", "hcx_queries_pick", "hcx_queries_pickpeaks",
                     "hcx_queries_info", "hcx_queries_export", "hcx_queries_facts", "hcx_part_export", "hcx_combine",
                     "hcx_pepdb_create", "hcx_pepdb_counts", "hcx_pepdb_generate", "hcx_pepdb_export",
                     "hcx_pepdb_ms", "hcx_pepdb_free", "hcx_index_build_pepdb", "hcx_sort_mass"]:
            getattr(lib, name).restype = C.c_int
        _hcx = lib
    return _hcx


def host() -> C.CDLL:
    global _host
    if _host is None:
        lib = _load(LIBHOST, "libhcx_host.so")
        for name in ["hcx_db_create", "hcx_db_free", "hcx_db_lengths", "hcx_db_length_info",
                     "hcx_db_length_copy", "hcx_db_length_view", "hcx_spec_create",
                     "hcx_spec_free", "hcx_spec_info", "hcx_spec_copy", "hcx_spec_create_raw",
                     "hcx_spec_raw_info", "hcx_spec_raw_copy", "hcx_ms2_read", "hcx_ms2_free",
                     "hcx_ms2_info", "hcx_ms2_copy"]:
            getattr(lib, name).restype = C.c_int
        _host = lib
    return _host


def check(rc: int, lib: C.CDLL | None = None) -> None:
    if rc != SUCCESS:
        msg = ""
        try:
            msg = (lib or hcx()).hcx_last_error().decode()
        except Exception:  # noqa: BLE001 - the host library has no error text
            pass
        raise HcxError(rc, msg)
