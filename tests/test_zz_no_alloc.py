"""hcx_search_noalloc (the allocation-free search: scratch in a static device
array) against hcx_search, through the C ABI: two calls in a row, the second
with a smaller batch, for each query form; and the size limit."""
import ctypes as C
import sys

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


def dev_alloc(nbytes):
    p = C.c_void_p()
    assert hip().hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 16))) == 0
    return p


def to_host(dptr, arr):
    assert hip().hipMemcpy(arr.ctypes.data_as(C.c_void_p), dptr, C.c_size_t(arr.nbytes), 2) == 0   # device to host
    return arr


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def noalloc(eng, batch, kernel):
    arr, n = eng._idx_array(None)
    ns = batch.nspec
    d_res, d_hist, d_st = dev_alloc(N.RESULT_DTYPE.itemsize * ns), dev_alloc(4 * eng.hbins * ns), \
        dev_alloc(N.STATS_DTYPE.itemsize * ns)
    try:
        N.check(N.hcx().hcx_search_noalloc(eng._h, arr, n, batch._h, C.c_int32(kernel), d_res, d_hist, d_st))
        res = to_host(d_res, np.zeros(ns, N.RESULT_DTYPE))
        hist = to_host(d_hist, np.zeros((ns, eng.hbins), np.uint32))
        st = to_host(d_st, np.zeros(ns, N.STATS_DTYPE))
    finally:
        for d in (d_res, d_hist, d_st):
            hip().hipFree(d)
    return res, hist, st


@pytest.mark.parametrize("kernel", [N.KERNEL_PLAIN, N.KERNEL_BATCHED])
def test_noalloc_matches_search_twice_smaller_second(kernel):
    p = W.scaled("small", npep=6000, nspec=400, dM=100.0)
    db = W.Database(p)
    q = W.make_queries(db, p)
    q2 = q.subset(np.arange(0, q.nspec, 3))   # the second call: a third of the spectra
    prm = W.make_params(p)
    log(f"test_noalloc_matches_search_twice_smaller_second kernel={kernel}")
    eng = Engine(prm)
    eng.construct_db(db)
    b1, b2 = eng.upload(q), eng.upload(q2)
    ref1 = eng.query_spectrum(b1, kernel=kernel)
    ref2 = eng.query_spectrum(b2, kernel=kernel)
    got1 = noalloc(eng, b1, kernel)
    got2 = noalloc(eng, b2, kernel)
    for (res, hist, st), ref, label in ((got1, ref1, "first"), (got2, ref2, "second")):
        valid = (ref.results["flags"] & N.RES_VALID) != 0
        assert valid.sum() > 50, label
        assert np.array_equal(res, ref.results), label
        assert np.array_equal(hist, ref.hist), label
        assert np.array_equal(st, ref.stats), label
    eng.close()


def test_noalloc_rejects_a_batch_beyond_its_scratch():
    p = W.scaled("small", npep=3000, nspec=10)
    db = W.Database(p)
    prm = W.make_params(p)
    eng = Engine(prm)
    eng.construct_db(db)
    n = 400_000   # ~400 bytes of scratch per spectrum over these lengths: beyond 64 MiB
    q = W.Queries(np.full(n, 50000, np.uint32), np.full(n, 1000, np.uint32), np.arange(n + 1, dtype=np.uint32),
                  np.full(n, 1500.0, np.float32), np.full(n, 2, np.int32), np.zeros(n, np.float32),
                  np.zeros(n, np.int32), np.zeros(n, np.int32))
    b = eng.upload(q)
    arr, ni = eng._idx_array(None)
    d = dev_alloc(16)
    try:
        rc = N.hcx().hcx_search_noalloc(eng._h, arr, ni, b._h, C.c_int32(N.KERNEL_PLAIN), d, d, None)
    finally:
        hip().hipFree(d)
    assert rc == N.ERR_INVLD_SIZE
    assert b"limit" in N.hcx().hcx_last_error()
    eng.close()
