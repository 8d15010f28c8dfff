"""Learned tail-parameter surrogate on the GPU: MFMA lane layout, gradients
against the CPU reference, training, audit and fallback to the plain fit."""
import sys

import numpy as np
import pytest

from hicops_amd import native as N
from hicops_amd import workload as W
from hicops_amd.search import Engine
from tests import surrogate_ref as S

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng_small():
    prm = W.make_params(W.PRESETS["small"])
    eng = Engine(prm)
    yield eng
    eng.close()


def test_mfma_lane_layout(eng_small):
    log("test_mfma_lane_layout")
    rng = np.random.default_rng(5)
    A = rng.normal(size=(16, 16)).astype(np.float32)
    B = rng.normal(size=(16, 16)).astype(np.float32)
    D = eng_small.selftest_mfma(A, B)
    ref = S.bf16(A).astype(np.float64) @ S.bf16(B).astype(np.float64)
    assert np.allclose(D, ref, rtol=1e-5, atol=1e-5)


def test_gradients_match_cpu_reference(eng_small):
    log("test_gradients_match_cpu_reference")
    eng_small.surrogate_init(seed=11)
    p, nrm = eng_small.surrogate_params()
    rng = np.random.default_rng(6)
    X = rng.uniform(-2, 0, (48, S.F)).astype(np.float32)
    X[:, 96:] = 0
    X[:, 96] = rng.uniform(0, 1, 48)
    T = np.concatenate([rng.normal(size=(48, 3)), (rng.random((48, 1)) > 0.25)], 1).astype(np.float32)
    g, loss = eng_small.surrogate_gradient(X, T)
    lref, gref, _ = S.forward_backward(p, X, T, emulate_bf16=True)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    for name, a, b in [("W1", S.oW1, S.ob1), ("b1", S.ob1, S.oW2), ("W2", S.oW2, S.ob2), ("b2", S.ob2, S.oW3),
                       ("W3", S.oW3, S.ob3), ("b3", S.ob3, S.NPARAM)]:
        gg, gr = g[a:b].astype(np.float64), gref[a:b].astype(np.float64)
        scale = np.abs(gr).max() + 1e-12
        err = np.linalg.norm(gg - gr) / (np.linalg.norm(gr) + 1e-12)
        log(f"  {name}: rel L2 err {err:.2e}, max |g| {scale:.3e}")
        assert err <= 1e-2, name
        assert np.all(np.abs(gg - gr) <= 2e-2 * scale + 1e-7), name


def test_train_predict_audit_and_fallback():
    p = W.scaled("rank_share", npep=30000, nspec=3000, maxz=2)
    db = W.Database(p)
    prm = W.make_params(p)
    q_train = W.make_queries(db, p)
    q_test = W.make_queries(db, p, nspec=1500, seed=0xBEEF)
    eng = Engine(prm)
    eng.construct_db(db)
    bt = eng.upload(q_train)
    plain_train = eng.query_spectrum(bt, kernel=N.KERNEL_BATCHED)
    bs = eng.upload(q_test)
    plain = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED)
    valid = (plain.results["flags"] & N.RES_VALID) != 0
    log(f"test_train_predict_audit_and_fallback: {int(valid.sum())} valid test spectra")
    assert valid.sum() > 300 and ((plain_train.results["flags"] & N.RES_VALID) != 0).sum() > 600

    def held_out():
        out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE,
                                 audit_every=1 << 30, audit_tol=1e300)
        r = out.results
        # the top hits never depend on the e-value method, and located or
        # not, every tail fit equals the plain fit's ints and e-value
        for f in ("psid", "idxoffset", "sharedions", "cpsms", "hyperscore", "mu", "beta", "minhypscore",
                  "nexthypscore", "maxhypscore"):
            assert np.array_equal(r[f][valid], plain.results[f][valid]), f
        assert np.array_equal(r["e_x"][valid], plain.results["e_x"][valid])
        assert np.array_equal(out.hist, plain.hist)
        a = eng.last_audit()
        sur = (r["flags"] & N.RES_SURROGATE) != 0
        # (the audited spectrum carries the plain fit's flags)
        assert 0 <= a["located"] - int(sur.sum()) <= a["audited"]
        assert a["located"] + a["plain_fit"] == int(valid.sum())
        return a["located"] / max(1, a["located"] + a["plain_fit"])

    # what training gains: the share of spectra whose fit window the
    # network's prediction locates (the rest take the plain fit), untrained
    # against trained
    eng.surrogate_init(seed=3)
    rate0 = held_out()
    loss = eng.surrogate_train(bt, steps=3000, batch_size=256, lr=2e-3)
    rate = held_out()
    log(f"  trained: final loss {loss:.4f}; windows located on held-out spectra {rate:.3f} (untrained {rate0:.3f})")
    assert rate > 0.8 and rate > rate0 + 0.3

    # audit sample every 8th spectrum at 1e-3: the audited spectra carry the
    # plain fit, and nothing disagrees
    out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE, audit_every=8, audit_tol=1e-3)
    a = eng.last_audit()
    log(f"  audit: {a}")
    r = out.results
    audited = (r["flags"] & N.RES_AUDITED) != 0
    assert a["audited"] == int(audited.sum()) and a["audited"] > 0
    assert np.all(np.arange(len(r))[audited] % 8 == 0)
    assert a["disagreed"] == 0 and not a["fell_back"] and a["max_rel_err"] == 0.0
    e_p = plain.results["e_x"]
    assert np.array_equal(r["e_x"][valid], e_p[valid])

    # parameters round-trip through the host
    params, nrm = eng.surrogate_params()
    eng2 = Engine(prm)
    eng2.surrogate_set(params, nrm)
    p2, n2 = eng2.surrogate_params()
    assert np.array_equal(p2, params) and np.array_equal(n2, nrm)
    eng2.close()
    eng.close()


@pytest.mark.timeout(900)
def test_surrogate_config3_500k_pairs():
    # trained on 500,000 (histogram, plain fit) pairs from a search of the
    # open-search workload (10M peptides, +-500 Da) with spectra of their own
    # seed, evaluated on 50,000 held-out spectra of another seed: the
    # e-value error against the plain fit, and the share of spectra whose
    # window the prediction missed (they take the plain fit)
    p = W.PRESETS["open10m"]
    db = W.Database(p)
    prm = W.make_params(p)
    eng = Engine(prm)
    eng.construct_db(db)
    qt = W.make_queries(db, p, nspec=500_000, seed=p.seed ^ 0x3A11)
    qh = W.make_queries(db, p, nspec=50_000, seed=p.seed ^ 0x3B22)
    log(f"test_surrogate_config3_500k_pairs: {db.nforms} forms, {qt.nspec} training and {qh.nspec} held-out spectra")
    bt = eng.upload(qt)
    eng.query_device(bt, kernel=N.KERNEL_BATCHED)
    bh = eng.upload(qh)
    plain = eng.query_spectrum(bh, kernel=N.KERNEL_BATCHED)
    k, t = eng.query_device(bh, kernel=N.KERNEL_BATCHED)
    plain_ms = eng.last_timing()["evalue_ms"]
    valid = (plain.results["flags"] & N.RES_VALID) != 0

    def evaluate():
        out = eng.query_spectrum(bh, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE, audit_every=64,
                                 audit_tol=1e-3)
        a = eng.last_audit()
        e_s, e_p = out.results["e_x"][valid], plain.results["e_x"][valid]
        rel = np.abs(e_s - e_p) / np.abs(e_p)
        miss = a["plain_fit"] / max(1, a["located"] + a["plain_fit"])
        return out, a, rel, miss

    eng.surrogate_init(seed=3)
    _, a0, _, miss0 = evaluate()
    loss = eng.surrogate_train(bt, steps=6000, batch_size=256, lr=2e-3)
    bt.free()
    out, a, rel, miss = evaluate()
    k, t = eng.query_device(bh, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE, audit_every=64, audit_tol=1e-3)
    sur_ms = eng.last_timing()["evalue_ms"]
    q = np.quantile(rel, [0.5, 0.95, 0.99, 1.0])
    log(f"  {int(valid.sum())} valid held-out spectra; loss {loss:.4f}; relative e-value error vs the plain fit: "
        f"median {q[0]:.3g}, 95% {q[1]:.3g}, 99% {q[2]:.3g}, max {q[3]:.3g}; within 1e-3: {(rel <= 1e-3).mean():.4f}")
    log(f"  prediction missed (plain fit): {miss:.4f} trained, {miss0:.4f} untrained; audit {a}")
    log(f"  e-value stage: surrogate {sur_ms:.3f} ms, plain fit {plain_ms:.3f} ms for {qh.nspec} spectra")
    assert (rel <= 1e-3).mean() >= 0.95
    assert a["disagreed"] == 0 and not a["fell_back"]
    assert miss < 0.1 and miss < miss0
    for f in ("psid", "mu", "beta", "minhypscore", "nexthypscore"):
        assert np.array_equal(out.results[f][valid], plain.results[f][valid]), f
    eng.close()
