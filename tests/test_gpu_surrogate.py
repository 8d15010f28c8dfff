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
    ok = valid & ((plain.results["flags"] & N.RES_FIT_UB) == 0)
    log(f"test_train_predict_audit_and_fallback: {int(valid.sum())} valid test spectra")
    assert ok.sum() > 300 and ((plain_train.results["flags"] & N.RES_VALID) != 0).sum() > 600

    def held_out_error():
        out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE,
                                 audit_every=1 << 30, audit_tol=1e300)
        r = out.results
        # the top hits never depend on the e-value method
        for f in ("psid", "idxoffset", "sharedions", "cpsms", "hyperscore"):
            assert np.array_equal(r[f], plain.results[f]), f
        assert np.array_equal(out.hist, plain.hist)
        sel = ok.copy()
        sel[0] = False                               # spectrum 0 is the audit sample
        assert np.all((r["flags"][sel] & N.RES_SURROGATE) != 0)
        d = np.abs(np.log10(r["e_x"][sel]) - np.log10(plain.results["e_x"][sel]))
        return float(np.median(d)), float(np.mean(d))

    # what training gains: the same network untrained, and the best predictor
    # without learning -- the training set's median tail slope and intercept
    # for every spectrum, e = 10^(mu/1e6 * hyp + beta/1e6) * cpsms
    eng.surrogate_init(seed=3)
    med0, _ = held_out_error()
    rt = plain_train.results
    okt = ((rt["flags"] & N.RES_VALID) != 0) & ((rt["flags"] & N.RES_FIT_UB) == 0)
    mu_c, beta_c = np.median(rt["mu"][okt]) / 1e6, np.median(rt["beta"][okt]) / 1e6
    rp = plain.results
    sel = ok.copy()
    sel[0] = False
    log_e_const = mu_c * rp["maxhypscore"][sel] + beta_c + np.log10(rp["cpsms"][sel].astype(np.float64))
    med_const = float(np.median(np.abs(log_e_const - np.log10(rp["e_x"][sel]))))

    loss = eng.surrogate_train(bt, steps=3000, batch_size=256, lr=2e-3)
    med, mean = held_out_error()
    log(f"  trained: final loss {loss:.4f}, held-out |dlog10 e| median {med:.4f} mean {mean:.4f}; "
        f"untrained {med0:.4f}, constant tail parameters {med_const:.4f}")
    # measured: untrained 1.7 (its standardised outputs sit near the training
    # means, i.e. the constant predictor), constant parameters 1.7, trained
    # 0.25-0.35: training cuts the error about five-fold
    assert med < 0.5 * med_const and med < 0.5 * med0
    # measured: loss ~0.5 (standardised).  The fit extrapolates its line from
    # the bulk of the score distribution to the top score, so a slope error is
    # amplified by that distance; the learned model gets e within a small
    # factor, not within 1e-3 (hence the audit)
    assert loss < 0.7 and med < 0.45

    # audit sample every 8th spectrum at 1e-3: audited spectra report the plain
    # fit; any disagreement sends the whole batch back to the plain fit
    out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE, audit_every=8, audit_tol=1e-3)
    a = eng.last_audit()
    log(f"  audit: {a}")
    r = out.results
    audited = (r["flags"] & N.RES_AUDITED) != 0
    assert a["audited"] == int(audited.sum()) and a["audited"] > 0
    assert np.all(np.arange(len(r))[audited] % 8 == 0)
    e_p = plain.results["e_x"]
    assert np.all(np.abs(r["e_x"][audited] - e_p[audited]) <= 1e-12 * np.abs(e_p[audited]))
    if a["fell_back"]:
        assert not np.any((r["flags"] & N.RES_SURROGATE) != 0)
        assert np.all(np.abs(r["e_x"][valid] - e_p[valid]) <= 1e-12 * np.abs(e_p[valid]))
        for f in ("mu", "beta", "minhypscore", "nexthypscore"):
            assert np.array_equal(r[f][valid], plain.results[f][valid]), f
    else:
        assert a["disagreed"] == 0
        assert np.all((r["flags"][valid & ~audited] & N.RES_SURROGATE) != 0)
    # with no tolerance at all the predictions stay (outliers reach e x 1e9)
    out = eng.query_spectrum(bs, kernel=N.KERNEL_BATCHED, evalue=N.EVAL_SURROGATE, audit_every=8, audit_tol=1e300)
    a = eng.last_audit()
    assert not a["fell_back"] and ((out.results["flags"] & N.RES_SURROGATE) != 0).sum() > 0

    # parameters round-trip through the host
    params, nrm = eng.surrogate_params()
    eng2 = Engine(prm)
    eng2.surrogate_set(params, nrm)
    p2, n2 = eng2.surrogate_params()
    assert np.array_equal(p2, params) and np.array_equal(n2, nrm)
    eng2.close()
    eng.close()
