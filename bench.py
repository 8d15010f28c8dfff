#!/usr/bin/env python3
"""Benchmark: open search (+-500 Da) over a 10M-peptide index on one MI355X.

One step = one full search of the query batch (scan + scoring + top hit +
tail-fit e-values) with the index and spectra already resident in HBM.
Prints one JSON line (see the fields below).

Algorithmic bytes of one search (query kernel):
    B_q = sum over spectra of 4*N_iA + 8*N_probe + 8*P + 8
N_iA: iA entries inside the precursor windows, N_probe: (length, peak, bin)
probes reading bA[bin], bA[bin+1], P: peaks (m/z + intensity), 8: the top hit.
roofline.achieved = B_q / (average query-kernel time, HIP events).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "spectra/s/GPU open search 10M peptides, plain vs batched+surrogate; % HBM BW"
HBM_PEAK = 8.0e12


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--preset", default="open10m")
    ap.add_argument("--npep", type=int, default=0)
    ap.add_argument("--nspec", type=int, default=0)
    ap.add_argument("--kernel", default="both", choices=["plain", "batched", "both"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0, help="peptides per index chunk (0: the library default)")
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="CPU work per timed oracle query run (three runs, median)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-only", action="store_true", help="the CPU baseline alone, no GPU")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched form with the library defaults instead of the autotuned tuning")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one batched search, then a 1 GiB device copy (run under rocprofv3 --pmc)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 --pmc runs of --pmc-probe into the traffic figure")
    ap.add_argument("--no-mirror", action="store_true",
                    help="skip the reference-shaped entry point leg (DSLIM_QuerySpectrum through the C++ driver)")
    ap.add_argument("--no-merge", action="store_true",
                    help="skip the shard merge leg (two cyclic index shards: export and CombineResults)")
    ap.add_argument("--pepdb-only", action="store_true",
                    help="time the peptide database stage (LBE + MODS) alone and print its JSON")
    args = ap.parse_args()

    if args.traffic:
        return traffic_from_pmc(args.traffic[0], args.traffic[1], args.preset)

    if args.cpu_only:
        return cpu_only(args)

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        return 2

    from hicops_amd import native as N
    from hicops_amd import workload as W
    from hicops_amd.search import Engine

    kw = {}
    if args.npep:
        kw["npep"] = args.npep
    if args.nspec:
        kw["nspec"] = args.nspec
    p = W.scaled(args.preset, **kw)
    t0 = time.perf_counter()
    db = W.Database(p)
    log(f"database: {db.nforms} forms over {len(db.lengths)} lengths ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    raw, q = W.make_raw_spectra(db, p)      # q: the same spectra normalised by the generator
    npeaks = int(q.idx[-1])
    log(f"queries: {q.nspec} spectra, {npeaks} peaks ({time.perf_counter() - t0:.1f}s)")
    prm = W.make_params(p)

    eng = Engine(prm, chunk=args.chunk)
    if args.pepdb_only:
        print(json.dumps({"preset": p.name, "pepdb": pepdb_stage(eng, db, p, prm, args.steps, args.warmup,
                                                                 not args.no_cpu)}), flush=True)
        eng.close()
        return 0
    t0 = time.perf_counter()
    idx = eng.construct_db(db)
    build_wall = time.perf_counter() - t0
    build_ms = sum(ix.build_ms for ix in idx)
    nions = sum(ix.nions for ix in idx)
    b_build = 4 * nions + 4 * (eng.NB + 1) * len(idx)
    log(f"index: {nions} ions, device build {build_ms:.1f} ms, wall {build_wall:.1f}s")
    batch = eng.upload(q)

    if args.pmc_probe:
        if not args.no_autotune:
            log(f"pmc probe autotune: {eng.autotune(batch, N.KERNEL_BATCHED)}")
        eng.query_device(batch, kernel=N.KERNEL_BATCHED)
        k, t = eng.query_device(batch, kernel=N.KERNEL_BATCHED)
        log(f"pmc probe: batched query kernel {k:.1f} ms")
        calibration_copy(1 << 30)
        eng.close()
        return 0

    # forms: plain (one workgroup per spectrum, plain tail fit); batched (the
    # persistent work-stealing kernel, autotuned, plain tail fit);
    # batched+surrogate (the same scan, e-values from the learned model with
    # the sampled audit and its fallback).  `value` is the last form run.
    # Each step is a separate search launch so that the scan kernel is timed
    # on its own (a graph replay, hcx_search_opts.use_graph, is timed whole).
    forms = {"plain": ["plain"], "batched": ["batched", "batched+surrogate"],
             "both": ["plain", "batched", "batched+surrogate"]}[args.kernel]
    runs = {}
    tuned = None
    sur_info = None
    for form in forms:
        opts = dict(kernel=N.KERNEL_PLAIN if form == "plain" else N.KERNEL_BATCHED, tile=args.tile)
        if form == "batched":
            if not args.no_autotune:
                t0 = time.perf_counter()
                tuned = eng.autotune(batch, N.KERNEL_BATCHED)
                log(f"autotune ({time.perf_counter() - t0:.1f}s): {tuned}")
        if form == "batched+surrogate":
            # trained on plain fits of a separate batch of 20 000 spectra
            t0 = time.perf_counter()
            qt = W.make_queries(db, p, nspec=min(20000, q.nspec), seed=p.seed ^ 0x7A11)
            bt = eng.upload(qt)
            eng.query_device(bt, kernel=N.KERNEL_BATCHED)
            eng.surrogate_init(seed=3)
            loss = eng.surrogate_train(bt, steps=3000, batch_size=256, lr=2e-3)
            bt.free()
            sur_info = {"train_spectra": qt.nspec, "train_steps": 3000, "train_loss": round(loss, 4),
                        "train_s": round(time.perf_counter() - t0, 2)}
            log(f"surrogate trained: {sur_info}")
            opts.update(evalue=N.EVAL_SURROGATE, audit_every=64, audit_tol=1e-3)
        for i in range(args.warmup):
            k, t = eng.query_device(batch, **opts)
            log(f"{form} warmup {i}: query kernel {k:.1f} ms, total {t:.1f} ms")
        kms, tms, ems, pms, tpm = [], [], [], [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            k, t = eng.query_device(batch, **opts)
            lt = eng.last_timing()
            kms.append(k)
            tms.append(t)
            ems.append(lt["evalue_ms"])
            pms.append(lt["packed_ms"])
            tpm.append(lt["two_pass_ms"])
            log(f"{form} step {i}: scan {k:.1f} ms (packed {pms[-1]:.1f} + two-pass {tpm[-1]:.1f}), "
                f"total {t:.1f} ms, e-values {ems[-1]:.2f} ms")
        wall = time.perf_counter() - t0
        out = eng.fetch(batch, want_hist=False, want_stats=True)
        runs[form] = dict(wall=wall, kms=float(np.mean(kms)), tms=float(np.mean(tms)), ems=float(np.mean(ems)),
                          pms=float(np.mean(pms)), tpm=float(np.mean(tpm)), out=out)
        if form == "batched+surrogate":
            sur_info["audit"] = eng.last_audit()

    head = runs[forms[-1]]
    st = head["out"].stats
    res = head["out"].results
    n_ia, n_probe = int(st["n_ia"].sum()), int(st["n_probe"].sum())
    b_q = 4 * n_ia + 8 * n_probe + 8 * npeaks + 8 * q.nspec
    # the dominant kernel: the packed scan over the spectra it scores (the
    # two-pass kernel takes the rest, timed separately)
    packed = (res["flags"] & N.RES_TWO_PASS) == 0
    pk_peaks = np.diff(q.idx.astype(np.int64))[packed].sum()
    b_pk = (4 * int(st["n_ia"][packed].sum()) + 8 * int(st["n_probe"][packed].sum()) + 8 * int(pk_peaks)
            + 8 * int(packed.sum()))
    ms_per_step = head["wall"] / args.steps * 1e3
    value = q.nspec * args.steps / head["wall"]
    dom_ms = head["pms"] if head["pms"] > 0 else head["kms"]
    dom_bytes = b_pk if head["pms"] > 0 else b_q
    achieved = dom_bytes / (dom_ms / 1e3)
    valid = int(((res["flags"] & N.RES_VALID) != 0).sum())
    tied = int(((res["flags"] & N.RES_TIE_REPLAYED) != 0).sum())

    # HBM bytes per launch of the dominant kernel from the committed counter
    # runs (bench.py --traffic), used only while the kernel sources they were
    # measured on are unchanged
    traffic, traffic_note = None, "no counter file for this workload"
    tname = "plain" if forms[-1] == "plain" else "batched"   # both batched forms run the same scan
    tfile = ROOT / "profiles" / f"traffic_{args.preset}_{tname}.json"
    if tfile.exists() and not kw:
        try:
            tj = json.loads(tfile.read_text())
            if tj.get("kernel_sources_sha") == kernel_sources_sha():
                traffic, traffic_note = tj.get("hbm_bytes_per_launch"), f"rocprofv3 --pmc, {tfile.name}"
            else:
                traffic_note = f"{tfile.name} was measured on other kernel sources: stale, not used"
        except Exception:  # noqa: BLE001
            traffic = None

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prm, db, q, st, args.cpu_seconds)
    pick = queries_pick(eng, raw, p, args.steps, args.warmup, not args.no_cpu)
    pdb = pepdb_stage(eng, db, p, prm, args.steps, args.warmup, not args.no_cpu)
    mirror = None if args.no_mirror else mirror_leg(db, q, prm, args.steps, args.warmup)
    merge = None if args.no_merge else merge_leg(eng, db, batch, prm, args.steps, args.warmup, not args.no_cpu)

    if sur_info is not None:
        sur_info["surrogate_evalue_ms"] = round(runs["batched+surrogate"]["ems"], 3)
        if "batched" in runs:
            sur_info["plain_fit_ms"] = round(runs["batched"]["ems"], 3)
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": (f"open search +-{p.dM:g} Da, {db.nforms} peptides (lengths {p.min_len}-{p.max_len}, "
                         f"fragment charge 1-{p.maxz}), {q.nspec} spectra ({p.peaks_min}-{p.peaks_max} peaks)"),
            "preset": p.name, "peptides": db.nforms, "spectra": q.nspec, "ions": nions,
            "form": forms[-1], "tile": args.tile, "tuning": eng.tuning(),
            "autotuned": tuned is not None,
            "evalue": "surrogate+audit" if forms[-1] == "batched+surrogate" else "tailfit",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic, "traffic_source": traffic_note,
            "kernel": DOM_KERNEL if head["pms"] > 0 else "scan", "algorithmic_bytes": dom_bytes,
            "kernel_ms": round(dom_ms, 3),
            "scan_stage": {"kernels": "packed + two-pass", "algorithmic_bytes": b_q, "ms": round(head["kms"], 3),
                           "frac": round(b_q / (head["kms"] / 1e3) / HBM_PEAK, 4)},
            "packed_spectra": int(packed.sum()),
        },
        "cpu_baseline": cpu,
        "forms": {f: {"spectra_per_s": round(q.nspec * args.steps / r["wall"], 2),
                      "scan_ms": round(r["kms"], 3), "packed_ms": round(r["pms"], 3),
                      "two_pass_ms": round(r["tpm"], 3), "search_ms": round(r["tms"], 3),
                      "evalue_ms": round(r["ems"], 3)}
                  for f, r in runs.items()},
        "surrogate": sur_info,
        "index_build": {"ions": nions, "device_ms": round(build_ms, 2),
                        "ions_per_s": round(nions / (build_ms / 1e3), 1),
                        "frac_hbm": round(b_build / (build_ms / 1e3) / HBM_PEAK, 4)},
        "queries_pick": pick,
        "pepdb": pdb,
        "mirror": mirror,
        "merge": merge,
        "psms_valid": valid,
        "tied_spectra": tied,
        "host_fits": int(((res["flags"] & N.RES_FIT_HOST) != 0).sum()),
        "work": {"n_ia": n_ia, "n_probe": n_probe, "peaks": npeaks},
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


def merge_leg(eng, db, batch, prm, steps: int, warmup: int, with_cpu: bool) -> dict:
    """The one-GPU part of the multi-rank path on the benchmark workload: the
    database dealt cyclically to G = 2 index shards (every 2nd pepEntry of each
    length, lbe_internal.cpp:75-103), the batch searched against each shard,
    each shard's partial result exported to host memory (hcx_part_export: the
    partRes records and u16 histogram slices a rank writes, one slot of hbins
    entries per spectrum), then CombineResults over the two (hcx_combine:
    the records and slices uploaded, merged, the merged tail fit, the fResult
    rows and keys back).  Timed: export per shard and the merge, host buffers
    on both sides (they stand for the reference's per-rank files)."""
    from hicops_amd import native as N

    class Shard:
        def __init__(self, g, G):
            self.g, self.G = g, G
            self.lengths = [L for L in db.lengths if len(db.entries(L)) > g]

        def seqs(self, L):
            return db.seqs(L)

        def entries(self, L):
            return np.ascontiguousarray(db.entries(L)[self.g::self.G])

    G = 2
    packs, slices, exp_ms = [], [], []
    for g in range(G):
        ix = eng.construct_db(Shard(g, G))
        eng.query_spectrum(batch, indices=ix, want_hist=False, want_stats=False, kernel=N.KERNEL_BATCHED)
        for ixx in ix:
            ixx.free()
        ts = []
        for _ in range(warmup + steps):
            t0 = time.perf_counter()
            pk, sl = eng.part_export(batch, qid0=0)
            ts.append(time.perf_counter() - t0)
        exp_ms.append(float(np.median(ts[warmup:])) * 1e3)
        packs.append(pk)
        slices.append(sl)
    ts = []
    for _ in range(warmup + steps):
        t0 = time.perf_counter()
        out, keys = eng.combine(packs, slices)
        ts.append(time.perf_counter() - t0)
    comb_ms = float(np.median(ts[warmup:])) * 1e3
    n = batch.nspec
    reported = int(np.count_nonzero(keys < G))   # keys[s] == G: not reported
    wire = sum(p.nbytes + s.nbytes for p, s in zip(packs, slices))
    log(f"merge: {G} shards, export {', '.join(f'{x:.1f}' for x in exp_ms)} ms, combine {comb_ms:.1f} ms "
        f"({wire / 1e9:.2f} GB of records), {reported} spectra reported")
    res = {"shards": G, "spectra": n, "slot": int(slices[0].shape[1]), "wire_bytes": int(wire),
           "export_ms": [round(x, 2) for x in exp_ms], "combine_ms": round(comb_ms, 2),
           "combine_spectra_per_s": round(n / (comb_ms / 1e3), 1), "reported": reported,
           "includes": "export: device search state to host records; combine: records from host, merge, "
                       "merged tail fit, rows and keys to host"}
    if with_cpu:
        from oracle import binding as O
        k = min(n, 20000)
        sub_p = [np.ascontiguousarray(p[:k]) for p in packs]
        sub_s = [np.ascontiguousarray(s[:k]) for s in slices]
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            out_o, keys_o = O.combine(prm, sub_p, sub_s, int(slices[0].shape[1]))
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        same = bool(np.array_equal(keys_o, keys[:k]) and np.array_equal(out_o, out[:k]))
        res["cpu_baseline"] = {"value": round(k / t, 1), "unit": "spectra/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's CombineResults over the first {k} spectra of the same "
                                         f"two shards' records; median of 3",
                               "same_rows_and_keys": same}
    return res


def mirror_leg(db, q, prm, steps: int, warmup: int) -> dict:
    """DSLIM_QuerySpectrum, the reference-shaped entry point, timed the way an
    application calls it: the C++ driver (tests/cpp/dslim_driver, linked to
    libhcx_dslim.so) builds each length's index with DSLIM_Construct from host
    pepEntries, then calls DSLIM_QuerySpectrum on the benchmark's batch from
    host memory, warmup + steps times.  Each call uploads the batch over PCIe,
    runs the batched search with tail-fit e-values, copies the results back
    and passes every PSM to the application's DFile_PrintScore (here a
    counter).  Wall time per call, from the driver's own clock."""
    import shutil
    import subprocess
    import tempfile
    from hicops_amd import native as N
    from hicops_amd import workload as W
    drv = ROOT / "tests" / "cpp" / "dslim_driver"
    if not drv.exists():
        return {"skipped": "tests/cpp/dslim_driver is not built"}
    d = Path(tempfile.mkdtemp(prefix="hcx_mirror_"))
    try:
        W.write_driver_case(d, db, q, prm, N.KERNEL_BATCHED, repeat=warmup + steps, quiet=1)
        t0 = time.perf_counter()
        r = subprocess.run([str(drv), str(d), "1"], capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
        if r.returncode:
            return {"error": f"driver exit {r.returncode}: {r.stderr[-400:]}"}
        lines = [x for x in (d / "timing.txt").read_text().splitlines() if x]
        ms = [float(x) for x in lines if not x.startswith("psms")]
        psms = int(next(x for x in lines if x.startswith("psms")).split()[1])
    finally:
        shutil.rmtree(d, ignore_errors=True)
    t = float(np.median(ms[warmup:])) if len(ms) > warmup else float(ms[-1])
    log(f"mirror: DSLIM_QuerySpectrum {t:.1f} ms per {q.nspec}-spectrum call ({psms} PSMs), driver wall {wall:.1f}s")
    return {"entry_point": "DSLIM_QuerySpectrum(Queries*, Index*, uint_t)", "kernel": "batched", "evalue": "tailfit",
            "spectra": q.nspec, "calls_ms": [round(x, 2) for x in ms], "ms_per_call": round(t, 2),
            "spectra_per_s": round(q.nspec / (t / 1e3), 1), "psms_per_call": psms,
            "includes": "Queries upload from host memory, search, results to host, PSM sink"}


def pepdb_stage(eng, db, p, prm, steps: int, warmup: int, with_cpu: bool) -> dict:
    """The peptide database stage (LBE_CountPeps + LBE_Initialize: mass
    filter, variant counts, entries, the exact std::sort order) for every
    length of the benchmark database, from sequence lines to sorted
    pepEntries in HBM.  Device time is the filter/count pass of each length
    plus one generate call over all lengths, with the lines in HBM; the
    lines' upload over PCIe is reported beside it.
    Algorithmic bytes: the sequence characters read once, 24-byte entries
    written once."""
    from hicops_amd import workload as W
    mc = W.modconditions(p)
    lines = {L: db.seqs(L) for L in db.lengths}

    def one():
        t0 = time.perf_counter()
        dbs = [eng.count_peps(L, lines[L], mc) for L in db.lengths]
        eng.initialize(dbs)
        wall = time.perf_counter() - t0
        um = sum(d.timing()[0] for d in dbs)
        cm = sum(d.timing()[1] for d in dbs)
        gm = dbs[0].timing()[2]
        forms = sum(d.lcltot for d in dbs)
        for d in dbs:
            d.free()
        return wall, cm, gm, forms, um

    for _ in range(warmup):
        one()
    rs = [one() for _ in range(steps)]
    wall = float(np.median([r[0] for r in rs]))
    cm = float(np.median([r[1] for r in rs]))
    gm = float(np.median([r[2] for r in rs]))
    um = float(np.median([r[4] for r in rs]))
    forms = rs[0][3]
    chars = sum(len(v) for v in lines.values())
    nbytes = chars + 24 * forms
    dev_ms = cm + gm
    log(f"pepdb: {forms} forms over {len(lines)} lengths; device {dev_ms:.2f} ms "
        f"(filter/count {cm:.2f} + generate/sort {gm:.2f}), upload {um:.2f} ms, wall {wall * 1e3:.1f} ms")
    out = {"forms": forms, "lengths": len(lines), "modconditions": mc, "device_ms": round(dev_ms, 3),
           "count_ms": round(cm, 3), "generate_ms": round(gm, 3), "upload_ms": round(um, 3),
           "with_upload_forms_per_s": round(forms / ((dev_ms + um) / 1e3), 1), "wall_ms": round(wall * 1e3, 2),
           "forms_per_s": round(forms / (dev_ms / 1e3), 1),
           "roofline": {"bound": "hbm", "achieved": round(nbytes / (dev_ms / 1e3) / 1e9, 2),
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(nbytes / (dev_ms / 1e3) / HBM_PEAK, 5), "algorithmic_bytes": nbytes}}
    if with_cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import binding as O
        threads = cpu_facts()["threads"]

        def orc(L):
            s = O.lbe_filter(lines[L], L, prm.min_mass, prm.max_mass)
            return len(O.lbe_generate(prm, s, L, mc)[0])

        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            with ThreadPoolExecutor(threads) as ex:
                n = sum(ex.map(orc, db.lengths))
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        assert n == forms
        out["cpu_baseline"] = {"value": round(forms / t, 1), "unit": "forms/s", "cores": threads, "kind": "port",
                               "sample": "the whole database: the oracle's LBE (filter, counts, variants, glibc "
                                         "qsort, std::sort) per length, lengths side by side; median of 3",
                               "runs_s": [round(x, 3) for x in ts]}
    return out


def queries_pick(eng, raw, p, steps: int, warmup: int, with_cpu: bool) -> dict:
    """The Queries producer: MSQuery::pickpeaks over the benchmark's raw
    spectra (before normalisation), raw batch resident in HBM.  Algorithmic
    bytes of the normalisation kernel per spectrum: 8*S raw peaks read +
    8*kept peaks written + 24 B of precursor fields read + 16 B written."""
    qalen = p.peaks_max
    rb = eng.upload_raw(raw)
    for _ in range(warmup):
        b, _k, _t = eng.pickpeaks_timed(rb, qalen, p.base_int, p.min_int)
        b.free()
    kms, tms = [], []
    t0 = time.perf_counter()
    for i in range(steps):
        b, k, t = eng.pickpeaks_timed(rb, qalen, p.base_int, p.min_int)
        kms.append(k)
        tms.append(t)
        if i < steps - 1:
            b.free()
    wall = time.perf_counter() - t0
    kept = int(b.export().idx[-1])
    b.free()
    rb.free()
    S = int(raw.idx[-1])
    nbytes = 8 * S + 8 * kept + 40 * raw.nspec
    km = float(np.mean(kms))
    log(f"pickpeaks: {raw.nspec} spectra, {S} raw -> {kept} kept peaks; kernel {km:.3f} ms, "
        f"pass {np.mean(tms):.3f} ms, wall {wall / steps * 1e3:.1f} ms per call")
    out = {"spectra_per_s": round(raw.nspec / (float(np.mean(tms)) / 1e3), 1),
           "kernel_ms": round(km, 4), "pass_ms": round(float(np.mean(tms)), 4),
           "call_ms": round(wall / steps * 1e3, 3), "raw_peaks": S, "kept_peaks": kept, "qalen": qalen,
           "roofline": {"bound": "hbm", "achieved": round(nbytes / (km / 1e3) / 1e9, 2),
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(nbytes / (km / 1e3) / HBM_PEAK, 4), "algorithmic_bytes": nbytes}}
    if with_cpu:
        from oracle import binding as O
        from hicops_amd import workload as W
        n = min(raw.nspec, 20000)
        sub = W.RawSpectra(raw.mz[:raw.idx[n]], raw.intensity[:raw.idx[n]], raw.idx[:n + 1].copy(),
                           raw.prec_mz[:n], raw.z[:n], raw.rtime[:n])
        t0 = time.perf_counter()
        O.pickpeaks(sub, qalen, p.base_int, p.min_int)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(n / dt, 1), "unit": "spectra/s", "cores": 1, "kind": "port",
                               "sample": f"first {n} spectra, oracle pickpeaks (serial, as extractbatch)"}
    return out


SCAN_KERNEL = "k_items_packed"
KERNEL_SOURCES = ["k_packed.hip", "k_scan_common.h", "hcx_internal.h"]


def kernel_sources_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        h.update((ROOT / "hicops_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:16]
DOM_KERNEL = SCAN_KERNEL
COPY_KERNEL = "copyBuffer"


def calibration_copy(nbytes: int) -> None:
    """A plain device-to-device copy of known size, for scaling FETCH_SIZE /
    WRITE_SIZE (they count in KiB per dispatch with a hardware-dependent
    granularity)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"calibration copy: {nbytes} bytes")


def traffic_from_pmc(fetch_csv: str, write_csv: str, preset: str) -> int:
    """HBM bytes per launch of the scan kernel from FETCH_SIZE and WRITE_SIZE
    (rocprofv3 counter_collection CSVs of a --pmc-probe run), each scaled by
    what the same counter reads for the 1 GiB calibration copy."""
    import csv

    def read(path, counter):
        scan, copy = [], []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            if SCAN_KERNEL in r["Kernel_Name"]:
                scan.append(float(r["Counter_Value"]))
            elif COPY_KERNEL in r["Kernel_Name"]:
                copy.append(float(r["Counter_Value"]))
        return scan, copy

    fs, fc = read(fetch_csv, "FETCH_SIZE")
    ws, wc = read(write_csv, "WRITE_SIZE")
    if not fs or not ws or not fc or not wc:
        print(json.dumps({"error": "scan or copy kernel missing from the counter files"}))
        return 1
    gib_kib = float(1 << 20)
    f_factor = gib_kib / max(fc)      # the copy reads 1 GiB
    w_factor = gib_kib / max(wc)      # and writes 1 GiB
    fetch = fs[-1] * 1024 * f_factor
    write = ws[-1] * 1024 * w_factor
    out = {"kernel": SCAN_KERNEL, "preset": preset, "fetch_kib_raw": fs[-1], "write_kib_raw": ws[-1],
           "copy_fetch_kib": max(fc), "copy_write_kib": max(wc), "fetch_factor": round(f_factor, 4),
           "write_factor": round(w_factor, 4), "fetch_bytes": round(fetch), "write_bytes": round(write),
           "hbm_bytes_per_launch": round(fetch + write), "kernel_sources_sha": kernel_sources_sha()}
    dst = ROOT / "profiles" / f"traffic_{preset}_batched.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out), flush=True)
    return 0


def cpu_facts() -> dict:
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    share = len(os.sched_getaffinity(0))
    threads = min(share, int(os.environ.get("OMP_NUM_THREADS", share) or share))
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": share, "threads": threads}


def cpu_baseline(prm, db, q, gpu_stats, seconds: float, runs: int = 3):
    """The oracle (a C restatement of the reference's DSLIM_Construct and
    query loop, OpenMP over lengths and over spectra) on the host cores of
    the machine that holds the GPU: index build ions/s over the whole database and query spectra/s
    over a bounded random sample of the same batch against that index, each
    the median of `runs` runs.  gpu_stats (or None): the GPU search's work
    counts, compared with the oracle's for the sample."""
    # thread placement as the reference's OpenMP runs are set up; read by
    # libgomp when the oracle library is first loaded
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import binding as O

    facts = cpu_facts()
    threads = facts["threads"]
    built, tb = None, []
    for r in range(runs):
        built, dt = O.index_build_all(prm, db, threads, out=built)
        tb.append(dt)
    nions = sum(len(a) for a, _ in built.values())
    t_build = float(np.median(tb))
    log(f"cpu index build: {nions} ions, {['%.2f' % t for t in tb]} s on {threads} threads")
    orc = O.OracleDB.from_built(prm, db, built)

    rng = np.random.default_rng(1)
    order = rng.permutation(q.nspec)
    probe = q.subset(np.sort(order[:64]))
    t0 = time.perf_counter()
    orc.search(probe, threads)
    rate = probe.nspec / max(time.perf_counter() - t0, 1e-6)
    n = int(min(q.nspec, max(64, rate * seconds)))
    sel = np.sort(order[:n])
    sample = q.subset(sel)
    ts = []
    for r in range(runs):
        t0 = time.perf_counter()
        _, _, st = orc.search(sample, threads)
        ts.append(time.perf_counter() - t0)
    t_q = float(np.median(ts))
    match = None
    if gpu_stats is not None:
        match = bool(np.array_equal(st["n_ia"], gpu_stats["n_ia"][sel]) and
                     np.array_equal(st["n_probe"], gpu_stats["n_probe"][sel]))
    log(f"cpu query: {n} spectra, {['%.2f' % t for t in ts]} s on {threads} threads; "
        f"work counts match GPU: {match}")
    return {"value": round(n / t_q, 2), "unit": "spectra/s", "cores": threads, "kind": "port",
            "sample": f"{n} random spectra of the same batch against the oracle's own index of the whole "
                      f"database, OpenMP over spectra; median of {runs}",
            "runs_s": [round(t, 3) for t in ts],
            "index_build": {"ions": nions, "ions_per_s": round(nions / t_build, 1), "median_s": round(t_build, 3),
                            "runs_s": [round(t, 3) for t in tb], "sample": "the whole database, OpenMP over lengths"},
            **facts, "omp_proc_bind": os.environ["OMP_PROC_BIND"], "omp_places": os.environ["OMP_PLACES"],
            "work_counts_match_gpu": match}


def cpu_only(args) -> int:
    """The CPU baseline alone (no GPU): the same synthetic database and
    spectra as the GPU run of this preset."""
    from hicops_amd import workload as W

    kw = {}
    if args.npep:
        kw["npep"] = args.npep
    if args.nspec:
        kw["nspec"] = args.nspec
    p = W.scaled(args.preset, **kw)
    db = W.Database(p)
    _raw, q = W.make_raw_spectra(db, p)
    prm = W.make_params(p)
    log(f"cpu-only: {db.nforms} forms, {q.nspec} spectra")
    cpu = cpu_baseline(prm, db, q, None, args.cpu_seconds)
    print(json.dumps({"preset": p.name, "peptides": db.nforms, "spectra": q.nspec, "cpu_baseline": cpu}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
