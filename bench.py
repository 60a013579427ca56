#!/usr/bin/env python3
"""Benchmark: RkNN QPS at k=10 at fixed recall on the 1,000,000 x 128 mixture
workload (1,000 centres, sigma 20, rounded and clipped to [0, 255]; 10,000
queries from the same centres), with the GPU build seconds of the index.

A step is one pass of the query path (lift, beam + flood range search over
the G48 lifted graph, LRQ postfilter) over the 10,000-query batch, with the
index, thresholds and queries resident in HBM.

Forms: "chosen" = bf16-MFMA build sweeps (exact top-K filter, RP-leaf) +
persistent work-queue query launch; "plain" = FP32 build sweeps + one
launch per batch with one block per query.  Both produce the same index
bit for bit (checked here), so recall is shared; build seconds and QPS are
measured for each.

Fixed recall: the beam width L is the smallest swept L whose recall against
the exact RkNN ground truth is >= 0.95; when no swept L reaches 0.95 (the
search-tree entry limits recall on this clustered workload, for the
reference as well), L is the smallest one within 0.5 points of the recall at
the largest swept L.  The reference's query loop on the same graph (cpu
baseline) returns the same candidates, so its recall at that L is the same.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(HERE, "anqi-implementation_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
TARGET_RECALL = 0.95


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def query_bytes(ix, bufs):
    """Algorithmic HBM bytes of one persistent-search batch: 4 d' per graph
    distance (4 (d + 1) with the split lifted layout: main row + lift term),
    4 (G+1) per expanded node, (4 d + 4) per verified candidate, 4 d per query."""
    d, dl, G = ix.base.shape[1], ix.lifted.shape[1], ix.links.shape[1]
    if ix.handle.lmain:
        dl = d + 1
    nd = bufs["nd"].double().sum().item()
    nx = bufs["nexp"].double().sum().item()
    nc = bufs["cnt"].double().sum().item()
    nq = bufs["cnt"].shape[0]
    return 4 * dl * nd + 4 * (G + 1) * nx + (4 * d + 4) * nc + 4 * d * nq


def layout_bytes(ix, bufs):
    """Bytes the index layout in use reads for the same batch: with byte rows
    d + 4 per graph distance (row bytes + lift term) and d + 8 per verified
    candidate (row bytes + norm + threshold); with sentinel adjacency rows 4 G
    per expanded node (no degree); otherwise the float-row figures."""
    h = ix.handle
    d, G = ix.base.shape[1], ix.links.shape[1]
    nd = bufs["nd"].double().sum().item()
    nx = bufs["nexp"].double().sum().item()
    nc = bufs["cnt"].double().sum().item()
    nq = bufs["cnt"].shape[0]
    br = bool(getattr(h, "byte_rows", False))
    per_d = (d + 4) if br else 4 * (d + 1)
    per_x = 4 * G if h.plinks else 4 * (G + 1)
    per_c = (d + 8) if br else (4 * d + 8)
    return per_d * nd + per_x * nx + per_c * nc + 4 * d * nq


# scripts/run_final.sh's pinned configuration for the reference's rknn_query
RUN_FINAL_ENV = {
    "ANQI_GRAPH_GEOMETRY": "anyk_lift", "ANQI_RADIUS_MODE": "rankm_resid_u8_lrq_floor",
    "ANQI_VERIFIER_PLACEMENT": "postfilter", "ANQI_RANKM": "8", "ANQI_NK": "100", "ANQI_GRAPH_NK": "100",
    "ANQI_RANKM_SLACK": "0", "ANQI_RECHECK": "0", "ANQI_RQ_OBJECTIVE": "radius_reconstruction",
    "ANQI_ALPHA": "1.2", "ANQI_LIFTNNK": "50", "ANQI_GRAPHNND": "6", "ANQI_PREBUILT_TOPK_K": "100",
    "ANQI_BUILD_BREAKDOWN": "1", "ANQI_BENCH_CALIBRATED_TIMING": "0",
}


def host_cpu():
    """Host processor as lscpu reports it, and the cores this process may use."""
    info = {"threads_available": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
                info["smt"] = "on" if int(v) > 1 else "off"
            elif k in ("Socket(s)", "Core(s) per socket", "CPU(s)"):
                info[k.lower().replace("(s)", "s").replace(" ", "_")] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError) as e:
        info["lscpu_error"] = repr(e)
    return info


def cpu_build_baseline(ctx, base_np, queries_np, n_sub, threads):
    """The reference's own build, compiled from its sources (oracle/_ref), on
    the first n_sub rows of the bench base at the reference's thread count
    (nndindex uses min(64, hardware threads)): build_aknn_source (the AKNN
    source), then rknn_query's index build (verifier encode, lift, graph and
    its cache files), each timed by the tool's own clock (the "[aknn] DONE
    elapsed_s" and "BUILD_BREAKDOWN index_total_s" lines,
    tools/build_aknn_source.cpp:157-161, src/rknn_query.cpp:2339, 2368-2375).
    The verifier files and the ground truth that rknn_query needs as inputs
    are made here and are not timed.  Beside it, this build's chosen form on
    the same rows."""
    import torch
    import formats
    import lrq_fit
    import pipeline
    refbin = os.path.join(HERE, "oracle", "_ref")
    aknn_bin, query_bin = os.path.join(refbin, "build_aknn_source"), os.path.join(refbin, "rknn_query")
    if not (os.access(aknn_bin, os.X_OK) and os.access(query_bin, os.X_OK)):
        return None
    sub = np.ascontiguousarray(base_np[:n_sub])
    q = np.ascontiguousarray(queries_np[:256])
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        pfx = os.path.join(td, "sub")
        formats.write_fbin(pfx + "_base.bin", sub)
        formats.write_fbin(pfx + "_query.bin", q)
        with open(pfx + "_lifted.meta", "w") as f:
            f.write("0 128 100 l2 1.0\n")
        r = subprocess.run([aknn_bin, pfx + "_base.bin", pfx, "100", "500", "64", "600", str(threads), "10", "l2"],
                           env=env, capture_output=True, text=True, timeout=900)
        m = re.search(r"elapsed_s=([0-9.eE+-]+)", r.stdout)
        if r.returncode != 0 or not m:
            log("reference build_aknn_source failed", r.returncode, r.stderr[-300:])
            return None
        aknn_s = float(m.group(1))
        _, keys = formats.read_rec(pfx + "_baseknn_aknn.rec", n_sub)
        mu, G, cb, _ = lrq_fit.fit(ctx.to_dev(np.maximum(keys, 0.0)), log=lambda *a: None,
                                   codebook_dtype=torch.float32)
        lrq_fit.write_basis(pfx + "_radbasis_M8.bin", mu, G)
        lrq_fit.write_codebook(pfx + "_learned_rq_codebook_M8_u8.bin", cb)
        bd = ctx.to_dev(sub)
        _, ek = ctx.exact_knn(bd, 10, precision=1)
        off, gids, gkeys = ctx.rknn_membership(bd, ctx.to_dev(q), ek[:, 9].contiguous())
        tot = int(off[-1])
        formats.write_csr(pfx + "_rknn_gt_k10.bin", off, gids.cpu().numpy().view(np.uint32)[:tot],
                          gkeys.cpu().numpy()[:tot])
        qenv = dict(env, **RUN_FINAL_ENV)
        qenv.update({"ANQI_PREBUILT_TOPK_REC": pfx + "_baseknn_aknn.rec",
                     "ANQI_GRAPH_CACHE_DIR": td + "/graph", "ANQI_VERIFIER_CACHE_DIR": td + "/verifier"})
        os.makedirs(td + "/graph")
        os.makedirs(td + "/verifier")
        r = subprocess.run([query_bin, "64", pfx, "48", "500", "64", "600", "10"], env=qenv, capture_output=True,
                           text=True, timeout=900)
        m = re.search(r"BUILD_BREAKDOWN .*index_total_s=([0-9.eE+-]+)", r.stdout)
        if r.returncode != 0 or not m:
            log("reference rknn_query failed", r.returncode, r.stderr[-300:])
            return None
        index_s = float(m.group(1))
    gix = pipeline.build(ctx, sub, precision=1, log=lambda *a: None)
    gt = gix.times
    ctx.tree_free(gix.tree)
    del gix
    gpu_no_train = gt["build_s"] - gt["verifier_train_s"]
    return {"value": aknn_s + index_s, "unit": "s", "cores": threads, "kind": "reference",
            "sample": f"first {n_sub} rows of the bench base; build_aknn_source 100 500 64 600 {threads} 10 l2, "
                      f"then rknn_query's index build under scripts/run_final.sh's pins (verifier training "
                      f"is a separate tool in the reference and is not included)",
            "aknn_source_s": aknn_s, "index_s": index_s,
            "gpu_chosen_build_s_same_rows": gt["build_s"],
            "gpu_chosen_build_s_same_rows_without_verifier_training": gpu_no_train,
            "ratio_reference_over_gpu": (aknn_s + index_s) / gpu_no_train}


def cpu_baseline(ctx, ix, queries_np, thr, L, seconds_target=12.0):
    """The reference's own query loop (range_search + postfilter, compiled
    from the reference sources) on this host's cores, over the same graph."""
    import pipeline
    sys.path.insert(0, os.path.join(HERE, "tests"))
    from anqi_testlib import ref, ref_available
    if not ref_available():
        return None
    R = ref()
    R.ref_rknn_batch.restype = ctypes.c_double
    R.ref_rknn_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                 ctypes.POINTER(ctypes.c_uint64)]
    import formats
    cores = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        lp = os.path.join(td, "lifted.bin")
        gp = os.path.join(td, "graph.bin")
        formats.write_fbin(lp, ix.lifted.cpu().numpy())
        pipeline.export_reference_graph(ctx, ix, gp)
        n, dl = ix.lifted.shape
        h = R.ref_index_open(lp.encode(), n, dl, ix.links.shape[1], 50, 100, 8.0, 1.2, (td + "/").encode())
        rc = R.ref_index_load_graph(h, gp.encode())
        if rc != 0:
            log("reference graph load failed", rc)
            return None
        base = np.ascontiguousarray(ix.base_np)
        q = np.ascontiguousarray(queries_np)
        thr_h = np.ascontiguousarray(thr.cpu().numpy())
        c, p, dd = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        nq = q.shape[0]
        R2 = ix.mstar2 + 1.0
        probe = min(nq, 1000)
        s = R.ref_rknn_batch(h, base.ctypes.data, base.shape[1], q.ctypes.data, probe, ix.scale, R2, L,
                             thr_h.ctypes.data, cores, 1, ctypes.byref(c), ctypes.byref(p), ctypes.byref(dd))
        # per-query check on the same graph: candidate / accepted / distance
        # counts, and the tree ties (|margin| < EPS) where the reference draws
        # a random side from its thread's engine and this build goes left
        det = {k: np.zeros(nq, np.uint32) for k in ("cnt", "pred", "nd", "entry", "ties")}
        R.ref_rknn_detail.restype = None
        R.ref_rknn_detail(h, base.ctypes.data, base.shape[1], q.ctypes.data, nq, ctypes.c_double(ix.scale),
                          ctypes.c_double(R2), ctypes.c_uint64(L), thr_h.ctypes.data, cores,
                          *(det[k].ctypes.data for k in ("cnt", "pred", "nd", "entry", "ties")))
        reps = max(1, int(seconds_target / max(s * nq / probe, 1e-3)))
        s = R.ref_rknn_batch(h, base.ctypes.data, base.shape[1], q.ctypes.data, nq, ix.scale, R2, L,
                             thr_h.ctypes.data, cores, reps, ctypes.byref(c), ctypes.byref(p), ctypes.byref(dd))
        R.ref_index_close(h)
        tot = nq * reps
        return {"value": nq * reps / s, "unit": "queries/s", "cores": cores, "kind": "reference",
                "_totals": {"ref_avg_cand": c.value / tot, "ref_avg_pred": p.value / tot,
                            "ref_avg_dist": dd.value / tot},
                "_detail": det,
                "sample": f"{nq} queries x {reps} passes ({s:.1f} s), reference range_search + postfilter "
                          f"(rknn_query.cpp:2705-2776, ground-truth evaluation excluded) on the GPU-built graph, "
                          f"L={L}, {cores} OpenMP threads"}


def index_digest(ix):
    """Digest of the built index (graph rows, lift meta, verifier codes)."""
    import hashlib
    h = hashlib.sha256()
    for t in (ix.links, ix.deg, ix.codes, ix.coeff, ix.lifted):
        h.update(t.cpu().numpy().tobytes())
    h.update(np.float64([ix.mstar2, ix.scale]).tobytes())
    h.update(np.uint32([ix.tree.nodes]).tobytes())
    return h.hexdigest()[:16]


def load_traffic(kernel, workload_tag):
    """HBM bytes per launch measured with rocprofv3 PMC counters for this
    kernel and workload (profiles/traffic_<kernel>_<workload>.json), or None."""
    path = os.path.join(HERE, "profiles", f"traffic_{kernel}_{workload_tag}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def stage_build(ctx, a):
    """The chosen-form index build of the bench base (precision 1: bf16 MFMA
    leaf sweep, batched gathers), after a warm-up build of 20,000 rows, with
    each stage's host time and the gathers' work counted on the device: the
    algorithmic bytes of a gather are 4 d per distance computed plus 4 per
    adjacency id read (d = 128 for the AKNN gather, d' = 130 for the graph
    gather)."""
    import pipeline
    import workload
    base = workload.million_base(a.n)
    w = pipeline.build(ctx, base[:20000], precision=1, log=lambda *x: None)
    ctx.tree_free(w.tree)
    del w
    ix = pipeline.build(ctx, base, precision=1, log=log)
    t = ix.times
    d = base.shape[1]
    ad, gd = t["aknn_detail"], t["graph_detail"]
    out = {"n": a.n, "times": {k: v for k, v in t.items() if not isinstance(v, dict)}, "aknn": ad, "graph": gd}
    for name, det, dim in (("aknn_gather", ad, d), ("graph_gather", gd, d + 2)):
        b = 4.0 * dim * det["gather_d2"] + 4.0 * det["gather_adj"]
        out[name] = {"distances": det["gather_d2"], "adjacency_ids": det["gather_adj"], "algorithmic_bytes": b,
                     "per_row_distances": det["gather_d2"] / a.n, "per_row_adjacency_ids": det["gather_adj"] / a.n,
                     "stage_s": det["gather_s"], "stage_GBps": b / det["gather_s"] / 1e9}
    ctx.tree_free(ix.tree)
    return out


def stage_exact(ctx, a):
    """The exact self-excluded top-K of the bench base (tools/dataset_gt.cpp:
    206-245: the bf16 MFMA sweep with FP32 re-scoring, precision 1) for
    K = --exact-k, after a 20,000-row warm-up; host time around the
    synchronised call and the algorithmic work 2 d n (n - 1) flop."""
    import torch
    import workload
    base = workload.million_base(a.n)
    xd = ctx.to_dev(base)
    ctx.exact_knn(xd[:20000].contiguous(), a.exact_k, precision=1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ctx.exact_knn(xd, a.exact_k, precision=1)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t
    n, d = base.shape
    flop = 2.0 * d * n * (n - 1)
    return {"n": n, "d": d, "K": a.exact_k, "seconds": secs, "flop": flop, "TFLOPs": flop / secs / 1e12,
            "frac_bf16_peak": flop / secs / 2.5e15}


def stage_lrq(ctx, a):
    """Verifier kernels on the 1M x 128 mixture's AKNN cutoffs: encode
    (coefficients + codes) and the threshold decode for k = 10 and for
    k = 10, 50, 100 in one pass, HIP-event timed on torch's stream, with
    algorithmic bytes (encode: 400 B cutoffs read, 32 B coefficients and
    2 x 100 B codes written per row; decode: 32 B coefficients + 1 code byte
    read and 4 B written per row and rank, the coefficients once)."""
    import torch
    import lrq_fit
    import workload
    if a.n <= 1_000_000:
        base = workload.million_base(a.n)
        x = ctx.to_dev(base)
        ids, keys, _ = ctx.aknn_build(x, K=100, leaf=500, trees=64, ef=600, precision=1)
        cut, _ = ctx.aknn_cutoffs(keys)
        del ids, keys, x
        source = "AKNN cutoffs of the 1M x 128 mixture (first n rows)"
    else:
        # larger n: radius-like synthetic rows (ascending, 2e4..1e5), seed 5
        g = torch.Generator(device="cuda:0").manual_seed(5)
        cut = (torch.rand((a.n, 1), device="cuda:0", generator=g) * 4e4 + 2e4 +
               torch.cumsum(torch.rand((a.n, 100), device="cuda:0", generator=g) * 400, dim=1)).float().contiguous()
        source = "synthetic radius-like rows (torch generator seed 5)"
    mu, G, cb, _ = lrq_fit.fit(cut[:1_000_000].contiguous(), log=lambda *z: None, codebook_dtype=torch.float32)
    mu, G, cb = ctx.to_dev(mu), ctx.to_dev(G), ctx.to_dev(cb)
    n, nk = cut.shape
    reps = a.steps

    def timed(fn):
        """Mean GPU time of one call: events around each call (host launch
        gaps between calls are not counted; the call's counter reset is)."""
        fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for e0, e1 in ev:
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        return sum(e0.elapsed_time(e1) for e0, e1 in ev) / reps

    out = {"stage": "lrq", "n": n, "nk": nk, "reps": reps, "cutoffs": source}
    res = {}

    def enc():
        res["enc"] = ctx.lrq_encode(cut, mu, G, cb)

    ms = timed(enc)
    coeff, codes, codes_km = res["enc"]
    byts = n * (4 * nk + 32 + 2 * nk)
    out["encode"] = {"ms": ms, "bytes": byts, "GBps": byts / ms / 1e6, "frac": byts / ms / 1e6 / HBM_PEAK_GBS,
                     "soft_fallbacks": ctx.lrq_fallbacks()[0]}
    for ks in ((10,), (10, 50, 100)):
        thr = ctx.empty((len(ks), n), torch.float32)
        ms = timed(lambda: ctx.lrq_materialize_multi(coeff, codes_km, mu, G, cb, list(ks), out=thr))
        byts = n * (32 + 5 * len(ks))
        out["decode_k" + "_".join(map(str, ks))] = {
            "ms": ms, "bytes": byts, "GBps": byts / ms / 1e6, "frac": byts / ms / 1e6 / HBM_PEAK_GBS,
            "soft_fallbacks": ctx.lrq_fallbacks()[1]}
    # the all-soft binary64 form on the same inputs, for the record
    ctx.set_lrq_form(1)
    try:
        ms = timed(enc)
        same = all(torch.equal(u, v) for u, v in zip(res["enc"], (coeff, codes, codes_km)))
        out["encode_soft_form"] = {"ms": ms, "same_output": bool(same)}
        thr1 = ctx.empty((1, n), torch.float32)
        ms = timed(lambda: ctx.lrq_materialize_multi(coeff, codes_km, mu, G, cb, [10], out=thr1))
        out["decode_k10_soft_form"] = {"ms": ms}
    finally:
        ctx.set_lrq_form(0)
    thr_c = ctx.lrq_materialize_multi(coeff, codes_km, mu, G, cb, [10])
    out["decode_k10_soft_form"]["same_output"] = bool(torch.equal(thr_c, thr1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--L", type=int, default=0, help="fixed beam width (0: recall sweep)")
    ap.add_argument("--form", choices=["chosen", "plain"], default="chosen", help="form reported as value")
    ap.add_argument("--no-plain-build", action="store_true", help="skip the plain-form build timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-build-n", type=int, default=100_000,
                    help="rows of the bench base the reference's own build is timed on")
    ap.add_argument("--detail", default="")
    ap.add_argument("--exact-k", type=int, default=100, help="K of --stage exact")
    ap.add_argument("--stage", choices=["lrq", "build", "exact"], default="",
                    help="time one stage's kernels instead of the headline measurement (prints its own JSON)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed region, one 1 GiB device-to-device copy (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported by this build "
              f"(the query path is data-parallel over replicated indexes; no multi-GPU run exists)",
              file=sys.stderr)
        sys.exit(2)
    os.environ.setdefault("ANQI_PROGRESS", "1")
    import torch
    import pipeline
    import pyanqi
    import workload
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    ctx = pyanqi.Context(0)
    if a.stage:
        print(json.dumps({"lrq": stage_lrq, "build": stage_build, "exact": stage_exact}[a.stage](ctx, a)),
              flush=True)
        return
    t = time.perf_counter()
    base = workload.million_base(a.n)
    queries = workload.million_query(a.nq)
    sha = workload.sha256(base)
    log(f"data n={a.n} nq={a.nq} in {time.perf_counter() - t:.1f}s sha256 base={sha[:16]}")
    # one-time device/library initialisation (code objects, solver handles)
    # outside the timed builds
    wix = pipeline.build(ctx, base[:20000], precision=1, log=lambda *x: None)
    ctx.tree_free(wix.tree)
    del wix
    plain_build = None
    plain_digest = None
    if not a.no_plain_build:
        ixp = pipeline.build(ctx, base, precision=0, log=log)
        plain_build = ixp.times
        plain_digest = index_digest(ixp)
        ctx.tree_free(ixp.tree)
        del ixp
        torch.cuda.empty_cache()
        log("plain build", json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in plain_build.items()}))
    ix = pipeline.build(ctx, base, precision=1, log=log)
    chosen_digest = index_digest(ix)
    log("chosen build", json.dumps({k: (round(v, 3) if isinstance(v, float) else v) for k, v in ix.times.items()}))
    if plain_digest is not None and plain_digest != chosen_digest:
        raise SystemExit(f"plain and chosen builds differ: {plain_digest} vs {chosen_digest}")
    qd = ctx.to_dev(queries)
    t = time.perf_counter()
    off, gids, _ = pipeline.ground_truth(ctx, ix.base, qd, a.k)
    log(f"ground truth k={a.k}: {len(gids)} pairs in {time.perf_counter() - t:.1f}s")
    thr = pipeline.thresholds(ctx, ix, a.k)
    out_cap = 4096
    bufs = ctx.search_buffers(a.nq, out_cap)
    sweep = []
    Ls = [a.L] if a.L else [8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256]
    for L in Ls:
        ctx.rknn_search(ix.handle, qd, L, thr, mode=1, bufs=bufs)
        m = pipeline.metrics(bufs, off, gids)
        m["L"] = L
        sweep.append(m)
        log("sweep", json.dumps({k: round(v, 4) if isinstance(v, float) else v for k, v in m.items()}))
    if a.L:
        chosen_L, rule = a.L, "fixed by --L"
    elif any(m["recall"] >= TARGET_RECALL for m in sweep):
        chosen_L = min(m["L"] for m in sweep if m["recall"] >= TARGET_RECALL)
        rule = f"smallest L with recall >= {TARGET_RECALL}"
    else:
        best = sweep[-1]["recall"]
        chosen_L = min(m["L"] for m in sweep if m["recall"] >= best - 0.005)
        rule = f"recall {TARGET_RECALL} unreachable; smallest L within 0.5 pt of recall {best:.4f} at L={sweep[-1]['L']}"
    point = [m for m in sweep if m["L"] == chosen_L][0]
    ctx.rknn_search(ix.handle, qd, chosen_L, thr, mode=1, bufs=bufs)
    B = query_bytes(ix, bufs)
    B_layout = layout_bytes(ix, bufs)
    # chosen form: persistent launch with autotuned work areas and grid
    tune, tune_ms = ctx.search_autotune(ix.handle, thr, qd, None, None, chosen_L)
    ctx.set_search_tuning(tune)
    tuning = {"waves_per_cu": tune.waves_per_cu, "visited_slots": tune.visited_slots,
              "result_cap": tune.result_cap, "autotune_best_ms": tune_ms}
    log("autotune", json.dumps(tuning))

    def timed(mode):
        """Whole-call throughput (host clock around synchronised calls: the
        batch's entry descent and ordering, the search, the status copy) and
        the search kernel's own duration (HIP events the library records on
        its stream around that launch)."""
        for _ in range(a.warmup):
            ctx.rknn_search(ix.handle, qd, chosen_L, thr, mode=mode, bufs=bufs)
        torch.cuda.synchronize()
        ctx.time_kernels(True)
        kms = 0.0
        t0 = time.perf_counter()
        for _ in range(a.steps):
            ctx.rknn_search(ix.handle, qd, chosen_L, thr, mode=mode, bufs=bufs)
            kms += ctx.last_kernel_ms()
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        ctx.time_kernels(False)
        return a.nq * a.steps / secs, secs, kms / a.steps

    qps_c, secs_c, launch_c = timed(1)
    qps_p, secs_p, launch_p = timed(0)
    if a.calib_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst
    if a.form == "chosen":
        qps, secs, launch_ms, mode, kern = qps_c, secs_c, launch_c, 1, "k_range_search_persistent"
    else:
        qps, secs, launch_ms, mode, kern = qps_p, secs_p, launch_p, 0, "k_range_search"
    achieved = B / (launch_ms * 1e-3) / 1e9
    tag = f"mix{a.n // 1000}k-L{chosen_L}" if a.n % 1_000_000 else f"mix{a.n // 1_000_000}m-L{chosen_L}"
    tr = load_traffic(kern, tag)
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(ctx, ix, queries, thr, chosen_L)
            if cpu is not None:
                det = cpu.pop("_detail")
                gc, gp, gn = (bufs[k].cpu().numpy().astype(np.int64) for k in ("cnt", "pred", "nd"))
                diff = (gc != det["cnt"]) | (gp != det["pred"]) | (gn != det["nd"])
                cpu["same_graph_check"] = {
                    "gpu_avg_cand": float(gc.mean()), "gpu_avg_pred": float(gp.mean()),
                    "gpu_avg_dist": float((gn + gc).mean()),
                    **cpu.pop("_totals"),
                    "queries_differing": int(diff.sum()),
                    "queries_differing_with_tree_tie": int((diff & (det["ties"] > 0)).sum()),
                    "queries_with_tree_tie": int((det["ties"] > 0).sum())}
        except Exception as e:   # reported, never fatal for the GPU measurement
            log("cpu baseline failed:", repr(e))
        try:
            threads = min(64, len(os.sched_getaffinity(0)))
            t = time.perf_counter()
            bb = cpu_build_baseline(ctx, base, queries, min(a.cpu_build_n, a.n), threads)
            log(f"reference build baseline in {time.perf_counter() - t:.1f}s:", json.dumps(bb))
            if bb is not None:
                cpu = cpu if cpu is not None else {}
                cpu["build"] = bb
        except Exception as e:
            log("cpu build baseline failed:", repr(e))
        if cpu is not None:
            cpu["host"] = host_cpu()
    build_c = ix.times["build_s"]
    other = {"chosen_query_qps": qps_c, "plain_query_qps": qps_p, "query_qps_ratio": qps_c / qps_p,
             "chosen_build_s": build_c, "chosen_search_tuning": tuning}
    if plain_build is not None:
        other.update({"plain_build_s": plain_build["build_s"], "build_s_ratio": build_c / plain_build["build_s"],
                      "plain_build_breakdown_s": {k: plain_build[k] for k in ("aknn_source_s", "verifier_train_s",
                                                                             "encode_s", "lift_s", "graph_s")},
                      "index_digest": chosen_digest, "plain_index_identical": plain_digest == chosen_digest})
    res = {
        "metric": "RkNN QPS @k=10 at fixed recall (chosen vs plain GPU form); 1Mx128 build s",
        "value": qps, "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": secs / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"gaussian-mixture {a.n}x128 (1000 centres, sigma 20, rounded, clip 0..255), "
                               f"{a.nq} queries, k={a.k}, AKNN 100/500/64/600, G48, LocalJoin 6, M8+u8 LRQ",
                   "base_sha256": sha, "form": a.form, "L": chosen_L, "L_rule": rule,
                   "recall": point["recall"], "precision": point["precision"], "F1": point["F1"],
                   "avg_cand": point["cand"], "avg_dist": point["dist"], "build_s": build_c,
                   "build_breakdown_s": {k: ix.times[k] for k in ("aknn_source_s", "verifier_train_s", "encode_s",
                                                                  "lift_s", "graph_s")},
                   **other},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes_per_launch"] if tr else None,
                     "kernel": kern, "launch_ms": launch_ms, "bytes_per_launch": B,
                     # the plain form reads float rows, links and degrees
                     "layout": ({"byte_rows": bool(getattr(ix.handle, "byte_rows", False)),
                                 "sentinel_adjacency": bool(ix.handle.plinks), "bytes_per_launch": B_layout}
                                if mode == 1 else {"byte_rows": False, "sentinel_adjacency": False,
                                                   "bytes_per_launch": B}),
                     **({"traffic_source": tr["source"]} if tr else {})},
        "cpu_baseline": cpu,
    }
    if a.detail:
        with open(a.detail, "w") as f:
            json.dump({"result": res, "sweep": sweep, "times": ix.times, "plain_times": plain_build}, f, indent=1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
