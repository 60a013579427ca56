#!/usr/bin/env python3
"""Generate the reference golden fixtures under tests/golden/.

Runs in the development container only: it drives the reference's own tools
and harness, compiled from the reference sources by `make -C oracle ref`
into oracle/_ref/, and the reference verifier trainer script
(tools/train_final_m8_u8.py; its numpy eigh/quantile results are pinned by
nothing else).  Outputs are data only (inputs and expected outputs):

  sample/   the 256 x 128 N(0,1) sample (seeds 0/1), full reference pipeline:
            .rec (threads=1), exact top-100, RkNN CSR k=10/50/100, basis and
            codebook, graph cache + lifted file + graph meta, ANQIVF2
            verifier artifact, rknn_query stdout for k=10/50/100, per-query
            range-search candidate lists at L=16,32,64,128, materialized
            thresholds and postfilter decisions, distance bit vectors.
  trainer3k/ the reference trainer's radbasis and codebook for the synthetic
            3,000 x 100 cutoff matrix of anqi_testlib.trainer_cutoffs, with a
            2,000-row training sample (exercises the seeded row sampling and
            quantile interpolation that the 256-row sample cannot).
  mix20k/   first 20,000 rows of the million-row mixture + 200 queries:
            reference recall numbers only (AKNN recall vs exact, rknn_query
            recall/precision per L at k=10) -- the recall-parity targets.

  sample_ip/ a 1,000 x 32 N(0,1) sample under the inner-product metric: the reference's
            build_aknn_source / dataset_gt / rknn_query with metric ip, its
            lift, graph, verifier artifact and stdout, harness dumps (R2 per
            query, candidate lists, thresholds, post-filter decisions); the
            basis and codebook come from lrq_fit (pinned to the reference
            trainer by the sample and trainer3k fixtures).

Usage: python3 tests/golden/make_golden.py [--only sample|sample_ip|sample_variants|trainer|mix20k|mix1m]
       [--work DIR]
"""
from __future__ import annotations

import argparse
import ctypes
import glob
import json
import os
import shutil
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DIR = os.path.join(REPO, "oracle", "_ref")
REF_SRC = "/root/reference"
sys.path.insert(0, os.path.join(REPO, "anqi-implementation_amd"))
import workload  # noqa: E402

PIN_ENV = {
    "ANQI_GRAPH_GEOMETRY": "anyk_lift",
    "ANQI_RADIUS_MODE": "rankm_resid_u8_lrq_floor",
    "ANQI_VERIFIER_PLACEMENT": "postfilter",
    "ANQI_RANKM": "8",
    "ANQI_NK": "100",
    "ANQI_GRAPH_NK": "100",
    "ANQI_RANKM_SLACK": "0",
    "ANQI_RECHECK": "0",
    "ANQI_RQ_OBJECTIVE": "radius_reconstruction",
    "ANQI_ALPHA": "1.2",
    "ANQI_LIFTNNK": "50",
    "ANQI_GRAPHNND": "6",
    "ANQI_BUILD_BREAKDOWN": "1",
    "ANQI_PREBUILT_TOPK_K": "100",
    "ANQI_BENCH_CALIBRATED_TIMING": "0",
}


def run(cmd, env=None, log=None):
    e = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    if env:
        e.update(env)
    r = subprocess.run(cmd, env=e, capture_output=True, text=True)
    if log is not None:
        with open(log, "w") as f:
            f.write(r.stdout)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-4000:])
        raise SystemExit(f"command failed ({r.returncode}): {' '.join(cmd)}")
    return r.stdout


def fit_with_lrq_fit(pfx, n):
    """Basis and codebook files from this repository's trainer (lrq_fit, CPU,
    one torch thread), which tests/test_trainer_golden.py holds bit-exact to
    the reference trainer's files; used where the reference trainer was not
    run for the fixture."""
    import torch
    import lrq_fit
    ids, keys = read_rec(pfx + "_baseknn_aknn.rec", n)
    torch.set_num_threads(1)
    mean, basis, cb, _ = lrq_fit.fit(torch.from_numpy(np.ascontiguousarray(keys, np.float32)), log=lambda *a: None)
    lrq_fit.write_basis(pfx + "_radbasis_M8.bin", mean, basis)
    lrq_fit.write_codebook(pfx + "_learned_rq_codebook_M8_u8.bin", cb)


def pipeline(pfx, base, query, aknn_args, threads, L_list, ks, work, metric="l2", trainer="reference"):
    workload.write_fbin(pfx + "_base.bin", base)
    workload.write_fbin(pfx + "_query.bin", query)
    with open(pfx + "_lifted.meta", "w") as f:
        f.write(f"0 {base.shape[1]} 100 {metric} 1.0\n")
    omp = {"OMP_NUM_THREADS": str(threads)}
    run([os.path.join(REF_DIR, "build_aknn_source"), pfx + "_base.bin", pfx] + aknn_args, env=omp)
    thr = os.path.join(work, "thr")
    run([os.path.join(REF_DIR, "dataset_gt"), "--base", pfx + "_base.bin", "--K", "100",
         "--rk_k", "100", "--threads", str(threads), "--tile", "64", "--mode", "thresholds",
         "--metric", metric, "--out_prefix", thr], env=omp)
    for k in ks:
        t = os.path.join(work, f"gt{k}")
        run([os.path.join(REF_DIR, "dataset_gt"), "--base", pfx + "_base.bin", "--query",
             pfx + "_query.bin", "--K", "100", "--rk_k", str(k), "--threads", str(threads),
             "--tile", "64", "--mode", "ground-truth", "--metric", metric, "--threshold-table",
             thr + "_baseknn_gt.bin", "--out_prefix", t], env=omp)
        shutil.copy(t + "_rknn_gt.bin", f"{pfx}_rknn_gt_k{k}.bin")
    if trainer == "reference":
        run([sys.executable, os.path.join(REF_SRC, "tools", "train_final_m8_u8.py"), "--pfx", pfx,
             "--threshold-source", pfx + "_baseknn_aknn.rec", "--codebook-workers", "1"], env=omp)
    else:
        fit_with_lrq_fit(pfx, base.shape[0])
    runroot = os.path.join(work, "run")
    outs = {}
    for k in ks:
        env = dict(PIN_ENV)
        env.update(omp)
        env["ANQI_PREBUILT_TOPK_REC"] = pfx + "_baseknn_aknn.rec"
        env["ANQI_GRAPH_CACHE_DIR"] = os.path.join(runroot, "graph")
        env["ANQI_VERIFIER_CACHE_DIR"] = os.path.join(runroot, "verifier")
        env["ANQI_AKNN_BUILD_SECONDS"] = "0"
        env["ANQI_VERIFIER_TRAIN_SECONDS"] = "0"
        os.makedirs(env["ANQI_GRAPH_CACHE_DIR"], exist_ok=True)
        os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"], exist_ok=True)
        outs[k] = run([os.path.join(REF_DIR, "rknn_query"), ",".join(map(str, L_list)), pfx,
                       "48", "500", "64", "600", str(k)], env=env)
    return thr, runroot, outs


def parse_points(out):
    pts = []
    for line in out.splitlines():
        if line.startswith("[rknn] L_descent="):
            f = line.replace("|", " ").split()
            vals = {}
            for tok in f:
                if "=" in tok:
                    a, b = tok.split("=", 1)
                    vals[a] = b
            pts.append({"L": int(vals["[rknn] L_descent".split()[-1]] if False else line.split("=")[1].split()[0]),
                        "recall": float(vals["recall"]), "precision": float(vals["precision"]),
                        "F1": float(vals["F1"]), "cand": float(vals["cand"]),
                        "pred": float(vals["pred"]), "ndist": float(vals["#dist"])})
    return pts


def read_rec(path, n, k=100):
    r = np.fromfile(path, dtype=np.dtype([("id", "<u4"), ("key", "<f4")])).reshape(n, k)
    return r["id"].copy(), r["key"].copy()


def read_gt(path):
    with open(path, "rb") as f:
        n, K = struct.unpack("<II", f.read(8))
        ids = np.frombuffer(f.read(n * K * 4), dtype=np.uint32).reshape(n, K)
        keys = np.frombuffer(f.read(n * K * 4), dtype=np.float32).reshape(n, K)
    return ids, keys


def make_sample(out_dir):
    os.makedirs(out_dir, exist_ok=True)
    base, query = workload.sample_base(), workload.sample_query()
    n, d = base.shape
    L_list = [16, 32, 64, 128]
    with tempfile.TemporaryDirectory() as work:
        pfx = os.path.join(work, "sample")
        thr, runroot, outs = pipeline(pfx, base, query, ["100", "64", "8", "64", "1", "10", "l2"],
                                      1, L_list, [10, 50, 100], work)
        for suf in ["_base.bin", "_query.bin", "_baseknn_aknn.rec", "_radbasis_M8.bin",
                    "_learned_rq_codebook_M8_u8.bin", "_rknn_gt_k10.bin", "_rknn_gt_k50.bin",
                    "_rknn_gt_k100.bin", "_lifted.meta"]:
            shutil.copy(pfx + suf, os.path.join(out_dir, "sample" + suf))
        shutil.copy(thr + "_baseknn_gt.bin", os.path.join(out_dir, "sample_baseknn_gt.bin"))
        gdir = os.path.join(runroot, "graph")
        graph = [p for p in glob.glob(os.path.join(gdir, "fd_metric_l2_M_48*"))]
        lift = glob.glob(os.path.join(gdir, "_approxlift_graph_*.bin"))
        meta = glob.glob(os.path.join(gdir, "_anqi_meta_graph_*.txt"))
        ver = glob.glob(os.path.join(runroot, "verifier", "_anqi_verifier_*.bin"))
        assert len(graph) == 1 and len(lift) == 1 and len(meta) == 1 and len(ver) == 1, (graph, lift, meta, ver)
        shutil.copy(graph[0], os.path.join(out_dir, "sample_graph.bin"))
        shutil.copy(lift[0], os.path.join(out_dir, "sample_approxlift.bin"))
        shutil.copy(meta[0], os.path.join(out_dir, "sample_graph_meta.txt"))
        shutil.copy(ver[0], os.path.join(out_dir, "sample_verifier.anqivf2"))
        for k, o in outs.items():
            with open(os.path.join(out_dir, f"sample_rknn_stdout_k{k}.txt"), "w") as f:
                f.write("\n".join(l for l in o.splitlines() if l.startswith("[rknn]")) + "\n")
        points = {k: parse_points(o) for k, o in outs.items()}

        # ---- harness dumps on the reference graph ----
        H = ctypes.CDLL(os.path.join(REF_DIR, "libref_harness.so"))
        H.ref_index_open.restype = ctypes.c_void_p
        H.ref_index_open.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.c_uint64, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_char_p]
        H.ref_index_load_graph.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        H.ref_range_search.restype = ctypes.c_uint32
        H.ref_range_search.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint64,
                                       ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        H.ref_tree_entry.restype = ctypes.c_uint32
        H.ref_tree_entry.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        H.ref_recheck_key.restype = ctypes.c_float
        H.ref_recheck_key.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        H.ref_lrq_materialize.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_void_p]
        folder = os.path.join(work, "hx") + "/"
        os.makedirs(folder, exist_ok=True)
        h = H.ref_index_open(lift[0].encode(), n, d + 2, 48, 50, 100, 8.0, 1.2, folder.encode())
        rc = H.ref_index_load_graph(h, graph[0].encode())
        assert rc == 0, rc
        M2, scale = map(float, open(meta[0]).read().split())
        R2 = M2 + 1.0
        qb = np.zeros((len(query), d + 2), dtype=np.float32)
        qb[:, :d] = (query.astype(np.float64) * scale).astype(np.float32)
        qb[:, d] = 1.0
        cands = {}
        entries = np.array([H.ref_tree_entry(h, qb[i].ctypes.data) for i in range(len(query))], dtype=np.uint32)
        for L in L_list:
            offs, ids, nd = [0], [], []
            for i in range(len(query)):
                buf = np.zeros(n, dtype=np.uint32)
                ndist = ctypes.c_uint64(0)
                c = H.ref_range_search(h, qb[i].ctypes.data, R2, L, buf.ctypes.data, n, ctypes.byref(ndist))
                ids.extend(buf[:c].tolist())
                offs.append(len(ids))
                nd.append(ndist.value)
            cands[L] = (np.array(offs, dtype=np.uint32), np.array(ids, dtype=np.uint32), np.array(nd, dtype=np.uint64))
        # verifier artifact -> thresholds via the reference's materialize_k
        raw = open(ver[0], "rb").read()
        hdr = struct.unpack("<8sQIIIII", raw[:36])
        assert hdr[0] == b"ANQIVF2\0" and hdr[1] == n
        coeff = np.frombuffer(raw[36:36 + n * 8 * 4], dtype=np.float32)
        codes = np.frombuffer(raw[36 + n * 32: 36 + n * 32 + n * 100], dtype=np.uint8)
        rb = open(pfx + "_radbasis_M8.bin", "rb").read()
        mu = np.frombuffer(rb[8:8 + 400], dtype=np.float32)
        basis = np.frombuffer(rb[408:408 + 3200], dtype=np.float32)
        cb = np.frombuffer(open(pfx + "_learned_rq_codebook_M8_u8.bin", "rb").read()[24:], dtype=np.float32)
        thrs, dec = {}, {}
        for k in (10, 50, 100):
            t = np.zeros(n, dtype=np.float32)
            H.ref_lrq_materialize(n, 100, 8, mu.ctypes.data, basis.ctypes.data, cb.ctypes.data, 256,
                                  coeff.ctypes.data, codes.ctypes.data, k, 1.0, t.ctypes.data)
            thrs[k] = t
            o, ids, _ = cands[128]
            keys = np.array([H.ref_recheck_key(np.ascontiguousarray(base[ids[j]]).ctypes.data,
                                               np.ascontiguousarray(query[qi]).ctypes.data, d)
                             for qi in range(len(query)) for j in range(o[qi], o[qi + 1])], dtype=np.float32)
            dec[k] = (keys <= t[ids]).astype(np.uint8)
            if k == 10:
                pf_keys = keys
        np.savez_compressed(os.path.join(out_dir, "sample_harness.npz"),
                            entries=entries, R2=np.float64(R2), M2=np.float64(M2), scale=np.float64(scale),
                            qb=qb,
                            **{f"cand_off_L{L}": cands[L][0] for L in L_list},
                            **{f"cand_ids_L{L}": cands[L][1] for L in L_list},
                            **{f"ndist_L{L}": cands[L][2] for L in L_list},
                            **{f"thr_k{k}": thrs[k] for k in thrs},
                            **{f"dec_k{k}": dec[k] for k in dec},
                            postfilter_keys_L128=pf_keys)
    with open(os.path.join(out_dir, "sample_points.json"), "w") as f:
        json.dump({str(k): v for k, v in points.items()}, f, indent=1)
    print("[golden] sample fixtures written to", out_dir)


def make_sample_ip(out_dir):
    """A sample under the inner-product metric ("ip"), 1,000 x 32 N(0,1) (seed
    2) with 64 queries (seed 3): the reference's
    build_aknn_source and dataset_gt with metric ip, basis and codebook from
    lrq_fit on that .rec (the trainer is metric-blind and already pinned),
    then the reference rknn_query for k = 10, 50, 100 and harness dumps on its
    graph: per-query R2 and candidate lists, thresholds (materialize_k for
    ip), post-filter keys -<o, q> and decisions."""
    os.makedirs(out_dir, exist_ok=True)
    # 1,000 rows (n >> K = 100, so the flood does not cover the graph), d = 32
    base, query = workload.gaussian(1000, 32, 2), workload.gaussian(64, 32, 3)
    n, d = base.shape
    L_list = [16, 32, 64, 128]
    with tempfile.TemporaryDirectory() as work:
        pfx = os.path.join(work, "sample_ip")
        thr, runroot, outs = pipeline(pfx, base, query, ["100", "64", "8", "64", "1", "10", "ip"],
                                      1, L_list, [10, 50, 100], work, metric="ip", trainer="lrq_fit")
        for suf in ["_base.bin", "_query.bin", "_baseknn_aknn.rec", "_radbasis_M8.bin",
                    "_learned_rq_codebook_M8_u8.bin", "_rknn_gt_k10.bin", "_rknn_gt_k50.bin",
                    "_rknn_gt_k100.bin", "_lifted.meta"]:
            shutil.copy(pfx + suf, os.path.join(out_dir, "sample_ip" + suf))
        shutil.copy(thr + "_baseknn_gt.bin", os.path.join(out_dir, "sample_ip_baseknn_gt.bin"))
        gdir = os.path.join(runroot, "graph")
        graph = glob.glob(os.path.join(gdir, "fd_metric_l2_M_48*"))
        lift = glob.glob(os.path.join(gdir, "_approxlift_graph_*.bin"))
        meta = glob.glob(os.path.join(gdir, "_anqi_meta_graph_*.txt"))
        ver = glob.glob(os.path.join(runroot, "verifier", "_anqi_verifier_*.bin"))
        assert len(graph) == 1 and len(lift) == 1 and len(meta) == 1 and len(ver) == 1, (graph, lift, meta, ver)
        shutil.copy(graph[0], os.path.join(out_dir, "sample_ip_graph.bin"))
        shutil.copy(lift[0], os.path.join(out_dir, "sample_ip_approxlift.bin"))
        shutil.copy(meta[0], os.path.join(out_dir, "sample_ip_graph_meta.txt"))
        shutil.copy(ver[0], os.path.join(out_dir, "sample_ip_verifier.anqivf2"))
        for k, o in outs.items():
            with open(os.path.join(out_dir, f"sample_ip_rknn_stdout_k{k}.txt"), "w") as f:
                f.write("\n".join(l for l in o.splitlines() if l.startswith("[rknn]")) + "\n")
        H = ctypes.CDLL(os.path.join(REF_DIR, "libref_harness.so"))
        vp, u32, u64, f32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_double
        H.ref_index_open.restype = vp
        H.ref_index_open.argtypes = [ctypes.c_char_p, u64, u64, u64, u64, u64, f32, f32, ctypes.c_char_p]
        H.ref_index_load_graph.argtypes = [vp, ctypes.c_char_p]
        H.ref_range_search.restype = u32
        H.ref_range_search.argtypes = [vp, vp, f64, u64, vp, u32, vp]
        H.ref_tree_entry.restype = u32
        H.ref_tree_entry.argtypes = [vp, vp]
        H.ref_query_lift.restype = f64
        H.ref_query_lift.argtypes = [vp, u32, f64, f64, ctypes.c_int, vp]
        H.ref_recheck_key_ip.restype = f32
        H.ref_recheck_key_ip.argtypes = [vp, vp, u32]
        H.ref_lrq_materialize_metric.argtypes = [u64, u32, u32, vp, vp, vp, u32, vp, vp, u32, f64, ctypes.c_int, vp]
        folder = os.path.join(work, "hx") + "/"
        os.makedirs(folder, exist_ok=True)
        h = H.ref_index_open(lift[0].encode(), n, d + 2, 48, 50, 100, 8.0, 1.2, folder.encode())
        assert H.ref_index_load_graph(h, graph[0].encode()) == 0
        M2, scale = map(float, open(meta[0]).read().split())
        qb = np.zeros((len(query), d + 2), dtype=np.float32)
        R2 = np.array([H.ref_query_lift(np.ascontiguousarray(query[i]).ctypes.data, d, scale, M2, 1,
                                        qb[i].ctypes.data) for i in range(len(query))], dtype=np.float64)
        entries = np.array([H.ref_tree_entry(h, qb[i].ctypes.data) for i in range(len(query))], dtype=np.uint32)
        cands = {}
        for L in L_list:
            offs, ids, nd = [0], [], []
            for i in range(len(query)):
                buf = np.zeros(n, dtype=np.uint32)
                ndist = ctypes.c_uint64(0)
                c = H.ref_range_search(h, qb[i].ctypes.data, float(R2[i]), L, buf.ctypes.data, n, ctypes.byref(ndist))
                ids.extend(buf[:c].tolist())
                offs.append(len(ids))
                nd.append(ndist.value)
            cands[L] = (np.array(offs, dtype=np.uint32), np.array(ids, dtype=np.uint32), np.array(nd, dtype=np.uint64))
        raw = open(ver[0], "rb").read()
        hdr = struct.unpack("<8sQIIIII", raw[:36])
        assert hdr[0] == b"ANQIVF2\0" and hdr[1] == n
        coeff = np.frombuffer(raw[36:36 + n * 8 * 4], dtype=np.float32)
        codes = np.frombuffer(raw[36 + n * 32: 36 + n * 32 + n * 100], dtype=np.uint8)
        rb = open(pfx + "_radbasis_M8.bin", "rb").read()
        mu = np.frombuffer(rb[8:8 + 400], dtype=np.float32)
        basis = np.frombuffer(rb[408:408 + 3200], dtype=np.float32)
        cb = np.frombuffer(open(pfx + "_learned_rq_codebook_M8_u8.bin", "rb").read()[24:], dtype=np.float32)
        thrs, dec = {}, {}
        o, ids, _ = cands[128]
        keys = np.array([H.ref_recheck_key_ip(np.ascontiguousarray(base[ids[j]]).ctypes.data,
                                              np.ascontiguousarray(query[qi]).ctypes.data, d)
                         for qi in range(len(query)) for j in range(o[qi], o[qi + 1])], dtype=np.float32)
        for k in (10, 50, 100):
            t = np.zeros(n, dtype=np.float32)
            H.ref_lrq_materialize_metric(n, 100, 8, mu.ctypes.data, basis.ctypes.data, cb.ctypes.data, 256,
                                         coeff.ctypes.data, codes.ctypes.data, k, 1.0, 1, t.ctypes.data)
            thrs[k] = t
            dec[k] = (keys <= t[ids]).astype(np.uint8)
        np.savez_compressed(os.path.join(out_dir, "sample_ip_harness.npz"),
                            entries=entries, R2=R2, M2=np.float64(M2), scale=np.float64(scale), qb=qb,
                            **{f"cand_off_L{L}": cands[L][0] for L in L_list},
                            **{f"cand_ids_L{L}": cands[L][1] for L in L_list},
                            **{f"ndist_L{L}": cands[L][2] for L in L_list},
                            **{f"thr_k{k}": thrs[k] for k in thrs},
                            **{f"dec_k{k}": dec[k] for k in dec},
                            postfilter_keys_L128=keys)
        points = {k: parse_points(o) for k, o in outs.items()}
    with open(os.path.join(out_dir, "sample_ip_points.json"), "w") as f:
        json.dump({str(k): v for k, v in points.items()}, f, indent=1)
    print("[golden] sample_ip fixtures written to", out_dir)


# Build variants of the reference's rknn_query on the sample: without the
# lifted LocalJoin (ANQI_REUSE_ORIG_KNN, which leaves every build step
# deterministic whatever the thread count), with the Vamana passes of each
# style, and with every query starting at one entry (ANQI_QUERY_SINGLE_ENTRY).
SAMPLE_VARIANTS = {
    "reuse": {"ANQI_REUSE_ORIG_KNN": "1"},
    "vamana_global": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_VAMANA": "1", "ANQI_VAMANA_L": "64"},
    "vamana_batch": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_VAMANA": "1", "ANQI_VAMANA_STYLE": "batch",
                     "ANQI_VAMANA_L": "64", "ANQI_VAMANA_PASS": "3", "ANQI_VAMANA_ENTRY": "medoid"},
    "single_entry": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_QUERY_SINGLE_ENTRY": "1", "ANQI_VAMANA_ENTRY": "7"},
    # edge policies over a residual table and lsg weights written beside the
    # fixture ("{dir}" is the fixture directory)
    "edge_ball": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_EDGE_POLICY": "ball_hybrid",
                  "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin"},
    "edge_cover_vamana": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_EDGE_POLICY": "residual_cover", "ANQI_EDGE_KS": "5,50",
                          "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin", "ANQI_VAMANA": "1",
                          "ANQI_VAMANA_STYLE": "batch", "ANQI_VAMANA_L": "64"},
    "edge_lsg": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_EDGE_POLICY": "lsg_bubble_hybrid",
                 "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin", "ANQI_LSG_MU_FILE": "{dir}/edge_mu.bin"},
    "edge_bubble_auto": {"ANQI_REUSE_ORIG_KNN": "1", "ANQI_EDGE_POLICY": "bubble_auto", "ANQI_EDGE_AUTO_CANDIDATES": "24",
                         "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin"},
}


def write_edge_inputs(out_dir, n=256, nk=100):
    """A signed residual table on the scale of the sample's lifted distances,
    and lsg weights in [0.5, 2]."""
    rng = np.random.default_rng(3)
    (rng.standard_normal((n, nk)) * 2.0e4).astype(np.float32).tofile(os.path.join(out_dir, "edge_resid.bin"))
    rng.uniform(0.5, 2.0, n).astype(np.float32).tofile(os.path.join(out_dir, "edge_mu.bin"))


def make_sample_variants(out_dir, sample_dir):
    """For each variant: the reference's graph cache file and its [rknn]
    stdout lines at k = 10 on the sample fixture."""
    os.makedirs(out_dir, exist_ok=True)
    write_edge_inputs(out_dir)
    with tempfile.TemporaryDirectory() as work:
        pfx = os.path.join(work, "sample")
        for suf in ["_base.bin", "_query.bin", "_baseknn_aknn.rec", "_radbasis_M8.bin",
                    "_learned_rq_codebook_M8_u8.bin", "_rknn_gt_k10.bin", "_lifted.meta"]:
            shutil.copy(os.path.join(sample_dir, "sample" + suf), pfx + suf)
        for name, extra in SAMPLE_VARIANTS.items():
            extra = {k: v.replace("{dir}", out_dir) for k, v in extra.items()}
            env = dict(PIN_ENV, OMP_NUM_THREADS="4", ANQI_PREBUILT_TOPK_REC=pfx + "_baseknn_aknn.rec",
                       ANQI_AKNN_BUILD_SECONDS="0", ANQI_VERIFIER_TRAIN_SECONDS="0", **extra)
            env["ANQI_GRAPH_CACHE_DIR"] = os.path.join(work, name, "graph")
            env["ANQI_VERIFIER_CACHE_DIR"] = os.path.join(work, name, "verifier")
            os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
            os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
            out = run([os.path.join(REF_DIR, "rknn_query"), "16,32,64,128", pfx, "48", "500", "64", "600", "10"],
                      env=env)
            graph = glob.glob(os.path.join(env["ANQI_GRAPH_CACHE_DIR"], "fd_metric_l2_M_48*"))
            assert len(graph) == 1, graph
            shutil.copy(graph[0], os.path.join(out_dir, f"{name}_graph.bin"))
            with open(os.path.join(out_dir, f"{name}_stdout.txt"), "w") as f:
                f.write("\n".join(l for l in out.splitlines() if l.startswith("[rknn]")) + "\n")
    with open(os.path.join(out_dir, "variants.json"), "w") as f:
        json.dump(SAMPLE_VARIANTS, f, indent=1)
    print("[golden] sample build variants written to", out_dir)


def make_trainer(out_dir, n=3000, train_sample=2000):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from anqi_testlib import trainer_cutoffs
    os.makedirs(out_dir, exist_ok=True)
    keys = trainer_cutoffs(n)
    with tempfile.TemporaryDirectory() as work:
        pfx = os.path.join(work, "tr")
        workload.write_fbin(pfx + "_base.bin", np.zeros((n, 1), np.float32))   # the trainer reads n from it
        rec = np.zeros((n, 100), dtype=np.dtype([("id", "<u4"), ("key", "<f4")]))
        rec["id"] = (np.arange(n)[:, None] + 1 + np.arange(100)[None, :]) % n
        rec["key"] = keys
        rec.tofile(pfx + "_baseknn_aknn.rec")
        run([sys.executable, os.path.join(REF_SRC, "tools", "train_final_m8_u8.py"), "--pfx", pfx,
             "--train-sample", str(train_sample), "--codebook-workers", "1"])
        for suf in ("_radbasis_M8.bin", "_learned_rq_codebook_M8_u8.bin"):
            shutil.copy(pfx + suf, os.path.join(out_dir, "trainer3k" + suf))
    with open(os.path.join(out_dir, "trainer3k.json"), "w") as f:
        json.dump({"n": n, "train_sample": train_sample, "cutoffs": "anqi_testlib.trainer_cutoffs(3000, 77)",
                   "cutoffs_sha256": workload.sha256(keys)}, f, indent=1)
    print("[golden] trainer fixture written to", out_dir)


def make_mix(out_dir, n=20000, nq=200):
    os.makedirs(out_dir, exist_ok=True)
    base = workload.million_base(n)
    query = workload.million_query(nq)
    L_list = [16, 32, 64, 128]
    with tempfile.TemporaryDirectory() as work:
        pfx = os.path.join(work, "mix")
        thr, runroot, outs = pipeline(pfx, base, query, ["100", "500", "64", "600", "8", "10", "l2"],
                                      8, L_list, [10], work)
        ids, keys = read_rec(pfx + "_baseknn_aknn.rec", n)
        gids, gkeys = read_gt(thr + "_baseknn_gt.bin")
        hit = sum(len(np.intersect1d(ids[i], gids[i], assume_unique=True)) for i in range(n))
        res = {"n": n, "nq": nq, "aknn_recall_at_100": hit / (n * 100.0),
               "base_sha256": workload.sha256(base), "query_sha256": workload.sha256(query),
               "points_k10": parse_points(outs[10]),
               "rknn_gt_k10_total": int(struct.unpack("<I", open(pfx + "_rknn_gt_k10.bin", "rb").read(4))[0])}
        with open(os.path.join(out_dir, "mix20k_reference.json"), "w") as f:
            json.dump(res, f, indent=1)
        with open(os.path.join(out_dir, "mix20k_rknn_stdout_k10.txt"), "w") as f:
            f.write("\n".join(l for l in outs[10].splitlines() if l.startswith("[rknn]")) + "\n")
    print("[golden] mix20k reference numbers written to", out_dir, res["aknn_recall_at_100"])


def make_mix1m(out_dir, work_root=None):
    """The reference's own pipeline on the full 1,000,000 x 128 mixture
    (10,000 queries, k = 10): recall / precision / candidate / distance
    numbers per L, lift meta and build seconds.  About 30 minutes on 8 host
    threads; work_root must hold ~3 GB."""
    import re
    os.makedirs(out_dir, exist_ok=True)
    base = workload.million_base(1_000_000)
    query = workload.million_query(10_000)
    with tempfile.TemporaryDirectory(dir=work_root) as work:
        pfx = os.path.join(work, "mix")
        thr, runroot, outs = pipeline(pfx, base, query, ["100", "500", "64", "600", "8", "10", "l2"],
                                      8, [8, 16, 32, 64, 128, 256, 1024], [10], work)
        out = outs[10]
        pts = []
        for m in re.finditer(r"\[rknn\] L_descent=(\d+)\s+\| recall=([\d.]+) precision=([\d.]+) F1=([\d.]+) \| "
                             r"avg cand=(\d+) pred=(\d+) #dist=(\d+) \| QPS=(\d+)", out):
            L, r, p, f, c, pr, d, q = m.groups()
            pts.append(dict(L=int(L), recall=float(r), precision=float(p), F1=float(f), cand=int(c), pred=int(pr),
                            dist=int(d), qps_8_threads=int(q)))
        lift = re.search(r"BUILD lift M\*\^2=([\d.]+) scale=([\d.]+)", out).groups()
        meta = open(pfx + "_aknn_build_time.meta").read().split("\n")
        res = dict(n=1_000_000, nq=10_000, k=10, base_sha256=workload.sha256(base),
                   query_sha256=workload.sha256(query), aknn_args=["100", "500", "64", "600", "8", "10", "l2"],
                   threads=8,
                   aknn_build_seconds_8_threads=float([l for l in meta if l.startswith("seconds")][0].split()[1]),
                   lift_mstar2=float(lift[0]), lift_scale=float(lift[1]),
                   build_breakdown=re.search(r"BUILD_BREAKDOWN .*", out).group(0), points_k10=pts,
                   note="reference tools compiled from the reference sources (oracle/Makefile 'ref'), run by "
                        "pipeline() with the run_final.sh environment pins; 8 host threads")
        with open(os.path.join(out_dir, "mix1m_reference.json"), "w") as f:
            json.dump(res, f, indent=1)
    print("[golden] mix1m reference numbers written to", out_dir)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["sample", "sample_ip", "sample_variants", "trainer", "mix20k", "mix1m"])
    ap.add_argument("--work", default=None, help="scratch directory for the 1M run")
    a = ap.parse_args()
    if not os.path.exists(os.path.join(REF_DIR, "libref_harness.so")):
        raise SystemExit("build the reference first: make -C oracle ref")
    if a.only in (None, "sample"):
        make_sample(os.path.join(HERE, "sample"))
    if a.only in (None, "sample_ip"):
        make_sample_ip(os.path.join(HERE, "sample_ip"))
    if a.only in (None, "sample_variants"):
        make_sample_variants(os.path.join(HERE, "sample_variants"), os.path.join(HERE, "sample"))
    if a.only in (None, "trainer"):
        make_trainer(os.path.join(HERE, "trainer3k"))
    if a.only in (None, "mix20k"):
        make_mix(os.path.join(HERE, "mix20k"))
    if a.only == "mix1m":
        make_mix1m(os.path.join(HERE, "mix1m"), a.work)


if __name__ == "__main__":
    main()
