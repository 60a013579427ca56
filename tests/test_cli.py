"""Drop-in command-line tools (anqi-implementation_amd/bin/): argument and
input validation give the reference tools' exit codes.  These cases fail
before any GPU work, so they run here; where the reference binaries are built
(oracle/_ref) the same invocation is run through them and the exit codes
compared."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from anqi_testlib import REPO, GOLDEN, ref_available

BIN = os.path.join(REPO, "anqi-implementation_amd", "bin")
REFBIN = os.path.join(REPO, "oracle", "_ref")


def run(tool, args, env=None, ref=False):
    e = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    e.update(env or {})
    exe = os.path.join(REFBIN if ref else BIN, tool)
    return subprocess.run([exe] + args, env=e, capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    base = np.random.default_rng(0).standard_normal((50, 8)).astype(np.float32)
    with open(d / "small_base.bin", "wb") as f:
        f.write(np.array([50, 8], np.uint32).tobytes() + base.tobytes())
    (d / "bad_lifted.meta").write_text("nonsense\n")
    (d / "ok_lifted.meta").write_text("0 128 100 l2 1.0\n")
    return d


def aknn_cases(tmp):
    return [([], 2), (["x", "y", "100", "500", "64", "600", "8", "10", "cosine"], 2),
            ([str(tmp / "missing.bin"), str(tmp / "o")], 3),
            ([str(tmp / "small_base.bin"), str(tmp / "o"), "100"], 3)]


def query_cases(tmp):
    return [([], 2, {}), (["16,x", str(tmp / "bad")], 2, {}), (["16", str(tmp / "missing")], 3, {}),
            (["16", str(tmp / "bad")], 3, {}), (["16", str(tmp / "ok"), "48", "500", "64", "600", "101"], 2, {}),
            (["16", str(tmp / "ok"), "48", "500", "64", "600", "10"], 3, {}),
            (["0", str(tmp / "ok")], 2, {})]


def test_build_aknn_source_exit_codes(tmp):
    for args, code in aknn_cases(tmp):
        r = run("build_aknn_source", args)
        assert r.returncode == code, (args, r.returncode, r.stderr)
        if ref_available() and os.path.exists(os.path.join(REFBIN, "build_aknn_source")):
            rr = run("build_aknn_source", args, ref=True)
            assert rr.returncode == code, (args, rr.returncode)


def test_rknn_query_exit_codes(tmp):
    for args, code, env in query_cases(tmp):
        r = run("rknn_query", args, env)
        assert r.returncode == code, (args, r.returncode, r.stderr)
        if ref_available() and os.path.exists(os.path.join(REFBIN, "rknn_query")):
            rr = run("rknn_query", args, env, ref=True)
            assert rr.returncode == code, (args, rr.returncode)


def gt_cases(tmp):
    b = str(tmp / "small_base.bin")
    return [([], 2), (["--base", b], 2), (["--base", b, "--out_prefix", str(tmp / "g"), "--mode", "x"], 2),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--bogus", "1"], 2),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--mode", "thresholds", "--K", "0"], 2),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--mode", "thresholds", "--K", "50"], 2),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--mode", "thresholds", "--K", "10", "--rk_k", "11"], 2),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--metric", "cos", "--query", b], 2),
            (["--base", str(tmp / "missing.bin"), "--out_prefix", str(tmp / "g"), "--mode", "thresholds"], 3),
            (["--base", b, "--out_prefix", str(tmp / "g"), "--query", str(tmp / "missing.bin"),
              "--mode", "ground-truth", "--K", "10", "--threshold-table", str(tmp / "missing_gt.bin")], 3)]


def test_dataset_gt_exit_codes(tmp):
    for args, code in gt_cases(tmp):
        r = run("dataset_gt", args)
        assert r.returncode == code, (args, r.returncode, r.stderr)
        if ref_available() and os.path.exists(os.path.join(REFBIN, "dataset_gt")):
            rr = run("dataset_gt", args, ref=True)
            assert rr.returncode == code, (args, rr.returncode)


def test_rknn_query_rejects_unpinned_configuration():
    """A configuration the GPU build does not implement is refused (exit 2)
    instead of silently computing the pinned one."""
    pfx = os.path.join(GOLDEN, "sample", "sample")
    for env in ({"ANQI_RADIUS_MODE": "approx"}, {"ANQI_VERIFIER_PLACEMENT": "integrated"},
                {"ANQI_RECHECK": "1"}, {"ANQI_RANKM": "4"}, {"ANQI_VAMANA": "0", "ANQI_VAMANA_L": "2000"},
                {"ANQI_RADIUS_RELAX": "0.5"}):
        r = run("rknn_query", ["16", pfx, "48", "500", "64", "600", "10"], env)
        assert r.returncode == 2, (env, r.returncode, r.stderr)
        assert "not supported" in r.stderr
    # ANQI_SPECIFIC is on when set at all; the reference refuses it with a
    # post-filter verifier (src/rknn_query.cpp:1806-1811)
    r = run("rknn_query", ["16", pfx, "48", "500", "64", "600", "10"], {"ANQI_SPECIFIC": "0"})
    assert r.returncode == 2 and "unset ANQI_SPECIFIC" in r.stderr


# scripts/run_final.sh's exports
PIN_ENV = {
    "ANQI_GRAPH_GEOMETRY": "anyk_lift", "ANQI_RADIUS_MODE": "rankm_resid_u8_lrq_floor",
    "ANQI_VERIFIER_PLACEMENT": "postfilter", "ANQI_RANKM": "8", "ANQI_NK": "100", "ANQI_GRAPH_NK": "100",
    "ANQI_RANKM_SLACK": "0", "ANQI_RECHECK": "0", "ANQI_RQ_OBJECTIVE": "radius_reconstruction",
    "ANQI_ALPHA": "1.2", "ANQI_LIFTNNK": "50", "ANQI_GRAPHNND": "6", "ANQI_BUILD_BREAKDOWN": "1",
    "ANQI_PREBUILT_TOPK_K": "100", "ANQI_BENCH_CALIBRATED_TIMING": "0", "OMP_NUM_THREADS": "2",
}
SAMPLE_FILES = ("_base.bin", "_query.bin", "_lifted.meta", "_baseknn_aknn.rec", "_radbasis_M8.bin",
                "_learned_rq_codebook_M8_u8.bin", "_rknn_gt_k10.bin", "_rknn_gt_k50.bin")


@pytest.mark.skipif(not os.path.exists(os.path.join(REFBIN, "rknn_query")), reason="reference tools not built here")
def test_cache_names_equal_reference(tmp_path):
    """The reference's rknn_query builds its caches for a copy of the sample;
    rknn_query here, with ANQI_REQUIRE_GRAPH_CACHE=1 on the same files, prints
    the same graph cache key, finds every cache file under the names it
    computes (state file and its key, .built marker, lifted vectors, lift meta,
    generate_config graph index) and loads the verifier artifact the reference
    wrote, so it passes the cache check and goes on to the GPU.  A rewritten
    .rec (new mtime/ctime in its stat identity) changes the key in both tools
    the same way, and both then refuse the required cache (exit 3)."""
    pfx = str(tmp_path / "sample")
    for suf in SAMPLE_FILES:
        shutil.copy(os.path.join(GOLDEN, "sample", "sample" + suf), pfx + suf)
    env = dict(PIN_ENV, ANQI_PREBUILT_TOPK_REC=pfx + "_baseknn_aknn.rec",
               ANQI_GRAPH_CACHE_DIR=str(tmp_path / "graph"), ANQI_VERIFIER_CACHE_DIR=str(tmp_path / "verifier"))
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    args = ["16", pfx, "48", "500", "64", "600", "10"]
    rr = run("rknn_query", args, env, ref=True)
    assert rr.returncode == 0, rr.stderr
    key = re.search(r"graph cache key=(\S+)", rr.stdout).group(1)
    ver = re.search(r"built independent compact verifier mode=\S+ bits=8 (\S+) ", rr.stdout).group(1)
    assert sum(f.startswith("fd_metric_l2_M_48_") for f in os.listdir(env["ANQI_GRAPH_CACHE_DIR"])) == 1
    req = dict(env, ANQI_REQUIRE_GRAPH_CACHE="1")
    r = run("rknn_query", args, req)
    assert f"[rknn] graph cache key={key}\n" in r.stdout
    assert f"loaded independent compact verifier mode=rank-M+LRQ bits=8 {ver} " in r.stdout
    assert "required graph cache" not in r.stderr
    assert r.returncode == 0 or (r.returncode == 3 and "GPU failure in context" in r.stderr), r.stderr
    st = os.stat(pfx + "_baseknn_aknn.rec")
    os.utime(pfx + "_baseknn_aknn.rec", (st.st_atime + 60, st.st_mtime + 60))
    keys = []
    for ref in (True, False):
        r = run("rknn_query", args, req, ref=ref)
        assert r.returncode == 3 and "required graph cache is incomplete or incompatible" in r.stderr, (ref, r.stderr)
        keys.append(re.search(r"graph cache key=(\S+)", r.stdout).group(1))
    assert keys[0] == keys[1] != key


def _u8_sample(tmp_path):
    """A copy of the sample whose base values are integers 0..255, as the
    fbin and as a uint8 file (ANQI_BASE_U8BIN, src/rknn_query.cpp:55-123)."""
    pfx = str(tmp_path / "sample")
    for suf in SAMPLE_FILES:
        shutil.copy(os.path.join(GOLDEN, "sample", "sample" + suf), pfx + suf)
    base = np.fromfile(pfx + "_base.bin", dtype=np.float32, offset=8).reshape(256, 128)
    q8 = np.clip(np.rint(base * 40 + 128), 0, 255).astype(np.uint8)
    with open(pfx + "_base.bin", "wb") as f:
        f.write(np.array([256, 128], np.uint32).tobytes() + q8.astype(np.float32).tobytes())
    with open(pfx + "_base.u8bin", "wb") as f:
        f.write(np.array([256, 128], np.uint32).tobytes() + q8.tobytes())
    return pfx


@pytest.mark.skipif(not os.path.exists(os.path.join(REFBIN, "rknn_query")), reason="reference tools not built here")
def test_u8_base_cache_key_and_bad_files_equal_reference(tmp_path):
    """With ANQI_BASE_U8BIN the reference keys the graph cache by the uint8
    file's identity and prints its load line; this tool prints the same key
    and line, finds the reference's caches, and gives the same exit codes for
    a truncated uint8 file (3) and for compact warm-start ids combined with a
    warm-start command (2)."""
    pfx = _u8_sample(tmp_path)
    env = dict(PIN_ENV, ANQI_PREBUILT_TOPK_REC=pfx + "_baseknn_aknn.rec", ANQI_BASE_U8BIN=pfx + "_base.u8bin",
               ANQI_GRAPH_CACHE_DIR=str(tmp_path / "graph"), ANQI_VERIFIER_CACHE_DIR=str(tmp_path / "verifier"))
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    args = ["16", pfx, "48", "500", "64", "600", "10"]
    rr = run("rknn_query", args, env, ref=True)
    assert rr.returncode == 0, rr.stderr
    key = re.search(r"graph cache key=(\S+)", rr.stdout).group(1)
    line = [l for l in rr.stdout.splitlines() if l.startswith("[rknn] loaded uint8 base directly as float")]
    r = run("rknn_query", args, dict(env, ANQI_REQUIRE_GRAPH_CACHE="1"))
    assert f"[rknn] graph cache key={key}\n" in r.stdout and line and line[0] in r.stdout.splitlines()
    assert "required graph cache" not in r.stderr
    assert r.returncode == 0 or (r.returncode == 3 and "GPU failure in context" in r.stderr), r.stderr
    # the key differs from the fbin run's (the uint8 file's identity is in it)
    r2 = run("rknn_query", args, dict(env, ANQI_BASE_U8BIN="", ANQI_REQUIRE_GRAPH_CACHE="1"))
    assert re.search(r"graph cache key=(\S+)", r2.stdout).group(1) != key
    with open(pfx + "_bad.u8bin", "wb") as f:
        f.write(np.array([256, 128], np.uint32).tobytes() + bytes(1000))
    for ref in (True, False):
        r = run("rknn_query", args, dict(env, ANQI_BASE_U8BIN=pfx + "_bad.u8bin"), ref=ref)
        assert r.returncode == 3 and "bad uint8 base" in r.stderr, (ref, r.returncode, r.stderr)
        r = run("rknn_query", args, dict(env, ANQI_COMPACT_PREBUILT_TOPK_IDS_INPLACE="1",
                                         ANQI_PREBUILT_WARM_START_COMMAND="cat " + pfx + "_baseknn_aknn.rec",
                                         ANQI_GRAPH_CACHE_DIR=str(tmp_path / "g2") + "/"), ref=ref)
        assert r.returncode == 2 and "cannot be combined" in r.stderr, (ref, r.returncode, r.stderr)


@pytest.mark.skipif(not os.path.exists(os.path.join(REFBIN, "rknn_query")), reason="reference tools not built here")
@pytest.mark.parametrize("extra", [
    {"ANQI_REUSE_ORIG_KNN": "1"},
    {"ANQI_VAMANA": "1"},
    {"ANQI_VAMANA": "", "ANQI_VAMANA_STYLE": "batch", "ANQI_VAMANA_L": "64", "ANQI_VAMANA_ENTRY": "medoid",
     "ANQI_REUSE_ORIG_KNN": "0"},
    {"ANQI_VAMANA": "1", "ANQI_VAMANA_STYLE": "old", "ANQI_VAMANA_PASS": "1", "ANQI_VAMANA_BATCH_FRAC": "0.1",
     "ANQI_VAMANA_EARLY_ALPHA": "1.05", "ANQI_QUERY_SINGLE_ENTRY": "1"},
    {"ANQI_EDGE_POLICY": "ball_hybrid", "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin", "ANQI_EDGE_KS": "5,50"},
    {"ANQI_EDGE_POLICY": "lsg_bubble_hybrid", "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin",
     "ANQI_LSG_MU_FILE": "{dir}/edge_mu.bin", "ANQI_LSG_ALPHA": "0.75", "ANQI_REUSE_ORIG_KNN": "1"},
    {"ANQI_EDGE_POLICY": "lsg_rng", "ANQI_LSG_MU_FILE": "{dir}/edge_mu.bin", "ANQI_VAMANA": "1"},
    {"ANQI_EDGE_POLICY": "rng"},
    {"ANQI_EDGE_POLICY": "bubble_auto", "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin",
     "ANQI_EDGE_AUTO_CANDIDATES": "24", "ANQI_EDGE_RESID_BUDGET": "0"},
    {"ANQI_EDGE_POLICY": "bubble_auto_full", "ANQI_EDGE_RESID_FILE": "{dir}/edge_resid.bin",
     "ANQI_EDGE_AUTO_MAX_SWAPS": "3", "ANQI_REUSE_ORIG_KNN": "1"},
    {"ANQI_GRAPH_RADIUS_SOURCE_PFX": "{golden}/sample"},
], ids=["reuse", "vamana", "vamana_batch", "vamana_old", "edge_ball", "edge_lsg_hybrid", "edge_lsg", "edge_rng",
        "edge_bubble_auto", "edge_bubble_auto_full", "radius_source_pfx"])
def test_variant_cache_names_equal_reference(tmp_path, extra):
    """The build variants (ANQI_REUSE_ORIG_KNN, ANQI_VAMANA* in both styles)
    key the graph cache as the reference does (src/rknn_query.cpp:301-341,
    include/index.h:1704-1770): after the reference's run, this tool finds all
    its caches under the same key and file names."""
    pfx = str(tmp_path / "sample")
    for suf in SAMPLE_FILES:
        shutil.copy(os.path.join(GOLDEN, "sample", "sample" + suf), pfx + suf)
    vdir = os.path.join(GOLDEN, "sample_variants")
    extra = {k: v.replace("{dir}", vdir).replace("{golden}", os.path.join(GOLDEN, "sample")) for k, v in extra.items()}
    env = dict(PIN_ENV, ANQI_PREBUILT_TOPK_REC=pfx + "_baseknn_aknn.rec",
               ANQI_GRAPH_CACHE_DIR=str(tmp_path / "graph"), ANQI_VERIFIER_CACHE_DIR=str(tmp_path / "verifier"),
               **extra)
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    args = ["16", pfx, "48", "500", "64", "600", "10"]
    rr = run("rknn_query", args, env, ref=True)
    assert rr.returncode == 0, rr.stderr
    key = re.search(r"graph cache key=(\S+)", rr.stdout).group(1)
    assert ("_vam1_style" in key) == ("ANQI_VAMANA" in extra) and ("_reuseorig1" in key) == ("ANQI_REUSE_ORIG_KNN" in extra)
    assert ("_edge" in key) == ("ANQI_EDGE_POLICY" in extra)
    r = run("rknn_query", args, dict(env, ANQI_REQUIRE_GRAPH_CACHE="1"))
    assert f"[rknn] graph cache key={key}\n" in r.stdout
    assert "required graph cache" not in r.stderr
    assert r.returncode == 0 or (r.returncode == 3 and "GPU failure in context" in r.stderr), r.stderr
