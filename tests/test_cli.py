"""Drop-in command-line tools (anqi-implementation_amd/bin/): argument and
input validation give the reference tools' exit codes.  These cases fail
before any GPU work, so they run here; where the reference binaries are built
(oracle/_ref) the same invocation is run through them and the exit codes
compared."""
import os
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
                {"ANQI_RECHECK": "1"}, {"ANQI_RANKM": "4"}, {"ANQI_REUSE_ORIG_KNN": "1"}):
        r = run("rknn_query", ["16", pfx, "48", "500", "64", "600", "10"], env)
        assert r.returncode == 2, (env, r.returncode, r.stderr)
        assert "not supported" in r.stderr
