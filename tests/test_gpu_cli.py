"""Drop-in tools on the GPU against the reference's own outputs for the
sample inputs (tests/golden/sample): rknn_query given the reference's AKNN
source writes a verifier artifact byte-identical to the reference's, the same
lift (M*^2, scale) and, per L_descent, the reference's recall and precision;
build_aknn_source writes a valid .rec whose recall against the exact top-100
is not below the reference source's."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from anqi_testlib import GOLDEN, REPO, Sample, read_rec

pytestmark = pytest.mark.gpu
BIN = os.path.join(REPO, "anqi-implementation_amd", "bin")
PIN_ENV = {
    "ANQI_GRAPH_GEOMETRY": "anyk_lift", "ANQI_RADIUS_MODE": "rankm_resid_u8_lrq_floor",
    "ANQI_VERIFIER_PLACEMENT": "postfilter", "ANQI_RANKM": "8", "ANQI_NK": "100", "ANQI_GRAPH_NK": "100",
    "ANQI_RANKM_SLACK": "0", "ANQI_RECHECK": "0", "ANQI_RQ_OBJECTIVE": "radius_reconstruction",
    "ANQI_ALPHA": "1.2", "ANQI_LIFTNNK": "50", "ANQI_GRAPHNND": "6", "ANQI_BUILD_BREAKDOWN": "1",
    "ANQI_PREBUILT_TOPK_K": "100", "ANQI_BENCH_CALIBRATED_TIMING": "0",
}
POINT = re.compile(r"\[rknn\] L_descent=(\d+)\s+\| recall=([\d.]+) precision=([\d.]+) F1=([\d.]+) \| "
                   r"avg cand=(\d+) pred=(\d+) #dist=(\d+)")
LIFT = re.compile(r"BUILD lift M\*\^2=([\d.e+-]+) scale=([\d.e+-]+)")


def say(*a):
    print("[gpu-cli]", *a, flush=True)


def points(text):
    return {int(m.group(1)): tuple(float(g) for g in m.groups()[1:]) for m in POINT.finditer(text)}


def test_rknn_query_on_reference_source(tmp_path):
    pfx = os.path.join(GOLDEN, "sample", "sample")
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env.update(PIN_ENV)
    env["ANQI_PREBUILT_TOPK_REC"] = pfx + "_baseknn_aknn.rec"
    env["ANQI_GRAPH_CACHE_DIR"] = str(tmp_path / "graph")
    env["ANQI_VERIFIER_CACHE_DIR"] = str(tmp_path / "verifier")
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    say("rknn_query sample k=10")
    r = subprocess.run([os.path.join(BIN, "rknn_query"), "16,32,64,128", pfx, "48", "500", "64", "600", "10"],
                       env=env, capture_output=True, text=True, timeout=300)
    say(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr
    ref = open(pfx + "_rknn_stdout_k10.txt").read()
    mine, theirs = points(r.stdout), points(ref)
    assert sorted(mine) == sorted(theirs)
    for L in theirs:
        assert abs(mine[L][0] - theirs[L][0]) <= 0.005 and abs(mine[L][1] - theirs[L][1]) <= 0.005, L
    assert LIFT.search(r.stdout).groups() == LIFT.search(ref).groups()
    art = [f for f in os.listdir(env["ANQI_VERIFIER_CACHE_DIR"]) if f.endswith(".bin")]
    assert len(art) == 1
    a = open(os.path.join(env["ANQI_VERIFIER_CACHE_DIR"], art[0]), "rb").read()
    assert a == open(pfx + "_verifier.anqivf2", "rb").read()


def test_build_aknn_source_sample(tmp_path):
    S = Sample()
    out = str(tmp_path / "s")
    say("build_aknn_source sample")
    r = subprocess.run([os.path.join(BIN, "build_aknn_source"), S.prefix + "_base.bin", out, "100", "500", "64",
                        "600", "8", "10", "l2"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "[aknn] DONE n=256 d=128 K=100" in r.stdout
    ids, keys = read_rec(out + "_baseknn_aknn.rec", S.n)
    assert (keys >= 0).all() and (np.diff(keys, axis=1) >= 0).all()
    rec = np.mean([len(set(ids[i]) & set(S.gt_ids[i])) for i in range(S.n)]) / 100
    ref_rec = np.mean([len(set(S.rec_ids[i]) & set(S.gt_ids[i])) for i in range(S.n)]) / 100
    say("aknn recall", rec, "reference source", ref_rec)
    assert rec >= ref_rec - 0.005
    meta = open(out + "_aknn_build_time.meta").read().split("\n")
    assert meta[0] == "ANQI_AKNN_BUILD_V1" and meta[1].startswith("seconds ")


def test_dataset_gt_sample_files_equal_reference(tmp_path):
    """dataset_gt on the sample: the exact top-100 table and the RkNN CSR at
    k = 10, 50, 100 are byte-identical to the reference tool's files
    (tests/golden/sample, made with --tile 64); the ground-truth mode reads
    the threshold table the thresholds mode wrote."""
    pfx = os.path.join(GOLDEN, "sample", "sample")
    out = str(tmp_path / "g")
    tool = os.path.join(BIN, "dataset_gt")
    r = subprocess.run([tool, "--base", pfx + "_base.bin", "--K", "100", "--rk_k", "100", "--mode", "thresholds",
                        "--out_prefix", out], capture_output=True, text=True, timeout=300)
    say(r.stdout[-600:])
    assert r.returncode == 0, r.stderr
    assert open(out + "_baseknn_gt.bin", "rb").read() == open(pfx + "_baseknn_gt.bin", "rb").read()
    for k in (10, 50, 100):
        o = str(tmp_path / f"q{k}")
        r = subprocess.run([tool, "--base", pfx + "_base.bin", "--query", pfx + "_query.bin", "--K", "100",
                            "--rk_k", str(k), "--mode", "ground-truth", "--threshold-table",
                            out + "_baseknn_gt.bin", "--out_prefix", o], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert open(o + "_rknn_gt.bin", "rb").read() == open(pfx + f"_rknn_gt_k{k}.bin", "rb").read(), k



def answers(text):
    """The per-L answer lines (recall, precision, F1, candidates, distances),
    the QPS field cut off."""
    out = []
    for line in text.splitlines():
        if line.startswith("[rknn] L_descent="):
            out.append(line.split("| QPS")[0])
        elif line.startswith(("[rknn] DIST_BREAKDOWN", "[rknn] CANDIDATE_RECALL")):
            out.append(line)
    return out


def cli_env(tmp_path, name):
    pfx = os.path.join(GOLDEN, "sample", "sample")
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env.update(PIN_ENV)
    env["ANQI_PREBUILT_TOPK_REC"] = pfx + "_baseknn_aknn.rec"
    env["ANQI_GRAPH_CACHE_DIR"] = str(tmp_path / name / "graph")
    env["ANQI_VERIFIER_CACHE_DIR"] = str(tmp_path / name / "verifier")
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    return pfx, env


def rknn(pfx, env, k):
    say(f"rknn_query sample k={k} caches {env['ANQI_GRAPH_CACHE_DIR']}")
    r = subprocess.run([os.path.join(BIN, "rknn_query"), "16,32,64,128", pfx, "48", "500", "64", "600", str(k)],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


BUILD_LINES = ("original-space AKNN", "built independent compact verifier", "BUILD lift", "BUILD_TIME", "Index saved")


def test_rknn_query_second_run_loads_its_caches(tmp_path):
    """A run at k = 10 builds and writes the caches; a second run at k = 50
    loads the lift, graph and verifier (no build line) and answers exactly as
    a fresh build at k = 50 does (src/rknn_query.cpp:1871-1894, 2052-2077)."""
    pfx, env = cli_env(tmp_path, "shared")
    first = rknn(pfx, env, 10)
    assert all(s in first for s in BUILD_LINES)
    second = rknn(pfx, env, 50)
    say(second[-1500:])
    assert "[rknn] loaded cached index (M=48)" in second and "INDEX_LOAD_TIME" in second
    assert "loaded independent compact verifier" in second and "Loading index from:" in second
    assert not any(s in second for s in BUILD_LINES)
    _, fresh_env = cli_env(tmp_path, "fresh")
    fresh = rknn(pfx, fresh_env, 50)
    assert "BUILD_TIME" in fresh
    assert answers(second) == answers(fresh) and len(answers(fresh)) == 12


def test_rknn_query_loads_reference_caches(tmp_path):
    """The reference's own cache files for the sample (graph index, lifted
    vectors, lift meta and verifier artifact of the reference run that wrote
    tests/golden/sample) put under the names this tool computes: runs at
    k = 10, 50 and 100 load them and print the reference's lift line and, per
    L_descent, its recall, precision, F1, candidate and distance counts
    exactly."""
    pfx, env = cli_env(tmp_path, "ref")
    rknn(pfx, env, 10)
    g, v = env["ANQI_GRAPH_CACHE_DIR"], env["ANQI_VERIFIER_CACHE_DIR"]
    names = os.listdir(g)
    swap = {"fd_metric_l2_M_48_": "_graph.bin", "_approxlift_graph_": "_approxlift.bin",
            "_anqi_meta_graph_": "_graph_meta.txt"}
    for head, suf in swap.items():
        mine = [f for f in names if f.startswith(head)]
        assert len(mine) == 1, (head, names)
        shutil.copy(pfx + suf, os.path.join(g, mine[0]))
    ver = os.listdir(v)
    assert len(ver) == 1
    shutil.copy(pfx + "_verifier.anqivf2", os.path.join(v, ver[0]))
    for k in (10, 50, 100):
        out = rknn(pfx, env, k)
        ref = open(pfx + f"_rknn_stdout_k{k}.txt").read()
        assert not any(s in out for s in BUILD_LINES), k
        lift = re.search(r"M\*\^2=\S+ scale=\S+", ref.split("BUILD lift" if k == 10 else "loaded cached index")[1])
        assert f"[rknn] loaded cached index (M=48): {lift.group(0)}" in out, k
        assert answers(out) == answers(ref), (k, answers(out), answers(ref))


def test_ingestion_options_give_the_same_answers(tmp_path):
    """The reference's scale options change how the inputs reach the build,
    not the answers: the base read from a uint8 file (ANQI_BASE_U8BIN), the
    warm-start ids streamed from a command (ANQI_PREBUILT_WARM_START_COMMAND)
    or from the .rec compacted in place to ids and then removed
    (ANQI_COMPACT_PREBUILT_TOPK_IDS_INPLACE), the .rec removed after the
    preparation (ANQI_RELEASE_PREBUILT_TOPK_AFTER_PREP), the verifier and the
    base released while the graph is built and read back after
    (ANQI_RELEASE_*_DURING_GRAPH_INDEX; src/rknn_query.cpp:55-123, 988-1057,
    2101-2190, 2257-2320).  Every run prints the answer lines of the plain
    run on a base of integer values 0..255."""
    src = os.path.join(GOLDEN, "sample", "sample")
    pfx = str(tmp_path / "s")
    for suf in ("_query.bin", "_lifted.meta", "_baseknn_aknn.rec", "_radbasis_M8.bin",
                "_learned_rq_codebook_M8_u8.bin", "_rknn_gt_k10.bin"):
        shutil.copy(src + suf, pfx + suf)
    base = np.fromfile(src + "_base.bin", dtype=np.float32, offset=8).reshape(256, 128)
    q8 = np.clip(np.rint(base * 40 + 128), 0, 255).astype(np.uint8)
    with open(pfx + "_base.bin", "wb") as f:
        f.write(np.array([256, 128], np.uint32).tobytes() + q8.astype(np.float32).tobytes())
    with open(pfx + "_base.u8bin", "wb") as f:
        f.write(np.array([256, 128], np.uint32).tobytes() + q8.tobytes())

    def go(name, extra, rec=None):
        env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
        env.update(PIN_ENV)
        env["ANQI_PREBUILT_TOPK_REC"] = rec or pfx + "_baseknn_aknn.rec"
        env["ANQI_GRAPH_CACHE_DIR"] = str(tmp_path / name / "graph")
        env["ANQI_VERIFIER_CACHE_DIR"] = str(tmp_path / name / "verifier")
        os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
        os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
        env.update(extra)
        say("rknn_query", name)
        r = subprocess.run([os.path.join(BIN, "rknn_query"), "16,32,64,128", pfx, "48", "500", "64", "600", "10"],
                           env=env, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, (name, r.stderr[-2000:])
        return r.stdout

    plain = go("plain", {})
    want = answers(plain)
    assert len(want) == 12
    out = go("u8", {"ANQI_BASE_U8BIN": pfx + "_base.u8bin"})
    assert "[rknn] loaded uint8 base directly as float" in out and answers(out) == want
    out = go("stream", {"ANQI_PREBUILT_WARM_START_COMMAND": "cat " + pfx + "_baseknn_aknn.rec"})
    assert "deferred original KNN IDs to streamed graph warm-start" in out and answers(out) == want
    for name, flag, line in (("compact", "ANQI_COMPACT_PREBUILT_TOPK_IDS_INPLACE", "compacted regenerable prebuilt"),
                             ("release", "ANQI_RELEASE_PREBUILT_TOPK_AFTER_PREP", "released regenerable prebuilt")):
        rec = str(tmp_path / (name + ".rec"))
        shutil.copy(pfx + "_baseknn_aknn.rec", rec)
        out = go(name, {flag: "1"}, rec=rec)
        assert line in out and answers(out) == want and not os.path.exists(rec), name
    out = go("release_mem", {"ANQI_RELEASE_VERIFIER_DURING_GRAPH_INDEX": "1",
                             "ANQI_RELEASE_BASE_DURING_GRAPH_INDEX": "1"})
    assert "reloaded compact verifier after graph index construction" in out
    assert "reloaded original base after graph index construction" in out and answers(out) == want
    # a stream that ends early is refused
    env_bad = {"ANQI_PREBUILT_WARM_START_COMMAND": "head -c 1000 " + pfx + "_baseknn_aknn.rec"}
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env.update(PIN_ENV)
    env.update(env_bad)
    env["ANQI_PREBUILT_TOPK_REC"] = pfx + "_baseknn_aknn.rec"
    env["ANQI_GRAPH_CACHE_DIR"] = str(tmp_path / "bad" / "graph")
    env["ANQI_VERIFIER_CACHE_DIR"] = str(tmp_path / "bad" / "verifier")
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    r = subprocess.run([os.path.join(BIN, "rknn_query"), "16", pfx, "48", "500", "64", "600", "10"], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 3 and "streamed warm-start ended early" in r.stderr


@pytest.mark.parametrize("variant", ["reuse", "vamana_global", "vamana_batch", "single_entry", "edge_ball",
                                     "edge_cover_vamana", "edge_lsg", "edge_bubble_auto"])
def test_build_variants_equal_reference(tmp_path, variant):
    """Build variants of the reference (tests/golden/sample_variants, its own
    rknn_query on the sample): ANQI_REUSE_ORIG_KNN (no lifted LocalJoin,
    include/index.h:1202-1238), ANQI_VAMANA in both styles (:932-1063,
    1262-1282), ANQI_QUERY_SINGLE_ENTRY (:1606-1607) and ANQI_EDGE_POLICY
    (:283-829) over a residual table and lsg weights.  The graph written here
    equals the reference's link for link, and the answer lines are its."""
    import json
    from anqi_testlib import read_graph
    vdir = os.path.join(GOLDEN, "sample_variants")
    extra = {k: v.replace("{dir}", vdir) for k, v in json.load(open(os.path.join(vdir, "variants.json")))[variant].items()}
    src = os.path.join(GOLDEN, "sample", "sample")
    env = {k: v for k, v in os.environ.items() if not k.startswith("ANQI_")}
    env.update(PIN_ENV)
    env.update(extra)
    env["ANQI_PREBUILT_TOPK_REC"] = src + "_baseknn_aknn.rec"
    env["ANQI_GRAPH_CACHE_DIR"] = str(tmp_path / "graph")
    env["ANQI_VERIFIER_CACHE_DIR"] = str(tmp_path / "verifier")
    os.makedirs(env["ANQI_GRAPH_CACHE_DIR"])
    os.makedirs(env["ANQI_VERIFIER_CACHE_DIR"])
    say("rknn_query", variant)
    r = subprocess.run([os.path.join(BIN, "rknn_query"), "16,32,64,128", src, "48", "500", "64", "600", "10"],
                       env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = open(os.path.join(vdir, f"{variant}_stdout.txt")).read()
    assert answers(r.stdout) == answers(ref) and len(answers(ref)) == 12
    mine = [f for f in os.listdir(env["ANQI_GRAPH_CACHE_DIR"]) if f.startswith("fd_metric_l2_M_48")]
    assert len(mine) == 1
    d1, l1, _ = read_graph(os.path.join(env["ANQI_GRAPH_CACHE_DIR"], mine[0]), 256, 48)
    d2, l2, _ = read_graph(os.path.join(vdir, f"{variant}_graph.bin"), 256, 48)
    assert np.array_equal(d1, d2) and np.array_equal(l1, l2)
