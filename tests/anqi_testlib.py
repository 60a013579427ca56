"""Shared test helpers: ctypes bindings for the oracle (oracle/liboracle.so),
the reference harness (oracle/_ref/libref_harness.so, present only where the
reference tree was available to build it) and readers for the fixture files.
"""
from __future__ import annotations

import ctypes
import os
import struct
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
PKG = os.path.join(REPO, "anqi-implementation_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
F32 = ctypes.c_float
F64 = ctypes.c_double
I32 = ctypes.c_int


def ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def _load(path, sigs):
    lib = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


ORACLE_SIGS = {
    "or_ip": (F32, [P, P, U32]),
    "or_l2": (F32, [P, P, U32]),
    "or_norm": (F32, [P, U32]),
    "or_d2": (F32, [P, P, F32, F32, U32]),
    "or_norms": (None, [P, U64, U32, P]),
    "or_exact_topk": (I32, [P, U32, U32, U32, P, P, I32]),
    "or_rknn_membership": (U64, [P, U32, P, U32, U32, P, P, P, P, I32]),
    "or_heap_push": (I32, [P, P, U32, U32, F32]),
    "or_tree_entry": (U32, [P, P]),
    "or_range_search": (U32, [P, U32, U32, P, P, U32, U32, P, F64, U32, P, U32, P]),
    "or_lift_query": (None, [P, U32, F64, P]),
    "or_recheck_key": (F32, [P, P, U32]),
    "or_lrq_encode_value": (U32, [F64, P, U32]),
    "or_lrq_encode": (None, [P, U64, U32, U32, P, P, P, U32, P, P]),
    "or_lrq_materialize": (None, [P, P, U64, U32, U32, P, P, P, U32, U32, F64, P]),
    "or_lift": (None, [P, U64, U32, P, P, P, P]),
    "or_lift_metric": (None, [P, U64, U32, P, I32, P, P, P]),
    "or_query_lift": (F64, [P, U32, F64, F64, I32, P]),
    "or_recheck_key_ip": (F32, [P, P, U32]),
    "or_ipdist": (F32, [P, P, U32]),
    "or_lrq_materialize_metric": (None, [P, P, U64, U32, U32, P, P, P, U32, U32, F64, I32, P]),
    "or_exact_topk_metric": (I32, [P, U32, U32, U32, I32, P, P, I32]),
    "or_rknn_membership_metric": (U64, [P, U32, P, U32, U32, I32, P, P, P, P, I32]),
    "or_warm_start": (None, [P, P, U32, U32, P, U32, U32, P, P]),
    "or_nndescent_seq": (I32, [P, P, U32, U32, U32, P, P, I32, I32, P]),
    "or_nndescent_snapshot": (I32, [P, P, U32, U32, U32, P, P, I32, I32, P]),
    "or_reverse_csr": (None, [P, U32, U32, P, P]),
    "or_gather": (U32, [P, P, U32, U32, P, P, U32, P, P, U32, U32, U32, P, P, U32]),
    "or_pool_topk": (U32, [U32, P, P, U32, U32, P, P]),
    "or_prune": (U32, [P, P, U32, U32, U32, P, P, U32, U32, F32, P]),
    "or_commit": (None, [P, P, U32, U32, P, P, U32, F32, P, P]),
    "or_sort_pairs": (None, [P, P, U32]),
    "or_mt_stream": (None, [U64, P, U64]),
    "or_rp_tree_build": (U32, [P, U32, U32, U32, U32, U64, U32, P, P, P, P, P, P, P]),
    "or_vamana_collect": (U32, [P, P, U32, U32, P, P, U32, U32, U32, U32, P, P, U32]),
    "or_vamana_global_pass": (None, [P, P, U32, U32, P, P, U32, U32, U32, F32, P, P]),
    "or_vamana_batch": (None, [P, P, U32, U32, P, U32, U32, I32, F64, U64, F64, F32, F32, I32, U32, U64, U32,
                               P, P]),
    "or_vamana_medoid": (U32, [P, U32, U32, U64]),
    "or_mt64_seed": (None, [P, U64]),
    "or_mt64_next": (U64, [P]),
    "or_shuffle_u32": (None, [P, U32, P]),
    "or_prune_policy": (U32, [P, P, U32, U32, U32, P, P, U32, U32, F32, P, P]),
    "or_set_edge_policy": (None, [P]),
}

REF_SIGS = {
    "ref_ip": (F32, [P, P, U32]),
    "ref_ip_free": (F32, [P, P, U32]),
    "ref_l2": (F32, [P, P, U32]),
    "ref_norm": (F32, [P, U32]),
    "ref_recheck_key": (F32, [P, P, U32]),
    "ref_ipdist": (F32, [P, P, U32]),
    "ref_recheck_key_ip": (F32, [P, P, U32]),
    "ref_query_lift": (F64, [P, U32, F64, F64, I32, P]),
    "ref_heap_stream": (I32, [U32, P, P, U32, P, P]),
    "ref_queue_stream": (U32, [U32, P, P, P, U32, P, P, P, P]),
    "ref_index_open": (P, [ctypes.c_char_p, U64, U64, U64, U64, U64, F32, F32, ctypes.c_char_p]),
    "ref_index_close": (None, [P]),
    "ref_index_d2": (F32, [P, U64, U64]),
    "ref_index_norm": (F32, [P, U64]),
    "ref_index_load_graph": (I32, [P, ctypes.c_char_p]),
    "ref_tree_entry": (U32, [P, P]),
    "ref_range_search": (U32, [P, P, F64, U64, P, U32, P]),
    "ref_graph_arrays": (None, [P, P, P]),
    "ref_warm_start": (None, [P, P, U32, U32, P, P]),
    "ref_nndescent": (None, [P, P, P, U32, I32]),
    "ref_gather": (U32, [P, P, P, U32, U32, U64, P, P, U32]),
    "ref_prune": (U32, [P, U32, P, P, U32, P]),
    "ref_commit": (None, [P, P, P, U32, P, P]),
    "ref_build_graph": (None, [P, P, U32, P, P, P]),
    "ref_shuffle": (None, [U32, U64, P]),
    "ref_vamana_collect": (U32, [P, U32, U32, U64, P, P, U32]),
    "ref_vamana_entry": (U32, [P]),
    "ref_lrq_build": (None, [U64, U32, U32, P, P, P, U32, P, P, P]),
    "ref_lrq_materialize": (None, [U64, U32, U32, P, P, P, U32, P, P, U32, F64, P]),
    "ref_lrq_materialize_metric": (None, [U64, U32, U32, P, P, P, U32, P, P, U32, F64, I32, P]),
    "ref_lrq_encode_value": (U32, [F64, P, U32]),
    "ref_tree_build": (P, [P, U64, U64, U64, U64, U64]),
    "ref_tree_nodes": (U64, [P, P]),
    "ref_tree_export": (None, [P, P, P, P, P, U64, P, P]),
    "ref_tree_free": (None, [P]),
    "ref_rknn_detail": (None, [P, P, U32, P, U32, F64, F64, U64, P, I32, P, P, P, P, P]),
}

_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        path = os.path.join(REPO, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        _oracle = _load(path, ORACLE_SIGS)
    return _oracle


def ref_available():
    return os.path.exists(os.path.join(REPO, "oracle", "_ref", "libref_harness.so"))


def ref():
    global _ref
    if _ref is None:
        _ref = _load(os.path.join(REPO, "oracle", "_ref", "libref_harness.so"), REF_SIGS)
    return _ref


# ---- fixture readers --------------------------------------------------------
def read_fbin(path):
    with open(path, "rb") as f:
        n, d = struct.unpack("<II", f.read(8))
        return np.frombuffer(f.read(n * d * 4), dtype=np.float32).reshape(n, d).copy()


def read_rec(path, n, k=100):
    r = np.fromfile(path, dtype=np.dtype([("id", "<u4"), ("key", "<f4")])).reshape(n, k)
    return np.ascontiguousarray(r["id"]), np.ascontiguousarray(r["key"])


def read_gt(path):
    with open(path, "rb") as f:
        n, K = struct.unpack("<II", f.read(8))
        ids = np.frombuffer(f.read(n * K * 4), dtype=np.uint32).reshape(n, K).copy()
        keys = np.frombuffer(f.read(n * K * 4), dtype=np.float32).reshape(n, K).copy()
    return ids, keys


def read_csr(path):
    with open(path, "rb") as f:
        nq = struct.unpack("<I", f.read(4))[0]
        off = np.frombuffer(f.read((nq + 1) * 4), dtype=np.uint32).copy()
        ids = np.frombuffer(f.read(int(off[-1]) * 4), dtype=np.uint32).copy()
        keys = np.frombuffer(f.read(int(off[-1]) * 4), dtype=np.float32).copy()
    return off, ids, keys


def read_basis(path):
    raw = open(path, "rb").read()
    K, M = struct.unpack("<ii", raw[:8])
    mu = np.frombuffer(raw[8:8 + 4 * K], dtype=np.float32).copy()
    G = np.frombuffer(raw[8 + 4 * K:], dtype=np.float32).reshape(M, K).copy()
    return mu, G


def read_codebook(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"ANQIRQ1\0"
    nk, rm, bits, cs = struct.unpack("<iiii", raw[8:24])
    return np.frombuffer(raw[24:], dtype=np.float32).reshape(nk, cs).copy()


def read_verifier(path):
    raw = open(path, "rb").read()
    magic, n, nk, rankm, bits, cs, pol = struct.unpack("<8sQIIIII", raw[:36])
    assert magic == b"ANQIVF2\0"
    coeff = np.frombuffer(raw[36:36 + n * rankm * 4], dtype=np.float32).reshape(n, rankm).copy()
    codes = np.frombuffer(raw[36 + n * rankm * 4:36 + n * rankm * 4 + n * nk], dtype=np.uint8).reshape(n, nk).copy()
    return coeff, codes


def read_graph(path, n, M):
    """Reference graph cache (include/index.h:1810-1915) -> deg, links, tree."""
    raw = open(path, "rb").read()
    fn, fdim, fM = struct.unpack("<QQQ", raw[:24])
    assert fn == n and fM == M
    pos = 28
    deg = np.zeros(n, dtype=np.uint32)
    links = np.zeros((n, M), dtype=np.uint32)
    for i in range(n):
        dg = struct.unpack("<I", raw[pos:pos + 4])[0]
        pos += 4
        deg[i] = dg
        links[i, :dg] = np.frombuffer(raw[pos:pos + 4 * dg], dtype=np.uint32)
        pos += 4 * dg
    leaf_size, n_leaves, ns = struct.unpack("<QQQ", raw[pos:pos + 24])
    pos += 24
    left = np.zeros(ns, dtype=np.uint64)
    right = np.zeros(ns, dtype=np.uint64)
    off = np.zeros(ns, dtype=np.float32)
    hyper = np.zeros((ns, fdim), dtype=np.float32)
    idx_off = np.zeros(ns + 1, dtype=np.uint64)
    idx = []
    for k in range(ns):
        l, r = struct.unpack("<QQ", raw[pos:pos + 16])
        pos += 16
        o = struct.unpack("<f", raw[pos:pos + 4])[0]
        pos += 4
        hs = struct.unpack("<Q", raw[pos:pos + 8])[0]
        pos += 8
        hp = np.frombuffer(raw[pos:pos + 4 * hs], dtype=np.float32)
        pos += 4 * hs
        isz = struct.unpack("<Q", raw[pos:pos + 8])[0]
        pos += 8
        ii = np.frombuffer(raw[pos:pos + 4 * isz], dtype=np.uint32)
        pos += 4 * isz
        left[k], right[k], off[k] = l, r, o
        if hs:
            hyper[k] = hp
        idx.extend(ii.tolist())
        idx_off[k + 1] = len(idx)
    tree = dict(left=left, right=right, offset=off, hyper=hyper, idx_off=idx_off,
                idx=np.array(idx, dtype=np.uint32), nodes=ns, d=fdim)
    return deg, links, tree


class OrTree(ctypes.Structure):
    _fields_ = [("nodes", U32), ("d", U32), ("left", P), ("right", P), ("offset", P),
                ("hyper", P), ("idx_off", P), ("idx", P)]


def or_tree(tree):
    t = OrTree(tree["nodes"], tree["d"], ptr(tree["left"]), ptr(tree["right"]), ptr(tree["offset"]),
               ptr(tree["hyper"]), ptr(tree["idx_off"]), ptr(tree["idx"]))
    t._keep = tree
    return t


def trainer_cutoffs(n=3000, seed=77):
    """Synthetic AKNN cutoff matrix for the trainer fixture (tests/golden/
    trainer3k): rows ascending over the 100 ranks, an offset, a per-row scale
    and squared-uniform increments, all from the counter-based generator."""
    import workload
    u = workload.uniform(seed, 9, 0, n * 100).reshape(n, 100)
    scale = 20.0 + 200.0 * workload.uniform(seed, 10, 0, n)[:, None]
    off = 100.0 * workload.uniform(seed, 11, 0, n)[:, None]
    return (off + np.cumsum(scale * u * u, axis=1)).astype(np.float32)


class Sample:
    """The 256 x 128 reference fixtures (name "sample"), or the 1,000 x 32
    inner-product ones (name "sample_ip")."""

    def __init__(self, name="sample"):
        g = os.path.join(GOLDEN, name, name)
        self.metric = 1 if name == "sample_ip" else 0
        self.base = read_fbin(g + "_base.bin")
        self.query = read_fbin(g + "_query.bin")
        self.n, self.d = self.base.shape
        self.rec_ids, self.rec_keys = read_rec(g + "_baseknn_aknn.rec", self.n)
        self.gt_ids, self.gt_keys = read_gt(g + "_baseknn_gt.bin")
        self.csr = {k: read_csr(g + f"_rknn_gt_k{k}.bin") for k in (10, 50, 100)}
        self.mu, self.G = read_basis(g + "_radbasis_M8.bin")
        self.cb = read_codebook(g + "_learned_rq_codebook_M8_u8.bin")
        self.coeff, self.codes = read_verifier(g + "_verifier.anqivf2")
        self.lifted = read_fbin(g + "_approxlift.bin")
        self.M2, self.scale = map(float, open(g + "_graph_meta.txt").read().split())
        self.deg, self.links, self.tree = read_graph(g + "_graph.bin", self.n, 48)
        self.h = dict(np.load(g + "_harness.npz"))
        self.prefix = g


def oracle_rp_tree(x, leaf=100, seed=10, max_depth=100):
    """or_rp_tree_build on host data -> dict in the reference's node order."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    cap = 4 * n + 16
    t = dict(left=np.zeros(cap, np.uint64), right=np.zeros(cap, np.uint64), offset=np.zeros(cap, np.float32),
             hyper=np.zeros((cap, d), np.float32), idx_off=np.zeros(cap + 1, np.uint64), idx=np.zeros(n, np.uint32))
    draws = ctypes.c_uint64()
    k = oracle().or_rp_tree_build(ptr(x), n, d, leaf, max_depth, seed, cap, ptr(t["left"]), ptr(t["right"]),
                                  ptr(t["offset"]), ptr(t["hyper"]), ptr(t["idx_off"]), ptr(t["idx"]),
                                  ctypes.byref(draws))
    assert k > 0
    return dict(nodes=k, d=d, left=t["left"][:k], right=t["right"][:k], offset=t["offset"][:k],
                hyper=t["hyper"][:k], idx_off=t["idx_off"][:k + 1], idx=t["idx"][:int(t["idx_off"][k])],
                draws=draws.value)


def ref_rp_tree(x, leaf=100, seed=10, max_depth=100):
    """The reference's build_rptree through the harness (fresh engine)."""
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    R = ref()
    h = R.ref_tree_build(ptr(x), n, d, leaf, max_depth, seed)
    tot = ctypes.c_uint64()
    k = R.ref_tree_nodes(h, ctypes.byref(tot))
    t = dict(nodes=k, d=d, left=np.zeros(k, np.uint64), right=np.zeros(k, np.uint64),
             offset=np.zeros(k, np.float32), hyper=np.zeros((k, d), np.float32),
             idx_off=np.zeros(k + 1, np.uint64), idx=np.zeros(max(tot.value, 1), np.uint32))
    R.ref_tree_export(h, ptr(t["left"]), ptr(t["right"]), ptr(t["offset"]), ptr(t["hyper"]), d,
                      ptr(t["idx_off"]), ptr(t["idx"]))
    R.ref_tree_free(h)
    t["idx"] = t["idx"][:tot.value]
    return t


def tree_diff(a, b):
    """'' when two trees in the reference's node order are identical (float
    fields compared as bits), else the name of the first differing field."""
    if int(a["nodes"]) != int(b["nodes"]):
        return f"nodes {a['nodes']} != {b['nodes']}"
    for f in ("left", "right", "idx_off", "idx"):
        if not np.array_equal(np.asarray(a[f], np.uint64), np.asarray(b[f], np.uint64)):
            return f
    for f in ("offset", "hyper"):
        if not np.array_equal(np.asarray(a[f], np.float32).view(np.uint32),
                              np.asarray(b[f], np.float32).view(np.uint32)):
            return f
    return ""


def tree_datasets():
    """Inputs that exercise every branch of the tree build: lifted sample,
    mixture rows, integer rows with many ties, a two-point set that splits
    degenerately, and a depth cap."""
    import workload
    rng = np.random.default_rng(0)
    ties = np.floor(rng.random((4000, 20)) * 3).astype(np.float32)
    degen = np.zeros((500, 8), np.float32)
    degen[::2] = 1
    mix = workload.million_base(20000) * np.float32(0.01)
    return [("mixture-20000x128-leaf100", mix, 100, 100), ("mixture-5000-leaf7", mix[:5000], 7, 100),
            ("ties-4000x20-leaf10", ties, 10, 100), ("degenerate-500x8-leaf10", degen, 10, 100),
            ("depthcap-500x8-depth3", degen, 10, 3)]


class OrEdgePolicy(ctypes.Structure):
    """oracle/anqi_oracle.h or_edge_policy (the configured edge policy)."""
    _fields_ = [("resid_cover", ctypes.c_int), ("ball", ctypes.c_int), ("ball_pure", ctypes.c_int),
                ("occlude_all", ctypes.c_int), ("lsg", ctypes.c_int), ("lsg_alpha", ctypes.c_double),
                ("mu", ctypes.c_void_p), ("resid", ctypes.c_void_p), ("nk", ctypes.c_uint32),
                ("ks", ctypes.c_void_p), ("nks", ctypes.c_uint32), ("budget", ctypes.c_uint32),
                ("lambda_", ctypes.c_double), ("scale2", ctypes.c_double), ("bubble_auto", ctypes.c_int),
                ("bubble_full", ctypes.c_int), ("auto_candidates", ctypes.c_uint32), ("auto_max_swaps", ctypes.c_uint32)]


def edge_policy_config(env, n, M, folder):
    """configure_edge_residual_cover (include/index.h:305-438) restated from
    an environment dict: (flags dict, residual array or None, mu or None);
    None when the policy leaves every flag off.  Inputs are read as the
    reference reads them (residuals from ANQI_EDGE_RESID_FILE or
    folder + "_anqi_rankm_residual.bin", scale from folder + "_anqi_meta.txt")."""
    import numpy as _np
    pol = env.get("ANQI_EDGE_POLICY")
    if not pol:
        return None
    f = dict(resid_cover=0, ball=0, ball_pure=0, occlude_all=1, lsg=0, lsg_alpha=1.0, mu=None, resid=None, nk=100,
             ks=[], budget=16, lam=1.0, scale2=1.0, bubble_auto=0, bubble_full=0, auto_candidates=M, auto_max_swaps=M)
    if pol in ("lsg_rng", "lsg_bubble_hybrid"):
        f["lsg_alpha"] = float(env.get("ANQI_LSG_ALPHA", "1.0") or 0)
        f["mu"] = _np.fromfile(env["ANQI_LSG_MU_FILE"], dtype=_np.float32)
        f["lsg"] = 1
        if pol == "lsg_rng":
            return f
    needs = ("residual_cover", "ball_rng", "ball_hybrid", "ball_multik", "bubble_rng", "bubble_hybrid",
             "bubble_multik", "bubble_auto", "bubble_auto_full", "lsg_bubble_hybrid")
    if pol not in needs:
        return f if f["lsg"] else None
    nk = int(env.get("ANQI_EDGE_NK", "100") or 0) or 100
    f["nk"] = nk
    f["budget"] = min(max(0, int(env.get("ANQI_EDGE_RESID_BUDGET", "16") or 0)), M)
    f["lam"] = float(env.get("ANQI_EDGE_RESID_LAMBDA", "1.0") or 0)
    f["auto_candidates"] = max(1, int(env.get("ANQI_EDGE_AUTO_CANDIDATES", str(M)) or 0))
    f["auto_max_swaps"] = max(1, int(env.get("ANQI_EDGE_AUTO_MAX_SWAPS", str(M)) or 0))
    ks = [int(t) for t in env.get("ANQI_EDGE_KS", "10,50,100").split(",") if t and int(t) > 0]
    f["ks"] = [k for k in ks if 1 <= k <= nk]
    if not f["ks"]:
        return f if f["lsg"] else None
    auto = pol in ("bubble_auto", "bubble_auto_full")
    if f["budget"] == 0 and pol != "ball_rng" and not auto:
        return f if f["lsg"] else None
    try:
        mstar, scale = map(float, open(folder + "_anqi_meta.txt").read().split()[:2])
        f["scale2"] = scale * scale
    except (OSError, ValueError):
        pass
    f["resid"] = _np.fromfile(env.get("ANQI_EDGE_RESID_FILE") or folder + "_anqi_rankm_residual.bin",
                              dtype=_np.float32)
    assert f["resid"].size == n * nk
    f["resid_cover"] = 1
    f["bubble_auto"] = int(auto)
    f["bubble_full"] = int(pol == "bubble_auto_full")
    if pol != "residual_cover":
        f["resid_cover"] = 0
        f["ball"] = 1
        f["ball_pure"] = int(pol in ("ball_rng", "bubble_rng"))
        f["occlude_all"] = int(env.get("ANQI_EDGE_BALL_OCCLUSION", "all") != "any")
    return f


def or_edge_policy(f):
    """ctypes or_edge_policy for a flags dict (keeps its arrays alive)."""
    import numpy as _np
    ks = _np.array(f["ks"] or [0], dtype=_np.int32)
    E = OrEdgePolicy(f["resid_cover"], f["ball"], f["ball_pure"], f["occlude_all"], f["lsg"], f["lsg_alpha"],
                     ptr(f["mu"]) if f["mu"] is not None else None, ptr(f["resid"]) if f["resid"] is not None else None,
                     f["nk"], ptr(ks), len(f["ks"]), f["budget"], f["lam"], f["scale2"], f["bubble_auto"],
                     f["bubble_full"], f["auto_candidates"], f["auto_max_swaps"])
    E._keep = (f, ks)
    return E
