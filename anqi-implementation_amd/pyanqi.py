"""ctypes bindings for libanqi_hip.so (include/anqi.h).

PyTorch provides device memory and streams only: tensors are passed to the C
ABI as raw pointers.  Loading fails loudly when the library is missing; there
is no CPU path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libanqi_hip.so")

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int
F32 = ctypes.c_float
F64 = ctypes.c_double


class AknnParams(ctypes.Structure):
    _fields_ = [("K", U32), ("leaf", U32), ("trees", U32), ("ef", U32), ("ef_alpha", U32),
                ("max_depth", U32), ("seed", U64), ("precision", I32), ("metric", I32)]


class AknnStats(ctypes.Structure):
    _fields_ = [("forest_s", F64), ("leaf_s", F64), ("gather_s", F64), ("total_s", F64),
                ("leaf_pairs", U64), ("leaves", U32), ("max_leaf", U32), ("short_rows", U32),
                ("overflow_rows", U32), ("leaf_survivors", U64), ("gather_d2", U64), ("gather_adj", U64)]


class Tree(ctypes.Structure):
    _fields_ = [("nodes", U32), ("root", U32), ("d", U32), ("left", P), ("right", P), ("offset", P),
                ("hyper", P), ("entry", P), ("leaf_off", ctypes.POINTER(U64)),
                ("leaf_ids", ctypes.POINTER(U32))]


class EdgePolicy(ctypes.Structure):
    """include/anqi.h anqi_edge_policy (device pointers for mu and resid)."""
    _fields_ = [("resid_cover", I32), ("ball", I32), ("ball_pure", I32), ("occlude_all", I32), ("lsg", I32),
                ("lsg_alpha", F64), ("mu", P), ("resid", P), ("nk", U32), ("ks", I32 * 16), ("nks", U32),
                ("budget", U32), ("lambda_", F64), ("scale2", F64), ("bubble_auto", I32), ("bubble_full", I32),
                ("auto_candidates", U32), ("auto_max_swaps", U32)]


class GraphParams(ctypes.Structure):
    _fields_ = [("M", U32), ("nn_k", U32), ("leaf", U32), ("explore", U32), ("ef_alpha", U32),
                ("iters", I32), ("sample", I32), ("alpha", F32), ("seed", U64), ("form", I32),
                ("vamana", I32), ("vamana_L", U32), ("vamana_passes", I32), ("vamana_batch_frac", F64),
                ("vamana_max_batch", U64), ("vamana_batch_base", F64), ("vamana_early_alpha", F32),
                ("vamana_early_alpha_one", I32), ("vamana_entry", U32), ("edge", ctypes.POINTER(EdgePolicy))]


VAMANA_STYLES = {"global": 1, "batch": 2}


class GraphStats(ctypes.Structure):
    _fields_ = [("tree_s", F64), ("warm_s", F64), ("nnd_s", F64), ("gather_s", F64), ("commit_s", F64),
                ("total_s", F64), ("nnd_passes", I32), ("nnd_upd", ctypes.c_int64 * 16),
                ("overflow_rows", U32), ("gather_d2", U64), ("gather_adj", U64), ("vamana_s", F64)]


class SearchTuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", U32), ("visited_slots", U32), ("result_cap", U32)]


class Index(ctypes.Structure):
    _fields_ = [("base", P), ("bnorm", P), ("lifted", P), ("deg", P), ("links", P), ("n", U32), ("d", U32),
                ("dl", U32), ("M", U32), ("tree", ctypes.POINTER(Tree)), ("mstar2", F64), ("scale", F64),
                ("lmain", P), ("ltail", P), ("plinks", P), ("brows", P), ("metric", I32)]


METRIC_L2, METRIC_IP = 0, 1
METRICS = {"l2": METRIC_L2, "ip": METRIC_IP}


SIGS = {
    "anqi_ctx_create": (I32, [I32, ctypes.POINTER(P)]),
    "anqi_ctx_destroy": (None, [P]),
    "anqi_last_error": (ctypes.c_char_p, [P]),
    "anqi_ctx_sync": (I32, [P]),
    "anqi_ctx_set_stream": (I32, [P, P]),
    "anqi_ctx_stream": (P, [P]),
    "anqi_ctx_reset_stream": (I32, [P]),
    "anqi_ctx_time_kernels": (I32, [P, I32]),
    "anqi_ctx_last_kernel_ms": (F64, [P]),
    "anqi_ctx_set_search_tuning": (I32, [P, P]),
    "anqi_ctx_redo_counts": (I32, [P, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
    "anqi_batcher_redo_counts": (I32, [P, ctypes.POINTER(U32), ctypes.POINTER(U32)]),
    "anqi_dmalloc": (I32, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
    "anqi_dfree": (I32, [P, P]),
    "anqi_h2d": (I32, [P, P, P, ctypes.c_size_t]),
    "anqi_d2h": (I32, [P, P, P, ctypes.c_size_t]),
    "anqi_norms": (I32, [P, P, U64, U32, P]),
    "anqi_lift_split": (I32, [P, P, U32, U32, P, P]),
    "anqi_pack_links": (I32, [P, P, P, U32, U32, P]),
    "anqi_byte_rows": (I32, [P, P, P, U32, U32, F64, P, ctypes.POINTER(I32)]),
    "anqi_search_entries": (I32, [P, ctypes.POINTER(Index), P, U32, P]),
    "anqi_pair_d2": (I32, [P, P, P, U32, P, P, U64, P]),
    "anqi_exact_knn": (I32, [P, P, U32, U32, U32, I32, P, P]),
    "anqi_rknn_membership": (I32, [P, P, U32, P, U32, U32, P, P, ctypes.POINTER(U64), P, P]),
    "anqi_exact_knn_metric": (I32, [P, P, U32, U32, U32, I32, I32, P, P]),
    "anqi_rknn_membership_metric": (I32, [P, P, U32, P, U32, U32, I32, P, P, ctypes.POINTER(U64), P, P]),
    "anqi_aknn_build": (I32, [P, P, U32, U32, ctypes.POINTER(AknnParams), P, P, ctypes.POINTER(AknnStats)]),
    "anqi_lrq_encode": (I32, [P, P, U64, U32, U32, P, P, P, U32, P, P, P]),
    "anqi_transpose_codes": (I32, [P, P, U64, U32, P]),
    "anqi_lrq_materialize_multi": (I32, [P, P, P, U64, U32, U32, P, P, P, U32, P, U32, F64, P]),
    "anqi_lrq_materialize_metric": (I32, [P, P, P, U64, U32, U32, P, P, P, U32, P, U32, F64, I32, P]),
    "anqi_ctx_set_lrq_form": (I32, [P, I32]),
    "anqi_ctx_lrq_fallbacks": (I32, [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
    "anqi_lrq_materialize": (I32, [P, P, P, U64, U32, U32, P, P, P, U32, U32, F64, P]),
    "anqi_aknn_cutoffs": (I32, [P, P, U64, U32, P, P]),
    "anqi_aknn_cutoffs_metric": (I32, [P, P, U64, U32, I32, P, P]),
    "anqi_lift_host": (I32, [P, U64, U32, P, P, ctypes.POINTER(F64), ctypes.POINTER(F64)]),
    "anqi_lift": (I32, [P, P, U64, U32, P, P, ctypes.POINTER(F64), ctypes.POINTER(F64)]),
    "anqi_lift_host_metric": (I32, [P, U64, U32, P, I32, P, ctypes.POINTER(F64), ctypes.POINTER(F64)]),
    "anqi_lift_metric": (I32, [P, P, U64, U32, P, I32, P, ctypes.POINTER(F64), ctypes.POINTER(F64)]),
    "anqi_lrq_relax": (I32, [P, P, U64, U32, U32, P, P, P, U32, F32, I32, F32, ctypes.POINTER(F64), P, P]),
    "anqi_warm_start": (I32, [P, P, P, U32, U32, P, U32, U32, P, P]),
    "anqi_nndescent": (I32, [P, P, P, U32, U32, U32, P, P, I32, I32, P, ctypes.POINTER(I32), I32]),
    "anqi_reverse_csr": (I32, [P, P, U32, U32, P, U32, P, P]),
    "anqi_gather_topk": (I32, [P, P, P, U32, U32, P, P, U32, P, P, U32, U32, U32, P, P, P, I32]),
    "anqi_gather_prune": (I32, [P, P, P, U32, U32, P, P, U32, P, P, U32, U32, U32, F32, P, P, I32]),
    "anqi_commit": (I32, [P, P, P, U32, U32, P, P, U32, F32, P, P]),
    "anqi_rp_tree": (I32, [P, P, U32, U32, U32, U32, U64, ctypes.POINTER(Tree)]),
    "anqi_tree_upload": (I32, [P, U32, U32, U32, P, P, P, P, P, ctypes.POINTER(Tree)]),
    "anqi_tree_free": (I32, [P, ctypes.POINTER(Tree)]),
    "anqi_vamana_medoid": (I32, [P, P, U32, U32, U64, P]),
    "anqi_graph_build": (I32, [P, P, U32, U32, P, U32, ctypes.POINTER(GraphParams), P, P, ctypes.POINTER(Tree),
                               ctypes.POINTER(GraphStats)]),
    "anqi_rknn_search": (I32, [P, ctypes.POINTER(Index), P, U32, U32, P, U32, P, P, U32, P, P, P, P, P, P, P, I32]),
    "anqi_rknn_search_entries": (I32, [P, ctypes.POINTER(Index), P, U32, U32, P, U32, P, P, P, U32, P, P, P, P, P, P,
                                       P, I32]),
    "anqi_rknn_search_noalloc": (I32, [P, ctypes.POINTER(Index), P, U32, U32, P, U32, P, P, U32, P, P, P, P, P, P,
                                       P]),
    "anqi_compact_candidates": (I32, [P, U32, U32, P, P, P, P, P, P, U64, ctypes.POINTER(U64)]),
    "anqi_batcher_create": (I32, [P, ctypes.POINTER(Index), P, U32, U32, U32, U32, ctypes.POINTER(SearchTuning),
                                  ctypes.POINTER(P)]),
    "anqi_batcher_run": (I32, [P, P, P, P, U32, P, P, U64, ctypes.POINTER(U64)]),
    "anqi_batcher_destroy": (None, [P]),
    "anqi_search_autotune": (I32, [P, ctypes.POINTER(Index), P, P, U32, P, P, U32, ctypes.POINTER(SearchTuning),
                                   ctypes.POINTER(F64)]),
    "anqi_build_info": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C anqi-implementation_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class AnqiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"anqi status {code}: {msg}")
        self.code = code


def dp(t):
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return P(t.data_ptr())


def hp(a):
    return None if a is None else a.ctypes.data_as(P)


class Context:
    """One anqi_ctx bound to a device; runs on torch's current stream."""

    def __init__(self, device: int = 0):
        import torch
        self.torch = torch
        self.L = lib()
        self.ctx = P()
        rc = self.L.anqi_ctx_create(device, ctypes.byref(self.ctx))
        if rc != 0:
            raise AnqiError(rc, "anqi_ctx_create failed")
        self.device = device
        self.use_torch_stream()

    def use_torch_stream(self):
        s = self.torch.cuda.current_stream(self.device)
        self.L.anqi_ctx_set_stream(self.ctx, P(s.cuda_stream))

    def close(self):
        if self.ctx:
            self.L.anqi_ctx_destroy(self.ctx)
            self.ctx = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            raise AnqiError(rc, self.L.anqi_last_error(self.ctx).decode())

    def call(self, name, *args):
        self.check(getattr(self.L, name)(self.ctx, *args))

    def sync(self):
        self.call("anqi_ctx_sync")

    def time_kernels(self, enable=True):
        """HIP-event timing of each range-search kernel launch (anqi_ctx_time_kernels)."""
        self.call("anqi_ctx_time_kernels", 1 if enable else 0)

    def last_kernel_ms(self):
        return float(self.L.anqi_ctx_last_kernel_ms(self.ctx))

    # ---- helpers returning torch tensors -------------------------------------
    def empty(self, shape, dtype):
        return self.torch.empty(shape, dtype=dtype, device=f"cuda:{self.device}")

    def to_dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{self.device}")

    def norms(self, x):
        n, d = x.shape
        out = self.empty((n,), self.torch.float32)
        self.call("anqi_norms", dp(x), n, d, dp(out))
        return out

    def exact_knn(self, x, K, precision=0, metric=METRIC_L2):
        n, d = x.shape
        ids = self.empty((n, K), self.torch.int32)
        keys = self.empty((n, K), self.torch.float32)
        self.call("anqi_exact_knn_metric", dp(x), n, d, K, precision, metric, dp(ids), dp(keys))
        return ids, keys

    def rknn_membership(self, base, q, rk2, metric=METRIC_L2):
        n, d = base.shape
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint32)
        tot = U64(0)
        self.call("anqi_rknn_membership_metric", dp(base), n, dp(q), nq, d, metric, dp(rk2), hp(off),
                  ctypes.byref(tot), None, None)
        ids = self.empty((max(1, tot.value),), self.torch.int32)
        keys = self.empty((max(1, tot.value),), self.torch.float32)
        self.call("anqi_rknn_membership_metric", dp(base), n, dp(q), nq, d, metric, dp(rk2), hp(off),
                  ctypes.byref(tot), dp(ids), dp(keys))
        return off, ids[:tot.value], keys[:tot.value]

    def aknn_build(self, x, K=100, leaf=500, trees=64, ef=600, ef_alpha=8, seed=10, precision=0, max_depth=100,
                   metric=METRIC_L2):
        n, d = x.shape
        ids = self.empty((n, K), self.torch.int32)
        keys = self.empty((n, K), self.torch.float32)
        p = AknnParams(K, leaf, trees, ef, ef_alpha, max_depth, seed, precision, metric)
        st = AknnStats()
        self.call("anqi_aknn_build", dp(x), n, d, ctypes.byref(p), dp(ids), dp(keys), ctypes.byref(st))
        return ids, keys, st

    def aknn_cutoffs(self, keys, metric=METRIC_L2):
        """(cut = max(0, keys) -- keys - 1 for "ip" --, horizon = cut[:, K-1]) on
        the device (anqi_aknn_cutoffs_metric)."""
        n, K = keys.shape
        cut = self.empty((n, K), self.torch.float32)
        hz = self.empty((n,), self.torch.float32)
        self.call("anqi_aknn_cutoffs_metric", dp(keys), n, K, metric, dp(cut), dp(hz))
        return cut, hz

    def lrq_encode(self, cut, mu, basis, cb):
        n, nk = cut.shape
        rankm = basis.shape[0]
        coeff = self.empty((n, rankm), self.torch.float32)
        codes = self.empty((n, nk), self.torch.uint8)
        codes_km = self.empty((nk, n), self.torch.uint8)
        self.call("anqi_lrq_encode", dp(cut), n, nk, rankm, dp(mu), dp(basis), dp(cb), cb.shape[1], dp(coeff),
                  dp(codes), dp(codes_km))
        return coeff, codes, codes_km

    def lrq_materialize_multi(self, coeff, codes_km, mu, basis, cb, ks, slack=1.0, out=None, metric=METRIC_L2):
        """Thresholds for every k in ks, one pass (anqi_lrq_materialize_metric): (len(ks), n)."""
        n, rankm = coeff.shape
        kh = np.ascontiguousarray(ks, np.uint32)
        out = out if out is not None else self.empty((len(kh), n), self.torch.float32)
        self.call("anqi_lrq_materialize_metric", dp(coeff), dp(codes_km), n, mu.shape[0], rankm, dp(mu), dp(basis),
                  dp(cb), cb.shape[1], hp(kh), len(kh), slack, metric, dp(out))
        return out

    def set_lrq_form(self, form):
        """0: certified float-float (default); 1: soft binary64 throughout."""
        self.call("anqi_ctx_set_lrq_form", form)

    def lrq_fallbacks(self):
        """(encode, decode) elements the last calls recomputed in soft binary64."""
        e, d = U64(), U64()
        self.call("anqi_ctx_lrq_fallbacks", ctypes.byref(e), ctypes.byref(d))
        return e.value, d.value

    def transpose_codes(self, codes):
        n, nk = codes.shape
        out = self.empty((nk, n), self.torch.uint8)
        self.call("anqi_transpose_codes", dp(codes), n, nk, dp(out))
        return out

    def lrq_materialize(self, coeff, codes_km, mu, basis, cb, k, slack=1.0, out=None, metric=METRIC_L2):
        n, rankm = coeff.shape
        nk = mu.shape[0]
        thr = out if out is not None else self.empty((n,), self.torch.float32)
        if metric == METRIC_L2:
            self.call("anqi_lrq_materialize", dp(coeff), dp(codes_km), n, nk, rankm, dp(mu), dp(basis), dp(cb),
                      cb.shape[1], k, slack, dp(thr))
        else:
            kh = np.array([k], np.uint32)
            self.call("anqi_lrq_materialize_metric", dp(coeff), dp(codes_km), n, nk, rankm, dp(mu), dp(basis),
                      dp(cb), cb.shape[1], hp(kh), 1, slack, metric, dp(thr))
        return thr

    def warm_start(self, x, nrm, ids, cap):
        n, d = x.shape
        width = ids.shape[1]
        hidx = self.empty((n, cap), self.torch.int32)
        hval = self.empty((n, cap), self.torch.float32)
        self.call("anqi_warm_start", dp(x), dp(nrm), n, d, dp(ids), width, cap, dp(hidx), dp(hval))
        return hidx, hval

    def nndescent(self, x, nrm, hidx, hval, iters=6, sample=20, form=0):
        n, d = x.shape
        cap = hidx.shape[1]
        upd = np.zeros(max(iters, 1), np.int64)
        passes = I32(0)
        self.call("anqi_nndescent", dp(x), dp(nrm), n, d, cap, dp(hidx), dp(hval), iters, sample, hp(upd),
                  ctypes.byref(passes), form)
        return passes.value, upd[:passes.value]

    def reverse_csr(self, ids, n, rowcnt=None):
        rows, width = ids.shape
        roff = self.empty((n + 1,), self.torch.int32)
        rids = self.empty((max(1, rows * width),), self.torch.int32)
        self.call("anqi_reverse_csr", dp(ids), rows, width, dp(rowcnt), n, dp(roff), dp(rids))
        return roff, rids

    def gather_topk(self, x, nrm, hidx, hval, roff, rids, explore, pool_cap, K, form=0):
        n, d = x.shape
        ids = self.empty((n, K), self.torch.int32)
        keys = self.empty((n, K), self.torch.float32)
        cnt = self.empty((n,), self.torch.int32)
        self.call("anqi_gather_topk", dp(x), dp(nrm), n, d, dp(hidx), dp(hval), hidx.shape[1], dp(roff), dp(rids),
                  explore, pool_cap, K, dp(ids), dp(keys), dp(cnt), form)
        return ids, keys, cnt

    def gather_prune(self, x, nrm, hidx, hval, roff, rids, explore, pool_cap, M, alpha, form=0):
        n, d = x.shape
        O = self.empty((n, M), self.torch.int32)
        ocnt = self.empty((n,), self.torch.int32)
        self.call("anqi_gather_prune", dp(x), dp(nrm), n, d, dp(hidx), dp(hval), hidx.shape[1], dp(roff), dp(rids),
                  explore, pool_cap, M, alpha, dp(O), dp(ocnt), form)
        return O, ocnt

    def commit(self, x, nrm, O, ocnt, M, alpha):
        n, d = x.shape
        links = self.empty((n, M), self.torch.int32)
        deg = self.empty((n,), self.torch.int32)
        self.call("anqi_commit", dp(x), dp(nrm), n, d, dp(O), dp(ocnt), M, alpha, dp(links), dp(deg))
        return links, deg

    def rp_tree(self, x, leaf=100, max_depth=100, seed=10):
        n, d = x.shape
        t = Tree()
        self.call("anqi_rp_tree", dp(x), n, d, leaf, max_depth, seed, ctypes.byref(t))
        return t

    def lrq_relax(self, cut, mu, basis, cb, tau, mode=0, grad_scale=1.0, want_loss=True, grads=True):
        """Relaxed radius-reconstruction loss (mode 0 softmax/nearest, 1
        sigmoid/floor) and, optionally, its gradients (hand-written HIP
        backward).  Returns (loss or None, gG or None, gcb or None)."""
        n, nk = cut.shape
        rankm = basis.shape[0]
        gG = self.empty((rankm, nk), self.torch.float32) if grads else None
        gcb = self.empty(tuple(cb.shape), self.torch.float32) if grads else None
        loss = F64(0.0)
        self.call("anqi_lrq_relax", dp(cut), n, nk, rankm, dp(mu), dp(basis), dp(cb), cb.shape[1], float(tau),
                  int(mode), float(grad_scale), ctypes.byref(loss) if want_loss else None, dp(gG), dp(gcb))
        return (loss.value if want_loss else None), gG, gcb

    def search_autotune(self, ix, thr_tables, queries, Lq, kidx, Lmax):
        t = SearchTuning()
        ms = F64(0.0)
        self.call("anqi_search_autotune", ctypes.byref(ix), dp(thr_tables), dp(queries), queries.shape[0], dp(Lq),
                  dp(kidx), Lmax, ctypes.byref(t), ctypes.byref(ms))
        return t, ms.value

    def redo_counts(self):
        """(retry tier, global tier) queries of the last search call."""
        r, g = U32(), U32()
        self.call("anqi_ctx_redo_counts", ctypes.byref(r), ctypes.byref(g))
        return r.value, g.value

    def set_search_tuning(self, tuning=None):
        self.call("anqi_ctx_set_search_tuning", ctypes.byref(tuning) if tuning is not None else None)

    def batcher(self, ix, thr_tables, max_batch, Lmax, out_cap=4096, tuning=None):
        return Batcher(self, ix, thr_tables, max_batch, Lmax, out_cap, tuning)

    def tree_host(self, t):
        """Host copy of a device tree in the reference's node order (children
        before parents, root last; rptrees.h:99-152): dict of left/right
        (uint64, UINT32_MAX for leaves), offset, hyper, idx_off, idx."""
        import formats
        nodes = t.nodes
        left = np.zeros(nodes, np.uint32)
        right = np.zeros(nodes, np.uint32)
        off = np.zeros(nodes, np.float32)
        hyper = np.zeros((nodes, t.d), np.float32)
        for arr, src in ((left, t.left), (right, t.right), (off, t.offset), (hyper, t.hyper)):
            if arr.nbytes:
                self.check(self.L.anqi_d2h(self.ctx, hp(arr), ctypes.c_void_p(src), arr.nbytes))
        loff = np.ctypeslib.as_array(t.leaf_off, shape=(nodes + 1,)).copy()
        lids = np.ctypeslib.as_array(t.leaf_ids, shape=(max(int(loff[-1]), 1),))[:int(loff[-1])].copy()
        tn = formats.tree_postorder(left, right, off, hyper, loff, lids, root=t.root)
        k = len(tn)
        out = dict(nodes=k, d=t.d, left=np.zeros(k, np.uint64), right=np.zeros(k, np.uint64),
                   offset=np.zeros(k, np.float32), hyper=np.zeros((k, t.d), np.float32),
                   idx_off=np.zeros(k + 1, np.uint64))
        ids = []
        for i, (lk, rk, o, h, members) in enumerate(tn):
            out["left"][i], out["right"][i], out["offset"][i] = lk, rk, o
            if h is not None:
                out["hyper"][i] = h
            if members is not None:
                ids.extend(np.asarray(members).tolist())
            out["idx_off"][i + 1] = len(ids)
        out["idx"] = np.array(ids, np.uint32)
        return out

    def tree_upload(self, tree):
        """tree: dict of host arrays (nodes, root, d, left, right, offset, hyper, entry)."""
        t = Tree()
        keep = {k: np.ascontiguousarray(v) for k, v in tree.items() if isinstance(v, np.ndarray)}
        self.call("anqi_tree_upload", tree["nodes"], tree["root"], tree["d"], hp(keep["left"]), hp(keep["right"]),
                  hp(keep["offset"]), hp(keep["hyper"]), hp(keep["entry"]), ctypes.byref(t))
        return t

    def tree_free(self, t):
        self.call("anqi_tree_free", ctypes.byref(t))

    def graph_build(self, lifted, warm_ids, M=48, nn_k=50, leaf=100, explore=100, ef_alpha=8, iters=6, sample=20,
                    alpha=1.2, seed=10, form=0, vamana=None, vamana_L=128, vamana_passes=2, vamana_batch_frac=0.02,
                    vamana_max_batch=1000000, vamana_batch_base=2.0, vamana_early_alpha=1.0,
                    vamana_early_alpha_one=1, vamana_entry=0, edge=None):
        """anqi_graph_build; vamana: None, "global" or "batch" (ANQI_VAMANA_STYLE);
        iters=0 skips the lifted LocalJoin (ANQI_REUSE_ORIG_KNN)."""
        n, dl = lifted.shape
        links = self.empty((n, M), self.torch.int32)
        deg = self.empty((n,), self.torch.int32)
        t = Tree()
        st = GraphStats()
        p = GraphParams(M, nn_k, leaf, explore, ef_alpha, iters, sample, alpha, seed, form,
                        VAMANA_STYLES[vamana] if vamana else 0, vamana_L, vamana_passes, vamana_batch_frac,
                        vamana_max_batch, vamana_batch_base, vamana_early_alpha, vamana_early_alpha_one, vamana_entry,
                        ctypes.pointer(edge) if edge is not None else None)
        self.call("anqi_graph_build", dp(lifted), n, dl, dp(warm_ids), warm_ids.shape[1], ctypes.byref(p),
                  dp(links), dp(deg), ctypes.byref(t), ctypes.byref(st))
        return links, deg, t, st

    def vamana_medoid(self, x, sample=None):
        """choose_vamana_entry "medoid" over device rows (anqi_vamana_medoid)."""
        n, d = x.shape
        e = U32()
        self.call("anqi_vamana_medoid", dp(x), n, d, min(n, 100000) if sample is None else sample, ctypes.byref(e))
        return e.value

    def lift(self, base, rk_horizon, metric=METRIC_L2):
        """Lift of device base rows on the device (anqi_lift_metric): (lifted, mstar2, scale)."""
        n, d = base.shape
        lifted = self.empty((n, d + 2), self.torch.float32)
        m2, sc = F64(), F64()
        self.call("anqi_lift_metric", dp(base), n, d, dp(rk_horizon), metric, dp(lifted), ctypes.byref(m2),
                  ctypes.byref(sc))
        return lifted, m2.value, sc.value

    def lift_split(self, lifted):
        """Split lifted layout (anqi_lift_split): (lmain n x d, ltail n)."""
        n, dl = lifted.shape
        d = dl - 2
        lmain = self.empty((n, d), self.torch.float32)
        ltail = self.empty((n,), self.torch.float32)
        self.call("anqi_lift_split", dp(lifted), n, d, dp(lmain), dp(ltail))
        return lmain, ltail

    def pack_links(self, links, deg):
        """Sentinel adjacency (anqi_pack_links): n x M ids, UINT32_MAX past the degree."""
        n, M = links.shape
        plinks = self.empty((n, M), self.torch.int32)
        self.call("anqi_pack_links", dp(links), dp(deg), n, M, dp(plinks))
        return plinks

    def byte_rows(self, base, lmain, scale):
        """Byte rows (anqi_byte_rows): the n x d uint8 array, or None when the
        base is not all floats 0..255 (the bytes would not be exact)."""
        n, d = base.shape
        brows = self.empty((n, d), self.torch.uint8)
        exact = I32()
        self.call("anqi_byte_rows", dp(base), dp(lmain), n, d, scale, dp(brows), ctypes.byref(exact))
        return brows if exact.value else None

    def make_index(self, base, bnorm, lifted, deg, links, tree, mstar2, scale, split=True, packed=True,
                   byte_rows=True, metric=METRIC_L2):
        """split: give the persistent search the split lifted layout (d % 16 == 0);
        packed: and the sentinel adjacency rows (no degree read);
        byte_rows: and byte rows, when the base values are all floats 0..255;
        metric: METRIC_L2 or METRIC_IP (the lift must be of the same metric)."""
        lmain = ltail = plinks = brows = None
        if split and base.shape[1] % 16 == 0:
            lmain, ltail = self.lift_split(lifted)
            if packed:
                plinks = self.pack_links(links, deg)
            if byte_rows:
                brows = self.byte_rows(base, lmain, scale)
        ix = Index(dp(base), dp(bnorm), dp(lifted), dp(deg), dp(links), base.shape[0], base.shape[1],
                   lifted.shape[1], links.shape[1], ctypes.pointer(tree), mstar2, scale, dp(lmain), dp(ltail),
                   dp(plinks), dp(brows), metric)
        ix._keep = (base, bnorm, lifted, deg, links, tree, lmain, ltail, plinks, brows)
        ix.byte_rows = brows is not None
        return ix

    def search_buffers(self, nq, out_cap):
        T = self.torch
        return dict(cand=self.empty((nq, out_cap), T.int32), acc=self.empty((nq, out_cap), T.uint8),
                    cnt=self.empty((nq,), T.int32), nd=self.empty((nq,), T.int32),
                    pred=self.empty((nq,), T.int32), nexp=self.empty((nq,), T.int32),
                    status=self.empty((nq,), T.int32))

    def search_entries(self, ix, queries):
        """Search-tree entry point of each query (anqi_search_entries)."""
        out = self.empty((queries.shape[0],), self.torch.int32)
        self.call("anqi_search_entries", ctypes.byref(ix), dp(queries), queries.shape[0], dp(out))
        return out

    def rknn_search(self, ix, queries, L, thr, out_cap=2048, mode=0, Lq=None, Lmax=0, kidx=None, bufs=None,
                    entries=None):
        nq = queries.shape[0]
        b = bufs if bufs is not None else self.search_buffers(nq, out_cap)
        self.call("anqi_rknn_search_entries", ctypes.byref(ix), dp(queries), nq, L, dp(Lq), Lmax, dp(thr), dp(kidx),
                  dp(entries), b["cand"].shape[1], dp(b["cand"]), dp(b["acc"]), dp(b["cnt"]), dp(b["nd"]),
                  dp(b["pred"]), dp(b["nexp"]), dp(b["status"]), mode)
        return b


    def compact_candidates(self, bufs):
        """(off[nq+1], ids, accept) of a search's buffers as device CSR (anqi_compact_candidates)."""
        nq, cap = bufs["cand"].shape
        off = self.empty((nq + 1,), self.torch.int32)
        tot = U64()
        self.call("anqi_compact_candidates", nq, cap, dp(bufs["cand"]), dp(bufs["acc"]), dp(bufs["cnt"]), dp(off),
                  None, None, 0, ctypes.byref(tot))
        ids = self.empty((max(1, tot.value),), self.torch.int32)
        acc = self.empty((max(1, tot.value),), self.torch.uint8)
        self.call("anqi_compact_candidates", nq, cap, dp(bufs["cand"]), dp(bufs["acc"]), dp(bufs["cnt"]), dp(off),
                  dp(ids), dp(acc), tot.value, ctypes.byref(tot))
        return off, ids[:tot.value], acc[:tot.value]

    def rknn_search_noalloc(self, ix, queries, L, thr, bufs, Lq=None, Lmax=0, kidx=None):
        """anqi_rknn_search_noalloc into caller buffers (search_buffers)."""
        b = bufs
        self.call("anqi_rknn_search_noalloc", ctypes.byref(ix), dp(queries), queries.shape[0], L, dp(Lq), Lmax,
                  dp(thr), dp(kidx), b["cand"].shape[1], dp(b["cand"]), dp(b["acc"]), dp(b["cnt"]), dp(b["nd"]),
                  dp(b["pred"]), dp(b["nexp"]), dp(b["status"]))
        return b


def lift_host(base: np.ndarray, rk_horizon: np.ndarray, metric=METRIC_L2):
    base = np.ascontiguousarray(base, np.float32)
    rk = np.ascontiguousarray(rk_horizon, np.float32)
    n, d = base.shape
    out = np.zeros((n, d + 2), np.float32)
    m2, sc = F64(), F64()
    rc = lib().anqi_lift_host_metric(hp(base), n, d, hp(rk), metric, hp(out), ctypes.byref(m2), ctypes.byref(sc))
    if rc != 0:
        raise AnqiError(rc, "lift failed")
    return out, m2.value, sc.value


class Batcher:
    """anqi_batcher: one captured graph serving query batches of mixed k
    (threshold table index) and L, answers returned as host CSR."""

    def __init__(self, ctx, ix, thr_tables, max_batch, Lmax, out_cap=4096, tuning=None):
        self.ctx = ctx
        self.keep = (ix, thr_tables)
        self.h = P()
        self.max_batch = max_batch
        self.out_cap = out_cap
        ctx.check(ctx.L.anqi_batcher_create(ctx.ctx, ctypes.byref(ix), dp(thr_tables), thr_tables.shape[0],
                                            max_batch, Lmax, out_cap,
                                            ctypes.byref(tuning) if tuning is not None else None,
                                            ctypes.byref(self.h)))
        self.ids = np.zeros(max_batch * 256, np.uint32)

    def run(self, queries, L, kidx):
        """queries: host (nq, d) float32; L, kidx: host uint32 (nq,).
        Returns (off[nq+1], ids[off[-1]])."""
        q = np.ascontiguousarray(queries, np.float32)
        Lh = np.ascontiguousarray(L, np.uint32)
        kh = np.ascontiguousarray(kidx, np.uint32)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint32)
        tot = U64(0)
        for _ in range(2):
            rc = self.ctx.L.anqi_batcher_run(self.h, hp(q), hp(Lh), hp(kh), nq, hp(off), hp(self.ids),
                                             self.ids.size, ctypes.byref(tot))
            if rc == 2 and tot.value > self.ids.size:       # answers outgrew the host buffer
                self.ids = np.zeros(int(tot.value) * 2, np.uint32)
                continue
            self.ctx.check(rc)
            break
        return off, self.ids[:int(tot.value)].copy()

    def redo_counts(self):
        """(retry tier, global tier) queries of the last run."""
        r, g = U32(), U32()
        self.ctx.check(self.ctx.L.anqi_batcher_redo_counts(self.h, ctypes.byref(r), ctypes.byref(g)))
        return r.value, g.value

    def close(self):
        if self.h:
            self.ctx.L.anqi_batcher_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
