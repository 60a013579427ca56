"""Readers and writers for the reference's on-disk formats (little-endian).

  fbin                 [u32 n][u32 d][n*d f32]                 rknn_query.cpp:23-53
  <PFX>_baseknn_aknn.rec  n*K {u32 id, f32 key}               build_aknn_source.cpp:119-137
  _aknn_build_time.meta   ANQI_AKNN_BUILD_V1 text             build_aknn_source.cpp:139-155
  _radbasis_M8.bin     [i32 K][i32 M][f32 mu[K]][f32 G[M][K]]  train_final_m8_u8.py:176-180
  codebook             "ANQIRQ1\\0"[i32 nk,rankm,bits,size][f32]  train_final_m8_u8.py:183-187
  verifier artifact    "ANQIVF2\\0" header + coeff + codes     rknn_query.cpp:755-838
  graph cache          save_index/save_tree                    index.h:1810-1915
  RkNN CSR             [u32 nq][u32 off[nq+1]][ids][keys]      dataset_gt.cpp:302-313
"""
from __future__ import annotations

import struct

import numpy as np

U32MAX = 0xFFFFFFFF
FLT_MAX = np.float32(3.4028234663852886e38)


def write_fbin(path, a):
    a = np.ascontiguousarray(a, np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<II", a.shape[0], a.shape[1]))
        f.write(a.tobytes())


def read_fbin(path):
    with open(path, "rb") as f:
        n, d = struct.unpack("<II", f.read(8))
        return np.frombuffer(f.read(n * d * 4), np.float32).reshape(n, d).copy()


REC = np.dtype([("id", "<u4"), ("key", "<f4")])


def write_rec(path, ids, keys):
    r = np.empty(ids.shape, REC)
    r["id"] = ids
    r["key"] = keys
    r.tofile(path)


def read_rec(path, n, k=100):
    r = np.fromfile(path, REC).reshape(n, k)
    return np.ascontiguousarray(r["id"]), np.ascontiguousarray(r["key"])


def write_aknn_meta(path, seconds, metric, n, d, k, leaf, trees, ef, threads, seed):
    with open(path, "w") as f:
        f.write("ANQI_AKNN_BUILD_V1\n")
        f.write(f"seconds {seconds:.17g}\nmetric {metric}\nn {n}\nd {d}\nk {k}\nleaf {leaf}\n"
                f"trees {trees}\nef {ef}\nthreads {threads}\nseed {seed}\n")


def write_verifier_meta(path, seconds, n):
    with open(path, "w") as f:
        f.write(f"ANQI_VERIFIER_TRAIN_V1\nseconds {seconds:.17g}\nn {n}\nk 100\nrank_m 8\nbits 8\n")


def read_basis(path):
    raw = open(path, "rb").read()
    K, M = struct.unpack("<ii", raw[:8])
    mu = np.frombuffer(raw[8:8 + 4 * K], np.float32).copy()
    G = np.frombuffer(raw[8 + 4 * K:8 + 4 * K + 4 * K * M], np.float32).reshape(M, K).copy()
    return mu, G


def read_codebook(path):
    raw = open(path, "rb").read()
    if raw[:8] != b"ANQIRQ1\0":
        raise ValueError("bad learned RQ codebook magic")
    nk, rm, bits, cs = struct.unpack("<iiii", raw[8:24])
    return np.frombuffer(raw[24:24 + 4 * nk * cs], np.float32).reshape(nk, cs).copy()


def write_verifier(path, coeff, codes, policy=1):
    n, rankm = coeff.shape
    nk = codes.shape[1]
    with open(path + ".tmp", "wb") as f:
        f.write(b"ANQIVF2\0")
        f.write(struct.pack("<QIIIII", n, nk, rankm, 8, 256, policy))
        f.write(np.ascontiguousarray(coeff, np.float32).tobytes())
        f.write(np.ascontiguousarray(codes, np.uint8).tobytes())
    import os
    os.replace(path + ".tmp", path)


def read_verifier(path):
    raw = open(path, "rb").read()
    magic, n, nk, rankm, bits, cs, pol = struct.unpack("<8sQIIIII", raw[:36])
    if magic != b"ANQIVF2\0":
        raise ValueError("bad verifier magic")
    coeff = np.frombuffer(raw[36:36 + n * rankm * 4], np.float32).reshape(n, rankm).copy()
    off = 36 + n * rankm * 4
    codes = np.frombuffer(raw[off:off + n * nk], np.uint8).reshape(n, nk).copy()
    return coeff, codes


def tree_postorder(left, right, offset, hyper, leaf_off, leaf_ids, root=0):
    """Renumber a tree so children precede parents (the reference's node
    order, rptrees.h:99-152; root last).  Returns per-node lists."""
    order = []
    stack = [(root, False)]
    while stack:
        k, done = stack.pop()
        if left[k] == U32MAX or done:
            order.append(k)
            continue
        stack.append((k, True))
        stack.append((int(right[k]), False))
        stack.append((int(left[k]), False))
    newid = {k: i for i, k in enumerate(order)}
    nodes = []
    for k in order:
        if left[k] == U32MAX:
            nodes.append((U32MAX, U32MAX, FLT_MAX, None, leaf_ids[leaf_off[k]:leaf_off[k + 1]]))
        else:
            nodes.append((newid[int(left[k])], newid[int(right[k])], np.float32(offset[k]), hyper[k], None))
    return nodes


def write_graph(path, n, dim, M, deg, links, nodes, leaf_size, construction_time=0.0):
    """save_index + save_tree layout (index.h:1810-1863)."""
    n_leaves = sum(1 for nd in nodes if nd[0] == U32MAX)
    with open(path, "wb") as f:
        f.write(struct.pack("<QQQf", n, dim, M, construction_time))
        deg = np.asarray(deg, np.uint32)
        links = np.asarray(links, np.uint32)
        buf = bytearray()
        for i in range(n):
            dg = int(deg[i])
            buf += struct.pack("<I", dg)
            buf += links[i, :dg].tobytes()
        f.write(bytes(buf))
        f.write(struct.pack("<QQQ", leaf_size, n_leaves, len(nodes)))
        for l, r, off, hp, idx in nodes:
            f.write(struct.pack("<QQf", l, r, float(off)))
            if hp is None:
                f.write(struct.pack("<Q", 0))
            else:
                f.write(struct.pack("<Q", dim))
                f.write(np.ascontiguousarray(hp, np.float32).tobytes())
            if idx is None:
                f.write(struct.pack("<Q", 0))
            else:
                f.write(struct.pack("<Q", len(idx)))
                f.write(np.ascontiguousarray(idx, np.uint32).tobytes())


def write_csr(path, off, ids, keys):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(off) - 1))
        f.write(np.asarray(off, np.uint32).tobytes())
        f.write(np.asarray(ids, np.uint32).tobytes())
        f.write(np.asarray(keys, np.float32).tobytes())


def read_csr(path):
    with open(path, "rb") as f:
        nq = struct.unpack("<I", f.read(4))[0]
        off = np.frombuffer(f.read((nq + 1) * 4), np.uint32).copy()
        ids = np.frombuffer(f.read(int(off[-1]) * 4), np.uint32).copy()
        keys = np.frombuffer(f.read(int(off[-1]) * 4), np.float32).copy()
    return off, ids, keys
