// formats.hpp -- the reference's on-disk formats for the drop-in tools
// (little-endian, no padding; layouts per tools/build_aknn_source.cpp:19-23,
// src/rknn_query.cpp:23-53, 407-454, 755-838, 1896-1934, include/index.h:
// 1810-1863, tools/dataset_gt.cpp:302-313, tools/train_final_m8_u8.py:176-187).
#pragma once
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/anqi.h"

namespace anqi_cli {

// GPU form of the kernels (same outputs either way): ANQI_GPU_FORM=plain
// selects the FP32 sweeps, per-chunk gathers and per-query search launch;
// anything else (default) the bf16-MFMA sweeps, batched gathers and
// persistent register-beam search.
inline bool gpu_form_chosen() {
    const char* v = std::getenv("ANQI_GPU_FORM");
    return !(v && std::strcmp(v, "plain") == 0);
}

[[noreturn]] inline void die(int code, const char* fmt, const char* a = "", const char* b = "") {
    std::fprintf(stderr, fmt, a, b);
    std::fprintf(stderr, "\n");
    std::exit(code);
}

inline uint64_t file_size(const std::string& p) {
    struct stat st {};
    if (::stat(p.c_str(), &st) != 0) return UINT64_MAX;
    return (uint64_t)st.st_size;
}

inline bool exists(const std::string& p) { return file_size(p) != UINT64_MAX; }

// fbin: [u32 n][u32 d][n*d f32], size checked exactly (rknn_query.cpp:23-53).
inline bool read_fbin(const std::string& p, std::vector<float>& x, uint32_t& n, uint32_t& d) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return false;
    if (!f.read((char*)&n, 4) || !f.read((char*)&d, 4) || n == 0 || d == 0) return false;
    uint64_t values = (uint64_t)n * d;
    if (file_size(p) != 8 + values * 4) return false;
    x.resize(values);
    return (bool)f.read((char*)x.data(), values * 4);
}

// uint8 base file [u32 n][u32 d][n*d u8] read as floats (load_u8bin_as_float,
// src/rknn_query.cpp:55-107): 0 ok, 1 cannot open or read the header, 2 bad
// size; the values are float(u8).
inline int read_u8bin_as_float(const std::string& p, std::vector<float>& x, uint32_t& n, uint32_t& d) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return 1;
    if (!f.read((char*)&n, 4) || !f.read((char*)&d, 4)) return 1;
    const uint64_t values = (uint64_t)n * d;
    if (n == 0 || d == 0 || file_size(p) != 8 + values) return 2;
    std::vector<uint8_t> b(values);
    if (!f.read((char*)b.data(), (std::streamsize)values)) return 1;
    x.resize(values);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)values; i++) x[(size_t)i] = (float)b[(size_t)i];
    return 0;
}

inline bool write_fbin(const std::string& p, const float* x, uint32_t n, uint32_t d) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f.write((const char*)&n, 4);
    f.write((const char*)&d, 4);
    f.write((const char*)x, (std::streamsize)((uint64_t)n * d * 4));
    return (bool)f;
}

struct Rec { uint32_t id; float key; };
static_assert(sizeof(Rec) == 8, "record layout");

inline bool read_rec(const std::string& p, uint32_t n, uint32_t k, std::vector<uint32_t>& ids,
                     std::vector<float>& keys) {
    if (file_size(p) != (uint64_t)n * k * 8) return false;
    std::ifstream f(p, std::ios::binary);
    std::vector<Rec> r((size_t)n * k);
    if (!f.read((char*)r.data(), (std::streamsize)(r.size() * 8))) return false;
    ids.resize(r.size());
    keys.resize(r.size());
    for (size_t t = 0; t < r.size(); t++) { ids[t] = r[t].id; keys[t] = r[t].key; }
    return true;
}

// [i32 K][i32 M][f32 mu[K]][f32 G[M][K]]
inline bool read_basis(const std::string& p, uint32_t K, uint32_t M, std::vector<float>& mu,
                       std::vector<float>& G) {
    if (file_size(p) != 8 + 4ull * K + 4ull * M * K) return false;
    std::ifstream f(p, std::ios::binary);
    int32_t kb = 0, mb = 0;
    f.read((char*)&kb, 4);
    f.read((char*)&mb, 4);
    if (!f || (uint32_t)kb != K || (uint32_t)mb != M) return false;
    mu.resize(K);
    G.resize((size_t)M * K);
    f.read((char*)mu.data(), 4ull * K);
    f.read((char*)G.data(), 4ull * M * K);
    return (bool)f;
}

// "ANQIRQ1\0"[i32 nk][i32 rankm][i32 bits][i32 code_size][f32 cb[nk][code_size]]
inline int read_codebook(const std::string& p, uint32_t nk, uint32_t rankm, uint32_t bits,
                         std::vector<float>& cb) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return 1;
    char magic[8];
    f.read(magic, 8);
    if (f.gcount() != 8 || std::memcmp(magic, "ANQIRQ1\0", 8) != 0) return 2;
    int32_t h[4];
    f.read((char*)h, 16);
    if (!f || (uint32_t)h[0] != nk || (uint32_t)h[1] != rankm || (uint32_t)h[2] != bits || h[3] <= 0) return 3;
    cb.resize((size_t)nk * h[3]);
    f.read((char*)cb.data(), (std::streamsize)(cb.size() * 4));
    if (f.gcount() != (std::streamsize)(cb.size() * 4)) return 4;
    char extra;
    if (f.read(&extra, 1)) return 5;
    for (uint32_t k = 0; k < nk; k++)
        for (int32_t c = 1; c < h[3]; c++)
            if (cb[(size_t)k * h[3] + c] < cb[(size_t)k * h[3] + c - 1]) return 6;
    return 0;
}

// "ANQIVF2\0"[u64 n][u32 nk][u32 rankm][u32 bits][u32 code_size][u32 policy]
// [f32 coeff[n][rankm]][u8 codes[n*nk]]  (rknn_query.cpp:755-838), atomic publish.
inline bool write_verifier(const std::string& p, const std::vector<float>& coeff,
                           const std::vector<uint8_t>& codes, uint64_t n, uint32_t nk, uint32_t rankm) {
    std::string tmp = p + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        const char magic[8] = {'A', 'N', 'Q', 'I', 'V', 'F', '2', '\0'};
        uint32_t h[5] = {nk, rankm, 8u, 256u, 1u};
        f.write(magic, 8);
        f.write((const char*)&n, 8);
        f.write((const char*)h, 20);
        f.write((const char*)coeff.data(), (std::streamsize)(coeff.size() * 4));
        f.write((const char*)codes.data(), (std::streamsize)codes.size());
        if (!f) return false;
    }
    return std::rename(tmp.c_str(), p.c_str()) == 0;
}

inline bool read_verifier(const std::string& p, uint64_t n, uint32_t nk, uint32_t rankm,
                          std::vector<float>& coeff, std::vector<uint8_t>& codes) {
    if (file_size(p) != 36 + n * rankm * 4 + n * nk) return false;
    std::ifstream f(p, std::ios::binary);
    char magic[8];
    uint64_t fn = 0;
    uint32_t h[5];
    f.read(magic, 8);
    f.read((char*)&fn, 8);
    f.read((char*)h, 20);
    if (!f || std::memcmp(magic, "ANQIVF2\0", 8) != 0 || fn != n || h[0] != nk || h[1] != rankm || h[2] != 8 ||
        h[3] != 256 || h[4] != 1)
        return false;
    coeff.resize(n * rankm);
    codes.resize(n * nk);
    f.read((char*)coeff.data(), (std::streamsize)(coeff.size() * 4));
    f.read((char*)codes.data(), (std::streamsize)codes.size());
    return (bool)f;
}

// RkNN ground truth CSR (dataset_gt.cpp:302-313; validated as in
// rknn_query.cpp:1690-1718).
inline int read_gt_csr(const std::string& p, uint32_t nq, uint32_t n, std::vector<uint32_t>& off,
                       std::vector<uint32_t>& ids) {
    std::ifstream f(p, std::ios::binary);
    uint32_t gnq = 0;
    if (!f || !f.read((char*)&gnq, 4) || gnq != nq) return 1;
    off.resize((size_t)nq + 1);
    if (!f.read((char*)off.data(), off.size() * 4) || off[0] != 0) return 2;
    for (size_t i = 1; i < off.size(); i++)
        if (off[i] < off[i - 1]) return 2;
    if (file_size(p) != 4 + off.size() * 4 + (uint64_t)off.back() * 8) return 3;
    ids.resize(off.back());
    if (!f.read((char*)ids.data(), ids.size() * 4)) return 4;
    for (uint32_t id : ids)
        if (id >= n) return 4;
    return 0;
}

// Graph cache in the reference's save_index/save_tree layout
// (include/index.h:1810-1863): nodes renumbered so children precede parents.
inline bool write_graph(const std::string& p, uint64_t n, uint64_t dim, uint64_t M, const std::vector<uint32_t>& deg,
                        const std::vector<uint32_t>& links, const std::vector<uint32_t>& left,
                        const std::vector<uint32_t>& right, const std::vector<float>& offset,
                        const std::vector<float>& hyper, const uint64_t* leaf_off, const uint32_t* leaf_ids,
                        uint32_t root, uint64_t leaf_size, float construction_time) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    if (!f) return false;
    f.write((const char*)&n, 8);
    f.write((const char*)&dim, 8);
    f.write((const char*)&M, 8);
    f.write((const char*)&construction_time, 4);
    for (uint64_t i = 0; i < n; i++) {
        f.write((const char*)&deg[i], 4);
        f.write((const char*)&links[i * M], 4ull * deg[i]);
    }
    // post-order numbering
    const uint32_t nodes = (uint32_t)left.size();
    std::vector<uint32_t> order, newid(nodes);
    std::vector<std::pair<uint32_t, bool>> st{{root, false}};
    while (!st.empty()) {
        auto [k, done] = st.back();
        st.pop_back();
        if (left[k] == UINT32_MAX || done) { newid[k] = (uint32_t)order.size(); order.push_back(k); continue; }
        st.push_back({k, true});
        st.push_back({right[k], false});
        st.push_back({left[k], false});
    }
    uint64_t n_leaves = 0;
    for (uint32_t k : order) n_leaves += left[k] == UINT32_MAX;
    uint64_t ns = order.size();
    f.write((const char*)&leaf_size, 8);
    f.write((const char*)&n_leaves, 8);
    f.write((const char*)&ns, 8);
    for (uint32_t k : order) {
        bool leaf = left[k] == UINT32_MAX;
        uint64_t l = leaf ? UINT32_MAX : newid[left[k]], r = leaf ? UINT32_MAX : newid[right[k]];
        float off = leaf ? 3.402823466e+38f : offset[k];
        f.write((const char*)&l, 8);
        f.write((const char*)&r, 8);
        f.write((const char*)&off, 4);
        uint64_t hs = leaf ? 0 : dim;
        f.write((const char*)&hs, 8);
        if (!leaf) f.write((const char*)&hyper[(size_t)k * dim], 4 * dim);
        uint64_t is = leaf ? leaf_off[k + 1] - leaf_off[k] : 0;
        f.write((const char*)&is, 8);
        if (is) f.write((const char*)&leaf_ids[leaf_off[k]], 4 * is);
    }
    return (bool)f;
}

// Graph cache reader: adjacency and the tree in host arrays (root = last node).
inline bool read_graph(const std::string& p, uint64_t n, uint64_t dim, uint64_t M, std::vector<uint32_t>& deg,
                       std::vector<uint32_t>& links, std::vector<uint32_t>& left, std::vector<uint32_t>& right,
                       std::vector<float>& offset, std::vector<float>& hyper, std::vector<uint32_t>& entry) {
    std::ifstream f(p, std::ios::binary);
    uint64_t fn = 0, fd = 0, fM = 0;
    float ct = 0;
    if (!f.read((char*)&fn, 8) || !f.read((char*)&fd, 8) || !f.read((char*)&fM, 8) || !f.read((char*)&ct, 4))
        return false;
    if (fn != n || fd != dim || fM != M) return false;
    deg.assign(n, 0);
    links.assign(n * M, 0);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t d = 0;
        if (!f.read((char*)&d, 4) || d > M || !f.read((char*)&links[i * M], 4ull * d)) return false;
        for (uint32_t j = 0; j < d; j++)
            if (links[i * M + j] >= n) return false;
        deg[i] = d;
    }
    uint64_t leaf_size = 0, n_leaves = 0, ns = 0;
    if (!f.read((char*)&leaf_size, 8) || !f.read((char*)&n_leaves, 8) || !f.read((char*)&ns, 8) || ns == 0 ||
        ns > 2 * n)
        return false;
    left.assign(ns, UINT32_MAX);
    right.assign(ns, UINT32_MAX);
    offset.assign(ns, 0);
    hyper.assign(ns * dim, 0);
    entry.assign(ns, 0);
    for (uint64_t k = 0; k < ns; k++) {
        uint64_t l, r, hs, is;
        float o;
        if (!f.read((char*)&l, 8) || !f.read((char*)&r, 8) || !f.read((char*)&o, 4) || !f.read((char*)&hs, 8) ||
            hs > dim)
            return false;
        if (hs && !f.read((char*)&hyper[k * dim], 4 * hs)) return false;
        if (!f.read((char*)&is, 8) || is > n) return false;
        std::vector<uint32_t> idx(is);
        if (is && !f.read((char*)idx.data(), 4 * is)) return false;
        bool leaf = l == UINT32_MAX && r == UINT32_MAX;
        if (!leaf && !(l < k && r < k && hs == dim)) return false;
        if (leaf && idx.empty()) return false;
        left[k] = leaf ? UINT32_MAX : (uint32_t)l;
        right[k] = leaf ? UINT32_MAX : (uint32_t)r;
        offset[k] = o;
        entry[k] = leaf ? idx[0] : 0;
    }
    return (bool)f;
}

inline uint64_t fnv1a(const void* data, size_t len, uint64_t h = 1469598103934665603ull) {
    const unsigned char* p = (const unsigned char*)data;
    for (size_t i = 0; i < len; i++) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}

inline std::string hex16(uint64_t v) {
    char b[17];
    std::snprintf(b, sizeof(b), "%016llx", (unsigned long long)v);
    return b;
}

// size + FNV-1a of a file (rknn_query.cpp:208-224 fingerprint).
inline std::string fingerprint(const std::string& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) return p + ":missing";
    uint64_t h = 1469598103934665603ull, bytes = 0;
    std::vector<char> buf(1 << 20);
    while (f) {
        f.read(buf.data(), (std::streamsize)buf.size());
        std::streamsize got = f.gcount();
        h = fnv1a(buf.data(), (size_t)got, h);
        bytes += (uint64_t)got;
    }
    return p + ":sz" + std::to_string(bytes) + ":fnv" + hex16(h);
}

inline const char* env_or(const char* k, const char* d) {
    const char* v = std::getenv(k);
    return (v && *v) ? v : d;
}

inline bool env_flag(const char* k) {
    const char* v = std::getenv(k);
    if (!v || !*v) return false;
    std::string s(v);
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s != "0" && s != "false" && s != "no" && s != "off";
}

}  // namespace anqi_cli
