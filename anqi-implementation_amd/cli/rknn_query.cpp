// rknn_query -- drop-in for the reference's src/rknn_query.cpp on one MI355X.
//
// Same command line (L_DESCENT_LIST PREFIX [GRAPH_DEGREE] [LEAF_SIZE]
// [TREE_COUNT] [BUILD_EF] [QUERY_K], rknn_query.cpp:1547-1596), same input
// files (PREFIX_lifted.meta, _base.bin, _query.bin, _rknn_gt_k<K>.bin or
// _rknn_gt.bin, the AKNN .rec named by ANQI_PREBUILT_TOPK_REC, the rank-M basis
// PREFIX_radbasis_M8.bin and the u8 codebook), same stdout lines for the
// build and for every L_descent point, same exit codes (2 usage, 3 input).
//
// It implements the configuration that scripts/run_final.sh pins (any-k
// lifted graph, rank-8 + u8 LRQ floor verifier as a post-filter, no recheck,
// slack 0, LocalJoin 6, alpha 1.2, LIFTNNK 50).  Those ANQI_* variables may be
// unset or set to the pinned values; any other value is rejected (exit 2)
// rather than silently computing something else.
//
// Caches follow the reference's protocol (rknn_query.cpp:1841-1894, 2244-2297):
// the lifted vectors, lift meta, state file and graph index (generate_config
// name) under ANQI_GRAPH_CACHE_DIR, the verifier artifact under
// ANQI_VERIFIER_CACHE_DIR, all named by the same FNV keys, so a second run (at
// another k) loads the index instead of building it, and caches move between
// this tool and the reference's.
//
// Everything from the AKNN cutoffs on runs on the GPU through libanqi_hip.so:
// LRQ encode + threshold decode, lift, graph build, search tree, range search
// and post-filter.  Recall/precision/F1 and distance counts are computed on
// the host exactly as the reference does (rknn_query.cpp:2776-2798, 2891-2913).
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <cmath>
#include <iomanip>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "cache_keys.hpp"
#include "formats.hpp"

using namespace anqi_cli;

namespace {

anqi_ctx* g_ctx = nullptr;

[[noreturn]] void device_fail(const char* what) {
    std::fprintf(stderr, "[rknn] GPU failure in %s: %s\n", what, g_ctx ? anqi_last_error(g_ctx) : "no context");
    std::exit(3);
}

void ck(int rc, const char* what) {
    if (rc == ANQI_OK) return;
    if (rc == ANQI_E_USAGE) {
        std::fprintf(stderr, "[rknn] %s: %s\n", what, anqi_last_error(g_ctx));
        std::exit(2);
    }
    device_fail(what);
}

template <class T> T* dalloc(size_t count) {
    void* p = nullptr;
    ck(anqi_dmalloc(g_ctx, count * sizeof(T) + 16, &p), "device allocation");
    return static_cast<T*>(p);
}

template <class T> T* upload(const std::vector<T>& v) {
    T* p = dalloc<T>(v.size());
    ck(anqi_h2d(g_ctx, p, v.data(), v.size() * sizeof(T)), "upload");
    return p;
}

template <class T> void download(std::vector<T>& v, const T* p, size_t count) {
    v.resize(count);
    ck(anqi_d2h(g_ctx, v.data(), p, count * sizeof(T)), "download");
}

bool parse_size(const char* text, const char* name, size_t& out) {
    if (!text || !*text) {
        std::fprintf(stderr, "[rknn] invalid %s (expected a positive integer)\n", name);
        return false;
    }
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(text, &end, 10);
    if (errno == ERANGE || end == text || *end != '\0' || v == 0) {
        std::fprintf(stderr, "[rknn] invalid %s=%s (expected a positive integer)\n", name, text);
        return false;
    }
    out = (size_t)v;
    return true;
}

// The pinned configuration (scripts/run_final.sh exports): variable -> the
// accepted spellings.  Unset means the pinned value.
struct Pin {
    const char* name;
    std::vector<std::string> ok;
};
const std::vector<Pin>& pins() {
    static const std::vector<Pin> p = {
        {"ANQI_GRAPH_GEOMETRY", {"anyk_lift", "anyk", "lift"}},
        {"ANQI_RADIUS_MODE", {"rankm_resid_u8_lrq_floor"}},
        {"ANQI_VERIFIER_PLACEMENT", {"postfilter"}},
        {"ANQI_RANKM", {"8"}},
        {"ANQI_NK", {"100"}},
        {"ANQI_GRAPH_NK", {"100"}},
        {"ANQI_RANKM_SLACK", {"0", "0.0"}},
        {"ANQI_RECHECK", {"0", "false", "no", "off"}},
        {"ANQI_RQ_OBJECTIVE", {"radius_reconstruction"}},
        {"ANQI_ALPHA", {"1.2"}},
        {"ANQI_LIFTNNK", {"50"}},
        {"ANQI_GRAPHNND", {"6"}},
        {"ANQI_PREBUILT_TOPK_K", {"100"}},
        {"ANQI_SEARCH_SCORE", {"center"}},
        {"ANQI_BETA", {"0.5"}},
        {"ANQI_RCEPS", {"0", "0.0"}},
        {"ANQI_CANDIDATE_ENVELOPE", {"aknn", "aknn_k100", "default"}},
    };
    return p;
}

// A set variable is read with atoi / atof, even when empty (index.h:178-186).
int envi(const char* k, int d) {
    const char* v = std::getenv(k);
    return v ? std::atoi(v) : d;
}
double envd(const char* k, double d) {
    const char* v = std::getenv(k);
    return v ? std::atof(v) : d;
}

// A file of exactly `count` floats into v; the reference's edge-policy
// messages on failure (exit 2).
bool read_floats_exact(const std::string& path, size_t count, std::vector<float>& v, const char* what) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        std::fprintf(stderr, "[edge-policy] %s requested but missing %s\n", what, path.c_str());
        return false;
    }
    f.seekg(0, std::ios::end);
    const std::streamoff got = f.tellg();
    if (got < 0 || (size_t)got != count * sizeof(float)) {
        std::fprintf(stderr, "[edge-policy] %s size mismatch %s: expected %zu bytes, got %lld\n", what, path.c_str(),
                     count * sizeof(float), (long long)got);
        return false;
    }
    f.seekg(0, std::ios::beg);
    v.resize(count);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(count * sizeof(float)));
    if ((size_t)f.gcount() != count * sizeof(float)) {
        std::fprintf(stderr, "[edge-policy] short %s %s\n", what, path.c_str());
        return false;
    }
    return true;
}

// configure_edge_residual_cover (include/index.h:305-438): ANQI_EDGE_POLICY
// and its inputs as the lifted index reads them (folder: the graph cache
// directory).  The flags stay zero when the policy leaves the plain prune.
// Host copies of the tables in resid / mu.  Returns 0, or 2 after a message.
int configure_edge_policy(uint32_t n, uint32_t M, const std::string& folder, anqi_edge_policy& E,
                          std::vector<float>& resid, std::vector<float>& mu) {
    E = anqi_edge_policy{};
    E.lsg_alpha = 1.0; E.lambda = 1.0; E.scale2 = 1.0; E.occlude_all = 1;
    const char* pv = std::getenv("ANQI_EDGE_POLICY");
    if (!pv || !*pv) return 0;
    const std::string pol(pv);
    const bool lsg_policy = pol == "lsg_rng" || pol == "lsg_bubble_hybrid";
    if (lsg_policy) {
        E.lsg_alpha = envd("ANQI_LSG_ALPHA", 1.0);
        if (!std::isfinite(E.lsg_alpha) || E.lsg_alpha < 0.0) {
            std::fprintf(stderr, "[edge-policy] invalid ANQI_LSG_ALPHA=%.17g\n", E.lsg_alpha);
            return 2;
        }
        const char* mp = std::getenv("ANQI_LSG_MU_FILE");
        if (!mp || !*mp) {
            std::fprintf(stderr, "[edge-policy] lsg_rng requires ANQI_LSG_MU_FILE\n");
            return 2;
        }
        if (!read_floats_exact(mp, n, mu, "lsg mu table")) return 2;
        double mn = INFINITY, mx = 0.0;
        long double sum = 0.0;
        for (size_t i = 0; i < n; i++) {
            const double m = mu[i];
            if (!std::isfinite(m) || m <= 0.0) {
                std::fprintf(stderr, "[edge-policy] invalid lsg mu[%zu]=%.9g\n", i, m);
                return 2;
            }
            mn = std::min(mn, m);
            mx = std::max(mx, m);
            sum += m;
        }
        E.lsg = 1;
        std::fprintf(stderr, "[edge-policy] %s lsg enabled alpha=%.6g mu=%s min=%.6g mean=%.6g max=%.6g\n",
                     pol.c_str(), E.lsg_alpha, mp, mn, n ? (double)(sum / n) : 0.0, mx);
        if (pol == "lsg_rng") return 0;
    }
    static const char* const kResid[] = {"residual_cover", "ball_rng",      "ball_hybrid",   "ball_multik",
                                         "bubble_rng",     "bubble_hybrid", "bubble_multik", "bubble_auto",
                                         "bubble_auto_full", "lsg_bubble_hybrid"};
    if (std::find(std::begin(kResid), std::end(kResid), pol) == std::end(kResid)) return 0;
    uint32_t nk = (uint32_t)std::max(0, envi("ANQI_EDGE_NK", 100));
    if (nk == 0) nk = 100;
    E.budget = (uint32_t)std::min<int64_t>(std::max(0, envi("ANQI_EDGE_RESID_BUDGET", 16)), (int64_t)M);
    E.auto_candidates = (uint32_t)std::max(1, envi("ANQI_EDGE_AUTO_CANDIDATES", (int)M));
    E.auto_max_swaps = (uint32_t)std::max(1, envi("ANQI_EDGE_AUTO_MAX_SWAPS", (int)M));
    E.lambda = envd("ANQI_EDGE_RESID_LAMBDA", 1.0);
    {
        const char* kv = std::getenv("ANQI_EDGE_KS");
        std::stringstream ss(kv ? kv : "10,50,100");
        std::string tok;
        while (std::getline(ss, tok, ',')) {
            if (tok.empty()) continue;
            const int k = std::atoi(tok.c_str());
            if (k < 1 || (uint32_t)k > nk) continue;
            if (E.nks == 16) {
                std::fprintf(stderr, "[rknn] ANQI_EDGE_KS with more than 16 ranks is not supported by the GPU build\n");
                return 2;
            }
            E.ks[E.nks++] = k;
        }
    }
    if (E.nks == 0) return 0;
    const bool bubble_auto = pol == "bubble_auto" || pol == "bubble_auto_full";
    if (E.budget == 0 && pol != "ball_rng" && !bubble_auto) return 0;
    {
        std::ifstream mf(folder + "_anqi_meta.txt");
        double mstar = 0.0, scale = 1.0;
        if (mf >> mstar >> scale) E.scale2 = scale * scale;
    }
    const char* rf = std::getenv("ANQI_EDGE_RESID_FILE");
    const std::string rpath = rf ? std::string(rf) : folder + "_anqi_rankm_residual.bin";
    if (!read_floats_exact(rpath, (size_t)n * nk, resid, "residual table")) return 2;
    E.nk = nk;
    E.resid_cover = 1;
    if (pol != "residual_cover") {
        E.resid_cover = 0;
        E.ball = 1;
        E.ball_pure = pol == "ball_rng" || pol == "bubble_rng";
        const char* occ = std::getenv("ANQI_EDGE_BALL_OCCLUSION");
        E.occlude_all = !(occ && std::string(occ) == "any");
    }
    E.bubble_auto = bubble_auto;
    E.bubble_full = pol == "bubble_auto_full";
    std::string auto_note;
    if (bubble_auto)
        auto_note = " auto_candidates=" + (E.bubble_full ? std::string("all") : std::to_string(E.auto_candidates)) +
                    " auto_max_swaps=" + std::to_string(E.auto_max_swaps);
    std::fprintf(stderr, "[edge-policy] %s enabled ks=%s budget=%u lambda=%.4g scale2=%.6g nk=%u%s%s%s\n", pol.c_str(),
                 std::getenv("ANQI_EDGE_KS") ? std::getenv("ANQI_EDGE_KS") : "10,50,100", E.budget, E.lambda,
                 E.scale2, E.nk, E.ball ? " ball_occlusion=" : "", E.ball ? (E.occlude_all ? "all" : "any") : "",
                 auto_note.c_str());
    return 0;
}

// Base vectors: the fbin, or ANQI_BASE_U8BIN's uint8 file read as floats
// (load_base_vectors, rknn_query.cpp:55-123).  Exits 3 on a bad file.
void load_base(const std::string& PFX, std::vector<float>& base, uint32_t& n, uint32_t& d) {
    const char* u8 = std::getenv("ANQI_BASE_U8BIN");
    if (u8 && *u8) {
        const int rc = read_u8bin_as_float(u8, base, n, d);
        if (rc == 1) {
            std::fprintf(stderr, "[rknn] cannot read uint8 base header %s\n", u8);
            std::exit(3);
        }
        if (rc == 2) {
            std::fprintf(stderr, "[rknn] bad uint8 base %s: n=%u d=%u bytes=%lld expected=%llu\n", u8, n, d,
                         (long long)file_size(u8), (unsigned long long)(8 + (uint64_t)n * d));
            std::exit(3);
        }
        std::printf("[rknn] loaded uint8 base directly as float: %s n=%u d=%u %.3fGB\n", u8, n, d,
                    (double)n * d * 4 / 1e9);
        return;
    }
    if (!read_fbin(PFX + "_base.bin", base, n, d)) {
        std::fprintf(stderr, "[rknn] cannot read base vectors %s_base.bin\n", PFX.c_str());
        std::exit(3);
    }
}

// Warm-start ids from a stream of n rows of `stride` records -- {u32 id, f32
// distance}, or u32 ids alone (ids_only) -- keeping the first `width` ids of
// each row (nndindex::build_index's streamed warm start, include/index.h:
// 1127-1203): from a command's stdout, or from a file.  Returns false when
// the stream ends early, has extra data or the command fails.
bool read_warm_stream(FILE* in, bool ids_only, uint32_t n, uint32_t stride, uint32_t width,
                      std::vector<uint32_t>& ids) {
    ids.assign((size_t)n * width, 0u);
    const size_t rec = ids_only ? 4 : 8;
    std::vector<char> row((size_t)stride * rec);
    for (uint32_t i = 0; i < n; i++) {
        if (std::fread(row.data(), rec, stride, in) != stride) return false;
        for (uint32_t k = 0; k < width; k++) std::memcpy(&ids[(size_t)i * width + k], row.data() + k * rec, 4);
    }
    return std::fgetc(in) == EOF;
}

// ANQI_COMPACT_PREBUILT_TOPK_IDS_INPLACE: rewrites the .rec in place as its
// rows' first `width` ids (u32) and truncates it
// (compact_prebuilt_topk_ids_in_place, rknn_query.cpp:988-1057).
bool compact_rec_in_place(const std::string& path, uint32_t n, uint32_t source_k, uint32_t width) {
    std::FILE* f = std::fopen(path.c_str(), "r+b");
    if (!f) return false;
    std::vector<Rec> r(source_k);
    std::vector<uint32_t> ids(width);
    for (uint32_t i = 0; i < n; i++) {   // row i's ids land at or before its records: read before written
        if (std::fseek(f, (long)((uint64_t)i * source_k * sizeof(Rec)), SEEK_SET) ||
            std::fread(r.data(), sizeof(Rec), source_k, f) != source_k) {
            std::fclose(f);
            return false;
        }
        for (uint32_t k = 0; k < width; k++) {
            if (r[k].id >= n) {
                std::fprintf(stderr, "[rknn] invalid warm-start id at row=%u k=%u\n", i, k);
                std::exit(3);
            }
            ids[k] = r[k].id;
        }
        if (std::fseek(f, (long)((uint64_t)i * width * 4), SEEK_SET) ||
            std::fwrite(ids.data(), 4, width, f) != width) {
            std::fclose(f);
            return false;
        }
    }
    const bool ok = std::fflush(f) == 0 && ::ftruncate(fileno(f), (off_t)((uint64_t)n * width * 4)) == 0 &&
                    ::fsync(fileno(f)) == 0;
    return std::fclose(f) == 0 && ok;
}

void check_configuration() {
    if (std::getenv("ANQI_SPECIFIC")) {   // rknn_query.cpp:1806-1811
        std::fprintf(stderr, "[rknn] independent postfilter verifier requires a shared graph; unset ANQI_SPECIFIC\n");
        std::exit(2);
    }
    for (const Pin& p : pins()) {
        const char* v = std::getenv(p.name);
        if (!v || !*v) continue;
        if (std::find(p.ok.begin(), p.ok.end(), std::string(v)) == p.ok.end()) {
            std::fprintf(stderr,
                         "[rknn] %s=%s is not supported by the GPU build (supported: %s; this build implements the "
                         "configuration pinned by scripts/run_final.sh)\n",
                         p.name, v, p.ok[0].c_str());
            std::exit(2);
        }
    }
    if (std::getenv("ANQI_VAMANA")) {
        const int L = envi("ANQI_VAMANA_L", 128);
        if (L <= 0 || L > 1024) {
            std::fprintf(stderr, "[rknn] ANQI_VAMANA_L=%d is not supported by the GPU build (1..1024)\n", L);
            std::exit(2);
        }
    }
    const char* relax = std::getenv("ANQI_RADIUS_RELAX");   // off when it reads as 0 (rknn_query.cpp:841-845)
    if (relax && *relax && std::atof(relax) != 0.0) {
        std::fprintf(stderr, "[rknn] ANQI_RADIUS_RELAX is not supported by the GPU build\n");
        std::exit(2);
    }
}

double external_seconds(const char* name) {
    const char* raw = std::getenv(name);
    if (!raw || !*raw) return 0.0;
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(raw, &end);
    if (errno != 0 || end == raw || *end != '\0' || !std::isfinite(v) || v < 0.0) {
        std::fprintf(stderr, "[rknn] invalid %s=%s\n", name, raw);
        std::exit(2);
    }
    return v;
}

double positive_seconds_env(const char* name, double fallback) {
    const char* v = std::getenv(name);
    if (!v || !*v) return fallback;
    char* end = nullptr;
    errno = 0;
    const double p = std::strtod(v, &end);
    if (errno == ERANGE || end == v || *end != '\0' || !std::isfinite(p) || p <= 0.0) {
        std::fprintf(stderr, "[rknn] invalid %s=%s\n", name, v);
        std::exit(2);
    }
    return p;
}

std::string cache_dir(const char* env, const std::string& dflt) {
    const char* v = std::getenv(env);
    if (!v || !*v) return dflt;
    std::string s(v);
    if (s.back() != '/') s.push_back('/');
    return s;
}

double since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct PointStats {
    double sum_recall = 0, sum_prec = 0, sum_f1 = 0, sum_candidate_recall = 0;
    double sum_cand = 0, sum_pred = 0, sum_ndist = 0, sum_graph_ndist = 0, sum_verify_ndist = 0;
    size_t nonempty = 0, tasks = 0, batches = 0;
    double search_secs = 0;
    double secs = 0;
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3 || argc > 8) {
        std::fprintf(stderr,
                     "usage: %s L_DESCENT_LIST PREFIX [GRAPH_DEGREE] [LEAF_SIZE] [TREE_COUNT] [BUILD_EF] [QUERY_K]\n",
                     argv[0]);
        return 2;
    }
    std::vector<size_t> Lds;
    {
        std::string list(argv[1]);
        size_t start = 0;
        while (start <= list.size()) {
            const size_t comma = list.find(',', start);
            const std::string item = list.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
            size_t v = 0;
            if (!parse_size(item.c_str(), "L_DESCENT", v)) return 2;
            Lds.push_back(v);
            if (comma == std::string::npos) break;
            start = comma + 1;
        }
    }
    size_t Marg = 48, leafA = 1000, ntA = 64, efA = 400;
    if (argc > 3 && !parse_size(argv[3], "GRAPH_DEGREE", Marg)) return 2;
    if (argc > 4 && !parse_size(argv[4], "LEAF_SIZE", leafA)) return 2;
    if (argc > 5 && !parse_size(argv[5], "TREE_COUNT", ntA)) return 2;
    if (argc > 6 && !parse_size(argv[6], "BUILD_EF", efA)) return 2;
    const std::string PFX(argv[2]);
    const std::string meta_p = PFX + "_lifted.meta";
    std::string gt_p = PFX + "_rknn_gt.bin";

    double Mstar2_meta = 0.0, scale_meta = 0.0;
    size_t d0 = 0, kk = 0;
    std::string metric;
    {
        std::ifstream mf(meta_p);
        if (!(mf >> Mstar2_meta >> d0 >> kk >> metric >> scale_meta) || !std::isfinite(Mstar2_meta) ||
            Mstar2_meta < 0.0 || d0 == 0 || kk == 0 || (metric != "l2" && metric != "ip") ||
            !std::isfinite(scale_meta) || scale_meta <= 0.0) {
            std::fprintf(stderr, "[rknn] invalid or missing lift metadata: %s\n", meta_p.c_str());
            return 3;
        }
    }
    std::printf("[rknn] metric=%s M*^2=%.6g d=%zu k=%zu L_descent_points=%zu\n", metric.c_str(), Mstar2_meta, d0, kk,
                Lds.size());
    // "l2" or "ip": the lift coordinate, R2, the post-filter key and the
    // threshold clamp depend on it (rknn_query.cpp:751, 2227, 2555-2561, 2728)
    const int anqi_metric = metric == "ip" ? ANQI_METRIC_IP : ANQI_METRIC_L2;
    check_configuration();
    const size_t NK = 100, GRAPH_NK = 100;
    const int RANKM = 8;
    size_t Kq = kk;
    if (argc > 7 && !parse_size(argv[7], "QUERY_K", Kq)) return 2;
    if (Kq < 1 || Kq > NK) {
        std::fprintf(stderr, "[rknn] requested Kq=%zu is outside the represented rank range [1,%zu]\n", Kq, NK);
        return 2;
    }
    {
        std::string gp = PFX + "_rknn_gt_k" + std::to_string(Kq) + ".bin";
        if (exists(gp)) {
            gt_p = gp;
        } else if (Kq != kk) {
            std::fprintf(stderr,
                         "[rknn] missing k-specific reverse-kNN GT %s; the generic GT is only valid for metadata k=%zu\n",
                         gp.c_str(), kk);
            return 3;
        } else if (!exists(gt_p)) {
            std::fprintf(stderr, "[rknn] missing reverse-kNN GT %s\n", gt_p.c_str());
            return 3;
        }
    }
    std::printf("[rknn] shared-index query: Kq=%zu NK=%zu graph_NK=%zu GT=%s\n", Kq, NK, GRAPH_NK, gt_p.c_str());

    std::vector<float> base, query;
    uint32_t n = 0, d = 0, nq = 0, dq = 0;
    load_base(PFX, base, n, d);
    if (!read_fbin(PFX + "_query.bin", query, nq, dq)) {
        std::fprintf(stderr, "[rknn] cannot read query vectors %s_query.bin\n", PFX.c_str());
        return 3;
    }
    if (d != d0 || dq != d0) {
        std::fprintf(stderr, "[rknn] vector dimension does not match lift metadata: base=%u query=%u metadata=%zu\n", d,
                     dq, d0);
        return 3;
    }
    std::vector<uint32_t> goff, gids;
    switch (read_gt_csr(gt_p, nq, n, goff, gids)) {
        case 0: break;
        case 1: std::fprintf(stderr, "[rknn] invalid reverse-kNN GT header: %s\n", gt_p.c_str()); return 3;
        case 2: std::fprintf(stderr, "[rknn] invalid reverse-kNN GT offsets: %s\n", gt_p.c_str()); return 3;
        case 3: std::fprintf(stderr, "[rknn] invalid reverse-kNN GT size: %s\n", gt_p.c_str()); return 3;
        default: std::fprintf(stderr, "[rknn] invalid reverse-kNN GT ids: %s\n", gt_p.c_str()); return 3;
    }

    const std::string DSDIR = PFX.substr(0, PFX.rfind('/') + 1);
    const std::string graph_dir = cache_dir("ANQI_GRAPH_CACHE_DIR", DSDIR);
    const std::string verifier_dir = cache_dir("ANQI_VERIFIER_CACHE_DIR", DSDIR);
    const char* rec_env = std::getenv("ANQI_PREBUILT_TOPK_REC");
    const std::string rec_p = (rec_env && *rec_env) ? std::string(rec_env) : std::string();
    // warm-start id sources for a prebuilt .rec (rknn_query.cpp:2101-2135,
    // 2170-2190): a command streaming the rows, or the .rec compacted in place
    // to ids after the preparation and consumed by the warm start; or the .rec
    // removed once the cutoffs are read
    const char* stream_env = std::getenv("ANQI_PREBUILT_WARM_START_COMMAND");
    const std::string warm_cmd = (!rec_p.empty() && stream_env && *stream_env) ? std::string(stream_env) : "";
    const bool compact_ids = !rec_p.empty() && env_flag("ANQI_COMPACT_PREBUILT_TOPK_IDS_INPLACE");
    const bool release_rec = !rec_p.empty() && !compact_ids && env_flag("ANQI_RELEASE_PREBUILT_TOPK_AFTER_PREP");
    if (compact_ids && !warm_cmd.empty()) {
        std::fprintf(stderr, "[rknn] compact local warm-start IDs cannot be combined with "
                             "ANQI_PREBUILT_WARM_START_COMMAND\n");
        return 2;
    }
    const std::string basis_p = PFX + "_radbasis_M8.bin";
    const char* cbenv = std::getenv("ANQI_RQ_CODEBOOK");
    const std::string cb_p = (cbenv && *cbenv) ? std::string(cbenv) : PFX + "_learned_rq_codebook_M8_u8.bin";
    const uint32_t dl = d + 2;
    const auto tb0 = std::chrono::steady_clock::now();

    // cache keys, state and the reuse decision (rknn_query.cpp:1841-1894)
    const CacheNames cn =
        cache_names({PFX, graph_dir, verifier_dir, rec_p, basis_p, cb_p, Marg, leafA, ntA, efA, 50, n, dl});
    std::printf("[rknn] graph geometry=anyk_lift graph_rankm=0 verification_rankm=%d "
                "radius_mode=rankm_resid_u8_lrq_floor placement=postfilter\n", RANKM);
    std::printf("[rknn] graph cache key=%s\n", cn.graph_key.c_str());
    const bool can_load = exists(cn.state) && first_line(cn.state) == cn.state_key && exists(cn.marker) &&
                          exists(cn.approxlift) && exists(cn.meta) && exists(cn.graph_index);
    std::printf("[rknn] component caches: graph=%s verifier=%s\n", graph_dir.c_str(), verifier_dir.c_str());
    if (env_flag("ANQI_REQUIRE_GRAPH_CACHE") && !can_load) {
        std::fprintf(stderr, "[rknn] required graph cache is incomplete or incompatible: dir=%s key=%s\n",
                     graph_dir.c_str(), cn.graph_key.c_str());
        return 3;
    }
    const std::string radius_desc = cn.radius_key + ":postfilter";

    // rank-M basis and codebook (rknn_query.cpp:1896-1934, 407-454)
    std::vector<float> mu, G, cb;
    if (!read_basis(basis_p, NK, RANKM, mu, G)) {
        std::fprintf(stderr, "[rknn] invalid or missing rank-M basis %s\n", basis_p.c_str());
        return 3;
    }
    std::printf("[rknn] loaded rank-M basis: M=%d path=%s\n", RANKM, basis_p.c_str());
    int cbr = read_codebook(cb_p, NK, RANKM, 8, cb);
    if (cbr) {
        std::fprintf(stderr, "[rknn] invalid learned residual codebook %s (code %d)\n", cb_p.c_str(), cbr);
        return 3;
    }
    const uint32_t code_size = (uint32_t)(cb.size() / NK);
    // CompactLrqVerifier::persistent_bytes (rknn_query.cpp:631-636): payload,
    // mean, basis, codebook and the 68 header bytes
    const double verifier_bytes =
        (double)n * (RANKM * 4 + NK) + 4.0 * ((double)NK + (double)RANKM * NK + (double)NK * code_size) + 68.0;

    // a verifier artifact with this radius key is reused whatever the graph
    // cache holds (rknn_query.cpp:1999-2012)
    std::vector<float> coeff;
    std::vector<uint8_t> codes;
    bool verifier_ready = false;
    double verifier_load_ms = 0.0, encode_save_ms = 0.0;
    if (exists(cn.verifier)) {
        const auto tl = std::chrono::steady_clock::now();
        if (!read_verifier(cn.verifier, n, NK, RANKM, coeff, codes)) {
            std::fprintf(stderr, "[rknn] invalid LRQ verifier artifact %s\n", cn.verifier.c_str());
            return 3;
        }
        verifier_load_ms = since(tl) * 1e3;
        verifier_ready = true;
        std::printf("[rknn] loaded independent compact verifier mode=rank-M+LRQ bits=8 %s (persistent=%.3fGB "
                    "load_ms=%.3f)\n", cn.verifier.c_str(), verifier_bytes / 1e9, verifier_load_ms);
    }

    // HIP runtime start-up is left out of the build time
    const auto tc0 = std::chrono::steady_clock::now();
    ck(anqi_ctx_create(0, &g_ctx), "context");
    const double ctx_s = since(tc0);

    float* d_base = upload(base);
    float* d_mu = upload(mu);
    float* d_G = upload(G);
    float* d_cb = upload(cb);
    float* d_coeff = dalloc<float>((size_t)n * RANKM);
    uint8_t* d_codes_km = dalloc<uint8_t>((size_t)n * NK);
    if (verifier_ready) {
        ck(anqi_h2d(g_ctx, d_coeff, coeff.data(), coeff.size() * 4), "upload");
        uint8_t* d_codes = upload(codes);
        ck(anqi_transpose_codes(g_ctx, d_codes, n, NK, d_codes_km), "code transpose");
        ck(anqi_ctx_sync(g_ctx), "sync");
        anqi_dfree(g_ctx, d_codes);
    }

    // AKNN source: the prebuilt .rec (rknn_query.cpp:890-1010, 1101-1133) or a
    // GPU build; needed to build the graph or an absent verifier
    std::vector<uint32_t> aknn_ids;
    std::vector<float> aknn_keys;
    uint32_t* d_aknn = nullptr;
    float* d_keys = nullptr;
    double source_prepare_s = 0.0;
    auto load_source = [&]() -> int {
        if (d_keys) return 0;
        const auto t0 = std::chrono::steady_clock::now();
        std::printf("[rknn] original-space AKNN: leaf=%zu n_trees=%zu ef=%zu metric=%s\n", leafA, ntA, efA,
                    metric.c_str());
        if (!rec_p.empty()) {
            if (file_size(rec_p) != (uint64_t)n * 100 * 8) {
                std::fprintf(stderr, "[rknn] bad prebuilt top-k table %s: bytes=%lld expected=%llu\n", rec_p.c_str(),
                             (long long)(exists(rec_p) ? (long long)file_size(rec_p) : -1LL),
                             (unsigned long long)((uint64_t)n * 100 * 8));
                return 3;
            }
            if (!read_rec(rec_p, n, 100, aknn_ids, aknn_keys)) {
                std::fprintf(stderr, "[rknn] cannot read prebuilt top-k %s\n", rec_p.c_str());
                return 3;
            }
            for (size_t t = 0; t < aknn_ids.size(); t++)
                if (aknn_ids[t] >= n) {
                    std::fprintf(stderr, "[rknn] invalid prebuilt top-k id at row=%zu\n", t / 100);
                    return 3;
                }
            if (compact_ids)
                std::printf("[rknn] deferred original KNN IDs to compact local graph warm-start (source_k=100 "
                            "width=100)\n");
            else if (!warm_cmd.empty())
                std::printf("[rknn] deferred original KNN IDs to streamed graph warm-start (source_k=100 "
                            "width=100)\n");
            else
                std::printf("[rknn] loaded contiguous prebuilt original KNN %s (n=%u source_k=100 width=100)\n",
                            rec_p.c_str(), n);
            d_aknn = upload(aknn_ids);
            d_keys = upload(aknn_keys);
        } else {
            anqi_aknn_params ap{100, (uint32_t)leafA, (uint32_t)ntA, (uint32_t)efA, 8, 100, 10, 0, anqi_metric};
            anqi_aknn_stats st{};
            d_aknn = dalloc<uint32_t>((size_t)n * 100);
            d_keys = dalloc<float>((size_t)n * 100);
            ck(anqi_aknn_build(g_ctx, d_base, n, d, &ap, d_aknn, d_keys, &st), "AKNN source");
            // .rec keys: max(0, d2), or the "ip" distance minus 1 (tools/build_aknn_source.cpp:119-137)
            ck(anqi_aknn_cutoffs_metric(g_ctx, d_keys, n, 100, anqi_metric, d_keys, nullptr), "AKNN cutoffs");
            download(aknn_keys, d_keys, (size_t)n * 100);
            std::printf("[rknn] built original-space AKNN on the GPU (n=%u K=100 in %.2fs)\n", n, st.total_s);
        }
        std::printf("[rknn] original-space AKNN warm start ready: rows=%u source=%s\n", n,
                    rec_p.empty() ? "gpu" : (compact_ids || !warm_cmd.empty() ? "prebuilt-stream"
                                                                              : "prebuilt-contiguous"));
        if (!rec_p.empty()) {
            std::printf("[rknn] loaded AKNN cutoffs from interleaved top-k %s (n=%u K=%zu)\n", rec_p.c_str(), n, NK);
            std::printf("[rknn] loaded AKNN cutoff column from interleaved top-k %s (n=%u k=%zu)\n", rec_p.c_str(), n,
                        GRAPH_NK);
        }
        source_prepare_s = since(t0);
        return 0;
    };
    // verifier: rank-8 coefficients + u8 floor codes (rknn_query.cpp:603-753),
    // saved under the radius key (ensure_compact_lrq, :2028-2049)
    auto encode_verifier = [&]() -> int {
        if (verifier_ready) return 0;
        const auto te = std::chrono::steady_clock::now();
        uint8_t* d_codes = dalloc<uint8_t>((size_t)n * NK);
        ck(anqi_lrq_encode(g_ctx, d_keys, n, NK, RANKM, d_mu, d_G, d_cb, code_size, d_coeff, d_codes, d_codes_km),
           "LRQ encode");
        download(coeff, d_coeff, (size_t)n * RANKM);
        download(codes, d_codes, (size_t)n * NK);
        anqi_dfree(g_ctx, d_codes);
        if (!write_verifier(cn.verifier, coeff, codes, n, NK, RANKM)) {
            std::fprintf(stderr, "[rknn] failed writing verifier %s\n", cn.verifier.c_str());
            return 3;
        }
        encode_save_ms = since(te) * 1e3;
        verifier_ready = true;
        std::printf("[rknn] built independent compact verifier mode=rank-M+LRQ bits=8 %s (persistent=%.3fGB "
                    "encode_save_ms=%.3f)\n", cn.verifier.c_str(), verifier_bytes / 1e9, encode_save_ms);
        return 0;
    };

    // memory released while the graph is loaded or built and read back after
    // (ANQI_RELEASE_VERIFIER_DURING_GRAPH_INDEX / ANQI_RELEASE_BASE_DURING_GRAPH_INDEX,
    // rknn_query.cpp:2257-2280, 2303-2320): the verifier payload (host and
    // device) reloaded from its artifact, the base vectors from their file
    bool released = false, verifier_released = false, base_released = false;
    auto release_for_index = [&]() {
        if (released) return;
        released = true;
        if (env_flag("ANQI_RELEASE_VERIFIER_DURING_GRAPH_INDEX") && verifier_ready) {
            const double bytes = (double)coeff.capacity() * 4 + (double)codes.capacity();
            std::vector<float>().swap(coeff);
            std::vector<uint8_t>().swap(codes);
            anqi_dfree(g_ctx, d_coeff);
            anqi_dfree(g_ctx, d_codes_km);
            d_coeff = nullptr;
            d_codes_km = nullptr;
            verifier_ready = false;
            verifier_released = true;
            std::printf("[rknn] released compact verifier payload during graph index construction: %.3fGB\n",
                        bytes / 1e9);
        }
        if (env_flag("ANQI_RELEASE_BASE_DURING_GRAPH_INDEX")) {
            std::vector<float>().swap(base);
            anqi_dfree(g_ctx, d_base);
            d_base = nullptr;
            base_released = true;
            std::printf("[rknn] released original base before graph index construction: %.3fGB\n",
                        (double)n * d * 4 / 1e9);
        }
    };
    auto reload_after_index = [&]() -> int {
        if (verifier_released) {
            if (!read_verifier(cn.verifier, n, NK, RANKM, coeff, codes)) {
                std::fprintf(stderr, "[rknn] invalid LRQ verifier artifact %s\n", cn.verifier.c_str());
                return 3;
            }
            d_coeff = upload(coeff);
            d_codes_km = dalloc<uint8_t>((size_t)n * NK);
            uint8_t* d_codes = upload(codes);
            ck(anqi_transpose_codes(g_ctx, d_codes, n, NK, d_codes_km), "code transpose");
            ck(anqi_ctx_sync(g_ctx), "sync");
            anqi_dfree(g_ctx, d_codes);
            verifier_ready = true;
            std::printf("[rknn] reloaded compact verifier after graph index construction: %.3fGB\n",
                        verifier_bytes / 1e9);
        }
        if (base_released) {
            uint32_t rn = 0, rd = 0;
            load_base(PFX, base, rn, rd);
            if (rn != n || rd != d) {
                std::fprintf(stderr, "[rknn] reloaded base shape mismatch: got n=%u d=%u expected n=%u d=%u\n", rn,
                             rd, n, d);
                return 3;
            }
            d_base = upload(base);
            std::printf("[rknn] reloaded original base after graph index construction: %.3fGB\n",
                        (double)n * d * 4 / 1e9);
        }
        return 0;
    };

    double mstar2 = 0, scale = 0;
    float* d_lifted = nullptr;
    uint32_t* d_links = dalloc<uint32_t>((size_t)n * Marg);
    uint32_t* d_deg = dalloc<uint32_t>(n);
    anqi_tree tree{};
    double lift_verifier_prepare_s = 0.0, graph_index_s = 0.0;
    const std::string index_name = cn.graph_index.substr(graph_dir.size());
    bool graph_loaded = false;
    if (can_load) {
        // cached lift and graph (rknn_query.cpp:2052-2077; index.h:1826-1860)
        {
            std::ifstream f(cn.meta);
            if (!(f >> mstar2 >> scale)) {
                std::fprintf(stderr, "[rknn] bad meta cache %s\n", cn.meta.c_str());
                return 3;
            }
        }
        std::vector<float> lifted;
        uint32_t ln = 0, ld = 0;
        if (!read_fbin(cn.approxlift, lifted, ln, ld) || ln != n || ld != dl) {
            std::fprintf(stderr, "[rknn] bad lifted vector cache %s\n", cn.approxlift.c_str());
            return 3;
        }
        d_lifted = upload(lifted);
        if (!verifier_ready)
            if (int rc = load_source()) return rc;
        if (int rc = encode_verifier()) return rc;
        std::printf("[rknn] loaded cached index (M=%zu): M*^2=%.6g scale=%.4g radius=%s\n", Marg, mstar2, scale,
                    radius_desc.c_str());
        std::printf("[rknn] graph degree M=%zu\n", Marg);
        std::printf("Index config: %s\n", index_name.c_str());
        release_for_index();
        std::vector<uint32_t> hdeg, hlinks, left, right, entry;
        std::vector<float> offset, hyper;
        if (read_graph(cn.graph_index, n, dl, Marg, hdeg, hlinks, left, right, offset, hyper, entry)) {
            std::printf("Loading index from: %s\n", cn.graph_index.c_str());
            for (uint64_t i = 0; i < n; i++)   // device rows: slots past deg hold UINT32_MAX
                for (uint64_t j = hdeg[i]; j < Marg; j++) hlinks[i * Marg + j] = UINT32_MAX;
            ck(anqi_h2d(g_ctx, d_links, hlinks.data(), hlinks.size() * 4), "upload");
            ck(anqi_h2d(g_ctx, d_deg, hdeg.data(), hdeg.size() * 4), "upload");
            ck(anqi_tree_upload(g_ctx, (uint32_t)left.size(), (uint32_t)left.size() - 1, dl, left.data(),
                                right.data(), offset.data(), hyper.data(), entry.data(), &tree),
               "tree upload");
            graph_loaded = true;
        } else {
            std::fprintf(stderr, "incompatible or truncated cached index: %s\n", cn.graph_index.c_str());
        }
    } else {
        if (int rc = load_source()) return rc;
        const auto tv0 = std::chrono::steady_clock::now();
        if (int rc = encode_verifier()) return rc;
        std::printf("[rknn] RADIUS_MODE=%s conceptual_radius_store=%.3fGB\n", radius_desc.c_str(),
                    verifier_bytes / 1e9);
        // lift (rknn_query.cpp:2203-2236) with the AKNN 100th key as horizon
        std::vector<float> rk(n);
        for (uint32_t i = 0; i < n; i++) rk[i] = aknn_keys[(size_t)i * 100 + (GRAPH_NK - 1)];
        std::vector<float> lifted;
        if (anqi_cli::gpu_form_chosen()) {
            // device lift (integer soft binary64), the same bits as the host lift
            float* d_rk = upload(rk);
            d_lifted = dalloc<float>((size_t)n * dl);
            ck(anqi_lift_metric(g_ctx, d_base, n, d, d_rk, anqi_metric, d_lifted, &mstar2, &scale), "lift");
            anqi_dfree(g_ctx, d_rk);
            download(lifted, d_lifted, (size_t)n * dl);
        } else {
            lifted.resize((size_t)n * dl);
            ck(anqi_lift_host_metric(base.data(), n, d, rk.data(), anqi_metric, lifted.data(), &mstar2, &scale),
               "lift");
            d_lifted = upload(lifted);
        }
        // lift, meta and state files (rknn_query.cpp:2231-2245)
        if (!write_fbin(cn.approxlift, lifted.data(), n, dl))
            std::fprintf(stderr, "[rknn] cannot write lifted vector cache %s\n", cn.approxlift.c_str());
        { std::ofstream f(cn.meta); f << std::setprecision(17) << mstar2 << " " << scale << "\n"; }
        { std::ofstream f(cn.state); f << cn.state_key << "\n"; }
        lift_verifier_prepare_s = since(tv0);
        std::printf("[rknn] BUILD lift M*^2=%.6g scale=%.4g (graph radius=aknn_k100; verifier=%s)\n", mstar2, scale,
                    radius_desc.c_str());
        std::printf("[rknn] graph degree M=%zu\n", Marg);
        std::printf("Index config: %s\n", index_name.c_str());
    }
    // choose_vamana_entry (index.h:190-222) over the lifted rows: the Vamana
    // batch style's entry and, with ANQI_QUERY_SINGLE_ENTRY or
    // ANQI_VAMANA_QUERY_ENTRY, every query's (:1606-1607)
    const bool single_entry = std::getenv("ANQI_QUERY_SINGLE_ENTRY") || std::getenv("ANQI_VAMANA_QUERY_ENTRY");
    uint32_t graph_entry = 0;
    if (single_entry || std::getenv("ANQI_VAMANA")) {
        const char* ev = std::getenv("ANQI_VAMANA_ENTRY");
        const std::string sv = ev ? ev : "";
        if (sv == "medoid" || sv == "centroid") {
            const uint64_t sample = (uint64_t)(size_t)envi("ANQI_VAMANA_ENTRY_SAMPLE", (int)std::min<uint64_t>(n, 100000));
            ck(anqi_vamana_medoid(g_ctx, d_lifted, n, dl, sample, &graph_entry), "vamana entry");
        } else if (!sv.empty()) {
            char* endp = nullptr;
            const unsigned long long id = std::strtoull(ev, &endp, 10);
            graph_entry = (endp != ev && n > 0) ? (uint32_t)(id % n) : 0;
        }
    }
    if (!graph_loaded) {
        // graph + search tree (index.h:1071-1295), saved with its marker (:1293;
        // rknn_query.cpp:2297)
        if (int rc = load_source()) return rc;
        release_for_index();
        // the warm start's ids: streamed, read back from the compacted .rec, or
        // the source's own (index.h:1127-1218)
        uint32_t* d_warm = d_aknn;
        if (!warm_cmd.empty() || compact_ids) {
            if (compact_ids) {
                if (!compact_rec_in_place(rec_p, n, 100, 100)) {
                    std::fprintf(stderr, "[rknn] cannot finalize compact warm-start IDs %s\n", rec_p.c_str());
                    return 3;
                }
                std::printf("[rknn] compacted regenerable prebuilt top-k to IDs in place: %s source_k=100 width=100 "
                            "bytes=%.3fGB\n", rec_p.c_str(), (double)n * 100 * 4 / 1e9);
            }
            std::FILE* in = compact_ids ? std::fopen(rec_p.c_str(), "rb") : ::popen(warm_cmd.c_str(), "r");
            if (!in) {
                std::fprintf(stderr, "[rknn] cannot start streamed warm-start %s\n",
                             compact_ids ? rec_p.c_str() : "command");
                return 3;
            }
            std::vector<uint32_t> wids;
            const bool ok = read_warm_stream(in, compact_ids, n, 100, 100, wids);
            const int st = compact_ids ? std::fclose(in) : ::pclose(in);
            if (!ok || st != 0) {
                std::fprintf(stderr, "[rknn] streamed warm-start ended early, produced extra data, or failed\n");
                return 3;
            }
            if (compact_ids && ::unlink(rec_p.c_str()) != 0) {
                std::fprintf(stderr, "[rknn] cannot release consumed streamed warm-start file %s\n", rec_p.c_str());
                return 3;
            }
            d_warm = upload(wids);
        } else if (release_rec) {
            if (::unlink(rec_p.c_str()) != 0) {
                std::fprintf(stderr, "[rknn] cannot release regenerable prebuilt top-k %s: %s\n", rec_p.c_str(),
                             std::strerror(errno));
                return 3;
            }
            std::printf("[rknn] released regenerable prebuilt top-k after verifier/lift preparation: %s\n",
                        rec_p.c_str());
        }
        const auto tg0 = std::chrono::steady_clock::now();
        // ANQI_REUSE_ORIG_KNN: no lifted LocalJoin (index.h:1202-1238); ANQI_VAMANA:
        // the refinement passes after the commit (:1262-1282)
        anqi_graph_params gp{(uint32_t)Marg, 50, 100, 100, 8, std::getenv("ANQI_REUSE_ORIG_KNN") ? 0 : 6, 20, 1.2f, 10,
                             anqi_cli::gpu_form_chosen() ? 1 : 0};
        if (std::getenv("ANQI_VAMANA")) {
            const char* sv = std::getenv("ANQI_VAMANA_STYLE");
            const std::string style = sv ? sv : "global";
            gp.vamana = (style == "global" || style == "old") ? 1 : 2;
            gp.vamana_L = (uint32_t)envi("ANQI_VAMANA_L", 128);
            gp.vamana_passes = envi("ANQI_VAMANA_PASS", 2);
            gp.vamana_batch_frac = envd("ANQI_VAMANA_BATCH_FRAC", 0.02);
            gp.vamana_max_batch = (uint64_t)(size_t)envi("ANQI_VAMANA_MAX_BATCH", 1000000);
            gp.vamana_batch_base = envd("ANQI_VAMANA_BATCH_BASE", 2.0);
            const char* ea = std::getenv("ANQI_VAMANA_EARLY_ALPHA");
            gp.vamana_early_alpha = ea ? (float)std::atof(ea) : 1.0f;
            gp.vamana_early_alpha_one = envi("ANQI_VAMANA_EARLY_ALPHA_ONE", 1) != 0 ? 1 : 0;
            gp.vamana_entry = graph_entry;
            if (gp.vamana == 2)
                std::fprintf(stderr, "[vamana-batch] entry=%u L=%u pass=%d batch_frac=%.4g max_batch=%llu base=%.3g "
                                     "alpha=%.3g rev=csr\n", graph_entry, gp.vamana_L, gp.vamana_passes,
                             gp.vamana_batch_frac > 0 ? gp.vamana_batch_frac : 0.02,
                             (unsigned long long)std::min<uint64_t>(
                                 (uint64_t)std::max(1.0, (gp.vamana_batch_frac > 0 ? gp.vamana_batch_frac : 0.02) *
                                                             (double)n), gp.vamana_max_batch),
                             gp.vamana_batch_base > 1.0 ? gp.vamana_batch_base : 2.0, 1.2);
        }
        // the edge policy of every prune (configure_edge_residual_cover runs
        // when the index is built, index.h:1079)
        anqi_edge_policy edge{};
        std::vector<float> edge_resid, edge_mu;
        if (int rc = configure_edge_policy(n, (uint32_t)Marg, graph_dir, edge, edge_resid, edge_mu)) return rc;
        float* d_edge_resid = edge_resid.empty() ? nullptr : upload(edge_resid);
        float* d_edge_mu = edge_mu.empty() ? nullptr : upload(edge_mu);
        edge.resid = d_edge_resid;
        edge.mu = d_edge_mu;
        if (edge.resid_cover || edge.ball || edge.lsg) gp.edge = &edge;
        anqi_graph_stats gst{};
        ck(anqi_graph_build(g_ctx, d_lifted, n, dl, d_warm, 100, &gp, d_links, d_deg, &tree, &gst), "graph build");
        if (d_edge_resid) anqi_dfree(g_ctx, d_edge_resid);
        if (d_edge_mu) anqi_dfree(g_ctx, d_edge_mu);
        if (d_warm != d_aknn) anqi_dfree(g_ctx, d_warm);
        std::vector<uint32_t> hdeg, hlinks, left(tree.nodes), right(tree.nodes);
        std::vector<float> offset(tree.nodes), hyper((size_t)tree.nodes * dl);
        download(hdeg, d_deg, n);
        download(hlinks, d_links, (size_t)n * Marg);
        ck(anqi_d2h(g_ctx, left.data(), tree.left, left.size() * 4), "download");
        ck(anqi_d2h(g_ctx, right.data(), tree.right, right.size() * 4), "download");
        ck(anqi_d2h(g_ctx, offset.data(), tree.offset, offset.size() * 4), "download");
        ck(anqi_d2h(g_ctx, hyper.data(), tree.hyper, hyper.size() * 4), "download");
        if (write_graph(cn.graph_index, n, dl, Marg, hdeg, hlinks, left, right, offset, hyper, tree.leaf_off,
                        tree.leaf_ids, tree.root, 100, (float)gst.total_s)) {
            std::printf("Index saved to: %s\n", cn.graph_index.c_str());
            std::ofstream(cn.marker) << "1\n";
        } else {
            std::fprintf(stderr, "[rknn] cannot write graph cache %s\n", cn.graph_index.c_str());
        }
        graph_index_s = since(tg0);
    }
    if (int rc = reload_after_index()) return rc;
    float* d_bnorm = dalloc<float>(n);
    ck(anqi_norms(g_ctx, d_base, n, d, d_bnorm), "norms");
    ck(anqi_ctx_sync(g_ctx), "sync");
    const double build_s = since(tb0) - ctx_s;
    const double idx_gb = ((double)n * Marg * 4 + (double)n * dl * 4 + verifier_bytes) / 1e9;
    if (can_load && graph_loaded) {
        std::printf("[rknn] INDEX_LOAD_TIME=%.1fs INDEX_SIZE=%.2fGB (graph M=%zu + graph dim=%u + verification=%.3fGB "
                    "+ envelope=%.3fGB)\n", build_s, idx_gb, Marg, dl, verifier_bytes / 1e9, 0.0);
        if (env_flag("ANQI_BUILD_BREAKDOWN"))
            std::printf("[rknn] LOAD_BREAKDOWN index_load_s=%.6f verifier_load_ms=%.3f\n", build_s, verifier_load_ms);
    } else {
        const double aknn_build_s = external_seconds("ANQI_AKNN_BUILD_SECONDS");
        const double verifier_train_s = external_seconds("ANQI_VERIFIER_TRAIN_SECONDS");
        const double end_to_end = aknn_build_s + verifier_train_s + build_s;
        std::printf("[rknn] BUILD_TIME=%.1fs INDEX_SIZE=%.2fGB (aknn_source=%.1fs + verifier_training=%.1fs + "
                    "index=%.1fs, graph M=%zu + graph dim=%u + verification=%.3fGB + envelope=%.3fGB)\n",
                    end_to_end, idx_gb, aknn_build_s, verifier_train_s, build_s, Marg, dl, verifier_bytes / 1e9, 0.0);
        if (env_flag("ANQI_BUILD_BREAKDOWN"))
            std::printf("[rknn] BUILD_BREAKDOWN aknn_source_s=%.6f verifier_train_s=%.6f source_prepare_s=%.6f "
                        "lift_verifier_prepare_s=%.6f graph_index_s=%.6f index_total_s=%.6f end_to_end_s=%.6f "
                        "verifier_encode_save_ms=%.3f\n",
                        aknn_build_s, verifier_train_s, source_prepare_s, lift_verifier_prepare_s, graph_index_s,
                        build_s, end_to_end, encode_save_ms);
    }
    std::fflush(stdout);

    // thresholds for Kq (materialize_k, rknn_query.cpp:720-753, slack 1)
    const auto tm0 = std::chrono::steady_clock::now();
    float* d_thr = dalloc<float>(n);
    const uint32_t kq32 = (uint32_t)Kq;
    ck(anqi_lrq_materialize_metric(g_ctx, d_coeff, d_codes_km, n, NK, RANKM, d_mu, d_G, d_cb, code_size, &kq32, 1, 1.0,
                                   anqi_metric, d_thr),
       "threshold decode");
    ck(anqi_ctx_sync(g_ctx), "sync");
    std::printf("[rknn] independent verifier ready: mode=rank-M+LRQ k=%zu slack=0 persistent=%.3fGB "
                "temporary_thresholds=%.3fGB materialize_ms=%.3f bits=8\n",
                Kq, verifier_bytes / 1e9, n * 4 / 1e9, since(tm0) * 1e3);
    std::printf("[rknn] k-adaptive flood: Kq=%zu radj=off beta=0.50\n", Kq);
    std::printf("[rknn] candidate generator: AKNN-k100 envelope, center graph search; verifier is post-filter only\n");
    std::printf("[rknn] search score: mode=center lambda=1.0\n");
    std::printf("[rknn] query visited mode=lds-hash aggregate_store=0.000GB\n");

    size_t bench_repeats = 1;
    if (const char* r = std::getenv("ANQI_BENCH_REPEATS")) {
        size_t v = 0;
        if (!parse_size(r, "ANQI_BENCH_REPEATS", v)) return 2;
        bench_repeats = v;
    }
    const bool calibrated = env_flag("ANQI_BENCH_CALIBRATED_TIMING");
    const double min_total_s = positive_seconds_env("ANQI_BENCH_MIN_TOTAL_S", 30.0);
    const double min_repeat_s = positive_seconds_env("ANQI_BENCH_MIN_REPEAT_S", 10.0);
    std::printf("[rknn] benchmark repeats=%zu calibrated=%s (timing uses steady_clock sub-millisecond precision)\n",
                bench_repeats, calibrated ? "on" : "off");
    std::printf("[rknn] query parallelism: nq=%u device=MI355X mode=%s\n", nq,
                anqi_cli::gpu_form_chosen() ? "persistent" : "per-query");
    if (calibrated)
        std::printf("[rknn] calibrated timing: min_total_s=%.3f min_repeat_s=%.3f windows=%zu\n", min_total_s,
                    min_repeat_s, bench_repeats);
    else
        std::printf("[rknn] flattened_repeats=1 tasks=%zu\n", (size_t)nq * bench_repeats);

    anqi_index ix{d_base, d_bnorm, d_lifted, d_deg, d_links, n, d, dl, (uint32_t)Marg, &tree, mstar2, scale,
                  nullptr, nullptr, nullptr, nullptr, anqi_metric};
    if (anqi_cli::gpu_form_chosen() && d % 16 == 0) {
        // persistent search: aligned main rows, precomputed lift terms and
        // sentinel adjacency rows
        float* d_lmain = dalloc<float>((size_t)n * d);
        float* d_ltail = dalloc<float>(n);
        uint32_t* d_plinks = dalloc<uint32_t>((size_t)n * Marg);
        ck(anqi_lift_split(g_ctx, d_lifted, n, d, d_lmain, d_ltail), "lift split");
        ck(anqi_pack_links(g_ctx, d_links, d_deg, n, (uint32_t)Marg, d_plinks), "sentinel adjacency");
        ix.lmain = d_lmain;
        ix.ltail = d_ltail;
        ix.plinks = d_plinks;
        // byte rows when every base value is a float 0..255 (certified)
        uint8_t* d_brows = dalloc<uint8_t>((size_t)n * d);
        int exact = 0;
        ck(anqi_byte_rows(g_ctx, d_base, d_lmain, n, d, scale, d_brows, &exact), "byte rows");
        if (exact) ix.brows = d_brows;
        else anqi_dfree(g_ctx, d_brows);
    }
    float* d_query = upload(query);
    uint32_t out_cap = 4096;
    uint32_t* d_cand = nullptr;
    uint8_t* d_acc = nullptr;
    uint32_t* d_cnt = dalloc<uint32_t>(nq);
    uint32_t* d_nd = dalloc<uint32_t>(nq);
    uint32_t* d_pred = dalloc<uint32_t>(nq);
    uint32_t* d_status = dalloc<uint32_t>(nq);
    auto alloc_out = [&](uint32_t cap) {
        if (d_cand) anqi_dfree(g_ctx, d_cand);
        if (d_acc) anqi_dfree(g_ctx, d_acc);
        out_cap = cap;
        d_cand = dalloc<uint32_t>((size_t)nq * cap);
        d_acc = dalloc<uint8_t>((size_t)nq * cap);
    };
    alloc_out(out_cap);
    // candidates on the host as CSR: coff (nq + 1), cids / cacc (coff[nq])
    std::vector<uint32_t> cnt, nd, coff, cids;
    std::vector<uint8_t> cacc;
    uint32_t* d_coff = dalloc<uint32_t>((size_t)nq + 1);
    uint32_t* d_cids = nullptr;
    uint8_t* d_cacc = nullptr;
    uint64_t ccap = 0;
    // one batch: search + post-filter on the device; the candidates (ids and
    // accept bytes) compacted on the device and copied to the host
    uint32_t* d_entries = nullptr;   // single-entry queries: every query starts at graph_entry
    if (single_entry) d_entries = upload(std::vector<uint32_t>(nq, graph_entry));
    auto search = [&](size_t L) {
        for (;;) {
            if (d_entries)
                ck(anqi_rknn_search_entries(g_ctx, &ix, d_query, nq, (uint32_t)L, nullptr, (uint32_t)L, d_thr,
                                            nullptr, d_entries, out_cap, d_cand, d_acc, d_cnt, d_nd, d_pred, nullptr,
                                            d_status, anqi_cli::gpu_form_chosen() ? 1 : 0),
                   "range search");
            else
                ck(anqi_rknn_search(g_ctx, &ix, d_query, nq, (uint32_t)L, nullptr, (uint32_t)L, d_thr, nullptr,
                                    out_cap, d_cand, d_acc, d_cnt, d_nd, d_pred, nullptr, d_status,
                                    anqi_cli::gpu_form_chosen() ? 1 : 0),
                   "range search");
            download(cnt, d_cnt, nq);
            uint32_t mx = *std::max_element(cnt.begin(), cnt.end());
            if (mx <= out_cap) break;
            alloc_out(mx + 1024);
        }
        download(nd, d_nd, nq);
        uint64_t total = 0;
        for (uint32_t c : cnt) total += c;
        if (total > ccap) {
            if (d_cids) anqi_dfree(g_ctx, d_cids);
            if (d_cacc) anqi_dfree(g_ctx, d_cacc);
            ccap = total + total / 4 + 1024;
            d_cids = dalloc<uint32_t>(ccap);
            d_cacc = dalloc<uint8_t>(ccap);
        }
        ck(anqi_compact_candidates(g_ctx, nq, out_cap, d_cand, d_acc, d_cnt, d_coff, d_cids, d_cacc, ccap, &total),
           "candidate compaction");
        download(coff, d_coff, (size_t)nq + 1);
        download(cids, d_cids, total);
        download(cacc, d_cacc, total);
    };
    std::vector<std::vector<uint32_t>> truth(nq);
    for (uint32_t i = 0; i < nq; i++) {
        truth[i].assign(gids.begin() + goff[i], gids.begin() + goff[i + 1]);
        std::sort(truth[i].begin(), truth[i].end());
        truth[i].erase(std::unique(truth[i].begin(), truth[i].end()), truth[i].end());
    }
    // QPS as the reference times it (src/rknn_query.cpp:2705-2806): each
    // batch's search and its ground-truth evaluation inside the timer;
    // search_secs keeps the search and candidate copy alone
    auto run_point = [&](size_t L, size_t batches) {
        PointStats s;
        const auto t0 = std::chrono::steady_clock::now();
        for (size_t b = 0; b < batches; b++) {
            const auto ts = std::chrono::steady_clock::now();
            search(L);
            s.search_secs += since(ts);
            for (uint32_t i = 0; i < nq; i++) {
                const auto& T = truth[i];
                if (T.empty()) continue;
                s.nonempty++;
                const uint32_t c = cnt[i];
                std::vector<uint32_t> pred;
                size_t chit = 0;
                for (uint32_t t = 0; t < c; t++) {
                    const uint32_t id = cids[coff[i] + t];
                    if (std::binary_search(T.begin(), T.end(), id)) chit++;
                    if (cacc[coff[i] + t]) pred.push_back(id);
                }
                std::sort(pred.begin(), pred.end());
                pred.erase(std::unique(pred.begin(), pred.end()), pred.end());
                size_t hit = 0;
                for (uint32_t id : pred) hit += std::binary_search(T.begin(), T.end(), id);
                const double rec = (double)hit / T.size();
                const double prec = pred.empty() ? 1.0 : (double)hit / pred.size();
                s.sum_recall += rec;
                s.sum_prec += prec;
                s.sum_f1 += (rec + prec > 0 ? 2 * rec * prec / (rec + prec) : 0);
                s.sum_candidate_recall += (double)chit / T.size();
                s.sum_cand += c;
                s.sum_pred += pred.size();
                s.sum_ndist += (double)nd[i] + c;
                s.sum_graph_ndist += nd[i];
                s.sum_verify_ndist += c;
            }
        }
        s.secs = since(t0);
        s.tasks = (size_t)nq * batches;
        s.batches = batches;
        return s;
    };
    auto print_point = [&](size_t L, const PointStats& s, double qps) {
        if (s.nonempty == 0 || s.tasks == 0 || s.secs <= 0.0) {
            std::fprintf(stderr, "[rknn] empty benchmark statistics for L_descent=%zu\n", L);
            std::exit(3);
        }
        const double dn = (double)s.nonempty;
        std::printf("[rknn] L_descent=%-5zu | recall=%.4f precision=%.4f F1=%.4f | avg cand=%.0f pred=%.0f #dist=%.0f | "
                    "QPS=%.0f (parallel)\n",
                    L, s.sum_recall / dn, s.sum_prec / dn, s.sum_f1 / dn, s.sum_cand / dn, s.sum_pred / dn,
                    s.sum_ndist / dn, qps);
        std::printf("[rknn] DIST_BREAKDOWN L_descent=%zu #graph_dist=%.0f #verify_dist=%.0f #total_dist=%.0f\n", L,
                    s.sum_graph_ndist / dn, s.sum_verify_ndist / dn, s.sum_ndist / dn);
        std::printf("[rknn] CANDIDATE_RECALL L_descent=%zu recall=%.6f\n", L, s.sum_candidate_recall / dn);
        if (s.search_secs > 0.0)
            std::printf("[rknn] SEARCH_QPS L_descent=%zu qps=%.0f (device search, post-filter and candidate copy; "
                        "ground-truth evaluation excluded)\n", L, (double)s.tasks / s.search_secs);
        std::fflush(stdout);
    };
    for (size_t L : Lds) {
        if (!calibrated) {
            PointStats s = run_point(L, bench_repeats);
            print_point(L, s, (double)s.tasks / s.secs);
            continue;
        }
        const PointStats warm = run_point(L, 1);
        const PointStats calib = run_point(L, 1);
        const double target = std::max(min_repeat_s, min_total_s / (double)bench_repeats);
        const size_t planned = std::max<size_t>(1, (size_t)std::ceil(target / std::max(calib.secs, 1e-9)));
        PointStats agg;
        std::vector<double> qs;
        double min_obs = 1e300;
        for (size_t r = 0; r < bench_repeats; r++) {
            size_t batches = planned;
            PointStats m;
            for (int attempt = 0; attempt < 4; attempt++) {
                m = run_point(L, batches);
                if (m.secs + 1e-9 >= target) break;
                size_t next = std::max(batches + 1, (size_t)std::ceil((double)batches * target / std::max(m.secs, 1e-9) * 1.05));
                std::printf("[rknn] HOT_RETRY L_descent=%zu repeat=%zu seconds=%.6f target_s=%.6f batches=%zu "
                            "next_batches=%zu\n", L, r + 1, m.secs, target, batches, next);
                batches = next;
            }
            if (m.secs + 1e-9 < target) {
                std::fprintf(stderr, "[rknn] calibrated repeat too short for L_descent=%zu repeat=%zu seconds=%.6f "
                             "target_s=%.6f\n", L, r + 1, m.secs, target);
                return 3;
            }
            const double q = (double)m.tasks / m.secs;
            qs.push_back(q);
            min_obs = std::min(min_obs, m.secs);
            agg.sum_recall += m.sum_recall; agg.sum_prec += m.sum_prec; agg.sum_f1 += m.sum_f1;
            agg.sum_candidate_recall += m.sum_candidate_recall; agg.sum_cand += m.sum_cand;
            agg.sum_pred += m.sum_pred; agg.sum_ndist += m.sum_ndist; agg.sum_graph_ndist += m.sum_graph_ndist;
            agg.sum_verify_ndist += m.sum_verify_ndist; agg.nonempty += m.nonempty; agg.tasks += m.tasks;
            agg.batches += m.batches; agg.secs += m.secs; agg.search_secs += m.search_secs;
            std::printf("[rknn] HOT_REPEAT L_descent=%zu repeat=%zu/%zu batches=%zu tasks=%zu seconds=%.6f QPS=%.3f\n",
                        L, r + 1, bench_repeats, m.batches, m.tasks, m.secs, q);
        }
        std::vector<double> sq = qs;
        std::sort(sq.begin(), sq.end());
        const size_t mid = sq.size() / 2;
        const double med = sq.size() % 2 ? sq[mid] : 0.5 * (sq[mid - 1] + sq[mid]);
        print_point(L, agg, med);
        std::printf("[rknn] HOT_TIMING L_descent=%zu warm_s=%.6f calibration_s=%.6f measured_s=%.6f repeats=%zu "
                    "planned_batches_per_repeat=%zu actual_batches_total=%zu min_repeat_s=%.6f qps_median=%.3f "
                    "qps_samples=", L, warm.secs, calib.secs, agg.secs, bench_repeats, planned, agg.batches, min_obs, med);
        for (size_t i = 0; i < qs.size(); i++) std::printf("%s%.3f", i ? "," : "", qs[i]);
        std::printf("\n");
    }
    anqi_tree_free(g_ctx, &tree);
    anqi_ctx_destroy(g_ctx);
    return 0;
}
