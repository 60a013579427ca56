// cache_keys.hpp -- the reference's cache keys and cache file names for the
// configuration rknn_query implements, so that a run reuses (and hands over)
// the graph, lift and verifier caches exactly as src/rknn_query.cpp does:
//
//   graph key      graph_cache_key (rknn_query.cpp:301-341) plus the
//                  post-filter graph-cutoff suffix (:1843-1854), sanitized
//   radius key     radius_cache_key, rank-M residual branch (:1217-1242)
//   file names     :1855-1861 (lift, meta, state), :1870-1873 (marker and
//                  state key), :1999 (verifier), include/index.h:1704-1770
//                  (generate_config: the graph index itself)
//
// Keys are FNV-1a 64 over bytes; file identities are stat() fields, so a
// rewritten input invalidates every cache built from it.
#pragma once
#include <sys/stat.h>

#include <string>

#include "formats.hpp"

namespace anqi_cli {

// path:dev<>:ino<>:sz<>:mt<>:ct<> (rknn_query.cpp:226-234)
inline std::string file_stat_identity(const std::string& p) {
    struct stat st {};
    if (::stat(p.c_str(), &st) != 0) return p + ":missing";
    return p + ":dev" + std::to_string((uint64_t)st.st_dev) + ":ino" + std::to_string((uint64_t)st.st_ino) +
           ":sz" + std::to_string((uint64_t)st.st_size) + ":mt" + std::to_string((int64_t)st.st_mtime) +
           ":ct" + std::to_string((int64_t)st.st_ctime);
}

// 16 hex digits of the FNV-1a 64 of a string (rknn_query.cpp:236-243)
inline std::string hash16(const std::string& s) { return hex16(fnv1a(s.data(), s.size())); }

// every byte outside [0-9A-Za-z_-] becomes '_' (rknn_query.cpp:271-278)
inline std::string sanitize_tag(std::string s) {
    for (char& c : s)
        if (!((c >= '0' && c <= '9') || (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || c == '-' || c == '_'))
            c = '_';
    return s;
}

// Present at all (even empty): the reference tests these with getenv alone.
inline bool env_set(const char* k) { return std::getenv(k) != nullptr; }

struct CacheInputs {
    std::string pfx;           // dataset prefix
    std::string graph_dir;     // ANQI_GRAPH_CACHE_DIR with a trailing '/'
    std::string verifier_dir;  // ANQI_VERIFIER_CACHE_DIR with a trailing '/'
    std::string rec;           // ANQI_PREBUILT_TOPK_REC; empty: AKNN source built on the GPU
    std::string basis, codebook;
    size_t M, leaf, trees, ef;  // command-line graph degree and AKNN parameters
    size_t lift_nnk;            // ANQI_LIFTNNK (50)
    uint64_t n;
    uint32_t dim;               // lifted dimension d + 2
};

struct CacheNames {
    std::string threshold_source, radius_key, graph_key, state_key;
    std::string approxlift, meta, state, marker, graph_index, verifier;
};

// The base vectors' file: ANQI_BASE_U8BIN when set, else the fbin
// (base_vector_source, rknn_query.cpp:109-112).
inline std::string base_vector_source(const std::string& fbin_path) {
    const char* u8 = std::getenv("ANQI_BASE_U8BIN");
    return (u8 && *u8) ? std::string(u8) : fbin_path;
}

// Identity of the AKNN cutoffs (threshold_source_identity, rknn_query.cpp:
// 1059-1063).  Without a prebuilt .rec the reference reads the cutoffs from
// PFX_baseknn_aknn.bin; this tool builds them on the GPU from the base and
// the AKNN parameters instead, so its identity names those, and no cache of
// a reference run is taken for one of these runs or the other way round.
inline std::string threshold_source_identity(const CacheInputs& in, const std::string& pfx) {
    if (in.rec.empty())
        return file_stat_identity(base_vector_source(pfx + "_base.bin")) + ":gpu_aknn:leaf" + std::to_string(in.leaf) +
               ":nt" + std::to_string(in.trees) + ":ef" + std::to_string(in.ef) + ":k100";
    return file_stat_identity(in.rec) + ":interleaved_k" + env_or("ANQI_PREBUILT_TOPK_K", "100");
}
inline std::string threshold_source_identity(const CacheInputs& in) { return threshold_source_identity(in, in.pfx); }

// The edge-policy part of the graph key (rknn_query.cpp:314-324).
inline std::string edge_suffix() {
    if (!env_set("ANQI_EDGE_POLICY")) return "";
    std::string k = std::string("_edge") + env_or("ANQI_EDGE_POLICY", "none") + "_eks" +
                    env_or("ANQI_EDGE_KS", "10,50,100") + "_enk" + env_or("ANQI_EDGE_NK", "100") + "_erb" +
                    env_or("ANQI_EDGE_RESID_BUDGET", "16") + "_erl" + env_or("ANQI_EDGE_RESID_LAMBDA", "1.0") +
                    "_ebo" + env_or("ANQI_EDGE_BALL_OCCLUSION", "all");
    if (env_set("ANQI_EDGE_RESID_FILE")) k += "_erf" + hash16(fingerprint(env_or("ANQI_EDGE_RESID_FILE", "")));
    return k;
}

// The edge-policy part of the index name (generate_config, include/index.h:1745-1767).
inline std::string edge_config_suffix() {
    if (!env_set("ANQI_EDGE_POLICY")) return "";
    const std::string pol = env_or("ANQI_EDGE_POLICY", "none");
    std::string c = "_edge_" + pol;
    if (pol == "lsg_rng" || pol == "lsg_bubble_hybrid") {
        std::string mu_key = env_or("ANQI_LSG_CACHE_KEY", "");
        if (mu_key.empty()) mu_key = env_or("ANQI_LSG_MU_FILE", "");
        c += std::string("_lsa_") + env_or("ANQI_LSG_ALPHA", "1.0") + "_lsgmu_" + hash16(mu_key);
    }
    if (pol != "lsg_rng") {
        c += std::string("_eks_") + env_or("ANQI_EDGE_KS", "10,50,100") + "_enk_" + env_or("ANQI_EDGE_NK", "100") +
             "_erb_" + env_or("ANQI_EDGE_RESID_BUDGET", "16") + "_erl_" + env_or("ANQI_EDGE_RESID_LAMBDA", "1.0") +
             "_ebo_" + env_or("ANQI_EDGE_BALL_OCCLUSION", "all");
        if (pol == "bubble_auto" || pol == "bubble_auto_full")
            c += std::string("_eac_") + env_or("ANQI_EDGE_AUTO_CANDIDATES", "M") + "_eas_" +
                 env_or("ANQI_EDGE_AUTO_MAX_SWAPS", "M");
    }
    return c;
}

// The Vamana part of the graph key (rknn_query.cpp:325-340): the style and
// parameters as the environment spells them, "_vam0" when it is off.
inline std::string vamana_suffix() {
    if (!env_set("ANQI_VAMANA")) return "_vam0";
    const std::string style = env_or("ANQI_VAMANA_STYLE", "global");
    std::string k = "_vam1_style" + style + "_L" + env_or("ANQI_VAMANA_L", "128") + "_P" +
                    env_or("ANQI_VAMANA_PASS", "2") + "_bf" + env_or("ANQI_VAMANA_BATCH_FRAC", "0.02") + "_mb" +
                    env_or("ANQI_VAMANA_MAX_BATCH", "1000000") + "_bb" + env_or("ANQI_VAMANA_BATCH_BASE", "2") +
                    "_ea" + env_or("ANQI_VAMANA_EARLY_ALPHA", "1.0") + "_eao" +
                    env_or("ANQI_VAMANA_EARLY_ALPHA_ONE", "1") + "_entry" + env_or("ANQI_VAMANA_ENTRY", "0");
    if (style == "batch") k += "_implcsr1";
    return k;
}

inline CacheNames cache_names(const CacheInputs& in) {
    CacheNames c;
    c.threshold_source = threshold_source_identity(in);
    // radius_cache_key, rank-M residual modes with a learned codebook
    const char* bias = std::getenv("ANQI_RESID_BIAS");
    const char* basis_id = std::getenv("ANQI_RQ_BASIS_ID");
    const char* cb_id = std::getenv("ANQI_RQ_CODEBOOK_ID");
    const std::string src_id = env_set("ANQI_RADIUS_SOURCE_ID")
                                   ? std::string(env_or("ANQI_RADIUS_SOURCE_ID", "missing_radius_source_id"))
                                   : c.threshold_source;
    c.radius_key = std::string("rankm_resid_u8_lrq_floor:M") + env_or("ANQI_RANKM", "8") + ":b" + (bias ? bias : "0") +
                   ":obj" + env_or("ANQI_RQ_OBJECTIVE", "radius_reconstruction") + ":basis" +
                   ((basis_id && *basis_id) ? std::string(basis_id) : fingerprint(in.basis)) + ":cb" +
                   ((cb_id && *cb_id) ? std::string(cb_id) : fingerprint(in.codebook)) + ":src" + hash16(src_id) +
                   ":rqpred_v3";
    // graph_cache_key for the any-k lifted geometry, graph rank 0
    std::string g = "M" + std::to_string(in.M) + "_leaf" + std::to_string(in.leaf) + "_nt" + std::to_string(in.trees) +
                    "_ef" + std::to_string(in.ef) + "_geomanyk_lift_rankm0_specific0_alpha" + env_or("ANQI_ALPHA", "1.2") +
                    "_liftnnk" + env_or("ANQI_LIFTNNK", "50") + "_graphnnd" + env_or("ANQI_GRAPHNND", "6") +
                    (env_set("ANQI_REUSE_ORIG_KNN") ? "_reuseorig1" : "") + edge_suffix() + vamana_suffix();
    g = sanitize_tag(g);
    // the post-filter verifier's graph is keyed by its cutoff source instead
    // of the radius key, the dataset named by ANQI_GRAPH_RADIUS_SOURCE_PFX
    // when it is set (rknn_query.cpp:1843-1853)
    const std::string gsrc = env_or("ANQI_GRAPH_RADIUS_SOURCE_PFX", in.pfx.c_str());
    const std::string radius_id = env_set("ANQI_GRAPH_RADIUS_ID")
                                      ? std::string(env_or("ANQI_GRAPH_RADIUS_ID", "aknn_k100_v1"))
                                      : file_stat_identity(base_vector_source(gsrc + "_base.bin")) + "|" +
                                            threshold_source_identity(in, gsrc) + ":k100:v2";
    c.graph_key = sanitize_tag(g + "_graph_cutoff_k100_" + hash16(radius_id));
    c.state_key = "graph=" + c.graph_key;
    const std::string tag = hash16(c.graph_key);
    c.approxlift = in.graph_dir + "_approxlift_graph_" + tag + ".bin";
    c.meta = in.graph_dir + "_anqi_meta_graph_" + tag + ".txt";
    c.state = in.graph_dir + "_anqi_graph_state_" + tag + ".txt";
    c.marker = in.graph_dir + "_anqi_" + c.graph_key + ".built";
    c.verifier = in.verifier_dir + "_anqi_verifier_" + hash16(c.radius_key) + ".bin";
    // generate_config of the lifted index (ef_alpha 8 and the prune alpha are
    // floats there, printed by std::to_string); the graph key reaches it
    // through ANQI_GRAPH_CACHE_KEY, which the reference sets before building
    const char* alpha = std::getenv("ANQI_ALPHA");
    float prune_alpha = alpha ? (float)std::atof(alpha) : 1.2f;
    if (!(prune_alpha > 0.0f)) prune_alpha = 1.2f;   // the index's own default (index.h:143)
    const std::string cfg = "fd_metric_l2_M_" + std::to_string(in.M) + "_nnk_" + std::to_string(in.lift_nnk) +
                            "_ntrees_64_n_" + std::to_string(in.n) + "_dim_" + std::to_string(in.dim) + "_efa_" +
                            std::to_string(8.0f) + "_pa_" + std::to_string(prune_alpha) + "_er_100_graphnnd_" +
                            env_or("ANQI_GRAPHNND", "6") + "_gk_" + hash16(c.graph_key) + edge_config_suffix() +
                            (env_set("ANQI_REUSE_ORIG_KNN") ? "_reuseorig1" : "");
    c.graph_index = in.graph_dir + sanitize_tag(cfg);
    return c;
}

// First line of a text file, "" when it cannot be read.
inline std::string first_line(const std::string& p) {
    std::ifstream f(p);
    std::string s;
    std::getline(f, s);
    return s;
}

}  // namespace anqi_cli
