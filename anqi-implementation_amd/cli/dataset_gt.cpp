// dataset_gt -- drop-in for the reference's tools/dataset_gt.cpp on one
// MI355X: same options (--base --query --out_prefix --K --rk_k --tile
// --metric --threads --mode both|thresholds|ground-truth --threshold-table),
// same files (PREFIX_baseknn_gt.bin, PREFIX_rknn_gt.bin.rk,
// PREFIX_threshold_build_time.meta, PREFIX_rknn_gt.bin), same stdout lines
// and exit codes (2 usage, 3 bad input or write failure).
//   thresholds: exact self-excluded top-K of every base row by (key, id),
//     key = fnmadd(2, <x_i, x_j>, |x_i|^2 + |x_j|^2) (tools/dataset_gt.cpp:
//     206-257), from anqi_exact_knn (bf16 MFMA filter + FP32 re-score: the
//     FP32 keys and ids bit for bit);
//   ground-truth: reverse-kNN membership key <= rk2[j], each query's rows
//     sorted by (key, id) (:283-313), from anqi_rknn_membership.
// --tile and --threads shape the reference's CPU loops and are recorded only.
#include <cerrno>
#include <chrono>
#include <climits>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <omp.h>

#include "formats.hpp"

using namespace anqi_cli;

namespace {
anqi_ctx* g_ctx = nullptr;

void ck(int rc, const char* what) {
    if (rc != ANQI_OK) {
        std::cerr << "[dgt] GPU " << what << " failed: " << anqi_last_error(g_ctx) << "\n";
        std::exit(3);
    }
}

size_t parse_positive(const char* text, const char* name) {
    if (!text || !*text) {
        std::cerr << "[dgt] missing value for " << name << "\n";
        std::exit(2);
    }
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(text, &end, 10);
    if (errno == ERANGE || end == text || *end != '\0' || v == 0) {
        std::cerr << "[dgt] invalid " << name << "=" << text << "\n";
        std::exit(2);
    }
    return (size_t)v;
}

// fbin with the reference's checks (exit 3 on a bad file)
std::vector<float> load_fbin(const std::string& p, uint32_t& n, uint32_t& d) {
    std::vector<float> x;
    if (!exists(p)) {
        std::cerr << "[dgt] cannot open " << p << "\n";
        std::exit(3);
    }
    if (!read_fbin(p, x, n, d) || n == 0 || d == 0) {
        std::cerr << "[dgt] invalid fbin size: " << p << "\n";
        std::exit(3);
    }
    return x;
}

void require_write(const std::ofstream& f, const std::string& p) {
    if (!f) {
        std::cerr << "[dgt] failed to write " << p << "\n";
        std::exit(3);
    }
}

template <class T> T* dalloc(size_t count) {
    void* p = nullptr;
    ck(anqi_dmalloc(g_ctx, count * sizeof(T) + 16, &p), "allocation");
    return static_cast<T*>(p);
}
}  // namespace

int main(int argc, char** argv) {
    std::string base_p, query_p, prefix, metric = "l2", mode = "both", threshold_table;
    size_t K = 100, rk_k = 10, tile = 512, nthreads = 0;
    auto next_value = [&](int& i, const std::string& option) -> const char* {
        if (i + 1 >= argc) {
            std::cerr << "[dgt] missing value for " << option << "\n";
            std::exit(2);
        }
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--base") base_p = next_value(i, a);
        else if (a == "--query") query_p = next_value(i, a);
        else if (a == "--K") K = parse_positive(next_value(i, a), "K");
        else if (a == "--rk_k") rk_k = parse_positive(next_value(i, a), "rk_k");
        else if (a == "--out_prefix") prefix = next_value(i, a);
        else if (a == "--tile") tile = parse_positive(next_value(i, a), "tile");
        else if (a == "--metric") metric = next_value(i, a);
        else if (a == "--threads") nthreads = parse_positive(next_value(i, a), "threads");
        else if (a == "--mode") mode = next_value(i, a);
        else if (a == "--threshold-table") threshold_table = next_value(i, a);
        else {
            std::cerr << "[dgt] unknown option: " << a << "\n";
            return 2;
        }
    }
    if (mode != "both" && mode != "thresholds" && mode != "ground-truth") {
        std::cerr << "[dgt] mode must be both, thresholds, or ground-truth\n";
        return 2;
    }
    const bool build_thresholds = mode != "ground-truth";
    const bool build_ground_truth = mode != "thresholds";
    if (base_p.empty() || prefix.empty() || (build_ground_truth && query_p.empty())) {
        std::cerr << "usage: " << argv[0]
                  << " --base BASE.bin [--query QUERY.bin] --out_prefix PREFIX"
                  << " [--K 100] [--rk_k 10] [--tile 512] [--metric l2|ip]"
                  << " [--threads N] [--mode both|thresholds|ground-truth]"
                  << " [--threshold-table FILE]\n";
        return 2;
    }
    if (metric != "l2" && metric != "ip") {
        std::cerr << "[dgt] metric must be l2 or ip\n";
        return 2;
    }
    // key = -<x, y> for "ip" (dataset_gt.cpp:225, 292)
    const int anqi_metric = metric == "ip" ? ANQI_METRIC_IP : ANQI_METRIC_L2;
    if (nthreads > (size_t)INT_MAX) {
        std::cerr << "[dgt] thread count is too large\n";
        return 2;
    }
    if (nthreads) omp_set_num_threads((int)nthreads);
    const int T = omp_get_max_threads();   // recorded, as the reference does
    const auto t_all = std::chrono::steady_clock::now();
    uint32_t n = 0, D = 0;
    std::vector<float> base = load_fbin(base_p, n, D);
    if (K >= n || rk_k > K || K > UINT32_MAX) {
        std::cerr << "[dgt] require rk_k <= K < base vector count\n";
        return 2;
    }
    std::cout << "[dgt] mode=" << mode << " n=" << n << " D=" << D << " K=" << K << " rk_k=" << rk_k
              << " tile=" << tile << " threads=" << T << "\n";
    ck(anqi_ctx_create(0, &g_ctx), "context");
    float* d_base = dalloc<float>(base.size());
    ck(anqi_h2d(g_ctx, d_base, base.data(), base.size() * 4), "upload");
    std::vector<float> rk2(n);
    double threshold_build_s = 0.0;
    if (build_thresholds) {
        const auto t1 = std::chrono::steady_clock::now();
        uint32_t* d_ids = dalloc<uint32_t>((size_t)n * K);
        float* d_keys = dalloc<float>((size_t)n * K);
        ck(anqi_exact_knn_metric(g_ctx, d_base, n, D, (uint32_t)K, 1, anqi_metric, d_ids, d_keys), "exact top-K");
        std::vector<uint32_t> ids((size_t)n * K);
        std::vector<float> keys((size_t)n * K);
        ck(anqi_d2h(g_ctx, ids.data(), d_ids, ids.size() * 4), "download");
        ck(anqi_d2h(g_ctx, keys.data(), d_keys, keys.size() * 4), "download");
        anqi_dfree(g_ctx, d_ids);
        anqi_dfree(g_ctx, d_keys);
        for (uint32_t i = 0; i < n; i++) rk2[i] = keys[(size_t)i * K + (rk_k - 1)];
        std::cout << "[dgt] base-to-base top-K done in "
                  << std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count() << "s\n";
        {
            const std::string p = prefix + "_baseknn_gt.bin";
            std::ofstream f(p, std::ios::binary);
            const uint32_t un = n, uk = (uint32_t)K;
            f.write((const char*)&un, 4);
            f.write((const char*)&uk, 4);
            f.write((const char*)ids.data(), (std::streamsize)ids.size() * 4);
            f.write((const char*)keys.data(), (std::streamsize)keys.size() * 4);
            require_write(f, p);
            std::cout << "[dgt] wrote " << p << " (n*K = " << n << "*" << K << ")\n";
        }
        {
            const std::string p = prefix + "_rknn_gt.bin.rk";
            std::ofstream f(p, std::ios::binary);
            const uint32_t un = n, uk = (uint32_t)rk_k;
            f.write((const char*)&un, 4);
            f.write((const char*)&uk, 4);
            f.write((const char*)rk2.data(), (std::streamsize)n * 4);
            require_write(f, p);
            std::cout << "[dgt] wrote " << p << " (rank-k threshold array)\n";
        }
        threshold_build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_all).count();
        {
            const std::string p = prefix + "_threshold_build_time.meta";
            std::ofstream f(p);
            f << "ANQI_THRESHOLD_BUILD_V1\n"
              << "seconds " << std::setprecision(17) << threshold_build_s << "\n"
              << "metric " << metric << "\n"
              << "n " << n << "\n"
              << "d " << D << "\n"
              << "k " << K << "\n"
              << "threads " << T << "\n";
            require_write(f, p);
            std::cout << "[dgt] wrote " << p << "\n";
        }
        std::cout << "[dgt] THRESHOLD_BUILD_TIME=" << threshold_build_s << "s\n";
    } else {
        // rank-k column of a threshold table (tools/dataset_gt.cpp:72-115)
        if (threshold_table.empty()) threshold_table = prefix + "_baseknn_gt.bin";
        std::ifstream in(threshold_table, std::ios::binary);
        uint32_t tn = 0, tk = 0;
        const uint64_t bytes = file_size(threshold_table);
        if (!in.read((char*)&tn, 4) || !in.read((char*)&tk, 4) || tn != n || rk_k > tk ||
            bytes != 8 + 8ull * tn * tk) {
            std::cerr << "[dgt] invalid threshold table: " << threshold_table << "\n";
            return 3;
        }
        std::vector<float> row(tk);
        in.seekg((std::streamoff)(8 + 4ull * tn * tk));
        for (uint32_t i = 0; i < n; i++) {
            if (!in.read((char*)row.data(), 4ll * tk)) {
                std::cerr << "[dgt] truncated threshold table: " << threshold_table << "\n";
                return 3;
            }
            rk2[i] = row[rk_k - 1];
        }
        std::cout << "[dgt] loaded rank-" << rk_k << " threshold column from " << threshold_table << "\n";
    }
    double ground_truth_s = 0.0;
    if (build_ground_truth) {
        uint32_t nq = 0, dq = 0;
        std::vector<float> query = load_fbin(query_p, nq, dq);
        if (dq != D) {
            std::cerr << "[dgt] base/query dimension mismatch: " << D << " != " << dq << "\n";
            return 3;
        }
        std::cout << "[dgt] nq=" << nq << "\n";
        const auto t2 = std::chrono::steady_clock::now();
        float* d_q = dalloc<float>(query.size());
        float* d_rk = dalloc<float>(n);
        ck(anqi_h2d(g_ctx, d_q, query.data(), query.size() * 4), "upload");
        ck(anqi_h2d(g_ctx, d_rk, rk2.data(), (size_t)n * 4), "upload");
        std::vector<uint32_t> off((size_t)nq + 1);
        uint64_t total = 0;
        ck(anqi_rknn_membership_metric(g_ctx, d_base, n, d_q, nq, D, anqi_metric, d_rk, off.data(), &total, nullptr, nullptr),
           "membership count");
        if (total > UINT32_MAX) {
            std::cerr << "[dgt] reverse-kNN output exceeds uint32 CSR capacity\n";
            return 3;
        }
        uint32_t* d_ids = dalloc<uint32_t>(total ? total : 1);
        float* d_keys = dalloc<float>(total ? total : 1);
        ck(anqi_rknn_membership_metric(g_ctx, d_base, n, d_q, nq, D, anqi_metric, d_rk, off.data(), &total, d_ids, d_keys),
           "membership");
        std::vector<uint32_t> ids(total);
        std::vector<float> keys(total);
        ck(anqi_d2h(g_ctx, ids.data(), d_ids, total * 4), "download");
        ck(anqi_d2h(g_ctx, keys.data(), d_keys, total * 4), "download");
        ground_truth_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t2).count();
        std::cout << "[dgt] reverse-kNN membership done in " << ground_truth_s << "s\n";
        {
            const std::string p = prefix + "_rknn_gt.bin";
            std::ofstream f(p, std::ios::binary);
            const uint32_t unq = nq;
            f.write((const char*)&unq, 4);
            f.write((const char*)off.data(), (std::streamsize)(nq + 1) * 4);
            f.write((const char*)ids.data(), (std::streamsize)total * 4);
            f.write((const char*)keys.data(), (std::streamsize)total * 4);
            require_write(f, p);
            std::cout << "[dgt] wrote " << p << " (total_pairs=" << off[nq]
                      << ", mean |R|=" << (double)off[nq] / nq << ")\n";
        }
        if (!build_thresholds) {
            const std::string p = prefix + "_rknn_gt.bin.rk";
            std::ofstream f(p, std::ios::binary);
            const uint32_t un = n, uk = (uint32_t)rk_k;
            f.write((const char*)&un, 4);
            f.write((const char*)&uk, 4);
            f.write((const char*)rk2.data(), (std::streamsize)n * 4);
            require_write(f, p);
            std::cout << "[dgt] wrote " << p << " (rank-k threshold array)\n";
        }
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_all).count();
    std::cout << "[dgt] summary mode=" << mode << " threads=" << T << " threshold_build_time=" << threshold_build_s
              << "s ground_truth_time=" << ground_truth_s << "s total_time=" << secs << "s\n";
    anqi_ctx_destroy(g_ctx);
    return 0;
}
