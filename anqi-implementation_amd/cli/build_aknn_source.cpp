// build_aknn_source -- drop-in for the reference's tools/build_aknn_source.cpp
// on one MI355X: same argv (BASE.fbin OUTPUT_PREFIX [K=100] [leaf=500]
// [trees=64] [ef=600] [threads=64] [seed=10] [metric=l2]), same outputs
// (PREFIX_baseknn_aknn.rec published through a .part rename, and
// PREFIX_aknn_build_time.meta), same row checks, stdout line and exit codes
// (2 usage, 3 bad input, 4 row count, 5 invalid row, 6 write failure).
// The RP forest, leaf brute force and neighbour gather run on the GPU
// (anqi_aknn_build); `threads` is recorded only.
#include <chrono>
#include <iomanip>
#include <iostream>
#include <sys/resource.h>

#include "formats.hpp"

using namespace anqi_cli;

static double peak_gb() {
    struct rusage u {};
    getrusage(RUSAGE_SELF, &u);
    return u.ru_maxrss / 1e6;
}

static bool parse_u64(const char* s, uint64_t& out) {
    if (!s || !*s) return false;
    char* end = nullptr;
    errno = 0;
    unsigned long long v = std::strtoull(s, &end, 10);
    if (errno == ERANGE || end == s || *end != '\0') return false;
    out = v;
    return true;
}

int main(int argc, char** argv) {
    if (argc < 3 || argc > 10) {
        std::cerr << "usage: " << argv[0]
                  << " BASE.fbin OUTPUT_PREFIX [K=100] [leaf=500]"
                     " [trees=64] [ef=600] [threads=64] [seed=10] [metric=l2]\n";
        return 2;
    }
    const std::string base_path = argv[1];
    const std::string prefix = argv[2];
    uint64_t k = 100, leaf = 500, trees = 64, ef = 600, threads = 64, seed = 10;
    uint64_t* slots[6] = {&k, &leaf, &trees, &ef, &threads, &seed};
    for (int a = 3; a < argc && a < 9; a++)
        if (!parse_u64(argv[a], *slots[a - 3])) {
            std::cerr << "[aknn] invalid numeric argument: " << argv[a] << "\n";
            return 2;
        }
    const std::string metric = argc > 9 ? argv[9] : "l2";
    if (metric != "l2" && metric != "ip") {
        std::cerr << "[aknn] metric must be l2 or ip\n";
        return 2;
    }
    const auto started = std::chrono::steady_clock::now();
    uint32_t n = 0, d = 0;
    {
        std::ifstream header(base_path, std::ios::binary);
        if (!header.read((char*)&n, 4) || !header.read((char*)&d, 4) || n <= k || d == 0 || k > UINT32_MAX) {
            std::cerr << "[aknn] invalid base file or K: " << base_path << "\n";
            return 3;
        }
    }
    std::vector<float> x;
    if (!read_fbin(base_path, x, n, d)) {
        std::cerr << "[aknn] invalid base file or K: " << base_path << "\n";
        return 3;
    }
    anqi_ctx* ctx = nullptr;
    if (anqi_ctx_create(0, &ctx) != ANQI_OK) {
        std::cerr << "[aknn] no HIP device: " << (ctx ? anqi_last_error(ctx) : "context creation failed") << "\n";
        return 3;
    }
    void *dx = nullptr, *dids = nullptr, *dkeys = nullptr;
    if (anqi_dmalloc(ctx, x.size() * 4, &dx) || anqi_dmalloc(ctx, (size_t)n * k * 4, &dids) ||
        anqi_dmalloc(ctx, (size_t)n * k * 4, &dkeys) || anqi_h2d(ctx, dx, x.data(), x.size() * 4)) {
        std::cerr << "[aknn] device allocation failed: " << anqi_last_error(ctx) << "\n";
        return 3;
    }
    const bool ip = metric == "ip";
    anqi_aknn_params p{(uint32_t)k, (uint32_t)leaf, (uint32_t)trees, (uint32_t)ef, 8, 100, seed,
                       anqi_cli::gpu_form_chosen() ? 1 : 0, ip ? ANQI_METRIC_IP : ANQI_METRIC_L2};
    anqi_aknn_stats st{};
    int rc = anqi_aknn_build(ctx, (const float*)dx, n, d, &p, (uint32_t*)dids, (float*)dkeys, &st);
    if (rc != ANQI_OK) {
        std::cerr << "[aknn] " << anqi_last_error(ctx) << "\n";
        return rc == ANQI_E_INPUT ? 3 : (rc == ANQI_E_USAGE ? 2 : 3);
    }
    std::vector<uint32_t> ids((size_t)n * k);
    std::vector<float> dist((size_t)n * k);
    if (anqi_d2h(ctx, ids.data(), dids, ids.size() * 4) || anqi_d2h(ctx, dist.data(), dkeys, dist.size() * 4)) {
        std::cerr << "[aknn] row-count mismatch\n";
        return 4;
    }
    // row checks of build_aknn_source.cpp:93-117
    for (size_t row = 0; row < n; ++row) {
        const uint32_t* r = &ids[row * k];
        const float* dr = &dist[row * k];
        for (size_t rank = 0; rank < k; ++rank) {
            if (r[rank] >= n || r[rank] == row) {
                std::cerr << "[aknn] invalid neighbor at row=" << row << " rank=" << rank << "\n";
                return 5;
            }
            for (size_t prev = 0; prev < rank; ++prev)
                if (r[prev] == r[rank]) {
                    std::cerr << "[aknn] duplicate neighbor at row=" << row << " rank=" << rank << "\n";
                    return 5;
                }
            if (rank != 0 && dr[rank] < dr[rank - 1]) {
                std::cerr << "[aknn] unsorted row=" << row << "\n";
                return 5;
            }
        }
    }
    const std::string record_path = prefix + "_baseknn_aknn.rec";
    const std::string record_tmp = record_path + ".part";
    {
        std::ofstream out(record_tmp, std::ios::binary | std::ios::trunc);
        std::vector<Rec> rowrec(k);
        for (size_t row = 0; row < n; ++row) {
            for (size_t rank = 0; rank < k; ++rank)
                rowrec[rank] = Rec{ids[row * k + rank],
                                   ip ? dist[row * k + rank] - 1.0f : std::max(0.0f, dist[row * k + rank])};
            out.write((const char*)rowrec.data(), (std::streamsize)(k * sizeof(Rec)));
        }
        out.close();
        if (!out) {
            std::cerr << "[aknn] failed writing " << record_tmp << "\n";
            return 6;
        }
    }
    if (std::rename(record_tmp.c_str(), record_path.c_str()) != 0) {
        std::perror("rename");
        return 6;
    }
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    const std::string timing_path = prefix + "_aknn_build_time.meta";
    {
        std::ofstream t(timing_path);
        t << "ANQI_AKNN_BUILD_V1\n"
          << "seconds " << std::setprecision(17) << elapsed << "\n"
          << "metric " << metric << "\n"
          << "n " << n << "\n"
          << "d " << d << "\n"
          << "k " << k << "\n"
          << "leaf " << leaf << "\n"
          << "trees " << trees << "\n"
          << "ef " << ef << "\n"
          << "threads " << threads << "\n"
          << "seed " << seed << "\n";
        t.close();
        if (!t) {
            std::cerr << "[aknn] failed writing " << timing_path << "\n";
            return 6;
        }
    }
    std::cout << "[aknn] DONE n=" << n << " d=" << d << " K=" << k << " leaf=" << leaf << " trees=" << trees
              << " ef=" << ef << " threads=" << threads << " seed=" << seed << " metric=" << metric
              << " elapsed_s=" << elapsed << " peak_gb=" << peak_gb() << " source=" << record_path << "\n";
    anqi_ctx_destroy(ctx);
    return 0;
}
