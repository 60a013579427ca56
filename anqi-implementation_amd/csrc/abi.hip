// abi.hip -- C ABI (include/anqi.h) over the HIP kernels: context, memory,
// and the host orchestration of the AKNN source build, the lifted-graph
// build, the verifier and the RkNN query.  Errors are returned as status
// codes with a message in the context; nothing throws across the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/anqi.h"
#include "common.hpp"
#include "kernels.hpp"

// ---- work area of the entry point that allocates nothing ---------------------------------------
// anqi_rknn_search_noalloc takes its counters, batch arrays (entries, keys,
// work order, redo list) and global-tier rows from this array instead of
// hipMalloc.  Each device holds its own copy until the process ends.  Host
// code never uses the symbol itself (on the host it names a stand-in, not
// device memory): anqi_ctx_create, with the device current, records the
// device address hipGetSymbolAddress returns, and every kernel argument and
// fill derives from that address.  Calls that overlap on one device share it
// and race.
#define NOALLOC_BYTES (64ull << 20)             // 64 MiB per device
#define NOALLOC_MAX_NQ 65536u                   // queries per call
#define NOALLOC_ENT_OFF 256ull                  // counters first
#define NOALLOC_GLOBAL_OFF (NOALLOC_ENT_OFF + 16ull * NOALLOC_MAX_NQ)   // then 4 words per query
#define NOALLOC_MAX_DEV 64
static __device__ __align__(256) uint8_t scratch[NOALLOC_BYTES];
static void* g_scratch_ptr[NOALLOC_MAX_DEV];
static hipError_t g_scratch_err[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

// Records the current device's scratch address (once per device).
static void scratch_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEV) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = p;
}

struct anqi_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    int cus = 0;
    uint32_t* counter = nullptr;   // device work counter of the persistent search
    anqi_search_tuning tune{0, 0, 0};   // persistent search work areas / grid (0 = defaults)
    uint32_t* ent = nullptr;       // persistent search's batch: entries, keys, work order, redo list (4 x ent_cap)
    size_t ent_cap = 0;
    // optional HIP-event timing of the range-search kernel launch itself
    bool time_kernels = false;
    hipEvent_t kev0 = nullptr, kev1 = nullptr;
    double last_kernel_ms = 0.0;
    uint32_t redo_retry = 0, redo_global = 0;   // last search: queries rerun per tier
    int lrq_form = 0;                           // 0 certified float-float, 1 soft binary64
    unsigned long long* lrq_fb = nullptr;       // device: encode / decode fallback counts
};

static int anqi_set_err(anqi_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int anqi_fail_hip(anqi_ctx* ctx, hipError_t e, const char* what, int line) {
    return anqi_set_err(ctx, ANQI_E_DEVICE, "HIP error %d (%s) at abi.hip:%d: %s", (int)e,
                        hipGetErrorString(e), line, what);
}

#define HCHK(expr) HIP_CHECK_RET(expr)

// Stage progress on stderr when ANQI_PROGRESS is set (long builds).
static void progress(const char* fmt, ...) {
    static int on = -1;
    if (on < 0) on = std::getenv("ANQI_PROGRESS") != nullptr;
    if (!on) return;
    va_list ap;
    va_start(ap, fmt);
    std::fprintf(stderr, "[anqi] ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
    va_end(ap);
    std::fflush(stderr);
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// RAII device buffer
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = b;
        return hipMalloc(&p, b ? b : 16);
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// MT19937 (Matsumoto & Nishimura 1998), the engine behind std::mt19937:
// seed recurrence x[i] = 1812433253 (x[i-1] ^ x[i-1] >> 30) + i, twist every
// 624 outputs, tempering (11, 7/0x9d2c5680, 15/0xefc60000, 18).
struct Mt19937 {
    uint32_t x[624];
    uint32_t i = 624;
    explicit Mt19937(uint32_t seed) {
        x[0] = seed;
        for (uint32_t k = 1; k < 624; k++) x[k] = 1812433253u * (x[k - 1] ^ (x[k - 1] >> 30)) + k;
    }
    uint32_t next() {
        if (i >= 624) {
            for (uint32_t k = 0; k < 624; k++) {
                uint32_t y = (x[k] & 0x80000000u) | (x[(k + 1) % 624] & 0x7fffffffu);
                x[k] = x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
            }
            i = 0;
        }
        uint32_t y = x[i++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
};

static uint32_t next_pow2(uint64_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

extern "C" {

const char* anqi_build_info(void) {
    return "libanqi_hip: hipcc --offload-arch=gfx950, FP32 distance contract, soft-binary64 verifier";
}

int anqi_ctx_create(int device, anqi_ctx** out) {
    if (!out) return ANQI_E_USAGE;
    *out = nullptr;
    anqi_ctx* ctx = new anqi_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { delete ctx; return ANQI_E_DEVICE; }
    e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e != hipSuccess) { delete ctx; return ANQI_E_DEVICE; }
    scratch_setup(device);
    ctx->stream = ctx->own;
    if (hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        ctx->cus <= 0) {
        (void)hipStreamDestroy(ctx->own);
        delete ctx;
        return ANQI_E_DEVICE;
    }
    *out = ctx;
    return ANQI_OK;
}

void anqi_ctx_destroy(anqi_ctx* ctx) {
    if (!ctx) return;
    if (ctx->own) { (void)hipStreamSynchronize(ctx->own); (void)hipStreamDestroy(ctx->own); }
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->counter) (void)hipFree(ctx->counter);
    if (ctx->lrq_fb) (void)hipFree(ctx->lrq_fb);
    if (ctx->ent) (void)hipFree(ctx->ent);
    if (ctx->kev0) (void)hipEventDestroy(ctx->kev0);
    if (ctx->kev1) (void)hipEventDestroy(ctx->kev1);
    delete ctx;
}

int anqi_ctx_time_kernels(anqi_ctx* ctx, int enable) {
    if (enable && !ctx->kev0) {
        HCHK(hipEventCreate(&ctx->kev0));
        HCHK(hipEventCreate(&ctx->kev1));
    }
    ctx->time_kernels = enable != 0;
    ctx->last_kernel_ms = 0.0;
    return ANQI_OK;
}

double anqi_ctx_last_kernel_ms(const anqi_ctx* ctx) { return ctx ? ctx->last_kernel_ms : 0.0; }

int anqi_ctx_redo_counts(const anqi_ctx* ctx, uint32_t* retry, uint32_t* global) {
    if (!ctx) return ANQI_E_USAGE;
    if (retry) *retry = ctx->redo_retry;
    if (global) *global = ctx->redo_global;
    return ANQI_OK;
}

const char* anqi_last_error(const anqi_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int anqi_ctx_set_search_tuning(anqi_ctx* ctx, const anqi_search_tuning* tune) {
    ctx->tune = tune ? *tune : anqi_search_tuning{0, 0, 0};
    return ANQI_OK;
}

int anqi_ctx_sync(anqi_ctx* ctx) {
    HCHK(hipStreamSynchronize(ctx->stream));
    return ANQI_OK;
}

int anqi_ctx_set_stream(anqi_ctx* ctx, void* stream) {
    ctx->stream = (hipStream_t)stream;   // NULL: the default (null) stream
    return ANQI_OK;
}

int anqi_ctx_reset_stream(anqi_ctx* ctx) {
    ctx->stream = ctx->own;
    return ANQI_OK;
}

void* anqi_ctx_stream(anqi_ctx* ctx) { return (void*)ctx->stream; }

int anqi_dmalloc(anqi_ctx* ctx, size_t bytes, void** out) {
    HCHK(hipMalloc(out, bytes ? bytes : 16));
    return ANQI_OK;
}
int anqi_dfree(anqi_ctx* ctx, void* p) {
    if (p) HCHK(hipFree(p));
    return ANQI_OK;
}
int anqi_h2d(anqi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    HCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    return ANQI_OK;
}
int anqi_d2h(anqi_ctx* ctx, void* dst, const void* src, size_t bytes) {
    HCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    return ANQI_OK;
}

int anqi_norms(anqi_ctx* ctx, const float* x, uint64_t n, uint32_t d, float* out) {
    if (d == 0 || d > 16 * QMAX_CHUNKS + 15) return anqi_set_err(ctx, ANQI_E_USAGE, "dimension %u unsupported", d);
    HCHK(launch_norms(x, n, d, out, ctx->stream));
    return ANQI_OK;
}

int anqi_pair_d2(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t d, const uint32_t* a,
                 const uint32_t* b, uint64_t m, float* out) {
    HCHK(launch_pair_d2(x, nrm, d, a, b, m, out, ctx->stream));
    return ANQI_OK;
}

// ---- exact -----------------------------------------------------------------------------
int anqi_exact_knn(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t K, int precision,
                   uint32_t* ids, float* keys) {
    return anqi_exact_knn_metric(ctx, x, n, d, K, precision, ANQI_METRIC_L2, ids, keys);
}

int anqi_exact_knn_metric(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t K, int precision,
                          int metric, uint32_t* ids, float* keys) {
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "exact_knn: metric ANQI_METRIC_L2 or ANQI_METRIC_IP");
    const int ipm = metric == ANQI_METRIC_IP;
    if (K == 0 || K >= n || K > 128) return anqi_set_err(ctx, ANQI_E_USAGE, "exact_knn: need 0 < K < n, K <= 128");
    if (d > 16 * QMAX_CHUNKS + 15) return anqi_set_err(ctx, ANQI_E_USAGE, "dimension %u unsupported", d);
    hipStream_t s = ctx->stream;
    DBuf nrm;
    HCHK(nrm.alloc(sizeof(float) * (size_t)n));
    HCHK(launch_norms(x, n, d, nrm.as<float>(), s));
    if (precision == 1 && exact_mfma_supported(d, K)) {
        // bf16 sweep on the matrix cores, exact FP32 re-scoring inside a
        // per-row error bound: same ids and keys as the FP32 sweep
        DBuf xb, nx, ex, marg;
        HCHK(xb.alloc(2ull * n * d));
        HCHK(nx.alloc(4ull * n));
        HCHK(ex.alloc(4ull * n));
        HCHK(marg.alloc(4ull * n));
        HCHK(launch_to_bf16(x, (uint64_t)n * d, xb.as<uint16_t>(), s));
        HCHK(launch_exact_norms2(x, nrm.as<float>(), n, d, nx.as<float>(), ex.as<float>(), s));
        std::vector<float> hn(n), he(n);
        HCHK(hipMemcpyAsync(hn.data(), nx.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(he.data(), ex.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        float Y = 0.0f, E = 0.0f;
        for (uint32_t i = 0; i < n; i++) { Y = std::max(Y, hn[i]); E = std::max(E, he[i]); }
        HCHK(launch_exact_margin(nx.as<float>(), ex.as<float>(), n, d, Y, E, marg.as<float>(), s));
        HCHK(launch_exact_mfma(x, xb.as<uint16_t>(), nrm.as<float>(), marg.as<float>(), n, d, K, 0, n, ipm, ids,
                               keys, s));
    } else {
        HCHK(launch_exact_topk(x, nrm.as<float>(), n, d, K, 0, n, ipm, ids, keys, s));
    }
    HCHK(hipStreamSynchronize(s));
    return ANQI_OK;
}

int anqi_rknn_membership(anqi_ctx* ctx, const float* base, uint32_t n, const float* q, uint32_t nq,
                         uint32_t d, const float* rk2, uint32_t* off, uint64_t* total, uint32_t* ids,
                         float* keys) {
    return anqi_rknn_membership_metric(ctx, base, n, q, nq, d, ANQI_METRIC_L2, rk2, off, total, ids, keys);
}

int anqi_rknn_membership_metric(anqi_ctx* ctx, const float* base, uint32_t n, const float* q, uint32_t nq,
                                uint32_t d, int metric, const float* rk2, uint32_t* off, uint64_t* total,
                                uint32_t* ids, float* keys) {
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "rknn_membership: metric ANQI_METRIC_L2 or ANQI_METRIC_IP");
    const int ipm = metric == ANQI_METRIC_IP;
    DBuf nrm, cnt;
    HCHK(nrm.alloc(sizeof(float) * (size_t)n));
    HCHK(cnt.alloc(sizeof(uint32_t) * (size_t)nq));
    HCHK(launch_norms(base, n, d, nrm.as<float>(), ctx->stream));
    if (!ids) {
        HCHK(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t) * nq, ctx->stream));
        HCHK(launch_rknn_count(base, nrm.as<float>(), n, q, nq, d, ipm, rk2, cnt.as<uint32_t>(), ctx->stream));
        std::vector<uint32_t> c(nq);
        HCHK(hipMemcpyAsync(c.data(), cnt.p, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, ctx->stream));
        HCHK(hipStreamSynchronize(ctx->stream));
        uint64_t t = 0;
        off[0] = 0;
        for (uint32_t i = 0; i < nq; i++) { t += c[i]; off[i + 1] = (uint32_t)t; }
        *total = t;
        return ANQI_OK;
    }
    DBuf doff;
    HCHK(doff.alloc(sizeof(uint32_t) * (size_t)(nq + 1)));
    HCHK(hipMemcpyAsync(doff.p, off, sizeof(uint32_t) * (nq + 1), hipMemcpyHostToDevice, ctx->stream));
    HCHK(launch_rknn_fill(base, nrm.as<float>(), n, q, nq, d, ipm, rk2, doff.as<uint32_t>(), ids, keys,
                          ctx->stream));
    // order each row by (key, id) on the host (rows are short)
    uint64_t tot = off[nq];
    std::vector<uint32_t> hid(tot);
    std::vector<float> hk(tot);
    HCHK(hipMemcpyAsync(hid.data(), ids, 4 * tot, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipMemcpyAsync(hk.data(), keys, 4 * tot, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    std::vector<std::pair<float, uint32_t>> row;
    for (uint32_t i = 0; i < nq; i++) {
        row.clear();
        for (uint32_t t = off[i]; t < off[i + 1]; t++) row.emplace_back(hk[t], hid[t]);
        std::sort(row.begin(), row.end());
        for (uint32_t t = off[i]; t < off[i + 1]; t++) { hk[t] = row[t - off[i]].first; hid[t] = row[t - off[i]].second; }
    }
    HCHK(hipMemcpyAsync(ids, hid.data(), 4 * tot, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipMemcpyAsync(keys, hk.data(), 4 * tot, hipMemcpyHostToDevice, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    if (total) *total = tot;
    return ANQI_OK;
}

// ---- verifier ------------------------------------------------------------------------------
// Fallback counter slot (0 encode, 1 decode) of the certified forms, zeroed.
static int lrq_counter(anqi_ctx* ctx, int slot, unsigned long long** out) {
    if (!ctx->lrq_fb) {
        HCHK(hipMalloc((void**)&ctx->lrq_fb, 2 * sizeof(unsigned long long)));
        HCHK(launch_fill_u32((uint32_t*)ctx->lrq_fb, 0u, 4, ctx->stream));
    }
    HCHK(launch_fill_u32((uint32_t*)(ctx->lrq_fb + slot), 0u, 2, ctx->stream));
    *out = ctx->lrq_fb + slot;
    return ANQI_OK;
}

int anqi_lrq_encode(anqi_ctx* ctx, const float* cutoffs, uint64_t n, uint32_t nk, uint32_t rankm,
                    const float* mu, const float* basis, const float* codebook, uint32_t code_size,
                    float* coeff, uint8_t* codes, uint8_t* codes_kmajor) {
    if (code_size != 256) return anqi_set_err(ctx, ANQI_E_INPUT, "only u8 codebooks (256 levels) are supported");
    if (ctx->lrq_form == 0 && lrq_fast_supported(nk, rankm, code_size)) {
        unsigned long long* fb = nullptr;
        if (int rc = lrq_counter(ctx, 0, &fb)) return rc;
        HCHK(launch_lrq_encode_fast(cutoffs, n, nk, rankm, mu, basis, codebook, coeff, codes, codes_kmajor, fb,
                                    ctx->stream));
        return ANQI_OK;
    }
    HCHK(launch_lrq_encode(cutoffs, n, nk, rankm, mu, basis, codebook, code_size, coeff, codes, codes_kmajor,
                           ctx->stream));
    return ANQI_OK;
}

int anqi_transpose_codes(anqi_ctx* ctx, const uint8_t* codes, uint64_t n, uint32_t nk, uint8_t* codes_kmajor) {
    HCHK(launch_transpose_codes(codes, n, nk, codes_kmajor, ctx->stream));
    return ANQI_OK;
}

int anqi_lrq_materialize_metric(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor, uint64_t n,
                                uint32_t nk, uint32_t rankm, const float* mu, const float* basis,
                                const float* codebook, uint32_t code_size, const uint32_t* ks, uint32_t nks,
                                double slack, int metric, float* thr) {
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "metric: ANQI_METRIC_L2 or ANQI_METRIC_IP");
    for (uint32_t t = 0; t < nks; t++)
        if (ks[t] < 1 || ks[t] > nk) return anqi_set_err(ctx, ANQI_E_USAGE, "LRQ verifier is not ready for k=%u", ks[t]);
    if (nks == 0 || n == 0) return ANQI_OK;
    const int clamp = metric == ANQI_METRIC_L2;
    if (ctx->lrq_form == 0 && slack == 1.0 && lrq_fast_supported(nk, rankm, code_size)) {
        unsigned long long* fb = nullptr;
        if (int rc = lrq_counter(ctx, 1, &fb)) return rc;
        HCHK(launch_lrq_decode(coeff, codes_kmajor, n, nk, rankm, mu, basis, codebook, code_size, ks, nks, clamp,
                               thr, fb, ctx->stream));
        return ANQI_OK;
    }
    uint64_t sb;
    std::memcpy(&sb, &slack, 8);
    for (uint32_t t = 0; t < nks; t++)
        HCHK(launch_lrq_materialize(coeff, codes_kmajor, n, nk, rankm, mu, basis, codebook, code_size, ks[t], sb,
                                    clamp, thr + (uint64_t)t * n, ctx->stream));
    return ANQI_OK;
}

int anqi_lrq_materialize_multi(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor, uint64_t n,
                               uint32_t nk, uint32_t rankm, const float* mu, const float* basis,
                               const float* codebook, uint32_t code_size, const uint32_t* ks, uint32_t nks,
                               double slack, float* thr) {
    return anqi_lrq_materialize_metric(ctx, coeff, codes_kmajor, n, nk, rankm, mu, basis, codebook, code_size, ks,
                                       nks, slack, ANQI_METRIC_L2, thr);
}

int anqi_lrq_materialize(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor, uint64_t n,
                         uint32_t nk, uint32_t rankm, const float* mu, const float* basis,
                         const float* codebook, uint32_t code_size, uint32_t k_one_based, double slack,
                         float* thr) {
    return anqi_lrq_materialize_multi(ctx, coeff, codes_kmajor, n, nk, rankm, mu, basis, codebook, code_size,
                                      &k_one_based, 1, slack, thr);
}

int anqi_ctx_set_lrq_form(anqi_ctx* ctx, int form) {
    if (form != 0 && form != 1) return anqi_set_err(ctx, ANQI_E_USAGE, "lrq form: 0 (certified) or 1 (soft)");
    ctx->lrq_form = form;
    return ANQI_OK;
}

int anqi_ctx_lrq_fallbacks(anqi_ctx* ctx, uint64_t* encode, uint64_t* decode) {
    unsigned long long h[2] = {0, 0};
    if (ctx->lrq_fb) {
        HCHK(hipMemcpyAsync(h, ctx->lrq_fb, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        HCHK(hipStreamSynchronize(ctx->stream));
    }
    if (encode) *encode = h[0];
    if (decode) *decode = h[1];
    return ANQI_OK;
}

int anqi_lift_split(anqi_ctx* ctx, const float* lifted, uint32_t n, uint32_t d, float* lmain, float* ltail) {
    if (d == 0 || (d & 15u) || d > 16 * QMAX_CHUNKS)
        return anqi_set_err(ctx, ANQI_E_USAGE, "lift_split: d must be a multiple of 16, at most %d", 16 * QMAX_CHUNKS);
    HCHK(launch_lift_split(lifted, n, d, lmain, ltail, ctx->stream));
    return ANQI_OK;
}

int anqi_byte_rows(anqi_ctx* ctx, const float* base, const float* lmain, uint32_t n, uint32_t d, double scale,
                   uint8_t* brows, int* exact) {
    if (!base || !lmain || !brows || !exact || d == 0 || (d & 15u) || d > 16 * QMAX_CHUNKS)
        return anqi_set_err(ctx, ANQI_E_USAGE, "byte_rows: null array, or d not a multiple of 16 up to %d",
                            16 * QMAX_CHUNKS);
    DBuf bad;
    HCHK(bad.alloc(256));
    HCHK(launch_fill_u32(bad.as<uint32_t>(), 0u, 1, ctx->stream));
    HCHK(launch_byte_rows(base, lmain, n, d, scale, brows, bad.as<uint32_t>(), ctx->stream));
    uint32_t nbad = 0;
    HCHK(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    *exact = nbad == 0;
    return ANQI_OK;
}

int anqi_pack_links(anqi_ctx* ctx, const uint32_t* links, const uint32_t* deg, uint32_t n, uint32_t M,
                    uint32_t* plinks) {
    if (!links || !deg || !plinks || M == 0 || M > 64)
        return anqi_set_err(ctx, ANQI_E_USAGE, "pack_links: null array or M outside 1..64");
    HCHK(launch_pack_links(links, deg, n, M, plinks, ctx->stream));
    return ANQI_OK;
}

// ---- lift (host FP64; src/rknn_query.cpp:2203-2236) ------------------------------------------
int anqi_lrq_relax(anqi_ctx* ctx, const float* cutoffs, uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                   const float* basis, const float* codebook, uint32_t code_size, float tau, int mode,
                   float grad_scale, double* loss, float* grad_basis, float* grad_codebook) {
    if (mode != 0 && mode != 1) return anqi_set_err(ctx, ANQI_E_USAGE, "lrq_relax: mode 0 or 1");
    if (n == 0 || n > 0xffffffffull || nk == 0 || nk > 128 || rankm == 0 || rankm > 16 || code_size == 0 ||
        code_size > 1024 || !(tau > 0.0f))
        return anqi_set_err(ctx, ANQI_E_USAGE, "lrq_relax: need n > 0, nk <= 128, rankm <= 16, tau > 0");
    hipStream_t s = ctx->stream;
    DBuf work, rows;
    const size_t nkz = (size_t)n * nk, nm = (size_t)n * rankm;
    HCHK(work.alloc(4 * (5 * nkz + 2 * nm)));
    HCHK(rows.alloc(4 * (size_t)n));
    float* w = work.as<float>();
    RelaxParams P{};
    P.cut = cutoffs; P.n = (uint32_t)n; P.nk = nk; P.rankm = rankm; P.ncode = code_size; P.mode = mode;
    P.mu = mu; P.G = basis; P.cb = codebook;
    P.inv_tau = 1.0f / tau;
    P.gscale = grad_scale / (float)((double)n * nk);
    P.e = w; P.q = w + nkz; P.mx = w + 2 * nkz; P.sm = w + 3 * nkz; P.ge = w + 4 * nkz;
    P.c = w + 5 * nkz; P.gc = w + 5 * nkz + nm;
    P.loss_rows = rows.as<float>();
    P.gG = grad_basis; P.gcb = grad_codebook;
    if (grad_basis) HCHK(launch_fill_u32((uint32_t*)grad_basis, 0u, (uint64_t)rankm * nk, s));
    if (grad_codebook) HCHK(launch_fill_u32((uint32_t*)grad_codebook, 0u, (uint64_t)nk * code_size, s));
    HCHK(launch_relax(P, s));
    if (loss) {
        std::vector<float> h(n);
        HCHK(hipMemcpyAsync(h.data(), rows.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        double t = 0.0;
        for (float v : h) t += v;
        *loss = t / ((double)n * nk);
    } else {
        HCHK(hipStreamSynchronize(s));
    }
    return ANQI_OK;
}

int anqi_lift_host_metric(const float* base, uint64_t n, uint32_t d, const float* rk_horizon, int metric,
                          float* lifted, double* mstar2, double* scale) {
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP) return ANQI_E_USAGE;
    const bool ip = metric == ANQI_METRIC_IP;
    // squared norms summed in row order (the reference's OpenMP reduction
    // with one thread); GCC contracts the multiply-adds into fma.
    double sum_n2 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        const float* o = base + i * d;
        double s = 0.0;
        for (uint32_t j = 0; j < d; j++) s = std::fma((double)o[j], (double)o[j], s);
        sum_n2 += s;
    }
    double s_scale = 1.0 / std::sqrt(sum_n2 / (double)n + 1e-12), sc2 = s_scale * s_scale;
    uint32_t dl = d + 1, dout = d + 2;
    double maxn2 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        const float* o = base + i * d;
        float* h = lifted + i * dout;
        double nr2 = 0.0;
        for (uint32_t j = 0; j < d; j++) {
            double v = (double)o[j] * s_scale;
            h[j] = (float)v;
            nr2 = std::fma(v, v, nr2);
        }
        h[d] = ip ? (float)((double)rk_horizon[i] * sc2) : (float)(0.5 * std::fma((double)rk_horizon[i], sc2, -nr2));
        double hn2 = std::fma((double)h[d], (double)h[d], nr2);
        if (hn2 > maxn2) maxn2 = hn2;
    }
    double M2 = maxn2 * 1.0001;
    for (uint64_t i = 0; i < n; i++) {
        float* h = lifted + i * dout;
        double hn2 = 0.0;
        for (uint32_t j = 0; j < dl; j++) hn2 = std::fma((double)h[j], (double)h[j], hn2);
        double pad = M2 - hn2;
        if (pad < 0) pad = 0;
        h[dl] = (float)std::sqrt(pad);
    }
    *mstar2 = M2;
    *scale = s_scale;
    return ANQI_OK;
}

int anqi_lift_host(const float* base, uint64_t n, uint32_t d, const float* rk_horizon, float* lifted,
                   double* mstar2, double* scale) {
    return anqi_lift_host_metric(base, n, d, rk_horizon, ANQI_METRIC_L2, lifted, mstar2, scale);
}

int anqi_aknn_cutoffs(anqi_ctx* ctx, const float* keys, uint64_t n, uint32_t K, float* cut, float* horizon) {
    return anqi_aknn_cutoffs_metric(ctx, keys, n, K, ANQI_METRIC_L2, cut, horizon);
}

int anqi_aknn_cutoffs_metric(anqi_ctx* ctx, const float* keys, uint64_t n, uint32_t K, int metric, float* cut,
                             float* horizon) {
    if (K == 0) return anqi_set_err(ctx, ANQI_E_USAGE, "aknn_cutoffs: K must be positive");
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "aknn_cutoffs: metric ANQI_METRIC_L2 or ANQI_METRIC_IP");
    HCHK(launch_aknn_cutoffs(keys, n, K, metric == ANQI_METRIC_IP, cut, horizon, ctx->stream));
    return ANQI_OK;
}

int anqi_lift_metric(anqi_ctx* ctx, const float* base, uint64_t n, uint32_t d, const float* rk_horizon,
                     int metric, float* lifted, double* mstar2, double* scale) {
    if (n == 0 || d == 0) return anqi_set_err(ctx, ANQI_E_USAGE, "lift: empty input");
    if (metric != ANQI_METRIC_L2 && metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "lift: metric ANQI_METRIC_L2 or ANQI_METRIC_IP");
    hipStream_t s = ctx->stream;
    DBuf rows, mx;
    HCHK(rows.alloc(8 * n));
    HCHK(mx.alloc(256));
    // squared norms per row on the device, summed in row order here (the
    // reference's reduction order)
    HCHK(launch_lift_norm2(base, n, d, rows.as<uint64_t>(), s));
    std::vector<double> sn(n);
    HCHK(hipMemcpyAsync(sn.data(), rows.p, 8 * n, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    double sum_n2 = 0.0;
    for (uint64_t i = 0; i < n; i++) sum_n2 += sn[i];
    const double s_scale = 1.0 / std::sqrt(sum_n2 / (double)n + 1e-12), sc2 = s_scale * s_scale;
    uint64_t sb, cb;
    std::memcpy(&sb, &s_scale, 8);
    std::memcpy(&cb, &sc2, 8);
    HCHK(hipMemsetAsync(mx.p, 0, 8, s));
    HCHK(launch_lift_rows(base, n, d, rk_horizon, sb, cb, metric == ANQI_METRIC_IP, lifted, mx.as<uint64_t>(), s));
    double maxn2 = 0.0;
    HCHK(hipMemcpyAsync(&maxn2, mx.p, 8, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    const double M2 = maxn2 * 1.0001;
    uint64_t mb;
    std::memcpy(&mb, &M2, 8);
    HCHK(launch_lift_pad(lifted, n, d, mb, s));
    HCHK(hipStreamSynchronize(s));
    *mstar2 = M2;
    *scale = s_scale;
    return ANQI_OK;
}

int anqi_lift(anqi_ctx* ctx, const float* base, uint64_t n, uint32_t d, const float* rk_horizon, float* lifted,
              double* mstar2, double* scale) {
    return anqi_lift_metric(ctx, base, n, d, rk_horizon, ANQI_METRIC_L2, lifted, mstar2, scale);
}

// ---- graph pieces -----------------------------------------------------------------------------
int anqi_warm_start(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* ids,
                    uint32_t width, uint32_t cap, uint32_t* hidx, float* hval) {
    HCHK(launch_warm_start(x, nrm, n, d, ids, width, cap, hidx, hval, ctx->stream));
    return ANQI_OK;
}

int anqi_nndescent(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, uint32_t cap,
                   uint32_t* hidx, float* hval, int iters, int sample, int64_t* upd, int* passes, int form) {
    if (sample > 120) sample = 120;
    if (sample < 2 || iters <= 0) { if (passes) *passes = 0; return ANQI_OK; }
    hipStream_t s = ctx->stream;
    // proposal buffers: the bound n s (s - 1) is ~6x the count seen in
    // practice, so start from a fraction of it and regrow (rerunning the
    // pass, whose proposals depend only on the pass-start heaps) if a pass
    // reports more
    const uint64_t bound = (uint64_t)n * sample * (sample - 1);
    uint64_t maxp = std::min<uint64_t>(bound, std::max<uint64_t>((uint64_t)n * 96, 1u << 20));
    DBuf samp, scnt, k0, k1, v0, v1, cnt, off, updd, tmp;
    HCHK(samp.alloc(4ull * n * sample));
    HCHK(scnt.alloc(4ull * n));
    HCHK(cnt.alloc(8));
    HCHK(updd.alloc(8));
    HCHK(off.alloc(4ull * (n + 1)));
    int end_bit = 32;
    while (end_bit < 64 && ((uint64_t)n >> (end_bit - 32))) end_bit++;
    size_t tb = 0;
    auto grow = [&](uint64_t cap) -> hipError_t {
        hipError_t e;
        maxp = cap;
        if ((e = k0.alloc(8 * maxp)) != hipSuccess) return e;
        if ((e = k1.alloc(8 * maxp)) != hipSuccess) return e;
        if ((e = v0.alloc(4 * maxp)) != hipSuccess) return e;
        if ((e = v1.alloc(4 * maxp)) != hipSuccess) return e;
        tb = 0;
        if ((e = prim_sort_pairs_u64_u32(nullptr, nullptr, nullptr, nullptr, maxp, end_bit, nullptr, tb, s)) !=
            hipSuccess)
            return e;
        return tmp.alloc(tb);
    };
    HCHK(grow(maxp));
    // chosen form: per-row bounds of the bf16 gram (k_exact_margin over the rows)
    const bool mfma = form && nnd_propose_mfma_supported(d, (uint32_t)sample);
    DBuf marg;
    if (mfma) {
        DBuf nx, ex;
        HCHK(nx.alloc(4ull * n));
        HCHK(ex.alloc(4ull * n));
        HCHK(marg.alloc(4ull * n));
        HCHK(launch_exact_norms2(x, nrm, n, d, nx.as<float>(), ex.as<float>(), s));
        std::vector<float> hn(n), he(n);
        HCHK(hipMemcpyAsync(hn.data(), nx.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(he.data(), ex.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        float Y = 0.0f, E = 0.0f;
        for (uint32_t i = 0; i < n; i++) { Y = std::max(Y, hn[i]); E = std::max(E, he[i]); }
        HCHK(launch_exact_margin(nx.as<float>(), ex.as<float>(), n, d, Y, E, marg.as<float>(), s));
        HCHK(hipStreamSynchronize(s));
    }
    int it = 0;
    for (; it < iters; it++) {
        HCHK(launch_nnd_sample(hidx, n, cap, (uint32_t)sample, samp.as<uint32_t>(), scnt.as<uint32_t>(), s));
        HCHK(hipMemsetAsync(cnt.p, 0, 8, s));
        HCHK(hipMemsetAsync(updd.p, 0, 8, s));
        if (mfma)
            HCHK(launch_nnd_propose_mfma(x, nrm, marg.as<float>(), n, d, samp.as<uint32_t>(), scnt.as<uint32_t>(),
                                         (uint32_t)sample, hval, cap, k0.as<uint64_t>(), v0.as<uint32_t>(),
                                         cnt.as<unsigned long long>(), maxp, s));
        else
            HCHK(launch_nnd_propose(x, nrm, n, d, samp.as<uint32_t>(), scnt.as<uint32_t>(), (uint32_t)sample, hval,
                                    cap, k0.as<uint64_t>(), v0.as<uint32_t>(), cnt.as<unsigned long long>(), maxp, s));
        unsigned long long m = 0;
        HCHK(hipMemcpyAsync(&m, cnt.p, 8, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        if (m > maxp) {
            if (m > bound) return anqi_set_err(ctx, ANQI_E_DEVICE, "LocalJoin proposal count above its bound");
            HCHK(grow(std::min<uint64_t>(bound, m + m / 8)));
            it--;
            continue;
        }
        if (m == 0) {   // nothing to apply: same stopping rule with zero updates
            if (upd) upd[it] = 0;
            if (0 < (int64_t)(n / 200)) { it++; break; }
            continue;
        }
        size_t tbm = tb;
        HCHK(prim_sort_pairs_u64_u32(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), m,
                                     end_bit, tmp.p, tbm, s));
        HCHK(launch_row_bounds(k1.as<uint64_t>(), m, n, off.as<uint32_t>(), s));
        HCHK(launch_nnd_apply(k1.as<uint64_t>(), v1.as<uint32_t>(), m, off.as<uint32_t>(), n, hidx, hval, cap,
                              updd.as<unsigned long long>(), s));
        unsigned long long u = 0;
        HCHK(hipMemcpyAsync(&u, updd.p, 8, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        if (upd) upd[it] = (int64_t)u;
        progress("localjoin pass %d: %llu proposals, %llu updates", it, (unsigned long long)m, u);
        if ((int64_t)u < (int64_t)(n / 200)) { it++; break; }
    }
    if (passes) *passes = it;
    return ANQI_OK;
}

// reverse CSR; also reports the maximum list length
static int reverse_csr(anqi_ctx* ctx, const uint32_t* ids, uint32_t rows, uint32_t width, const uint32_t* rowcnt,
                       uint32_t n, uint32_t* roff, uint32_t* rids, uint32_t* maxlen) {
    hipStream_t s = ctx->stream;
    DBuf cnt, tmp;
    HCHK(cnt.alloc(4ull * (n + 1)));
    HCHK(hipMemsetAsync(cnt.p, 0, 4ull * (n + 1), s));
    if (rowcnt) {
        // counts = cursors after a scatter into a scratch array
        DBuf scratch;
        HCHK(scratch.alloc(4ull * rows * width));
        HCHK(launch_scatter_rev(ids, rows, width, rowcnt, n, cnt.as<uint32_t>(), scratch.as<uint32_t>(), s));
        HCHK(hipStreamSynchronize(s));
    } else {
        HCHK(launch_count_refs(ids, (uint64_t)rows * width, n, cnt.as<uint32_t>(), s));
    }
    size_t tb = 0;
    HCHK(prim_exclusive_scan_u32(cnt.as<uint32_t>(), roff, (uint64_t)n + 1, nullptr, tb, s));
    HCHK(tmp.alloc(tb));
    HCHK(prim_exclusive_scan_u32(cnt.as<uint32_t>(), roff, (uint64_t)n + 1, tmp.p, tb, s));
    if (maxlen) {
        std::vector<uint32_t> c(n);
        HCHK(hipMemcpyAsync(c.data(), cnt.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        *maxlen = n ? *std::max_element(c.begin(), c.end()) : 0;
    }
    // cursors start at roff
    HCHK(hipMemcpyAsync(cnt.p, roff, 4ull * n, hipMemcpyDeviceToDevice, s));
    HCHK(launch_scatter_rev(ids, rows, width, rowcnt, n, cnt.as<uint32_t>(), rids, s));
    return ANQI_OK;
}

int anqi_reverse_csr(anqi_ctx* ctx, const uint32_t* ids, uint32_t rows, uint32_t width, const uint32_t* rowcnt,
                     uint32_t n, uint32_t* roff, uint32_t* rids) {
    return reverse_csr(ctx, ids, rows, width, rowcnt, n, roff, rids, nullptr);
}

// rows whose status is non-zero
static int failed_rows(anqi_ctx* ctx, const uint32_t* status, uint32_t n, std::vector<uint32_t>& rows) {
    std::vector<uint32_t> st(n);
    HCHK(hipMemcpyAsync(st.data(), status, 4ull * n, hipMemcpyDeviceToHost, ctx->stream));
    HCHK(hipStreamSynchronize(ctx->stream));
    rows.clear();
    for (uint32_t i = 0; i < n; i++)
        if (st[i]) rows.push_back(i);
    return ANQI_OK;
}

// work (optional, host): [0] distances computed, [1] adjacency ids read.
static int gather_run(anqi_ctx* ctx, GatherParams P, bool graph, uint32_t maxrev, uint32_t* overflow_rows,
                      uint64_t* work = nullptr) {
    hipStream_t s = ctx->stream;
    DBuf status, wk;
    HCHK(status.alloc(4ull * P.n));
    HCHK(hipMemsetAsync(status.p, 0, 4ull * P.n, s));
    P.status = status.as<uint32_t>();
    HCHK(wk.alloc(16));
    HCHK(hipMemsetAsync(wk.p, 0, 16, s));
    P.work = wk.as<unsigned long long>();
    auto read_work = [&]() -> int {
        if (!work) return ANQI_OK;
        HCHK(hipMemcpyAsync(work, wk.p, 16, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        return ANQI_OK;
    };
    // visits / pool entries bound; a Vamana collect has no pool cap, so its
    // slow path is sized for every row it can meet (up to 2^18 entries)
    uint64_t need = P.vmode ? std::min<uint64_t>((uint64_t)P.n + P.cap + P.vM, 1u << 18)
                            : (uint64_t)P.pool_cap + P.cap + maxrev;
    // fast path: LDS work area.  The register-list AKNN gather keeps its
    // visited set in 16-bit entries (half the LDS, so more rows per CU).
    // The register-list graph gather's pool holds 1024 entries (four rows per
    // CU instead of three); rows that outgrow it are redone on the slow path.
    // A Vamana collect (about a thousand visits at L = 128) starts at 2048.
    const bool rk = gather_register_lists(P, graph);
    P.hmask = std::min<uint32_t>(graph ? 4096 : 8192, next_pow2(need * 3 / 2 + 64)) - 1;
    P.h16 = (!graph && rk && P.hmask + 1 == H16_SLOTS) ? (P.n <= H16_MAX_ID ? 1u : P.n <= H24_MAX_ID ? 2u : 0u) : 0u;
    P.pmax = graph ? std::min<uint32_t>((rk && !P.vmode) ? 1024u : 2048u, next_pow2(need + 64)) : 0;
    if (!P.rows) P.nrows = P.n;   // else: the given row list
    std::vector<uint32_t> bad;
    int rc = ANQI_OK;
    if (graph && P.pol.on) {
        // an edge-policy prune needs more room than LDS holds: every row
        // takes the global work area
        bad.resize(P.nrows);
        if (P.rows) {
            HCHK(hipMemcpyAsync(bad.data(), P.rows, 4ull * P.nrows, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
        } else {
            for (uint32_t r = 0; r < P.nrows; r++) bad[r] = r;
        }
    } else {
        HCHK(launch_gather(P, graph, false, s));
        rc = failed_rows(ctx, status.as<uint32_t>(), P.n, bad);
        if (rc) return rc;
        if (overflow_rows) *overflow_rows = (uint32_t)bad.size();
        if (bad.empty()) return read_work();
    }
    if (graph && rk && !P.pol.on && P.pmax < 2048u && next_pow2(need + 64) > P.pmax) {
        // rows that outgrew the short pool: once more in LDS with the full one
        GatherParams R = P;
        R.pmax = std::min<uint32_t>(2048, next_pow2(need + 64));
        DBuf rrows;
        HCHK(rrows.alloc(4 * bad.size()));
        HCHK(hipMemcpyAsync(rrows.p, bad.data(), 4 * bad.size(), hipMemcpyHostToDevice, s));
        R.rows = rrows.as<uint32_t>();
        R.nrows = (uint32_t)bad.size();
        HCHK(launch_gather(R, graph, false, s));
        rc = failed_rows(ctx, status.as<uint32_t>(), P.n, bad);
        if (rc) return rc;
        if (bad.empty()) return read_work();
    }
    // slow path: global work area sized for the bound (rows in chunks)
    auto run_global = [&](uint32_t hmask, uint32_t pmax) -> int {
        GatherParams Q = P;
        Q.h16 = 0;
        Q.hmask = hmask;
        Q.pmax = graph ? pmax : 0;
        size_t words = (gather_lds_bytes(Q, graph, false) + 3) / 4;
        DBuf rows, ws;
        HCHK(rows.alloc(4 * bad.size()));
        HCHK(hipMemcpyAsync(rows.p, bad.data(), 4 * bad.size(), hipMemcpyHostToDevice, s));
        HCHK(hipMemsetAsync(status.p, 0, 4ull * P.n, s));
        const uint32_t chunk = 4096;
        HCHK(ws.alloc(4 * words * std::min<size_t>(chunk, bad.size())));
        for (size_t b = 0; b < bad.size(); b += chunk) {
            Q.rows = rows.as<uint32_t>() + b;
            Q.nrows = (uint32_t)std::min<size_t>(chunk, bad.size() - b);
            Q.gscratch = ws.as<uint32_t>();
            Q.gscratch_words = words;
            HCHK(launch_gather(Q, graph, true, s));
        }
        return failed_rows(ctx, status.as<uint32_t>(), P.n, bad);
    };
    if (graph && P.pol.on) {
        // every row of an edge-policy prune: first with the fast path's sizes
        rc = run_global(P.hmask, P.pmax);
        if (rc) return rc;
        if (bad.empty()) return read_work();
    }
    rc = run_global((uint32_t)next_pow2(need * 2 + 64) - 1, graph ? (uint32_t)next_pow2(need + 64) : 0);
    if (rc) return rc;
    if (!bad.empty()) return anqi_set_err(ctx, ANQI_E_DEVICE, "gather: %zu rows exceed the work area", bad.size());
    return read_work();
}

int anqi_gather_topk(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* hidx,
                     const float* hval, uint32_t cap, const uint32_t* roff, const uint32_t* rids, uint32_t explore,
                     uint32_t pool_cap, uint32_t K, uint32_t* ids, float* keys, uint32_t* cnt, int form) {
    if (K > 256) return anqi_set_err(ctx, ANQI_E_USAGE, "K <= 256");
    std::vector<uint32_t> ro(2);
    uint32_t maxrev = 0;
    {
        std::vector<uint32_t> h(n + 1);
        HCHK(hipMemcpyAsync(h.data(), roff, 4ull * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
        HCHK(hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < n; i++) maxrev = std::max(maxrev, h[i + 1] - h[i]);
    }
    GatherParams P{};
    P.x = x; P.nrm = nrm; P.n = n; P.d = d; P.hidx = hidx; P.hval = hval; P.cap = cap; P.roff = roff; P.rids = rids;
    P.explore = explore; P.pool_cap = pool_cap; P.K = K; P.out_ids = ids; P.out_keys = keys; P.out_cnt = cnt;
    P.batched = form ? 1u : 0u;
    return gather_run(ctx, P, false, maxrev, nullptr);
}

int anqi_gather_prune(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* hidx,
                      const float* hval, uint32_t cap, const uint32_t* roff, const uint32_t* rids, uint32_t explore,
                      uint32_t pool_cap, uint32_t M, float alpha, uint32_t* O, uint32_t* ocnt, int form) {
    if (M > 64) return anqi_set_err(ctx, ANQI_E_USAGE, "graph degree <= 64");
    uint32_t maxrev = 0;
    {
        std::vector<uint32_t> h(n + 1);
        HCHK(hipMemcpyAsync(h.data(), roff, 4ull * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
        HCHK(hipStreamSynchronize(ctx->stream));
        for (uint32_t i = 0; i < n; i++) maxrev = std::max(maxrev, h[i + 1] - h[i]);
    }
    GatherParams P{};
    P.x = x; P.nrm = nrm; P.n = n; P.d = d; P.hidx = hidx; P.hval = hval; P.cap = cap; P.roff = roff; P.rids = rids;
    P.explore = explore; P.pool_cap = pool_cap; P.M = M; P.alpha = alpha; P.out_ids = O; P.out_cnt = ocnt;
    P.batched = form ? 1u : 0u;
    HCHK(launch_fill_u32(O, U32_NONE, (uint64_t)n * M, ctx->stream));   // slots past ocnt[i]
    return gather_run(ctx, P, true, maxrev, nullptr);
}

// The prune of commit_edges over P.rows (or every row) in LDS, then the rows
// whose pool outgrew it with a global work area.  P.status: n words.
static int commit_run(anqi_ctx* ctx, PruneParams P, uint32_t maxin) {
    hipStream_t s = ctx->stream;
    const uint32_t n = P.n, M = P.M;
    HCHK(hipMemsetAsync(P.status, 0, 4ull * n, s));
    if (!P.rows) P.nrows = n;
    P.pmax = std::min<uint32_t>(2048, next_pow2((uint64_t)M + maxin + 64));
    std::vector<uint32_t> bad;
    int rc = ANQI_OK;
    if (P.pol.on) {
        // an edge-policy prune takes the global work area for every row
        bad.resize(P.nrows);
        if (P.rows) {
            HCHK(hipMemcpyAsync(bad.data(), P.rows, 4ull * P.nrows, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
        } else {
            for (uint32_t r = 0; r < P.nrows; r++) bad[r] = r;
        }
    } else {
        HCHK(launch_commit(P, false, s));
        rc = failed_rows(ctx, P.status, n, bad);
        if (rc) return rc;
        if (bad.empty()) return ANQI_OK;
    }
    PruneParams Q = P;
    Q.pmax = next_pow2((uint64_t)M + maxin + 64);
    size_t words = (commit_lds_bytes(Q) + 3) / 4;
    DBuf rows, ws;
    HCHK(rows.alloc(4 * bad.size()));
    HCHK(hipMemcpyAsync(rows.p, bad.data(), 4 * bad.size(), hipMemcpyHostToDevice, s));
    HCHK(hipMemsetAsync(P.status, 0, 4ull * n, s));
    const size_t chunk = 4096;
    HCHK(ws.alloc(4 * words * std::min(chunk, bad.size())));
    for (size_t b = 0; b < bad.size(); b += chunk) {
        Q.rows = rows.as<uint32_t>() + b;
        Q.nrows = (uint32_t)std::min(chunk, bad.size() - b);
        Q.gscratch = ws.as<uint32_t>();
        Q.gscratch_words = words;
        HCHK(launch_commit(Q, true, s));
    }
    rc = failed_rows(ctx, P.status, n, bad);
    if (rc) return rc;
    if (!bad.empty()) return anqi_set_err(ctx, ANQI_E_DEVICE, "commit: %zu rows exceed the work area", bad.size());
    return ANQI_OK;
}

// commit_edges (include/index.h:907-930) under an edge policy.
static int commit_policy(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* O,
                         const uint32_t* ocnt, uint32_t M, float alpha, uint32_t* links, uint32_t* deg,
                         const EdgePolicyDev& pol);

int anqi_commit(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* O,
                const uint32_t* ocnt, uint32_t M, float alpha, uint32_t* links, uint32_t* deg) {
    return commit_policy(ctx, x, nrm, n, d, O, ocnt, M, alpha, links, deg, EdgePolicyDev{});
}

static int commit_policy(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* O,
                         const uint32_t* ocnt, uint32_t M, float alpha, uint32_t* links, uint32_t* deg,
                         const EdgePolicyDev& pol) {
    if (M > 64) return anqi_set_err(ctx, ANQI_E_USAGE, "graph degree <= 64");
    hipStream_t s = ctx->stream;
    DBuf roff, rids, status;
    HCHK(roff.alloc(4ull * (n + 1)));
    HCHK(rids.alloc(4ull * n * M));
    HCHK(status.alloc(4ull * n));
    uint32_t maxin = 0;
    int rc = reverse_csr(ctx, O, n, M, ocnt, n, roff.as<uint32_t>(), rids.as<uint32_t>(), &maxin);
    if (rc) return rc;
    PruneParams P{};
    P.x = x; P.nrm = nrm; P.n = n; P.d = d; P.O = O; P.ocnt = ocnt; P.roff = roff.as<uint32_t>();
    P.rids = rids.as<uint32_t>(); P.M = M; P.alpha = alpha; P.links = links; P.deg = deg;
    P.status = status.as<uint32_t>();
    P.pol = pol;
    HCHK(launch_fill_u32(links, U32_NONE, (uint64_t)n * M, s));   // slots past deg[i]
    return commit_run(ctx, P, maxin);
}

// The device form of an edge policy; false (with the error set) when it is
// not one this build runs.
static bool edge_policy_dev(anqi_ctx* ctx, const anqi_edge_policy* e, uint32_t M, EdgePolicyDev& out) {
    out = EdgePolicyDev{};
    if (!e || !(e->resid_cover || e->ball || e->lsg)) return true;
    if (e->nks > 16 || ((e->resid_cover || e->ball) && (!e->resid || e->nk == 0)) || (e->lsg && !e->mu)) {
        anqi_set_err(ctx, ANQI_E_USAGE, "edge policy: need the residual table (nk > 0), at most 16 ranks, lsg weights");
        return false;
    }
    for (uint32_t t = 0; t < e->nks; t++)
        if (e->ks[t] < 1 || (uint32_t)e->ks[t] > e->nk) {
            anqi_set_err(ctx, ANQI_E_USAGE, "edge policy: rank %d outside 1..%u", e->ks[t], e->nk);
            return false;
        }
    out.on = 1;
    out.resid_cover = e->resid_cover; out.ball = e->ball; out.ball_pure = e->ball_pure;
    out.occlude_all = e->occlude_all; out.lsg = e->lsg;
    out.lsg_alpha = e->lsg_alpha; out.lambda = e->lambda; out.scale2 = e->scale2;
    out.mu = e->mu; out.resid = e->resid; out.nk = e->nk; out.nks = e->nks;
    out.budget = std::min(e->budget, M);
    for (uint32_t t = 0; t < e->nks; t++) out.ks[t] = e->ks[t];
    out.bubble_auto = e->bubble_auto && e->ball;
    out.bubble_full = e->bubble_full;
    out.auto_candidates = std::max(1u, e->auto_candidates);
    out.auto_max_swaps = std::max(1u, e->auto_max_swaps);
    return true;
}

}  // extern "C"

// ---- RP forest ---------------------------------------------------------------------------------
struct Forest {
    DBuf perm;                               // T x n (positions t*n + local)
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> leaves;   // per tree (start, size)
    // node structure for a single tree (want_nodes)
    std::vector<uint32_t> left, right, entry_pos;
    std::vector<float> offset;
    DBuf hyper;                              // nodes x d
    std::vector<uint64_t> leaf_off;
    std::vector<uint32_t> leaf_pos_start, leaf_pos_size;
    uint32_t nodes = 0;
};

static int build_forest(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t T, uint32_t leaf,
                        uint32_t max_depth, uint64_t seed, bool want_nodes, Forest& F, bool angular = false) {
    hipStream_t s = ctx->stream;
    uint64_t total = (uint64_t)T * n;
    DBuf perm2, seg_of, side, flags, lscan, tmp;
    HCHK(F.perm.alloc(4 * total));
    HCHK(perm2.alloc(4 * total));
    HCHK(seg_of.alloc(4 * total));
    HCHK(side.alloc(total));
    HCHK(flags.alloc(4 * total));
    HCHK(lscan.alloc(4 * total));
    {
        std::vector<uint32_t> iota(n);
        for (uint32_t i = 0; i < n; i++) iota[i] = i;
        for (uint32_t t = 0; t < T; t++)
            HCHK(hipMemcpyAsync(F.perm.as<uint32_t>() + (size_t)t * n, iota.data(), 4ull * n, hipMemcpyHostToDevice, s));
        HCHK(hipStreamSynchronize(s));
    }
    size_t tb = 0;
    HCHK(prim_exclusive_scan_u32(nullptr, nullptr, total, nullptr, tb, s));
    HCHK(tmp.alloc(tb));
    struct Seg { uint32_t start, size, tree, node; };
    std::vector<Seg> cur;
    F.leaves.assign(T, {});
    for (uint32_t t = 0; t < T; t++) cur.push_back({(uint32_t)((uint64_t)t * n), n, t, 0});
    // node bookkeeping (single tree)
    if (want_nodes) { F.left.assign(1, U32_NONE); F.right.assign(1, U32_NONE); F.offset.assign(1, F32_BIG); }
    std::vector<std::unique_ptr<DBuf>> level_hyper;
    std::vector<std::vector<uint32_t>> level_nodes;
    std::vector<std::pair<uint32_t, Seg>> leaf_nodes;
    for (uint32_t depth = 0;; depth++) {
        std::vector<Seg> active;
        for (auto& g : cur) {
            if (g.size <= leaf) {
                F.leaves[g.tree].push_back({g.start, g.size});
                if (want_nodes) leaf_nodes.push_back({g.node, g});
            } else if (depth >= max_depth) {
                Seg h = g;
                h.size = std::min(leaf, g.size);   // rptrees.h:131-135
                F.leaves[g.tree].push_back({h.start, h.size});
                if (want_nodes) leaf_nodes.push_back({g.node, h});
            } else {
                active.push_back(g);
            }
        }
        if (active.empty()) break;
        uint32_t ns = (uint32_t)active.size();
        progress("forest depth %u: %u segments to split", depth, ns);
        std::vector<uint32_t> hs(ns), hz(ns), hp(2 * ns);
        for (uint32_t k = 0; k < ns; k++) {
            hs[k] = active[k].start;
            hz[k] = active[k].size;
            uint64_t key = ((uint64_t)active[k].tree << 40) ^ ((uint64_t)depth << 32) ^ active[k].start;
            uint32_t r0 = (uint32_t)(rng_at(seed, key, 0) % active[k].size);
            uint32_t r1 = (uint32_t)(rng_at(seed, key, 1) % active[k].size);
            if (r0 == r1) r0 = (r1 + 1) % active[k].size;   // rptrees.h:33-35
            hp[2 * k] = r0;
            hp[2 * k + 1] = r1;
        }
        DBuf dstart, dsize, dpick, dleft, ddegen, doff;
        auto hyp = std::make_unique<DBuf>();
        HCHK(dstart.alloc(4ull * ns));
        HCHK(dsize.alloc(4ull * ns));
        HCHK(dpick.alloc(8ull * ns));
        HCHK(dleft.alloc(4ull * ns));
        HCHK(ddegen.alloc(ns));
        HCHK(doff.alloc(4ull * ns));
        HCHK(hyp->alloc(4ull * ns * d));
        HCHK(hipMemcpyAsync(dstart.p, hs.data(), 4ull * ns, hipMemcpyHostToDevice, s));
        HCHK(hipMemcpyAsync(dsize.p, hz.data(), 4ull * ns, hipMemcpyHostToDevice, s));
        HCHK(hipMemcpyAsync(dpick.p, hp.data(), 8ull * ns, hipMemcpyHostToDevice, s));
        HCHK(launch_fill_u32(seg_of.as<uint32_t>(), U32_NONE, total, s));
        HCHK(launch_seg_fill(seg_of.as<uint32_t>(), dstart.as<uint32_t>(), dsize.as<uint32_t>(), ns, s));
        ForestLevel L{};
        L.x = x; L.n = n; L.d = d; L.perm = F.perm.as<uint32_t>(); L.perm_out = perm2.as<uint32_t>();
        L.seg_of = seg_of.as<uint32_t>(); L.seg_start = dstart.as<uint32_t>(); L.seg_size = dsize.as<uint32_t>();
        L.seg_pick = dpick.as<uint32_t>(); L.seg_hyper = hyp->as<float>(); L.seg_off = doff.as<float>();
        L.side = side.as<uint8_t>(); L.nseg = ns; L.total = total; L.angular = angular ? 1 : 0;
        HCHK(launch_split_planes(L, s));
        HCHK(launch_split_sides(L, seed ^ ((uint64_t)depth * 0x100000001ull), s));
        HCHK(hipMemsetAsync(dleft.p, 0, 4ull * ns, s));
        HCHK(launch_seg_count_left(side.as<uint8_t>(), seg_of.as<uint32_t>(), total, dleft.as<uint32_t>(), s));
        std::vector<uint32_t> hl(ns);
        HCHK(hipMemcpyAsync(hl.data(), dleft.p, 4ull * ns, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        std::vector<uint8_t> dg(ns, 0);
        bool anyd = false;
        for (uint32_t k = 0; k < ns; k++)
            if (hl[k] == 0 || hl[k] == hz[k]) { dg[k] = 1; anyd = true; }
        if (anyd) {
            HCHK(hipMemcpyAsync(ddegen.p, dg.data(), ns, hipMemcpyHostToDevice, s));
            HCHK(launch_seg_redeal(side.as<uint8_t>(), seg_of.as<uint32_t>(), ddegen.as<uint8_t>(), total,
                                   seed ^ 0xA5A5A5A5ull ^ depth, s));
            HCHK(hipMemsetAsync(dleft.p, 0, 4ull * ns, s));
            HCHK(launch_seg_count_left(side.as<uint8_t>(), seg_of.as<uint32_t>(), total, dleft.as<uint32_t>(), s));
            HCHK(hipMemcpyAsync(hl.data(), dleft.p, 4ull * ns, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
        }
        HCHK(launch_left_flags(side.as<uint8_t>(), seg_of.as<uint32_t>(), total, flags.as<uint32_t>(), s));
        size_t tbb = tb;
        HCHK(prim_exclusive_scan_u32(flags.as<uint32_t>(), lscan.as<uint32_t>(), total, tmp.p, tbb, s));
        HCHK(launch_seg_partition(F.perm.as<uint32_t>(), perm2.as<uint32_t>(), side.as<uint8_t>(),
                                  seg_of.as<uint32_t>(), dstart.as<uint32_t>(), dleft.as<uint32_t>(),
                                  lscan.as<uint32_t>(), total, s));
        std::swap(F.perm.p, perm2.p);
        std::vector<float> hoff;
        if (want_nodes) {
            hoff.resize(ns);
            HCHK(hipMemcpyAsync(hoff.data(), doff.p, 4ull * ns, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
        }
        std::vector<Seg> next;
        std::vector<uint32_t> lvl_nodes(ns);
        for (uint32_t k = 0; k < ns; k++) {
            Seg g = active[k];
            uint32_t L0 = hl[k];
            uint32_t ln = 0, rn = 0;
            if (want_nodes) {
                ln = (uint32_t)F.left.size();
                rn = ln + 1;
                F.left.push_back(U32_NONE); F.right.push_back(U32_NONE); F.offset.push_back(F32_BIG);
                F.left.push_back(U32_NONE); F.right.push_back(U32_NONE); F.offset.push_back(F32_BIG);
                F.left[g.node] = ln;
                F.right[g.node] = rn;
                F.offset[g.node] = hoff[k];
                lvl_nodes[k] = g.node;
            }
            next.push_back({g.start, L0, g.tree, ln});
            next.push_back({g.start + L0, g.size - L0, g.tree, rn});
        }
        if (want_nodes) {
            level_nodes.push_back(lvl_nodes);
            level_hyper.push_back(std::move(hyp));
        }
        cur.swap(next);
    }
    if (want_nodes) {
        F.nodes = (uint32_t)F.left.size();
        HCHK(F.hyper.alloc(4ull * F.nodes * d));
        HCHK(hipMemsetAsync(F.hyper.p, 0, 4ull * F.nodes * d, s));
        for (size_t lv = 0; lv < level_nodes.size(); lv++)
            for (size_t k = 0; k < level_nodes[lv].size(); k++)
                HCHK(hipMemcpyAsync(F.hyper.as<float>() + (size_t)level_nodes[lv][k] * d,
                                    level_hyper[lv]->as<float>() + k * d, 4ull * d, hipMemcpyDeviceToDevice, s));
        F.entry_pos.assign(F.nodes, 0);
        F.leaf_pos_start.assign(F.nodes, 0);
        F.leaf_pos_size.assign(F.nodes, 0);
        for (auto& ln : leaf_nodes) {
            F.leaf_pos_start[ln.first] = ln.second.start;
            F.leaf_pos_size[ln.first] = ln.second.size;
        }
        HCHK(hipStreamSynchronize(s));
    }
    return ANQI_OK;
}

extern "C" {

int anqi_tree_free(anqi_ctx* ctx, anqi_tree* t) {
    (void)ctx;
    if (!t) return ANQI_OK;
    if (t->left) (void)hipFree(t->left);
    if (t->right) (void)hipFree(t->right);
    if (t->offset) (void)hipFree(t->offset);
    if (t->hyper) (void)hipFree(t->hyper);
    if (t->entry) (void)hipFree(t->entry);
    delete[] t->leaf_off;
    delete[] t->leaf_ids;
    std::memset(t, 0, sizeof(*t));
    return ANQI_OK;
}

int anqi_tree_upload(anqi_ctx* ctx, uint32_t nodes, uint32_t root, uint32_t d, const uint32_t* left,
                     const uint32_t* right, const float* offset, const float* hyper, const uint32_t* entry,
                     anqi_tree* out) {
    hipStream_t s = ctx->stream;
    std::memset(out, 0, sizeof(*out));
    out->nodes = nodes; out->root = root; out->d = d;
    HCHK(hipMalloc((void**)&out->left, 4ull * nodes));
    HCHK(hipMalloc((void**)&out->right, 4ull * nodes));
    HCHK(hipMalloc((void**)&out->offset, 4ull * nodes));
    HCHK(hipMalloc((void**)&out->hyper, 4ull * nodes * d));
    HCHK(hipMalloc((void**)&out->entry, 4ull * nodes));
    HCHK(hipMemcpyAsync(out->left, left, 4ull * nodes, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(out->right, right, 4ull * nodes, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(out->offset, offset, 4ull * nodes, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(out->hyper, hyper, 4ull * nodes * d, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(out->entry, entry, 4ull * nodes, hipMemcpyHostToDevice, s));
    HCHK(hipStreamSynchronize(s));
    return ANQI_OK;
}

// The reference's lifted search tree (rptrees.h:155-163 via index.h:1116): one
// std::mt19937 seeded with `seed` feeds every draw, depth first.  The engine's
// outputs do not depend on the data, so they are generated here and the
// device walks the chain (k_tree.hip).  A stream or node table that turns out
// too short is enlarged and the (deterministic) build rerun.
int anqi_rp_tree(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t leaf, uint32_t max_depth,
                 uint64_t seed, anqi_tree* out) {
    if (n == 0 || leaf == 0) return anqi_set_err(ctx, ANQI_E_INPUT, "search tree needs n > 0 and leaf > 0");
    if (d > 16 * QMAX_CHUNKS + 15) return anqi_set_err(ctx, ANQI_E_USAGE, "dimension %u unsupported", d);
    if (max_depth + 2 > TB_STACK) return anqi_set_err(ctx, ANQI_E_USAGE, "max_depth %u unsupported", max_depth);
    hipStream_t s = ctx->stream;
    std::memset(out, 0, sizeof(*out));
    uint64_t stream_len = 8ull * n / leaf + 8192;
    uint64_t cap = std::min<uint64_t>(2ull * n + 2, 4ull * n / leaf + 1024);
    std::vector<uint32_t> stream;
    Mt19937 mt((uint32_t)seed);
    DBuf dperm, dtmp, dside, dstatus;
    HCHK(dperm.alloc(4ull * n));
    HCHK(dtmp.alloc(4ull * n));
    HCHK(dside.alloc(n));
    HCHK(dstatus.alloc(16));
    uint32_t status[4] = {0, 0, 0, 0};
    for (int attempt = 0;; attempt++) {
        while (stream.size() < stream_len) stream.push_back(mt.next());
        DBuf dstream, dleft, dright, doff, dhyper, dls, dlz;
        HCHK(dstream.alloc(4ull * stream_len));
        HCHK(dleft.alloc(4ull * cap));
        HCHK(dright.alloc(4ull * cap));
        HCHK(doff.alloc(4ull * cap));
        HCHK(dhyper.alloc(4ull * cap * d));
        HCHK(dls.alloc(4ull * cap));
        HCHK(dlz.alloc(4ull * cap));
        HCHK(hipMemcpyAsync(dstream.p, stream.data(), 4ull * stream_len, hipMemcpyHostToDevice, s));
        TreeBuildParams P{};
        P.x = x; P.n = n; P.d = d; P.leaf = leaf; P.max_depth = max_depth;
        P.stream = (const uint32_t*)dstream.p; P.stream_len = stream_len;
        P.perm = (uint32_t*)dperm.p; P.tmp = (uint32_t*)dtmp.p; P.side = (uint8_t*)dside.p;
        P.left = (uint32_t*)dleft.p; P.right = (uint32_t*)dright.p; P.offset = (float*)doff.p;
        P.hyper = (float*)dhyper.p; P.leaf_start = (uint32_t*)dls.p; P.leaf_size = (uint32_t*)dlz.p;
        P.node_cap = (uint32_t)cap; P.status = (uint32_t*)dstatus.p;
        HCHK(launch_tree_build(P, s));
        HCHK(hipMemcpyAsync(status, dstatus.p, 16, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        if (status[1] == TB_ERR_STREAM && attempt < 40) { stream_len = stream_len * 2 + n; continue; }
        if (status[1] == TB_ERR_NODES && attempt < 40 && cap < 4ull * n + 4) {
            cap = std::min<uint64_t>(4ull * n + 4, cap * 2);
            continue;
        }
        if (status[1] != 0) return anqi_set_err(ctx, ANQI_E_DEVICE, "search tree build failed (code %u)", status[1]);
        const uint32_t nodes = status[0];
        std::vector<uint32_t> left(nodes), ls(nodes), lz(nodes), perm(n);
        HCHK(hipMemcpyAsync(left.data(), dleft.p, 4ull * nodes, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(ls.data(), dls.p, 4ull * nodes, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(lz.data(), dlz.p, 4ull * nodes, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(perm.data(), dperm.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipMalloc((void**)&out->left, 4ull * nodes));
        HCHK(hipMalloc((void**)&out->right, 4ull * nodes));
        HCHK(hipMalloc((void**)&out->offset, 4ull * nodes));
        HCHK(hipMalloc((void**)&out->hyper, 4ull * nodes * d));
        HCHK(hipMalloc((void**)&out->entry, 4ull * nodes));
        HCHK(hipMemcpyAsync(out->left, dleft.p, 4ull * nodes, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(out->right, dright.p, 4ull * nodes, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(out->offset, doff.p, 4ull * nodes, hipMemcpyDeviceToDevice, s));
        HCHK(hipMemcpyAsync(out->hyper, dhyper.p, 4ull * nodes * d, hipMemcpyDeviceToDevice, s));
        HCHK(launch_tree_entries(out->left, (const uint32_t*)dls.p, (const uint32_t*)dlz.p,
                                 (const uint32_t*)dperm.p, nodes, out->entry, s));
        HCHK(hipStreamSynchronize(s));
        out->nodes = nodes; out->root = 0; out->d = d;
        out->leaf_off = new uint64_t[nodes + 1];
        uint64_t tot = 0;
        out->leaf_off[0] = 0;
        for (uint32_t k = 0; k < nodes; k++) {
            if (left[k] == U32_NONE) tot += lz[k];
            out->leaf_off[k + 1] = tot;
        }
        out->leaf_ids = new uint32_t[tot ? tot : 1];
        for (uint32_t k = 0; k < nodes; k++)
            if (left[k] == U32_NONE) std::memcpy(out->leaf_ids + out->leaf_off[k], perm.data() + ls[k], 4ull * lz[k]);
        progress("search tree: %u nodes, %llu draws", nodes,
                 (unsigned long long)(((uint64_t)status[3] << 32) | status[2]));
        return ANQI_OK;
    }
}

// ---- AKNN source -------------------------------------------------------------------------------
int anqi_aknn_build(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, const anqi_aknn_params* p,
                    uint32_t* ids, float* keys, anqi_aknn_stats* st) {
    if (!p || p->K == 0 || p->K >= n || p->K > 256) return anqi_set_err(ctx, ANQI_E_INPUT, "invalid base file or K");
    if (d > 16 * QMAX_CHUNKS + 15) return anqi_set_err(ctx, ANQI_E_USAGE, "dimension %u unsupported", d);
    if (p->metric != ANQI_METRIC_L2 && p->metric != ANQI_METRIC_IP)
        return anqi_set_err(ctx, ANQI_E_USAGE, "aknn: metric ANQI_METRIC_L2 or ANQI_METRIC_IP");
    // "ip": angular splits and 1 - <x, y> distances (index.h:149, 164-176);
    // the leaf pairs run in the FP32 kernel
    const int ipm = p->metric == ANQI_METRIC_IP;
    hipStream_t s = ctx->stream;
    anqi_aknn_stats S{};
    double t0 = now_s();
    const uint32_t K = p->K;
    DBuf nrm, lid, lkey, lcnt;
    HCHK(nrm.alloc(4ull * n));
    HCHK(launch_norms(x, n, d, nrm.as<float>(), s));
    Forest F;
    int rc = build_forest(ctx, x, n, d, p->trees ? p->trees : 16, std::max<uint32_t>(p->leaf ? p->leaf : K, 10),
                          p->max_depth ? p->max_depth : 100, p->seed, false, F, ipm != 0);
    if (rc) return rc;
    HCHK(hipStreamSynchronize(s));
    double t1 = now_s();
    S.forest_s = t1 - t0;
    // leaf all-pairs (init_knn)
    HCHK(lid.alloc(4ull * n * K));
    HCHK(lkey.alloc(4ull * n * K));
    HCHK(lcnt.alloc(4ull * n));
    HCHK(hipMemsetAsync(lcnt.p, 0, 4ull * n, s));
    HCHK(launch_fill_u32(lid.as<uint32_t>(), U32_NONE, (uint64_t)n * K, s));
    uint32_t T = (uint32_t)F.leaves.size();
    // bf16 copy and per-row error bounds for the MFMA leaf sweep
    uint32_t max_leaf_all = 0;
    for (auto& lv : F.leaves)
        for (auto& pr : lv) max_leaf_all = std::max(max_leaf_all, pr.second);
    if (p->precision < 0 || p->precision > 2) return anqi_set_err(ctx, ANQI_E_USAGE, "precision %d", p->precision);
    const bool use_mfma = !ipm && p->precision >= 1 && leaf_mfma_supported(d, max_leaf_all, K);
    const bool use_fp8 = use_mfma && p->precision == 2;
    DBuf xb, lmarg, nsurv, cbuf, ccnt;
    float fp8_scale = 1.0f;
    if (use_mfma) {
        DBuf nx, ex;
        // candidate lists between the sweep and merge passes (a row's
        // survivors in one tree are fewer than its leaf's size)
        HCHK(cbuf.alloc(4ull * n * max_leaf_all));
        HCHK(ccnt.alloc(4ull * n));
        HCHK(hipMemsetAsync(ccnt.p, 0, 4ull * n, s));
        HCHK(xb.alloc((use_fp8 ? 1ull : 2ull) * n * d));
        HCHK(nx.alloc(4ull * n));
        HCHK(ex.alloc(4ull * n));
        HCHK(lmarg.alloc(4ull * n));
        HCHK(nsurv.alloc(8));
        HCHK(hipMemsetAsync(nsurv.p, 0, 8, s));
        std::vector<float> hn(n), he(n);
        if (use_fp8) {
            // power-of-two scale putting every |x_k| <= max row norm under e4m3's 448
            HCHK(launch_exact_norms2(x, nrm.as<float>(), n, d, nx.as<float>(), ex.as<float>(), s));
            HCHK(hipMemcpyAsync(hn.data(), nx.p, 4ull * n, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
            float Y = 0.0f;
            for (uint32_t i = 0; i < n; i++) Y = std::max(Y, hn[i]);
            int e2 = 0;
            std::frexp(std::max(Y, 1e-30f) / 448.0f, &e2);
            fp8_scale = std::ldexp(1.0f, e2);
            HCHK(launch_to_fp8(x, nrm.as<float>(), n, d, fp8_scale, xb.as<uint8_t>(), nx.as<float>(), ex.as<float>(),
                               s));
        } else {
            HCHK(launch_to_bf16(x, (uint64_t)n * d, xb.as<uint16_t>(), s));
            HCHK(launch_exact_norms2(x, nrm.as<float>(), n, d, nx.as<float>(), ex.as<float>(), s));
        }
        HCHK(hipMemcpyAsync(hn.data(), nx.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipMemcpyAsync(he.data(), ex.p, 4ull * n, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        float Y = 0.0f, E = 0.0f;
        for (uint32_t i = 0; i < n; i++) { Y = std::max(Y, hn[i]); E = std::max(E, he[i]); }
        HCHK(launch_exact_margin(nx.as<float>(), ex.as<float>(), n, d, Y, E, lmarg.as<float>(), s));
        HCHK(hipStreamSynchronize(s));
    }
    for (uint32_t t = 0; t < T; t++) {
        auto& lv = F.leaves[t];
        std::vector<uint32_t> hs(lv.size()), hz(lv.size());
        uint32_t mx = 0;
        for (size_t k = 0; k < lv.size(); k++) {
            hs[k] = lv[k].first - t * n;
            hz[k] = lv[k].second;
            mx = std::max(mx, hz[k]);
            S.leaf_pairs += (uint64_t)hz[k] * (hz[k] - (hz[k] ? 1 : 0)) / 2;
        }
        S.leaves += (uint32_t)lv.size();
        S.max_leaf = std::max(S.max_leaf, mx);
        DBuf ds, dz;
        HCHK(ds.alloc(4 * hs.size()));
        HCHK(dz.alloc(4 * hz.size()));
        HCHK(hipMemcpyAsync(ds.p, hs.data(), 4 * hs.size(), hipMemcpyHostToDevice, s));
        HCHK(hipMemcpyAsync(dz.p, hz.data(), 4 * hz.size(), hipMemcpyHostToDevice, s));
        LeafParams LP{};
        LP.x = x; LP.nrm = nrm.as<float>(); LP.n = n; LP.d = d; LP.perm = F.perm.as<uint32_t>() + (size_t)t * n;
        LP.leaf_start = ds.as<uint32_t>(); LP.leaf_size = dz.as<uint32_t>(); LP.nleaves = (uint32_t)lv.size();
        LP.K = K; LP.ids = lid.as<uint32_t>(); LP.keys = lkey.as<float>(); LP.cnt = lcnt.as<uint32_t>();
        LP.max_leaf = mx;
        LP.nsurv = use_mfma ? nsurv.as<unsigned long long>() : nullptr;
        LP.ipm = ipm;
        if (use_mfma)
            HCHK(launch_leaf_mfma(LP, use_fp8, xb.p, fp8_scale, lmarg.as<float>(), cbuf.as<uint32_t>(),
                                  ccnt.as<uint32_t>(), max_leaf_all, s));
        else
            HCHK(launch_leaf_knn(LP, false, s));
        HCHK(hipStreamSynchronize(s));   // keeps ds/dz alive until the launch ends
        if ((t & 7) == 7 || t + 1 == T) progress("leaf all-pairs: tree %u/%u", t + 1, T);
    }
    HCHK(launch_random_fill(x, nrm.as<float>(), n, d, K, p->seed ^ 0x5EEDull, ipm, lid.as<uint32_t>(),
                            lkey.as<float>(), lcnt.as<uint32_t>(), s));
    HCHK(hipStreamSynchronize(s));
    if (use_mfma) {
        HCHK(hipMemcpy(&S.leaf_survivors, nsurv.p, 8, hipMemcpyDeviceToHost));
        progress("leaf sweep (%s): %llu of %llu ordered pairs passed the filter", use_fp8 ? "fp8" : "bf16",
                 (unsigned long long)S.leaf_survivors, (unsigned long long)(2 * S.leaf_pairs));
    }
    double t2 = now_s();
    S.leaf_s = t2 - t1;
    HCHK(F.perm.alloc(16));
    // gather (build_approx_knn tail)
    DBuf roff, rids, ocnt;
    HCHK(roff.alloc(4ull * (n + 1)));
    HCHK(rids.alloc(4ull * n * K));
    HCHK(ocnt.alloc(4ull * n));
    rc = reverse_csr(ctx, lid.as<uint32_t>(), n, K, nullptr, n, roff.as<uint32_t>(), rids.as<uint32_t>(), nullptr);
    if (rc) return rc;
    progress("aknn: forest %.2fs leaves %.2fs; gathering", S.forest_s, S.leaf_s);
    uint32_t ef = p->ef ? p->ef : 100;
    uint32_t pool_cap = std::max<uint32_t>(K, (p->ef_alpha ? p->ef_alpha : 8) * ef);
    uint32_t maxrev = 0;
    {
        std::vector<uint32_t> h(n + 1);
        HCHK(hipMemcpyAsync(h.data(), roff.p, 4ull * (n + 1), hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < n; i++) maxrev = std::max(maxrev, h[i + 1] - h[i]);
    }
    GatherParams G{};
    G.x = x; G.nrm = nrm.as<float>(); G.n = n; G.d = d; G.hidx = lid.as<uint32_t>(); G.hval = lkey.as<float>();
    G.cap = K; G.roff = roff.as<uint32_t>(); G.rids = rids.as<uint32_t>(); G.explore = ef; G.pool_cap = pool_cap;
    G.K = K; G.out_ids = ids; G.out_keys = keys; G.out_cnt = ocnt.as<uint32_t>();
    G.batched = p->precision >= 1 ? 1u : 0u;
    G.ipm = ipm;
    uint64_t work[2] = {0, 0};
    rc = gather_run(ctx, G, false, maxrev, &S.overflow_rows, work);
    if (rc) return rc;
    S.gather_d2 = work[0];
    S.gather_adj = work[1];
    std::vector<uint32_t> oc(n);
    HCHK(hipMemcpyAsync(oc.data(), ocnt.p, 4ull * n, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; i++) S.short_rows += oc[i] < K;
    double t3 = now_s();
    S.gather_s = t3 - t2;
    S.total_s = t3 - t0;
    if (st) *st = S;
    return ANQI_OK;
}

// ---- graph -------------------------------------------------------------------------------------
// ---- Vamana refinement ----------------------------------------------------------------
// std::mt19937_64 (Nishimura 2000) and std::shuffle as libstdc++ (GCC 11) runs
// it with a 64-bit engine over fewer than 2^32 elements: after one draw in
// [0, 2) when the length is even, each draw in [0, b (b + 1)) places two
// successive elements (x / (b + 1), x % (b + 1)); a draw in [0, r) is
// Lemire's nearly-divisionless method on the 128-bit product.
struct Mt64 {
    uint64_t x[312];
    uint32_t i = 312;
    explicit Mt64(uint64_t seed) {
        x[0] = seed;
        for (uint32_t k = 1; k < 312; k++) x[k] = 6364136223846793005ull * (x[k - 1] ^ (x[k - 1] >> 62)) + k;
    }
    uint64_t next() {
        if (i >= 312) {
            for (uint32_t k = 0; k < 312; k++) {
                const uint64_t y = (x[k] & 0xFFFFFFFF80000000ull) | (x[(k + 1) % 312] & 0x7FFFFFFFull);
                x[k] = x[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1u) ? 0xB5026F5AA96619E9ull : 0ull);
            }
            i = 0;
        }
        uint64_t y = x[i++];
        y ^= (y >> 29) & 0x5555555555555555ull;
        y ^= (y << 17) & 0x71D67FFFEDA60000ull;
        y ^= (y << 37) & 0xFFF7EEE000000000ull;
        return y ^ (y >> 43);
    }
    uint64_t below(uint64_t r) {
        unsigned __int128 p = (unsigned __int128)next() * r;
        uint64_t low = (uint64_t)p;
        if (low < r) {
            const uint64_t t = (0ull - r) % r;
            while (low < t) {
                p = (unsigned __int128)next() * r;
                low = (uint64_t)p;
            }
        }
        return (uint64_t)(p >> 64);
    }
};

static void shuffle_u32(std::vector<uint32_t>& a, Mt64& g) {
    const uint64_t n = a.size();
    if (n == 0) return;
    uint64_t i = 1;
    if (n % 2 == 0) std::swap(a[i++], a[g.below(2)]);
    while (i != n) {
        const uint64_t b = i + 1, r = g.below(b * (b + 1));
        std::swap(a[i++], a[r / (b + 1)]);
        std::swap(a[i++], a[r % (b + 1)]);
    }
}

// The gather of a Vamana collect: graph mode, beam L from the entries, pool
// uncapped, pruned with alpha into O/ocnt (rows indexed by row id).
static GatherParams vamana_gather_params(const float* lifted, const float* nrm, uint32_t n, uint32_t dl,
                                         const uint32_t* hidx, const float* hval, uint32_t cap,
                                         const anqi_graph_params* p, const uint32_t* links, const uint32_t* deg,
                                         float alpha, uint32_t* O, uint32_t* ocnt) {
    GatherParams G{};
    G.x = lifted; G.nrm = nrm; G.n = n; G.d = dl; G.hidx = hidx; G.hval = hval; G.cap = cap;
    G.explore = p->vamana_L ? p->vamana_L : 128;
    G.pool_cap = UINT32_MAX;
    G.M = p->M; G.alpha = alpha; G.out_ids = O; G.out_cnt = ocnt;
    G.batched = 1;
    G.vlinks = links; G.vdeg = deg; G.vM = p->M;
    return G;
}

// ANQI_VAMANA after the commit (include/index.h:1262-1282): global passes
// (collect from each row's tree entry, heap slots, prune, commit_edges) or
// the batch insertion of vamana_batch_refine (:932-1063).
static int vamana_refine(anqi_ctx* ctx, const float* lifted, const float* nrm, uint32_t n, uint32_t dl,
                         const uint32_t* hidx, const float* hval, uint32_t cap, const anqi_graph_params* p,
                         const anqi_tree* tree, const EdgePolicyDev& pol, uint32_t* links, uint32_t* deg) {
    hipStream_t s = ctx->stream;
    const uint32_t M = p->M;
    const int passes = p->vamana_passes;
    if (n == 0 || passes <= 0) return ANQI_OK;
    DBuf O, ocnt;
    HCHK(O.alloc(4ull * n * M));
    HCHK(ocnt.alloc(4ull * n));
    if (p->vamana == 1) {
        DBuf ent;
        HCHK(ent.alloc(4ull * n));
        HCHK(launch_tree_entry_rows(lifted, n, dl, tree->left, tree->right, tree->offset, tree->hyper, tree->entry,
                                    tree->root, ent.as<uint32_t>(), s));
        for (int pass = 0; pass < passes; pass++) {
            GatherParams G = vamana_gather_params(lifted, nrm, n, dl, hidx, hval, cap, p, links, deg, p->alpha,
                                                  O.as<uint32_t>(), ocnt.as<uint32_t>());
            G.vmode = 1;
            G.ventry = ent.as<uint32_t>();
            G.pol = pol;
            int rc = gather_run(ctx, G, true, 0, nullptr);
            if (rc) return rc;
            rc = commit_policy(ctx, lifted, nrm, n, dl, O.as<uint32_t>(), ocnt.as<uint32_t>(), M, p->alpha, links,
                               deg, pol);
            if (rc) return rc;
            progress("vamana global pass %d", pass);
        }
        return ANQI_OK;
    }
    // batch style
    double frac = p->vamana_batch_frac > 0.0 ? p->vamana_batch_frac : 0.02;
    uint64_t max_batch = std::min<uint64_t>((uint64_t)std::max(1.0, frac * (double)n), p->vamana_max_batch);
    if (max_batch == 0) max_batch = n;
    const double base = p->vamana_batch_base > 1.0 ? p->vamana_batch_base : 2.0;
    const uint32_t entry = p->vamana_entry < n ? p->vamana_entry : 0;
    const uint64_t maxb = n;   // a batch never exceeds the rows (base^k (base - 1) may exceed max_batch)
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    DBuf dorder, Ob, obcnt, roff, rids, tl, tc, status;
    HCHK(dorder.alloc(4ull * n));
    HCHK(Ob.alloc(4ull * maxb * M));
    HCHK(obcnt.alloc(4ull * maxb));
    HCHK(roff.alloc(4ull * (n + 1)));
    HCHK(rids.alloc(4ull * maxb * M));
    HCHK(tl.alloc(4ull * n));
    HCHK(tc.alloc(4));
    HCHK(status.alloc(4ull * n));
    for (int pass = 0; pass < passes; pass++) {
        const float alpha = (p->vamana_early_alpha_one && pass + 1 < passes) ? p->vamana_early_alpha : p->alpha;
        Mt64 g(p->seed + 0x9e3779b97f4a7c15ull * (uint64_t)(pass + 1));
        shuffle_u32(order, g);
        HCHK(hipMemcpyAsync(dorder.p, order.data(), 4ull * n, hipMemcpyHostToDevice, s));
        uint64_t count = 0, inc = 0;
        while (count < n) {
            uint64_t fl = count, ce;
            const double p0 = std::pow(base, (double)inc);
            if (p0 <= (double)max_batch) {
                fl = (uint64_t)p0 - 1;
                ce = std::min<uint64_t>((uint64_t)std::pow(base, (double)(inc + 1)) - 1, n);
            } else {
                ce = std::min<uint64_t>(count + max_batch, n);
            }
            count = ce;
            inc++;
            if (ce <= fl || fl >= n) continue;
            const uint32_t nb = (uint32_t)(ce - fl);
            const uint32_t* brows = dorder.as<uint32_t>() + fl;
            GatherParams G = vamana_gather_params(lifted, nrm, n, dl, hidx, hval, cap, p, links, deg, alpha,
                                                  O.as<uint32_t>(), ocnt.as<uint32_t>());
            G.vmode = 2;
            G.ventry0 = entry;
            G.rows = brows;
            G.nrows = nb;
            G.pol = pol;
            int rc = gather_run(ctx, G, true, 0, nullptr);
            if (rc) return rc;
            HCHK(launch_vamana_write(brows, nb, O.as<uint32_t>(), ocnt.as<uint32_t>(), M, links, deg,
                                     Ob.as<uint32_t>(), obcnt.as<uint32_t>(), s));
            uint32_t maxin = 0;
            rc = reverse_csr(ctx, Ob.as<uint32_t>(), nb, M, obcnt.as<uint32_t>(), n, roff.as<uint32_t>(),
                             rids.as<uint32_t>(), &maxin);
            if (rc) return rc;
            HCHK(launch_nonempty_rows(roff.as<uint32_t>(), n, tl.as<uint32_t>(), tc.as<uint32_t>(), s));
            uint32_t nt = 0;
            HCHK(hipMemcpyAsync(&nt, tc.p, 4, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
            if (nt == 0) continue;
            // the rows the batch points to: their new edges plus the batch rows
            // pointing to them, pruned in place
            PruneParams P{};
            P.x = lifted; P.nrm = nrm; P.n = n; P.d = dl; P.O = links; P.ocnt = deg;
            P.roff = roff.as<uint32_t>(); P.rids = rids.as<uint32_t>(); P.rmap = brows;
            P.M = M; P.alpha = alpha; P.links = links; P.deg = deg; P.status = status.as<uint32_t>();
            P.rows = tl.as<uint32_t>(); P.nrows = nt;
            P.pol = pol;
            rc = commit_run(ctx, P, maxin);
            if (rc) return rc;
        }
        progress("vamana batch pass %d", pass);
    }
    return ANQI_OK;
}

int anqi_vamana_medoid(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint64_t sample, uint32_t* entry) {
    if (!ctx || !x || !entry || n == 0 || d == 0) return anqi_set_err(ctx, ANQI_E_USAGE, "vamana medoid: need rows");
    DBuf mean, best;
    HCHK(mean.alloc(8ull * d));
    HCHK(best.alloc(16));
    HCHK(launch_vamana_medoid(x, n, d, sample, mean.as<double>(), best.as<unsigned long long>(), entry, ctx->stream));
    return ANQI_OK;
}

int anqi_graph_build(anqi_ctx* ctx, const float* lifted, uint32_t n, uint32_t dl, const uint32_t* warm_ids,
                     uint32_t width, const anqi_graph_params* p, uint32_t* links, uint32_t* deg, anqi_tree* tree,
                     anqi_graph_stats* st) {
    if (!p || p->M == 0 || p->M > 64 || p->nn_k == 0 || p->nn_k > 256 || p->vamana < 0 || p->vamana > 2 ||
        (p->vamana && (p->vamana_L == 0 || p->vamana_L > 1024)))
        return anqi_set_err(ctx, ANQI_E_USAGE, "invalid graph parameters");
    hipStream_t s = ctx->stream;
    anqi_graph_stats S{};
    double t0 = now_s();
    DBuf nrm, hidx, hval, roff, rids, O, ocnt;
    HCHK(nrm.alloc(4ull * n));
    HCHK(launch_norms(lifted, n, dl, nrm.as<float>(), s));
    int rc = anqi_rp_tree(ctx, lifted, n, dl, p->leaf ? p->leaf : 100, 100, p->seed, tree);
    if (rc) return rc;
    double t1 = now_s();
    S.tree_s = t1 - t0;
    const uint32_t cap = p->nn_k;
    HCHK(hidx.alloc(4ull * n * cap));
    HCHK(hval.alloc(4ull * n * cap));
    HCHK(launch_warm_start(lifted, nrm.as<float>(), n, dl, warm_ids, width, cap, hidx.as<uint32_t>(),
                           hval.as<float>(), s));
    HCHK(hipStreamSynchronize(s));
    double t2 = now_s();
    S.warm_s = t2 - t1;
    int passes = 0;
    rc = anqi_nndescent(ctx, lifted, nrm.as<float>(), n, dl, cap, hidx.as<uint32_t>(), hval.as<float>(),
                        std::min(p->iters, 16), p->sample, S.nnd_upd, &passes, p->form);
    if (rc) return rc;
    S.nnd_passes = passes;
    HCHK(hipStreamSynchronize(s));
    double t3 = now_s();
    S.nnd_s = t3 - t2;
    HCHK(roff.alloc(4ull * (n + 1)));
    HCHK(rids.alloc(4ull * n * cap));
    uint32_t maxrev = 0;
    rc = reverse_csr(ctx, hidx.as<uint32_t>(), n, cap, nullptr, n, roff.as<uint32_t>(), rids.as<uint32_t>(), &maxrev);
    if (rc) return rc;
    HCHK(O.alloc(4ull * n * p->M));
    HCHK(ocnt.alloc(4ull * n));
    GatherParams G{};
    G.x = lifted; G.nrm = nrm.as<float>(); G.n = n; G.d = dl; G.hidx = hidx.as<uint32_t>(); G.hval = hval.as<float>();
    G.cap = cap; G.roff = roff.as<uint32_t>(); G.rids = rids.as<uint32_t>(); G.explore = p->explore ? p->explore : 100;
    G.pool_cap = std::max<uint32_t>(p->M, (p->ef_alpha ? p->ef_alpha : 8) * G.explore);
    G.M = p->M; G.alpha = p->alpha; G.out_ids = O.as<uint32_t>(); G.out_cnt = ocnt.as<uint32_t>();
    G.batched = p->form ? 1u : 0u;
    EdgePolicyDev pol;
    if (!edge_policy_dev(ctx, p->edge, p->M, pol)) return ANQI_E_USAGE;
    G.pol = pol;
    uint64_t work[2] = {0, 0};
    rc = gather_run(ctx, G, true, maxrev, &S.overflow_rows, work);
    if (rc) return rc;
    S.gather_d2 = work[0];
    S.gather_adj = work[1];
    HCHK(hipStreamSynchronize(s));
    double t4 = now_s();
    S.gather_s = t4 - t3;
    progress("graph: gather+prune %.2fs (overflow rows %u); commit", S.gather_s, S.overflow_rows);
    rc = commit_policy(ctx, lifted, nrm.as<float>(), n, dl, O.as<uint32_t>(), ocnt.as<uint32_t>(), p->M, p->alpha,
                       links, deg, pol);
    if (rc) return rc;
    HCHK(hipStreamSynchronize(s));
    double t5 = now_s();
    S.commit_s = t5 - t4;
    if (p->vamana) {
        rc = vamana_refine(ctx, lifted, nrm.as<float>(), n, dl, hidx.as<uint32_t>(), hval.as<float>(), cap, p, tree,
                           pol, links, deg);
        if (rc) return rc;
        HCHK(hipStreamSynchronize(s));
        S.vamana_s = now_s() - t5;
        t5 = now_s();
    }
    S.total_s = t5 - t0;
    if (st) *st = S;
    return ANQI_OK;
}

}  // extern "C"

// ---- search ------------------------------------------------------------------------------------
static float round_down_float(double v) {
    float f = (float)v;
    if ((double)f > v) f = std::nextafter(f, -INFINITY);
    return f;
}

// Work areas of one search call: the persistent form's counters and batch
// arrays, the redo list, and the global tier's rows.
struct SearchWork {
    uint32_t* counter = nullptr;   // [0] work counter, [1] overflow count, [2] list length
    uint32_t* ent = nullptr;       // 3 x nq: tree entries, entry keys, work order
    uint32_t* list = nullptr;      // nq: queries to redo
    uint32_t* gvis = nullptr;      // slots x global_stride(n) bitmap words
    uint32_t* gfifo = nullptr;     // slots x n FIFO entries
    uint32_t slots = 0;
};
// Supplies the global tier's rows for `need` queries (fewer slots: groups).
using GlobalArea = std::function<hipError_t(uint32_t need, SearchWork& w)>;
// Called after each global-tier group with the group's query list (device)
// and size; block b of the group ran query list[b], FIFO row b.
using GlobalSink = std::function<int(const uint32_t* dlist, uint32_t count, const SearchWork& w)>;

static uint32_t global_stride(uint32_t n) { return ((n + 31u) / 32u + 3u) & ~3u; }

// Indices of the queries whose status is in `mask` into W.list; returns the count.
static int status_list(anqi_ctx* ctx, const uint32_t* status, uint32_t nq, uint32_t mask, SearchWork& W,
                       uint32_t* count, hipStream_t s) {
    HCHK(launch_status_list(status, nq, mask, W.list, W.counter + 2, s));
    HCHK(hipMemcpyAsync(count, W.counter + 2, 4, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    return ANQI_OK;
}

// Reruns the queries that outgrew their work areas (status 1: visited set,
// 2: candidate FIFO): first with 16K-entry LDS areas, then the rest -- and,
// with gmask bit 3, the queries whose candidates exceed out_cap -- in the
// global tier, whose bitmap and FIFO cannot fill.  P: the first pass's
// parameters (its entries array, if any, indexed by query).
static int search_redo(anqi_ctx* ctx, const SearchParams& P, uint32_t nq, SearchWork& W, uint32_t gmask,
                       const GlobalArea& area, const GlobalSink* sink, hipStream_t s, uint32_t* n_retry,
                       uint32_t* n_global) {
    uint32_t cnt = 0;
    int rc = status_list(ctx, P.status, nq, (1u << 1) | (1u << 2), W, &cnt, s);
    if (rc) return rc;
    *n_retry = cnt;
    SearchParams Q = P;
    Q.nq_dev = nullptr;
    Q.ovf_count = nullptr;
    Q.work_counter = nullptr;
    if (cnt) {
        Q.hmask = 16383;
        Q.rcap = 16384;
        Q.qlist = W.list;
        Q.nq = cnt;
        HCHK(launch_range_search(Q, false, 0, s));
    }
    rc = status_list(ctx, P.status, nq, gmask, W, &cnt, s);
    *n_global = cnt;
    if (rc || cnt == 0) return rc;
    if (W.slots == 0) HCHK(area(cnt, W));
    Q.gvis = W.gvis;
    Q.gfifo = W.gfifo;
    Q.gstride = global_stride(P.n);
    Q.rcap = P.n;
    Q.hmask = 3;
    for (uint32_t g = 0; g < cnt; g += W.slots) {
        Q.qlist = W.list + g;
        Q.nq = std::min(W.slots, cnt - g);
        HCHK(launch_range_search_global(Q, s));
        if (sink && (rc = (*sink)(Q.qlist, Q.nq, W))) return rc;
    }
    HCHK(hipStreamSynchronize(s));
    return ANQI_OK;
}

// The search behind anqi_rknn_search(_entries): first pass (mode 0: one
// block per query; mode 1: tree descents, work order, persistent grid), then
// search_redo for the queries it could not finish.
static int search_run(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t L,
                      const uint32_t* Lq, uint32_t Lmax, const float* thr, const uint32_t* kidx,
                      const uint32_t* entries, uint32_t out_cap, uint32_t* cand_ids, uint8_t* accept,
                      uint32_t* cand_cnt, uint32_t* ndist, uint32_t* pred_cnt, uint32_t* nexp, uint32_t* status,
                      int mode, SearchWork& W, const GlobalArea& area) {
    hipStream_t s = ctx->stream;
    SearchParams P{};
    P.base = ix->base; P.bnorm = ix->bnorm; P.lifted = ix->lifted; P.deg = ix->deg; P.links = ix->links;
    P.lmain = ix->lmain; P.ltail = ix->ltail; P.plinks = ix->plinks; P.brows = ix->brows;
    P.n = ix->n; P.d = ix->d; P.dl = ix->dl; P.M = ix->M;
    P.t_left = ix->tree->left; P.t_right = ix->tree->right; P.t_off = ix->tree->offset;
    P.t_hyper = ix->tree->hyper; P.t_entry = ix->tree->entry; P.t_root = ix->tree->root;
    P.entries = entries;
    double sc = ix->scale;
    std::memcpy(&P.scale_bits, &sc, 8);
    P.R2f = round_down_float(ix->mstar2 + 1.0);
    P.ip = ix->metric == ANQI_METRIC_IP;
    P.queries = queries; P.nq = nq; P.L = L; P.Lq = Lq; P.Lmax = Lq ? std::max(Lmax, 1u) : std::max(L, 1u);
    P.thr = thr; P.kidx = kidx;
    // LDS work area sized to the beam: visited-set slots ~ 4 L + 1536 (at least
    // twice the typical visited count), results FIFO 1024.  Queries that
    // overflow it are rerun by search_redo.
    P.hmask = std::min<uint32_t>(8192, std::max<uint32_t>(2048, next_pow2(4ull * std::max(L, Lmax) + 1536))) - 1;
    P.rcap = 1024; P.out_cap = out_cap;
    if (mode == 1 && ctx->tune.visited_slots >= 2 * P.Lmax) P.hmask = ((ctx->tune.visited_slots + 3u) & ~3u) - 1;
    if (mode == 1 && ctx->tune.result_cap) P.rcap = ctx->tune.result_cap;
    P.cand_ids = cand_ids; P.accept = accept; P.cand_cnt = cand_cnt; P.ndist = ndist; P.pred_cnt = pred_cnt;
    P.nexp = nexp;
    P.status = status;
    if (P.Lmax > 1024) return anqi_set_err(ctx, ANQI_E_USAGE, "L <= 1024");
    HCHK(launch_fill_u32(W.counter, 0u, 2, s));   // work counter, overflow count
    P.work_counter = W.counter;
    P.ovf_count = W.counter + 1;
    if (mode == 1) {
        if (!P.entries) {
            // the batch's tree descents first, four queries per wavefront,
            // then the work order: nearest entries first
            float* key = (float*)(W.ent + nq);
            uint32_t* order = W.ent + 2ull * nq;
            HCHK(launch_tree_entry_batch(P, nq, W.ent, key, s));
            HCHK(launch_entry_order(P, key, order, s));
            P.entries = W.ent;
            P.qlist = order;
        }
    }
    if (P.ip) {
        // per-query R2f into the entry keys' slot, free once the work order is set
        float* r2 = (float*)(W.ent + nq);
        HCHK(launch_query_r2(P, nq, ix->mstar2, r2, s));
        P.qr2 = r2;
    }
    if (mode == 1) {
        int occ = std::max(1, search_persistent_occupancy(P));
        uint32_t wpc = ctx->tune.waves_per_cu ? std::min<uint32_t>(ctx->tune.waves_per_cu, (uint32_t)occ) : (uint32_t)occ;
        uint32_t grid = wpc * (uint32_t)std::max(1, ctx->cus);
        if (ctx->time_kernels) HCHK(hipEventRecord(ctx->kev0, s));
        HCHK(launch_range_search(P, true, grid, s));
    } else {
        if (ctx->time_kernels) HCHK(hipEventRecord(ctx->kev0, s));
        HCHK(launch_range_search(P, false, 0, s));
    }
    if (ctx->time_kernels) HCHK(hipEventRecord(ctx->kev1, s));
    ctx->redo_retry = ctx->redo_global = 0;
    // overflows are counted on the device: statuses are scanned only when
    // there is something to redo
    uint32_t novf = 0;
    HCHK(hipMemcpyAsync(&novf, P.ovf_count, 4, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    if (ctx->time_kernels) {
        float kms = 0.0f;
        HCHK(hipEventElapsedTime(&kms, ctx->kev0, ctx->kev1));
        ctx->last_kernel_ms = kms;
    }
    if (novf == 0) return ANQI_OK;
    P.qlist = nullptr;
    return search_redo(ctx, P, nq, W, (1u << 1) | (1u << 2), area, nullptr, s, &ctx->redo_retry,
                       &ctx->redo_global);
}

extern "C" {


int anqi_rknn_search(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t L,
                     const uint32_t* Lq, uint32_t Lmax, const float* thr, const uint32_t* kidx, uint32_t out_cap,
                     uint32_t* cand_ids, uint8_t* accept, uint32_t* cand_cnt, uint32_t* ndist, uint32_t* pred_cnt,
                     uint32_t* nexp, uint32_t* status, int mode) {
    return anqi_rknn_search_entries(ctx, ix, queries, nq, L, Lq, Lmax, thr, kidx, nullptr, out_cap, cand_ids,
                                    accept, cand_cnt, ndist, pred_cnt, nexp, status, mode);
}

int anqi_rknn_search_entries(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t L,
                             const uint32_t* Lq, uint32_t Lmax, const float* thr, const uint32_t* kidx,
                             const uint32_t* entries, uint32_t out_cap, uint32_t* cand_ids, uint8_t* accept,
                             uint32_t* cand_cnt, uint32_t* ndist, uint32_t* pred_cnt, uint32_t* nexp,
                             uint32_t* status, int mode) {
    if (!ix || !ix->tree || ix->M > 64 || (ix->metric != ANQI_METRIC_L2 && ix->metric != ANQI_METRIC_IP))
        return anqi_set_err(ctx, ANQI_E_USAGE, "invalid index");
    if (ix->dl != ix->d + 2 || ix->d > 16 * QMAX_CHUNKS) return anqi_set_err(ctx, ANQI_E_USAGE, "bad dimensions");
    if (nq == 0) return ANQI_OK;
    if (out_cap == 0) return anqi_set_err(ctx, ANQI_E_USAGE, "out_cap must be positive");
    hipStream_t s = ctx->stream;
    if (!ctx->counter) HCHK(hipMalloc((void**)&ctx->counter, 256));
    if (ctx->ent_cap < nq) {
        if (ctx->ent) { HCHK(hipStreamSynchronize(s)); HCHK(hipFree(ctx->ent)); ctx->ent = nullptr; }
        HCHK(hipMalloc((void**)&ctx->ent, 16ull * nq));
        ctx->ent_cap = nq;
    }
    SearchWork W;
    W.counter = ctx->counter;
    W.ent = ctx->ent;
    W.list = ctx->ent + 3ull * nq;
    DBuf gv, gf;
    GlobalArea area = [&](uint32_t need, SearchWork& w) -> hipError_t {
        // global tier rows for up to 1 GiB of bitmaps and FIFOs per group
        const size_t per = 4ull * global_stride(ix->n) + 4ull * ix->n;
        const uint32_t slots = (uint32_t)std::max<size_t>(1, std::min<size_t>(need, (1ull << 30) / per));
        hipError_t e;
        if ((e = gv.alloc(4ull * global_stride(ix->n) * slots)) != hipSuccess) return e;
        if ((e = gf.alloc(4ull * ix->n * slots)) != hipSuccess) return e;
        w.gvis = gv.as<uint32_t>();
        w.gfifo = gf.as<uint32_t>();
        w.slots = slots;
        return hipSuccess;
    };
    return search_run(ctx, ix, queries, nq, L, Lq, Lmax, thr, kidx, entries, out_cap, cand_ids, accept, cand_cnt,
                      ndist, pred_cnt, nexp, status, mode, W, area);
}

int anqi_rknn_search_noalloc(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t L,
                             const uint32_t* Lq, uint32_t Lmax, const float* thr, const uint32_t* kidx,
                             uint32_t out_cap, uint32_t* cand_ids, uint8_t* accept, uint32_t* cand_cnt,
                             uint32_t* ndist, uint32_t* pred_cnt, uint32_t* nexp, uint32_t* status) {
    if (!ix || !ix->tree || ix->M > 64 || (ix->metric != ANQI_METRIC_L2 && ix->metric != ANQI_METRIC_IP))
        return anqi_set_err(ctx, ANQI_E_USAGE, "invalid index");
    if (ix->dl != ix->d + 2 || ix->d > 16 * QMAX_CHUNKS) return anqi_set_err(ctx, ANQI_E_USAGE, "bad dimensions");
    if (nq == 0) return ANQI_OK;
    if (out_cap == 0) return anqi_set_err(ctx, ANQI_E_USAGE, "out_cap must be positive");
    if (nq > NOALLOC_MAX_NQ)
        return anqi_set_err(ctx, ANQI_E_USAGE, "no-alloc search: %u queries exceed its limit of %u per call", nq,
                            NOALLOC_MAX_NQ);
    const size_t gbytes = NOALLOC_BYTES - NOALLOC_GLOBAL_OFF;
    const size_t per = 4ull * global_stride(ix->n) + 4ull * ix->n;   // one global-tier row pair
    if (per > gbytes)
        return anqi_set_err(ctx, ANQI_E_USAGE, "no-alloc search: an index of %u rows exceeds its limit of %llu rows",
                            ix->n, (unsigned long long)(gbytes * 8 / 33));
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEV || !g_scratch_ptr[ctx->device])
        return anqi_set_err(ctx, ANQI_E_DEVICE, "no-alloc search: no scratch address for device %d (HIP error %d)",
                            ctx->device,
                            ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEV ? (int)g_scratch_err[ctx->device] : -1);
    uint8_t* base = (uint8_t*)g_scratch_ptr[ctx->device];
    SearchWork W;
    W.counter = (uint32_t*)base;
    W.ent = (uint32_t*)(base + NOALLOC_ENT_OFF);
    W.list = W.ent + 3ull * nq;
    const uint32_t n = ix->n;
    GlobalArea area = [&](uint32_t need, SearchWork& w) -> hipError_t {
        const uint32_t slots = (uint32_t)std::max<size_t>(1, std::min<size_t>(need, gbytes / per));
        w.gvis = (uint32_t*)(base + NOALLOC_GLOBAL_OFF);
        w.gfifo = w.gvis + (size_t)slots * global_stride(n);
        w.slots = slots;
        return hipSuccess;
    };
    return search_run(ctx, ix, queries, nq, L, Lq, Lmax, thr, kidx, nullptr, out_cap, cand_ids, accept, cand_cnt,
                      ndist, pred_cnt, nexp, status, 1, W, area);
}

}  // extern "C"

// ---- captured query batches (mixed k and L) ---------------------------------------------
// The index arrays, the k threshold tables and the batch buffers are fixed;
// the batch size lives in device memory, so one instantiated graph
// {counter reset -> persistent search -> scan -> compaction} serves every
// batch up to max_batch.  Search work areas and the persistent grid come from
// an anqi_search_tuning (anqi_search_autotune measures candidates).

static void fill_search_params(SearchParams& P, const anqi_index* ix) {
    P = SearchParams{};
    P.base = ix->base; P.bnorm = ix->bnorm; P.lifted = ix->lifted; P.deg = ix->deg; P.links = ix->links;
    P.lmain = ix->lmain; P.ltail = ix->ltail; P.plinks = ix->plinks; P.brows = ix->brows;
    P.n = ix->n; P.d = ix->d; P.dl = ix->dl; P.M = ix->M;
    P.t_left = ix->tree->left; P.t_right = ix->tree->right; P.t_off = ix->tree->offset;
    P.t_hyper = ix->tree->hyper; P.t_entry = ix->tree->entry; P.t_root = ix->tree->root;
    double sc = ix->scale;
    std::memcpy(&P.scale_bits, &sc, 8);
    P.R2f = round_down_float(ix->mstar2 + 1.0);
    P.ip = ix->metric == ANQI_METRIC_IP;
}

struct anqi_batcher {
    anqi_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    SearchParams P{};
    anqi_index ix{};
    uint32_t max_batch = 0, out_cap = 0, Lmax = 0, grid = 0, ntables = 0;
    uint32_t redo_retry = 0, redo_global = 0;   // last run: queries rerun per tier
    DBuf q, Lq, kidx, nq, cand, acc, cnt, nd, pred, status, off, ids, counter, ent;
    float* hq = nullptr;
    uint32_t *hL = nullptr, *hk = nullptr, *hnq = nullptr, *hstatus = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~anqi_batcher() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
        if (hq) (void)hipHostFree(hq);
        if (hL) (void)hipHostFree(hL);
        if (hk) (void)hipHostFree(hk);
        if (hnq) (void)hipHostFree(hnq);
        if (hstatus) (void)hipHostFree(hstatus);
    }
};

static hipError_t batch_enqueue(anqi_batcher* b) {
    hipError_t e;
    // work counter and overflow count
    if ((e = launch_fill_u32(b->counter.as<uint32_t>(), 0u, 2, b->s)) != hipSuccess) return e;
    SearchParams T = b->P;
    T.entries = nullptr;
    T.qlist = nullptr;
    float* key = (float*)(b->ent.as<uint32_t>() + b->max_batch);
    if ((e = launch_tree_entry_batch(T, b->max_batch, b->ent.as<uint32_t>(), key, b->s)) != hipSuccess) return e;
    if ((e = launch_entry_order(T, key, b->ent.as<uint32_t>() + 2ull * b->max_batch, b->s)) != hipSuccess) return e;
    // inner product: the queries' R2f replace the consumed entry keys (b->P.qr2)
    if (b->P.ip && (e = launch_query_r2(T, b->max_batch, b->ix.mstar2, key, b->s)) != hipSuccess) return e;
    if ((e = launch_range_search(b->P, true, b->grid, b->s)) != hipSuccess) return e;
    if ((e = launch_batch_scan(b->pred.as<uint32_t>(), b->nq.as<uint32_t>(), b->max_batch, b->off.as<uint32_t>(),
                               b->s)) != hipSuccess)
        return e;
    return launch_batch_compact(b->cand.as<uint32_t>(), b->acc.as<uint8_t>(), b->cnt.as<uint32_t>(),
                                b->off.as<uint32_t>(), b->nq.as<uint32_t>(), b->max_batch, b->out_cap,
                                b->ids.as<uint32_t>(), b->s);
}

extern "C" {

int anqi_compact_candidates(anqi_ctx* ctx, uint32_t nq, uint32_t out_cap, const uint32_t* cand_ids,
                            const uint8_t* accept, const uint32_t* cand_cnt, uint32_t* off, uint32_t* ids,
                            uint8_t* acc_out, uint64_t ids_cap, uint64_t* total) {
    hipStream_t s = ctx->stream;
    HCHK(launch_cand_scan(cand_cnt, nq, out_cap, off, s));
    uint32_t t = 0;
    HCHK(hipMemcpyAsync(&t, off + nq, 4, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    *total = t;
    if (!ids) return ANQI_OK;
    if (t > ids_cap) return anqi_set_err(ctx, ANQI_E_USAGE, "candidates need %u entries, capacity %llu", t,
                                         (unsigned long long)ids_cap);
    HCHK(launch_cand_copy(cand_ids, accept, off, nq, out_cap, ids, acc_out, s));
    return ANQI_OK;
}

int anqi_search_entries(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t* entries) {
    if (!ix || !ix->tree || ix->dl != ix->d + 2 || ix->d > 16 * QMAX_CHUNKS)
        return anqi_set_err(ctx, ANQI_E_USAGE, "invalid index");
    SearchParams P;
    fill_search_params(P, ix);
    P.queries = queries;
    P.nq = nq;
    HCHK(launch_tree_entry_batch(P, nq, entries, nullptr, ctx->stream));
    return ANQI_OK;
}

int anqi_batcher_create(anqi_ctx* ctx, const anqi_index* ix, const float* thr_tables, uint32_t ntables,
                        uint32_t max_batch, uint32_t Lmax, uint32_t out_cap, const anqi_search_tuning* tune,
                        anqi_batcher** out) {
    if (!ix || !ix->tree || ix->M > 64 || ix->dl != ix->d + 2 || ix->d > 16 * QMAX_CHUNKS)
        return anqi_set_err(ctx, ANQI_E_USAGE, "invalid index");
    if (max_batch == 0 || Lmax == 0 || Lmax > 1024 || out_cap == 0 || ntables == 0)
        return anqi_set_err(ctx, ANQI_E_USAGE, "batcher: need max_batch, 0 < Lmax <= 1024, out_cap, ntables > 0");
    std::unique_ptr<anqi_batcher> b(new anqi_batcher());
    b->ctx = ctx;
    b->ix = *ix;
    b->max_batch = max_batch; b->out_cap = out_cap; b->Lmax = Lmax; b->ntables = ntables;
    HCHK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
    const size_t mb = max_batch;
    HCHK(b->q.alloc(4 * mb * ix->d));
    HCHK(b->Lq.alloc(4 * mb));
    HCHK(b->kidx.alloc(4 * mb));
    HCHK(b->nq.alloc(256));
    HCHK(b->cand.alloc(4 * mb * out_cap));
    HCHK(b->acc.alloc(mb * out_cap));
    HCHK(b->cnt.alloc(4 * mb));
    HCHK(b->nd.alloc(4 * mb));
    HCHK(b->pred.alloc(4 * mb));
    HCHK(b->status.alloc(4 * mb));
    HCHK(b->off.alloc(4 * (mb + 1)));
    HCHK(b->ids.alloc(4 * mb * out_cap));
    HCHK(b->counter.alloc(256));
    HCHK(b->ent.alloc(16 * mb));   // entries, keys, work order, redo list
    HCHK(hipHostMalloc((void**)&b->hq, 4 * mb * ix->d, hipHostMallocDefault));
    HCHK(hipHostMalloc((void**)&b->hL, 4 * mb, hipHostMallocDefault));
    HCHK(hipHostMalloc((void**)&b->hk, 4 * mb, hipHostMallocDefault));
    HCHK(hipHostMalloc((void**)&b->hnq, 256, hipHostMallocDefault));
    HCHK(hipHostMalloc((void**)&b->hstatus, 4 * mb, hipHostMallocDefault));
    SearchParams& P = b->P;
    fill_search_params(P, ix);
    P.queries = b->q.as<float>(); P.nq = max_batch; P.nq_dev = b->nq.as<uint32_t>();
    P.L = Lmax; P.Lmax = Lmax; P.Lq = b->Lq.as<uint32_t>();
    P.thr = thr_tables; P.kidx = b->kidx.as<uint32_t>();
    anqi_search_tuning t{0, 0, 0};
    if (tune) t = *tune;
    const uint32_t slots = t.visited_slots ? ((t.visited_slots + 3u) & ~3u)
                                           : std::min<uint32_t>(8192, std::max<uint32_t>(2048, next_pow2(4ull * Lmax + 1536)));
    P.hmask = slots - 1;
    // an LDS FIFO longer than the output row would let a query finish with
    // candidates past out_cap (status 3); capped, such a query overflows
    // (status 2) and is redone
    P.rcap = std::min(t.result_cap ? t.result_cap : 1024u, out_cap);
    P.out_cap = out_cap;
    P.cand_ids = b->cand.as<uint32_t>(); P.accept = b->acc.as<uint8_t>(); P.cand_cnt = b->cnt.as<uint32_t>();
    P.ndist = b->nd.as<uint32_t>(); P.pred_cnt = b->pred.as<uint32_t>(); P.nexp = nullptr;
    P.status = b->status.as<uint32_t>(); P.work_counter = b->counter.as<uint32_t>();
    P.ovf_count = b->counter.as<uint32_t>() + 1;
    P.entries = b->ent.as<uint32_t>();
    P.qlist = b->ent.as<uint32_t>() + 2 * mb;
    if (P.ip) P.qr2 = (const float*)(b->ent.as<uint32_t>() + mb);   // batch_enqueue's k_query_r2
    const int occ = std::max(1, search_persistent_occupancy(P));
    const uint32_t wpc = t.waves_per_cu ? std::min<uint32_t>(t.waves_per_cu, (uint32_t)occ) : (uint32_t)occ;
    b->grid = wpc * (uint32_t)std::max(1, ctx->cus);
    // one uncaptured pass (sets kernel attributes, checks the launch), then capture
    b->hnq[0] = 0;
    HCHK(hipMemcpyAsync(b->nq.p, b->hnq, 4, hipMemcpyHostToDevice, b->s));
    HCHK(batch_enqueue(b.get()));
    HCHK(hipStreamSynchronize(b->s));
    HCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    hipError_t ce = batch_enqueue(b.get());
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(b->s, &g);
    if (ce != hipSuccess) { if (g) (void)hipGraphDestroy(g); return anqi_fail_hip(ctx, ce, "batch capture", __LINE__); }
    HCHK(ee);
    b->graph = g;
    HCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    *out = b.release();
    return ANQI_OK;
}

void anqi_batcher_destroy(anqi_batcher* b) { delete b; }

int anqi_batcher_redo_counts(const anqi_batcher* b, uint32_t* retry, uint32_t* global) {
    if (!b) return ANQI_E_USAGE;
    if (retry) *retry = b->redo_retry;
    if (global) *global = b->redo_global;
    return ANQI_OK;
}

int anqi_batcher_run(anqi_batcher* b, const float* h_queries, const uint32_t* h_L, const uint32_t* h_kidx,
                     uint32_t nq, uint32_t* h_off, uint32_t* h_ids, uint64_t ids_cap, uint64_t* total) {
    anqi_ctx* ctx = b->ctx;
    if (nq > b->max_batch) return anqi_set_err(ctx, ANQI_E_USAGE, "batch of %u exceeds capacity %u", nq, b->max_batch);
    const uint32_t d = b->ix.d;
    for (uint32_t i = 0; i < nq; i++) {
        if (h_L[i] == 0 || h_L[i] > b->Lmax) return anqi_set_err(ctx, ANQI_E_USAGE, "query %u: L outside 1..Lmax", i);
        if (h_kidx[i] >= b->ntables)
            return anqi_set_err(ctx, ANQI_E_USAGE, "query %u: table index %u outside 0..%u", i, h_kidx[i],
                                b->ntables - 1);
    }
    std::memcpy(b->hq, h_queries, 4ull * nq * d);
    std::memcpy(b->hL, h_L, 4ull * nq);
    std::memcpy(b->hk, h_kidx, 4ull * nq);
    b->hnq[0] = nq;
    hipStream_t s = b->s;
    HCHK(hipMemcpyAsync(b->q.p, b->hq, 4ull * nq * d, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(b->Lq.p, b->hL, 4ull * nq, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(b->kidx.p, b->hk, 4ull * nq, hipMemcpyHostToDevice, s));
    HCHK(hipMemcpyAsync(b->nq.p, b->hnq, 4, hipMemcpyHostToDevice, s));
    b->redo_retry = b->redo_global = 0;
    HCHK(hipGraphLaunch(b->exec, s));
    HCHK(hipMemcpyAsync(b->hstatus, b->counter.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost, s));
    HCHK(hipMemcpyAsync(h_off, b->off.p, 4ull * (nq + 1), hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    // queries redone in the global tier: their answers, read from the FIFO rows
    std::vector<uint32_t> gq;                  // query indices
    std::vector<std::vector<uint32_t>> gans;   // accepted ids, candidate order
    if (b->hstatus[0] != 0) {
        SearchParams Q = b->P;
        Q.nq = nq;
        Q.qlist = nullptr;
        SearchWork W;
        W.counter = b->counter.as<uint32_t>();
        W.ent = b->ent.as<uint32_t>();
        W.list = b->ent.as<uint32_t>() + 3ull * b->max_batch;
        DBuf gv, gf;
        const uint32_t n = b->ix.n;
        GlobalArea area = [&](uint32_t need, SearchWork& w) -> hipError_t {
            const size_t per = 4ull * global_stride(n) + 4ull * n;
            const uint32_t slots = (uint32_t)std::max<size_t>(1, std::min<size_t>(need, (1ull << 30) / per));
            hipError_t e;
            if ((e = gv.alloc(4ull * global_stride(n) * slots)) != hipSuccess) return e;
            if ((e = gf.alloc(4ull * n * slots)) != hipSuccess) return e;
            w.gvis = gv.as<uint32_t>();
            w.gfifo = gf.as<uint32_t>();
            w.slots = slots;
            return hipSuccess;
        };
        std::vector<uint32_t> hcnt(nq), row;
        GlobalSink sink = [&](const uint32_t* dlist, uint32_t count, const SearchWork& w) -> int {
            const size_t at = gq.size();
            gq.resize(at + count);
            HCHK(hipMemcpyAsync(gq.data() + at, dlist, 4ull * count, hipMemcpyDeviceToHost, s));
            HCHK(hipMemcpyAsync(hcnt.data(), b->cnt.p, 4ull * nq, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
            for (uint32_t j = 0; j < count; j++) {
                const uint32_t c = hcnt[gq[at + j]];
                row.resize(c);
                if (c) {
                    HCHK(hipMemcpyAsync(row.data(), w.gfifo + (size_t)j * n, 4ull * c, hipMemcpyDeviceToHost, s));
                    HCHK(hipStreamSynchronize(s));
                }
                std::vector<uint32_t> acc;
                for (uint32_t v : row)
                    if (v >> 31) acc.push_back(v & 0x7fffffffu);
                gans.push_back(std::move(acc));
            }
            return ANQI_OK;
        };
        // tiers: 16K-entry LDS areas, then the global tier for what is left
        // and for candidate lists longer than out_cap
        int rc = search_redo(ctx, Q, nq, W, (1u << 1) | (1u << 2) | (1u << 3), area, &sink, s, &b->redo_retry,
                             &b->redo_global);
        if (rc) return rc;
        HCHK(launch_batch_scan(b->pred.as<uint32_t>(), b->nq.as<uint32_t>(), b->max_batch, b->off.as<uint32_t>(), s));
        HCHK(launch_batch_compact(b->cand.as<uint32_t>(), b->acc.as<uint8_t>(), b->cnt.as<uint32_t>(),
                                  b->off.as<uint32_t>(), b->nq.as<uint32_t>(), b->max_batch, b->out_cap,
                                  b->ids.as<uint32_t>(), s));
        HCHK(hipMemcpyAsync(h_off, b->off.p, 4ull * (nq + 1), hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
    }
    if (gq.empty()) {
        *total = h_off[nq];
        if (*total > ids_cap) return anqi_set_err(ctx, ANQI_E_USAGE, "answers need %llu ids", (unsigned long long)*total);
        if (*total) {
            HCHK(hipMemcpyAsync(h_ids, b->ids.p, 4ull * *total, hipMemcpyDeviceToHost, s));
            HCHK(hipStreamSynchronize(s));
        }
        return ANQI_OK;
    }
    // splice the global tier's answers into the device CSR
    std::vector<uint32_t> dev_off(h_off, h_off + nq + 1), dev_ids(dev_off[nq]);
    if (!dev_ids.empty()) {
        HCHK(hipMemcpyAsync(dev_ids.data(), b->ids.p, 4ull * dev_ids.size(), hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
    }
    std::vector<int32_t> slot(nq, -1);
    for (size_t j = 0; j < gq.size(); j++) slot[gq[j]] = (int32_t)j;
    uint64_t t = 0;
    for (uint32_t i = 0; i < nq; i++) {
        h_off[i] = (uint32_t)t;
        t += slot[i] >= 0 ? gans[slot[i]].size() : (uint64_t)(dev_off[i + 1] - dev_off[i]);
    }
    h_off[nq] = (uint32_t)t;
    *total = t;
    if (t > ids_cap) return anqi_set_err(ctx, ANQI_E_USAGE, "answers need %llu ids", (unsigned long long)t);
    for (uint32_t i = 0; i < nq; i++) {
        if (slot[i] >= 0) std::copy(gans[slot[i]].begin(), gans[slot[i]].end(), h_ids + h_off[i]);
        else std::copy(dev_ids.begin() + dev_off[i], dev_ids.begin() + dev_off[i + 1], h_ids + h_off[i]);
    }
    return ANQI_OK;
}

int anqi_search_autotune(anqi_ctx* ctx, const anqi_index* ix, const float* thr_tables, const float* queries,
                         uint32_t nq, const uint32_t* Lq, const uint32_t* kidx, uint32_t Lmax,
                         anqi_search_tuning* out, double* best_ms) {
    if (!ix || !ix->tree || nq == 0 || Lmax == 0 || Lmax > 1024) return anqi_set_err(ctx, ANQI_E_USAGE, "autotune: bad input");
    hipStream_t s = ctx->stream;
    DBuf cand, acc, cnt, nd, pred, status, counter;
    const uint32_t out_cap = 4096;
    HCHK(cand.alloc(4ull * nq * out_cap));
    HCHK(acc.alloc(1ull * nq * out_cap));
    HCHK(cnt.alloc(4ull * nq));
    HCHK(nd.alloc(4ull * nq));
    HCHK(pred.alloc(4ull * nq));
    HCHK(status.alloc(4ull * nq));
    HCHK(counter.alloc(256));
    SearchParams P;
    fill_search_params(P, ix);
    P.queries = queries; P.nq = nq; P.L = Lmax; P.Lmax = Lmax; P.Lq = Lq; P.thr = thr_tables; P.kidx = kidx;
    P.out_cap = out_cap;
    P.cand_ids = cand.as<uint32_t>(); P.accept = acc.as<uint8_t>(); P.cand_cnt = cnt.as<uint32_t>();
    P.ndist = nd.as<uint32_t>(); P.pred_cnt = pred.as<uint32_t>(); P.status = status.as<uint32_t>();
    P.work_counter = counter.as<uint32_t>();
    DBuf ent;   // the sample's tree entries and work order, computed once as the persistent form does
    HCHK(ent.alloc(12ull * nq));
    HCHK(launch_tree_entry_batch(P, nq, ent.as<uint32_t>(), (float*)(ent.as<uint32_t>() + nq), s));
    HCHK(launch_entry_order(P, (const float*)(ent.as<uint32_t>() + nq), ent.as<uint32_t>() + 2ull * nq, s));
    P.entries = ent.as<uint32_t>();
    P.qlist = ent.as<uint32_t>() + 2ull * nq;
    if (P.ip) {
        HCHK(launch_query_r2(P, nq, ix->mstar2, (float*)(ent.as<uint32_t>() + nq), s));
        P.qr2 = (const float*)(ent.as<uint32_t>() + nq);
    }
    hipEvent_t e0, e1;
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    double best = 1e300;
    anqi_search_tuning bt{0, 0, 0};
    const uint32_t slot_opts[6] = {1536, 2048, 3072, 3584, 4096, 8192};
    const uint32_t rcap_opts[2] = {128, 512};   // later FIFO entries spill to the output rows
    std::vector<uint32_t> st(nq);
    for (uint32_t so : slot_opts) {
        if (so < 2 * Lmax) continue;
        for (uint32_t rc : rcap_opts) {
            P.hmask = so - 1;
            P.rcap = rc;
            const int occ = std::max(1, search_persistent_occupancy(P));
            for (int frac = 2; frac <= 4; frac++) {             // 1/2, 3/4, 4/4 of the resident waves
                uint32_t wpc = std::max<uint32_t>(1, (uint32_t)(occ * frac / 4));
                if (frac < 4 && wpc == (uint32_t)occ) continue;
                const uint32_t grid = wpc * (uint32_t)std::max(1, ctx->cus);
                double tot = 0.0;
                for (int rep = 0; rep < 3; rep++) {
                    HCHK(launch_fill_u32(counter.as<uint32_t>(), 0u, 1, s));
                    HCHK(hipEventRecord(e0, s));
                    HCHK(launch_range_search(P, true, grid, s));
                    HCHK(hipEventRecord(e1, s));
                    HCHK(hipEventSynchronize(e1));
                    float ms = 0.0f;
                    HCHK(hipEventElapsedTime(&ms, e0, e1));
                    if (rep) tot += ms;            // first pass warms the caches
                }
                HCHK(hipMemcpyAsync(st.data(), status.p, 4ull * nq, hipMemcpyDeviceToHost, s));
                HCHK(hipStreamSynchronize(s));
                uint32_t bad = 0;
                for (uint32_t v : st) bad += v != 0;
                // overflowing queries are rerun uncaptured; charge them
                const double cost = tot / 2.0 * (1.0 + 50.0 * (double)bad / nq);
                progress("autotune: slots %u fifo %u waves/CU %u -> %.3f ms (%u overflow)", so, rc, wpc, tot / 2.0, bad);
                if (cost < best) { best = cost; bt = anqi_search_tuning{wpc, so, rc}; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *out = bt;
    if (best_ms) *best_ms = best;
    return ANQI_OK;
}

}  // extern "C"
