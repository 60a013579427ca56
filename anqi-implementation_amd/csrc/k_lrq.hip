// k_lrq.hip -- Rank-M plus packed-u8 LRQ verifier: encode, materialize_k.
//
// CompactLrqVerifier::build / materialize_k (src/rknn_query.cpp:679-753,
// policy "floor", bits 8) accumulate in binary64 with GCC-contracted fma.
// Two forms give the reference's coefficients, codes and thresholds bit for
// bit without any FP64 instruction:
//   - soft: every double is carried as its bit pattern through
//     softdouble.hpp (k_lrq_coeff, k_lrq_codes, k_lrq_materialize);
//   - certified (default): FP32 float-float sums whose rounding to float is
//     certified against the binary64 evaluation's error bound, with the
//     soft form only for the rare element it cannot certify
//     (k_lrq_encode_rows, k_lrq_decode).
//
// Device layout: codes are kept k-major (nk x n) next to the row-major file
// layout, so a threshold decode streams 8 coefficients (32 B) + 1 code byte
// per row and rank and writes one float per rank: 37 B/row for one rank,
// HBM-bound.
#include "common.hpp"
#include "kernels.hpp"
#include "lrq_cert.hpp"

// c[i][m] = float(sum_k fma((rk - mu), G[m][k], c))  (rknn_query.cpp:697-704)
__global__ void k_lrq_coeff(const float* __restrict__ cut, uint64_t n, uint32_t nk,
                            uint32_t rankm, const float* __restrict__ mu,
                            const float* __restrict__ basis, float* __restrict__ coeff) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * rankm) return;
    uint64_t i = t / rankm;
    uint32_t m = (uint32_t)(t % rankm);
    const float* rk = cut + i * nk;
    const float* gm = basis + (size_t)m * nk;
    uint64_t c = 0;   // +0.0
    for (uint32_t k = 0; k < nk; k++) {
        uint64_t diff = sd_sub(sd_from_float(rk[k]), sd_from_float(mu[k]));
        c = sd_fma(diff, sd_from_float(gm[k]), c);
    }
    coeff[t] = sd_to_float(c);
}

// code[i][k] = upper_bound(cb_k, float(rk - recon)) - 1, clamped at 0
// (rknn_query.cpp:706-717, encode_learned_residual :482-513).
__global__ void k_lrq_codes(const float* __restrict__ cut, uint64_t n, uint32_t nk,
                            uint32_t rankm, const float* __restrict__ mu,
                            const float* __restrict__ basis, const float* __restrict__ cb,
                            uint32_t code_size, const float* __restrict__ coeff,
                            uint8_t* __restrict__ codes, uint8_t* __restrict__ codes_km) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nk) return;
    uint32_t k = (uint32_t)(t / n);
    uint64_t i = t % n;
    const float* ci = coeff + i * rankm;
    uint64_t recon = sd_from_float(mu[k]);
    for (uint32_t m = 0; m < rankm; m++)
        recon = sd_fma(sd_from_float(ci[m]), sd_from_float(basis[(size_t)m * nk + k]), recon);
    uint64_t res = sd_sub(sd_from_float(cut[i * nk + k]), recon);
    float fv = sd_to_float(res);
    const float* c = cb + (size_t)k * code_size;
    uint32_t lo = 0, hi = code_size;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (fv < c[mid]) hi = mid; else lo = mid + 1;
    }
    uint8_t code = (uint8_t)(lo == 0 ? 0 : lo - 1);
    codes[i * nk + k] = code;
    if (codes_km) codes_km[(uint64_t)k * n + i] = code;
}

// t_k(i) = float(max(0, slack * (mu_k + sum_m c_m G_mk + cb_k[code]))); for
// the inner-product metric (clamp 0) without the max (rknn_query.cpp:751)
__global__ void k_lrq_materialize(const float* __restrict__ coeff, const uint8_t* __restrict__ codes_km,
                                  uint64_t n, uint32_t nk, uint32_t rankm,
                                  const float* __restrict__ mu, const float* __restrict__ basis,
                                  const float* __restrict__ cb, uint32_t code_size, uint32_t k,
                                  uint64_t slack_bits, int clamp, float* __restrict__ thr) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t v = sd_from_float(mu[k]);
    if (rankm == 8) {
        const float4* c4 = reinterpret_cast<const float4*>(coeff + i * 8);
        float4 a = c4[0], b = c4[1];
        float cs[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int m = 0; m < 8; m++)
            v = sd_fma(sd_from_float(cs[m]), sd_from_float(basis[(size_t)m * nk + k]), v);
    } else {
        for (uint32_t m = 0; m < rankm; m++)
            v = sd_fma(sd_from_float(coeff[i * rankm + m]), sd_from_float(basis[(size_t)m * nk + k]), v);
    }
    uint8_t code = codes_km[(uint64_t)k * n + i];
    v = sd_add(v, sd_from_float(cb[(size_t)k * code_size + code]));
    uint64_t scaled = sd_mul(slack_bits, v);
    thr[i] = clamp && sd_lt_zero_or_zero(scaled) ? 0.0f : sd_to_float(scaled);
}

// ---- float-float forms with certified rounding ---------------------------
// The same values from FP32 arithmetic (lrq_cert.hpp): each binary64 sum is
// estimated as a float-float and its rounding to float is taken when it is
// certified; the rare element whose binary64 value could round either way
// (or whose terms fall outside the covered range) is recomputed in soft
// binary64, so every output keeps the reference's bits.  `fallbacks`
// (optional) counts those recomputed elements.
#define LRQ_ENC_ROWS 64

// Encode, one row per lane, LRQ_ENC_ROWS rows per block: coefficients
// c_m = float(sum_k ((double)r_k - mu_k) G_mk) and codes
// upper_bound(cb_k, float(r_k - (mu_k + sum_m c_m G_mk))) - 1.  A lane reads
// its own row (twice: coefficients, then codes); each rank's codewords are
// staged in LDS for the searches; codes leave through LDS as contiguous
// row-major bytes and as 64-byte k-major runs.  Rank 8, code_size 256.
// Lanes past n work on the block's first row and store nothing.
// The kernel is bound by its float-float arithmetic, not by memory: staging
// the rows through LDS in coalesced 32-rank chunks cut its HBM reads from
// 7.6 GB to 1.4 GB at 1M rows but ran 3.6 ms against 2.3 ms for this form.
__global__ void __launch_bounds__(LRQ_ENC_ROWS) k_lrq_encode_rows(
        const float* __restrict__ cut, uint64_t n, uint32_t nk, const float* __restrict__ mu,
        const float* __restrict__ basis, const float* __restrict__ cb, float* __restrict__ coeff,
        uint8_t* __restrict__ codes, uint8_t* __restrict__ codes_km, unsigned long long* fallbacks) {
    extern __shared__ uint8_t cbuf[];                  // LRQ_ENC_ROWS x nk codes
    __shared__ float4 lcb[64];                         // rank k's 256 codewords
    const uint32_t lane = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * LRQ_ENC_ROWS;
    const uint64_t i = r0 + lane;
    const uint32_t rows = (uint32_t)min<uint64_t>(LRQ_ENC_ROWS, n - r0);
    const bool live = lane < rows;
    const float* rk = cut + (live ? i : r0) * nk;
    uint32_t nfb = 0;
    const Coef8 c = lrq_coeffs(rk, nk, mu, basis, &nfb);
    for (uint32_t k = 0; k < nk; k++) {
        __syncthreads();                               // the previous rank's searches are done
        lcb[lane] = reinterpret_cast<const float4*>(cb + (size_t)k * 256)[lane];
        __syncthreads();
        const uint32_t code = lrq_code(rk[k], c, basis, nk, k, mu[k], (const float*)lcb, &nfb);
        if (live) cbuf[lane * nk + k] = (uint8_t)code;
    }
    if (live) {
        float4* c4 = reinterpret_cast<float4*>(coeff + i * LRQ_RANKM);
        c4[0] = make_float4(c.v[0], c.v[1], c.v[2], c.v[3]);
        c4[1] = make_float4(c.v[4], c.v[5], c.v[6], c.v[7]);
    }
    if (fallbacks && live && nfb) atomicAdd(fallbacks, (unsigned long long)nfb);
    __syncthreads();
    uint8_t* dst = codes + r0 * nk;
    for (uint32_t t = lane; t < rows * nk; t += LRQ_ENC_ROWS) dst[t] = cbuf[t];
    if (codes_km && live)
        for (uint32_t k = 0; k < nk; k++) codes_km[(uint64_t)k * n + i] = cbuf[lane * nk + k];
}

// Thresholds for up to LRQ_MAX_TABLES ranks in one pass over the rows (slack
// 1): t_k(i) = float(max(0, mu_k + sum_m c_m G_mk + cb_k[code_ik])), without
// the max for the inner-product metric (clamp 0)
// (rknn_query.cpp:722-753).  A block of 256 threads stages the ranks'
// codewords in LDS, then each thread takes LRQ_DEC_ROWS rows 256 apart: their
// coefficients (32 B) and code bytes are all requested before any is used,
// then one threshold per row and rank is written.
#define LRQ_DEC_ROWS 2
struct LrqRanks {
    uint32_t k[LRQ_MAX_TABLES];   // zero-based ranks
    uint32_t count;
};
__global__ void __launch_bounds__(256) k_lrq_decode(const float* __restrict__ coeff,
                                                    const uint8_t* __restrict__ codes_km, uint64_t n, uint32_t nk,
                                                    const float* __restrict__ mu,
                                                    const float* __restrict__ basis, const float* __restrict__ cb,
                                                    LrqRanks ranks, int clamp, float* __restrict__ thr,
                                                    unsigned long long* fallbacks) {
    __shared__ float lcb[LRQ_MAX_TABLES * 256];
    const uint32_t tid = threadIdx.x;
    const uint64_t r0 = (uint64_t)blockIdx.x * (256 * LRQ_DEC_ROWS) + tid;
    // the rows' coefficients and code bytes are requested first, then the
    // codewords are staged while those loads are in flight
    Coef8 c[LRQ_DEC_ROWS];
    uint32_t code[LRQ_MAX_TABLES][LRQ_DEC_ROWS];
#pragma unroll
    for (int j = 0; j < LRQ_DEC_ROWS; j++) {
        const uint64_t i = r0 + 256u * j;
        if (i < n) {
            const float4* c4 = reinterpret_cast<const float4*>(coeff + i * LRQ_RANKM);
            const float4 a = c4[0], b = c4[1];
            c[j] = Coef8{{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}};
        }
    }
#pragma unroll
    for (uint32_t t = 0; t < LRQ_MAX_TABLES; t++) {
        if (t >= ranks.count) break;
#pragma unroll
        for (int j = 0; j < LRQ_DEC_ROWS; j++) {
            const uint64_t i = r0 + 256u * j;
            code[t][j] = i < n ? codes_km[(uint64_t)ranks.k[t] * n + i] : 0u;
        }
    }
    for (uint32_t t = 0; t < ranks.count; t++) lcb[t * 256 + tid] = cb[(size_t)ranks.k[t] * 256 + tid];
    __syncthreads();
    uint32_t nfb = 0;
#pragma unroll
    for (uint32_t t = 0; t < LRQ_MAX_TABLES; t++) {
        if (t >= ranks.count) break;
        const uint32_t k = ranks.k[t];
#pragma unroll
        for (int j = 0; j < LRQ_DEC_ROWS; j++) {
            const uint64_t i = r0 + 256u * j;
            if (i < n)
                thr[(uint64_t)t * n + i] = lrq_thr(c[j], basis, nk, k, mu[k], lcb[t * 256 + code[t][j]], &nfb, clamp != 0);
        }
    }
    if (fallbacks && nfb) atomicAdd(fallbacks, (unsigned long long)nfb);
}

__global__ void k_transpose_codes(const uint8_t* __restrict__ codes, uint64_t n, uint32_t nk,
                                  uint8_t* __restrict__ out) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nk) return;
    uint32_t k = (uint32_t)(t / n);
    uint64_t i = t % n;
    out[t] = codes[i * nk + k];
}

static inline unsigned blocks_for(uint64_t m, unsigned b) { return (unsigned)((m + b - 1) / b); }

hipError_t launch_lrq_encode(const float* cut, uint64_t n, uint32_t nk, uint32_t rankm,
                             const float* mu, const float* basis, const float* cb,
                             uint32_t code_size, float* coeff, uint8_t* codes,
                             uint8_t* codes_kmajor, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lrq_coeff, dim3(blocks_for(n * rankm, 256)), dim3(256), 0, s, cut, n, nk,
                       rankm, mu, basis, coeff);
    hipLaunchKernelGGL(k_lrq_codes, dim3(blocks_for(n * nk, 256)), dim3(256), 0, s, cut, n, nk,
                       rankm, mu, basis, cb, code_size, coeff, codes, codes_kmajor);
    return hipGetLastError();
}

hipError_t launch_lrq_materialize(const float* coeff, const uint8_t* codes_kmajor, uint64_t n,
                                  uint32_t nk, uint32_t rankm, const float* mu,
                                  const float* basis, const float* cb, uint32_t code_size,
                                  uint32_t k_one_based, uint64_t slack_bits, int clamp, float* thr,
                                  hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lrq_materialize, dim3(blocks_for(n, 256)), dim3(256), 0, s, coeff,
                       codes_kmajor, n, nk, rankm, mu, basis, cb, code_size, k_one_based - 1,
                       slack_bits, clamp, thr);
    return hipGetLastError();
}

bool lrq_fast_supported(uint32_t nk, uint32_t rankm, uint32_t code_size) {
    return nk >= 1 && nk <= 256 && rankm == LRQ_RANKM && code_size == 256;
}

hipError_t launch_lrq_encode_fast(const float* cut, uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                                  const float* basis, const float* cb, float* coeff, uint8_t* codes,
                                  uint8_t* codes_kmajor, unsigned long long* fallbacks, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t lds = (size_t)LRQ_ENC_ROWS * nk;
    (void)rankm;
    hipLaunchKernelGGL(k_lrq_encode_rows, dim3(blocks_for(n, LRQ_ENC_ROWS)), dim3(LRQ_ENC_ROWS), lds, s, cut, n, nk,
                       mu, basis, cb, coeff, codes, codes_kmajor, fallbacks);
    return hipGetLastError();
}

hipError_t launch_lrq_decode(const float* coeff, const uint8_t* codes_kmajor, uint64_t n, uint32_t nk,
                             uint32_t rankm, const float* mu, const float* basis, const float* cb,
                             uint32_t code_size, const uint32_t* k_one_based, uint32_t nks, int clamp,
                             float* thr, unsigned long long* fallbacks, hipStream_t s) {
    if (n == 0 || nks == 0) return hipSuccess;
    for (uint32_t t0 = 0; t0 < nks; t0 += LRQ_MAX_TABLES) {
        LrqRanks r{};
        r.count = std::min<uint32_t>(LRQ_MAX_TABLES, nks - t0);
        for (uint32_t t = 0; t < r.count; t++) r.k[t] = k_one_based[t0 + t] - 1;
        (void)code_size;   // 256 (lrq_fast_supported)
        hipLaunchKernelGGL(k_lrq_decode, dim3(blocks_for(n, 256 * LRQ_DEC_ROWS)), dim3(256), 0, s, coeff,
                           codes_kmajor, n, nk, mu, basis, cb, r, clamp, thr + (uint64_t)t0 * n, fallbacks);
    }
    return hipGetLastError();
}

hipError_t launch_transpose_codes(const uint8_t* codes, uint64_t n, uint32_t nk,
                                  uint8_t* codes_kmajor, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_transpose_codes, dim3(blocks_for(n * nk, 256)), dim3(256), 0, s, codes, n,
                       nk, codes_kmajor);
    return hipGetLastError();
}
